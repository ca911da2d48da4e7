"""Loader and ctypes signatures of libkgmt.so (the HIP build of include/kgmt.h).

The library must have been built (python -m kgmt_amd.build); there is no
fallback path: a missing library raises ImportError-like KgmtLibraryMissing.
"""
from __future__ import annotations

import ctypes as C
import os

from .abi import KGMT_ABI_VERSION, KgmtAdjointParams, KgmtConfig, KgmtDeviceTree, KgmtProfile, KgmtResult

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libkgmt.so")


class KgmtLibraryMissing(RuntimeError):
    pass


class KgmtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"kgmt error {code}: {msg}")
        self.code = code


_P = C.c_void_p
_u32 = C.c_uint32
_i32 = C.c_int32
_cfgp = C.POINTER(KgmtConfig)
_resp = C.POINTER(KgmtResult)

# name -> (restype, argtypes); every symbol declared in include/kgmt.h
SIGNATURES = {
    "kgmt_abi_version": (C.c_int, []),
    "kgmt_last_error": (C.c_char_p, []),
    "kgmt_model_dims": (C.c_int, [_i32, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i32)]),
    "kgmt_config_preset": (C.c_int, [_i32, _cfgp]),
    "kgmt_config_check": (C.c_int, [_cfgp]),
    "kgmt_create": (C.c_int, [_cfgp, C.POINTER(_P)]),
    "kgmt_destroy": (C.c_int, [_P]),
    "kgmt_set_seed": (C.c_int, [_P, C.c_uint64]),
    "kgmt_set_graph": (C.c_int, [_P, _i32]),
    "kgmt_set_launch": (C.c_int, [_P, _u32, _u32, _u32, _u32]),
    "kgmt_set_stream": (C.c_int, [_P, _P]),
    "kgmt_synchronize": (C.c_int, [_P]),
    "kgmt_plan": (C.c_int, [_P, _P, _P, _P, _u32, _i32, _resp]),
    "kgmt_plan_begin": (C.c_int, [_P, _P, _P, _P, _u32, _i32]),
    "kgmt_plan_step": (C.c_int, [_P, _resp, C.POINTER(_i32)]),
    "kgmt_get_frontier": (C.c_int, [_P, _P, _P, _u32, C.POINTER(_u32)]),
    "kgmt_get_tree": (C.c_int, [_P, _P, _P, _P, _u32, C.POINTER(_u32)]),
    "kgmt_get_tree_device": (C.c_int, [_P, C.POINTER(KgmtDeviceTree)]),
    "kgmt_get_path": (C.c_int, [_P, _P, _u32, C.POINTER(_u32)]),
    "kgmt_get_regions": (C.c_int, [_P, _P, _P, _P, _u32, C.POINTER(_u32)]),
    "kgmt_set_profiling": (C.c_int, [_P, _i32]),
    "kgmt_get_profile": (C.c_int, [_P, C.POINTER(KgmtProfile)]),
    "kgmt_batch_set_profiling": (C.c_int, [_P, _i32]),
    "kgmt_batch_get_profile": (C.c_int, [_P, C.POINTER(KgmtProfile)]),
    "kgmt_batch_create": (C.c_int, [_cfgp, _u32, C.POINTER(_P)]),
    "kgmt_batch_destroy": (C.c_int, [_P]),
    "kgmt_batch_set_stream": (C.c_int, [_P, _P]),
    "kgmt_batch_plan": (C.c_int, [_P, _P, _P, _P, _P, _P, _i32, _resp, _P]),
    "kgmt_batch_get_path": (C.c_int, [_P, _u32, _P, _u32, C.POINTER(_u32)]),
    "kgmt_batch_get_tree": (C.c_int, [_P, _u32, _P, _P, _P, _u32, C.POINTER(_u32)]),
    "kgmt_batch_set_launch": (C.c_int, [_P, _u32, _u32, _u32, _u32]),
    "kgmt_draw_controls": (C.c_int, [_cfgp, _u32, _u32, _u32, _u32, _P, _P, _P, _P]),
    "kgmt_propagate_collide": (C.c_int, [_cfgp, _P, _P, _u32, _P, _P, _u32, _P, _u32, _P, _P, _P, _P, _P]),
    "kgmt_propagate_collide_noalloc": (C.c_int, [_cfgp, _P, _P, _u32, _P, _P, _u32, _P, _u32, _P, _P, _P, _P, _P]),
    "kgmt_propagate_collide_planner": (C.c_int, [_cfgp, _P, _P, _u32, _P, _P, _u32, _P, _u32, _P, _P, _P, _P, _P,
                                                 _P, _P]),
    "kgmt_frontier_update": (C.c_int, [_cfgp, _u32, _u32, _P, _P, _P, _P, _P, _P, _P, _u32, _P, _P, _P, _P, _P,
                                       _P]),
    "kgmt_region_update": (C.c_int, [_cfgp, _P, _P, _P, _P, _u32, _P, _P, _P, _P, _P, _P, _P]),
    "kgmt_score": (C.c_int, [_cfgp, _P, _P, _P, _P, _P]),
    "kgmt_compact": (C.c_int, [_P, _u32, _P, _P, _P]),
    "kgmt_repeat": (C.c_int, [_P, _P, _u32, _P, _u32, _P, _P]),
    "kgmt_build_scene": (C.c_int, [_P, _u32, _P, _u32, C.POINTER(_u32), _P]),
    "kgmt_rk4_adjoint": (C.c_int, [_cfgp, C.POINTER(KgmtAdjointParams), _u32, _u32, _P, _P, _P, _P, _P, _u32, _P, _P,
                                   _P]),
    "kgmt_refine": (C.c_int, [_cfgp, C.POINTER(KgmtAdjointParams), _u32, _u32, _P, _P, _P, _P, _P, _u32, _P, _P]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library with signatures applied (loaded once)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KgmtLibraryMissing(
                f"{LIB_PATH} is not built; run `python -m kgmt_amd.build` (HIP toolchain required)")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        found = handle.kgmt_abi_version()
        if found != KGMT_ABI_VERSION:
            raise KgmtLibraryMissing(
                f"{LIB_PATH} has ABI version {found}, the bindings expect {KGMT_ABI_VERSION}: rebuild it "
                "(`python -m kgmt_amd.build`)")
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().kgmt_last_error()
        raise KgmtError(rc, msg.decode() if msg else "")
