"""ctypes mirror of the structs and constants of include/kgmt.h.

Importing this module loads nothing; kgmt_amd._lib loads the HIP library.
"""
from __future__ import annotations

import ctypes as C

KGMT_ABI_VERSION = 5

KGMT_OK = 0
KGMT_E_ARG = -1
KGMT_E_HIP = -2
KGMT_E_NOMEM = -3
KGMT_E_CAPACITY = -4
KGMT_E_UNSUPPORTED = -5
KGMT_E_NODEVICE = -6

KGMT_MODEL_UNICYCLE = 0
KGMT_MODEL_DOUBLE_INTEGRATOR = 1
KGMT_MODEL_DUBINS_AIRPLANE = 2
KGMT_MODEL_QUADCOPTER = 3
KGMT_MODEL_DUBINS_5D = 4

KGMT_DI_REFERENCE = 0
KGMT_DI_EXACT = 1

KGMT_CHECK_SCENE = 0
KGMT_CHECK_CANDIDATES = 1
KGMT_CHECK_DRY_RUN = 2
KGMT_CHECK_ONE_PASS = 3

KGMT_PRESET_REFERENCE_QUAD = 0
KGMT_PRESET_DI6D = 1
KGMT_PRESET_DUBINS5D = 2
KGMT_PRESET_QUAD12D = 3
KGMT_PRESET_UNICYCLE2D = 4

KGMT_MAX_STATE = 12
KGMT_MAX_CONTROL = 4
KGMT_MAX_SAMPLE = 17

MODEL_DIMS = {  # model -> (S, C)
    KGMT_MODEL_DOUBLE_INTEGRATOR: (6, 3),
    KGMT_MODEL_DUBINS_AIRPLANE: (6, 3),
    KGMT_MODEL_QUADCOPTER: (12, 4),
    KGMT_MODEL_DUBINS_5D: (5, 3),
    KGMT_MODEL_UNICYCLE: (4, 2),
}


def dims(model: int):
    """(S, C, D) of a model."""
    s, c = MODEL_DIMS[model]
    return s, c, s + c + 1


def wdim(model: int) -> int:
    """Workspace dimension W: 2 for the planar unicycle, else 3. Obstacle
    rows hold 2W floats (min corner, max corner)."""
    return 2 if model == KGMT_MODEL_UNICYCLE else 3


class KgmtConfig(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("integrator", C.c_int32),
        ("capacity", C.c_uint32),
        ("max_iter", C.c_uint32),
        ("dt", C.c_float),
        ("steps_min", C.c_int32),
        ("steps_max", C.c_int32),
        ("goal_thresh", C.c_float),
        ("ws_min", C.c_float),
        ("ws_max", C.c_float),
        ("v_min", C.c_float),
        ("v_max", C.c_float),
        ("ctrl_min", C.c_float * 4),
        ("ctrl_max", C.c_float * 4),
        ("pitch_min", C.c_float),
        ("pitch_max", C.c_float),
        ("r1_len", C.c_int32),
        ("r2_len", C.c_int32),
        ("epsilon", C.c_float),
        ("children_per_entry", C.c_uint32),
        ("root_repeat", C.c_uint32),
        ("repeat_sparse", C.c_uint32),
        ("repeat_dense", C.c_uint32),
        ("sparse_valid_below", C.c_int32),
        ("quad_mass", C.c_float),
        ("quad_ix", C.c_float),
        ("quad_iy", C.c_float),
        ("quad_iz", C.c_float),
        ("quad_nu", C.c_float),
        ("quad_mu", C.c_float),
        ("gravity", C.c_float),
        ("query_id", C.c_uint32),
        ("uni_length", C.c_float),
        ("reserved", C.c_uint32),
        ("seed", C.c_uint64),
    ]

    def copy(self) -> "KgmtConfig":
        out = KgmtConfig()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(KgmtConfig))
        return out

    def replace(self, **kw) -> "KgmtConfig":
        out = self.copy()
        for k, v in kw.items():
            if k in ("ctrl_min", "ctrl_max"):
                arr = getattr(out, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(out, k, v)
        return out

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name in ("ctrl_min", "ctrl_max") else v
        return d

    @property
    def dims(self):
        return dims(self.model)

    @property
    def wdim(self) -> int:
        return wdim(self.model)

    @property
    def regions(self) -> int:
        """W_R1_LENGTH^W regions of the workspace grid."""
        return self.r1_len ** wdim(self.model)


class KgmtResult(C.Structure):
    _fields_ = [
        ("success", C.c_int32),
        ("goal_index", C.c_int32),
        ("iterations", C.c_uint32),
        ("tree_size", C.c_uint32),
        ("path_rows", C.c_uint32),
        ("path_cost", C.c_float),
        ("path_length", C.c_float),
        ("time_ms", C.c_float),
        ("expanded", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class KgmtAdjointParams(C.Structure):
    _fields_ = [
        ("w_effort", C.c_float),
        ("w_obstacle", C.c_float),
        ("w_goal", C.c_float),
        ("margin", C.c_float),
        ("tau", C.c_float),
        ("u_ref", C.c_float * 4),
        ("lr", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("adam_eps", C.c_float),
        ("iterations", C.c_uint32),
        ("clamp", C.c_int32),
    ]


class KgmtDeviceTree(C.Structure):
    _fields_ = [
        ("samples", C.c_void_p),
        ("parents", C.c_void_p),
        ("costs", C.c_void_p),
        ("tree_size", C.c_uint32),
        ("capacity", C.c_uint32),
        ("sample_dim", C.c_int32),
        ("reserved", C.c_uint32),
    ]


KGMT_PROFILE_KERNELS = 5
PROFILE_KERNEL_NAMES = ("propagate", "update", "finish")  # slots 3 and 4 are unused


class KgmtProfile(C.Structure):
    _fields_ = [
        ("launches", C.c_uint64 * KGMT_PROFILE_KERNELS),
        ("ms", C.c_double * KGMT_PROFILE_KERNELS),
        ("children", C.c_uint64),
        ("iterations", C.c_uint64),
        ("flagged", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {
            "kernels": {n: {"launches": int(self.launches[i]), "ms": float(self.ms[i])}
                        for i, n in enumerate(PROFILE_KERNEL_NAMES)},
            "children": int(self.children),
            "iterations": int(self.iterations),
            "flagged": int(self.flagged),
        }


assert C.sizeof(KgmtConfig) == 168, C.sizeof(KgmtConfig)
assert C.sizeof(KgmtResult) == 40, C.sizeof(KgmtResult)
assert C.sizeof(KgmtAdjointParams) == 60, C.sizeof(KgmtAdjointParams)
assert C.sizeof(KgmtProfile) == 104, C.sizeof(KgmtProfile)
assert C.sizeof(KgmtDeviceTree) == 40, C.sizeof(KgmtDeviceTree)
