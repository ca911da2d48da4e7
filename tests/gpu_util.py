"""Helpers of the GPU tests: torch tensors as device memory, C-ABI calls on
torch's current stream."""
import ctypes as C

import numpy as np
import torch

from kgmt_amd import _lib


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda:0")


def empty(shape, dtype):
    return torch.zeros(shape, dtype=dtype, device="cuda:0")


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def call(name, *args):
    fn = getattr(_lib.lib(), name)
    _lib.check(fn(*args))
    torch.cuda.synchronize()


def host(t):
    return t.cpu().numpy()


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a
