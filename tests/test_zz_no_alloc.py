"""athad_field_minmax_noalloc: searchMinMax_3D of a field in the caller's device memory, with
the reduction scratch in the library's fixed device array (no allocation), against the
context entry point athad_field_minmax on the same values: bit-equal extremes and positions,
on two calls in a row, the second on a smaller grid in FP32 storage."""
import ctypes

import numpy as np
import pytest

import athad_cond_amd as A
from athad_cond_amd import _lib as L

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    """The HIP runtime the library links (its device memory is the caller's here)."""
    global _hip
    if _hip is None:
        L.lib()
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def to_device(arr):
    p = ctypes.c_void_p()
    assert hip().hipMalloc(ctypes.byref(p), arr.nbytes) == 0
    assert hip().hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0   # host to device
    return p


def noalloc(ptr, precision, grid):
    lib = L.lib()
    fn = lib.athad_field_minmax_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    lo, hi = ctypes.c_double(), ctypes.c_double()
    pl, ph = (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
    st = fn(ptr, precision, *grid, None, ctypes.byref(lo), pl, ctypes.byref(hi), ph)
    return st, (lo.value, tuple(pl), hi.value, tuple(ph))


def planted(values):
    v = values.copy().reshape(-1)
    rng = np.random.default_rng(4)
    idx = rng.choice(v.size, 40, replace=False)
    v[idx[:10]] = np.nan
    v[idx[10:15]] = np.inf
    v[idx[15:20]] = -np.inf
    v[idx[20:25]] = v.max()       # ties of the maximum: the first occurrence is kept
    v[idx[25:30]] = v.min()
    return v.reshape(values.shape)


@pytest.mark.parametrize("case", [((21, 91, 181), "fp64"), ((11, 21, 41), "mixed")])
def test_noalloc_minmax_equals_context_entry_point(case):
    grid, precision = case
    print(f"case noalloc minmax {grid} {precision}", flush=True)
    with A.Integration(grid, precision=precision) as integ:
        integ.init_state(9)
        integ.run(6)
        for name in ("t", "u", "p_dyn"):
            vals = planted(integ.get(name))
            integ.set(name, vals)
            want = integ.field_minmax(name)
            host = integ.get(name)          # the values as the context stores them
            arr = np.ascontiguousarray(host if precision == "fp64" else host.astype(np.float32))
            ptr = to_device(arr)
            try:
                st, got = noalloc(ptr, L.FP64 if precision == "fp64" else L.MIXED, grid)
            finally:
                hip().hipFree(ptr)
            assert st == 0, L.lib().athad_last_error(None)
            assert got == want, (name, got, want)


def test_noalloc_refuses_out_of_range():
    L.lib()
    arr = np.zeros(4)
    ptr = to_device(arr)
    try:
        assert noalloc(ptr, L.FP64, (201, 2, 2))[0] == 4          # ATHAD_ERR_SHAPE: beyond the fixed scratch
        assert noalloc(None, L.FP64, (2, 1, 2))[0] == 1           # ATHAD_ERR_ARG
        assert noalloc(ptr, 7, (2, 1, 2))[0] == 1
        st, got = noalloc(ptr, L.FP64, (2, 1, 2))
        assert st == 0 and got == (0.0, (0, 0, 0), 0.0, (0, 0, 0))
    finally:
        hip().hipFree(ptr)
