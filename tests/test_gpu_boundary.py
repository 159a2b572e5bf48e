"""The boundary the reference will use: state held in its Array layout (double*** x, one
separately allocated km-row per (i, j), lib/Array.cpp:62-71) moved through athad_upload_rows /
athad_download_rows.

1. Every field uploaded and downloaded through deliberately non-contiguous row pointers is bit
   for bit what athad_set_field / athad_get_field move.
2. tests/capi_stub (a C++ consumer of include/athad_hip.h, compiled by __graft_entry__.build())
   runs the hook sequence of a maintainer's cAtmosphereModel: create (FP64, lex order), init,
   every field out to row storage and back in, set_iter, steps, every field out again. Its
   result equals the same sequence through the Python binding bit for bit.
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import athad_cond_amd as A
from athad_cond_amd import _lib as L

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STUB = os.path.join(ROOT, "tests", "capi_stub")
GRID = (11, 21, 41)


def scattered_rows(im, jm, km, rng):
    """Row storage with the rows in a shuffled order inside one buffer, with gaps between
    them, so no two consecutive (i, j) rows are adjacent."""
    nrow = im * jm
    stride = km + 5
    buf = np.full(nrow * stride + 3, np.nan)
    slots = rng.permutation(nrow)
    ptrs = (ctypes.POINTER(ctypes.c_double) * nrow)()
    base = buf.ctypes.data
    for r in range(nrow):
        ptrs[r] = ctypes.cast(base + (int(slots[r]) * stride + 2) * 8, ctypes.POINTER(ctypes.c_double))
    views = [buf[int(slots[r]) * stride + 2: int(slots[r]) * stride + 2 + km] for r in range(nrow)]
    return buf, ptrs, views


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rows_equal_contiguous_transfer(precision):
    print(f"case rows {precision}", flush=True)
    lib = L.lib()
    im, jm, km = GRID
    rng = np.random.default_rng(5)
    with A.Integration(GRID, precision=precision) as integ:
        integ.init_state(3)
        integ.run(4)
        for f, name in enumerate(L.FIELDS):
            want = integ.get(name)
            buf, ptrs, views = scattered_rows(im, jm, km, rng)
            L.check(lib.athad_download_rows(integ.h, 0, f, ptrs), integ.h)
            got = np.stack(views).reshape(im, jm, km)
            assert np.array_equal(got, want), ("download", name)
            # new values in, through rows; the contiguous read must return them
            new = want * 1.5 + 0.25 if name != "p_stat" else want * 1.25
            for r in range(im * jm):
                views[r][:] = new.reshape(im * jm, km)[r]
            L.check(lib.athad_upload_rows(integ.h, 0, f, ptrs), integ.h)
            back = integ.get(name)
            expect = new if precision == "fp64" or name == "p_stat" else new.astype(np.float32).astype(np.float64)
            assert np.array_equal(back, expect), ("upload", name)
            integ.set(name, want)


@pytest.mark.skipif(not os.path.exists(STUB), reason="tests/capi_stub not built (__graft_entry__.build())")
def test_cpp_consumer_matches_binding(tmp_path):
    print("case capi stub run", flush=True)
    steps = 8
    out = tmp_path / "stub.bin"
    r = subprocess.run([STUB, "run", *map(str, GRID), str(steps), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = json.loads(r.stdout)
    assert res["status"] == 0 and res["destroy"] == 0 and res["null_row"] == 1
    assert (res["iter_n"], res["total"]) == (steps, steps)
    stub = np.fromfile(out, dtype=np.float64).reshape(len(L.FIELDS), *GRID)
    with A.Integration(GRID, order="lex") as integ:
        integ.init_state(1234)
        integ.set_iter(0, 0)
        for _ in range(steps):
            integ.step()
        for f, name in enumerate(L.FIELDS):
            assert np.array_equal(stub[f], integ.get(name)), name
