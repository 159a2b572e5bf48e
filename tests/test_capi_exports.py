"""The C-ABI library loads and exports every entry point include/athad_hip.h declares (CPU;
no device calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "athad_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(athad_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("athad_create", "athad_step", "athad_poisson_sweep", "athad_satadj", "athad_densities",
                 "athad_turbulence", "athad_rk4", "athad_destroy"):
        assert must in names


def test_library_exports_all_declared_symbols():
    import athad_cond_amd as A
    L = A.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    assert L.athad_version().startswith(b"athad_hip")


def test_errors_without_context():
    import athad_cond_amd as A
    L = A.lib()
    assert L.athad_step(None) == 1          # ATHAD_ERR_ARG, no device touched
    assert L.athad_ice_scheme(None) == 1
    assert L.athad_set_ice_scheme(None, 1) == 1
    buf = (ctypes.c_double * 8)()
    assert L.athad_get_mp_field(None, 0, 0, buf) == 1
    assert L.athad_set_mp_field(None, 0, 0, buf) == 1
    assert L.athad_radiation(None) == 1
    assert L.athad_set_turbulence_model(None, 0) == 1
    assert L.athad_turbulence_init(None) == 1
    out = (ctypes.c_double * 64)()
    assert L.athad_get_conv_cape(None, 0, out) == 1
    g = A._lib.Grid(41, 180, 361)           # even jm is rejected before any device call
    h = ctypes.c_void_p()
    assert L.athad_create(1, ctypes.byref(g), None, 0, 1, ctypes.byref(h)) == 4
    assert b"unsupported grid" in L.athad_last_error(None)


def test_footprint_is_linear_in_cells():
    import athad_cond_amd as A
    a = A.footprint_bytes((41, 181, 361), "fp64")
    b = A.footprint_bytes((41, 181, 361), "mixed")
    assert 0.45 < b / a < 0.6
    assert A.footprint_bytes((61, 361, 721), "fp64") > 5 * a
