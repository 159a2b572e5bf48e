"""A C++ consumer of include/athad_hip.h (tests/capi_stub.cpp) compiles against the header and
links against the library the way the reference's Makefile would link it; on a null context
and bad arguments every entry point it calls returns ATHAD_ERR_ARG without touching a device,
and the host-only helpers answer (CPU, no GPU needed)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "athad_cond_amd", "libathad_hip.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")


def test_stub_compiles_links_and_refuses_null_context(tmp_path):
    sys.path.insert(0, ROOT)
    import __graft_entry__ as G
    exe = G.build_capi_stub(str(tmp_path / "capi_stub"))
    out = subprocess.run([exe, "null"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout)
    for key in ("create_bad", "upload_rows", "download_rows", "set_iter", "get_iter", "step", "rk4", "turb_model",
                "turb_init", "destroy"):
        assert r[key] == 1, key            # ATHAD_ERR_ARG
    assert r["ctx_null"] == 1
    assert r["last_error"] == "bad arguments"
    n = 11 * 21 * 41
    assert r["restart_bytes"] == 20 + 29 * n * 8
    assert abs(r["dew_point_60bar"] - 548.73) < 0.01      # boiling point at 60 bar (cond_column_selftest.cpp:74)
    assert r["member_from_surface"] == 0 and abs(r["x_H2O"] - 0.5578) < 1e-3
    import athad_cond_amd as A
    assert r["footprint"] == A.footprint_bytes((11, 21, 41), "fp64")
