"""Restart files in the reference's ATM1 layout (cAtmosphereModel::save_state / load_state,
atmosphere/FileIO_Atm.cpp:293-356) written and read by the HIP library."""
import numpy as np
import pytest

import athad_cond_amd as A
from athad_cond_amd import _lib as L

pytestmark = pytest.mark.gpu

GRID = (11, 21, 41)
HELD = L.RESTART_ARRAYS[:25]   # the state fields; the precipitation four are the ice scheme's


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_save_then_load_reproduces_the_state(precision):
    print(f"case restart round trip {GRID} {precision}", flush=True)
    with A.Integration(GRID, precision=precision) as a, A.Integration(GRID, precision=precision) as b:
        a.init_state(21)
        a.run(7)
        data = a.save_state()
        assert len(data) == A.restart_bytes(GRID)
        b.init_state(5)
        b.load_state(data)
        for name in HELD:
            assert np.array_equal(a.get(name), b.get(name)), name
        assert b.get_iter()[1] == a.get_iter()[1]
        assert b.save_state() == data


def test_file_holds_the_fields_in_reference_order(tmp_path):
    print(f"case restart file {GRID}", flush=True)
    path = tmp_path / "atm_restart_0Ma_6.bin"
    with A.Integration(GRID) as a:
        a.init_state(3)
        a.run(6)
        a.save_state(str(path))
        hdr, arrays = A.read_restart(path.read_bytes())
        assert (hdr["im"], hdr["jm"], hdr["km"]) == GRID
        assert hdr["total_iter_count"] == a.get_iter()[1]
        for name in HELD:
            assert np.array_equal(arrays[name], a.get(name)), name
        for name in L.RESTART_ARRAYS[25:]:
            assert not arrays[name].any(), name


def test_loads_a_restart_written_from_the_oracle(oracle):
    print(f"case restart from oracle {GRID}", flush=True)
    ref = oracle.OracleState(*GRID, seed=8)
    ref.run(5, "lex")
    data = A.write_restart({"im": GRID[0], "jm": GRID[1], "km": GRID[2], "total_iter_count": ref.total_iter},
                           {name: ref.field(name) for name in HELD})
    with A.Integration(GRID, order="lex") as b:
        b.init_state(1)
        b.load_state(data)
        for name in HELD:
            assert np.array_equal(b.get(name), ref.field(name)), name
        assert b.get_iter()[1] == ref.total_iter


def test_bad_restarts_are_refused(tmp_path):
    print("case restart errors", flush=True)
    with A.Integration(GRID) as a:
        a.init_state(2)
        data = bytearray(a.save_state())
        bad = bytearray(data)
        bad[0] ^= 1                                     # magic
        with pytest.raises(A.AthadError, match="status 4"):
            a.load_state(bytes(bad))
        other = A.write_restart({"im": 11, "jm": 21, "km": 43, "total_iter_count": 0}, {})
        with pytest.raises(A.AthadError, match="status 4"):   # grid mismatch
            a.load_state(other)
        with pytest.raises(A.AthadError, match="status 1"):   # truncated
            a.load_state(bytes(data[:-8]))
        with pytest.raises(A.AthadError, match="status 1"):   # no such file
            a.load_state(str(tmp_path / "missing.bin"))
        # a refused load leaves the state as it was
        assert bytes(data) == a.save_state()


def test_precipitation_arrays_travel_once_the_ice_scheme_has_run(oracle):
    """P_rain, P_snow, P_rainn, P_snown go into the file from the context and come back."""
    print(f"case restart with ice scheme {GRID}", flush=True)
    ref = oracle.OracleState(*GRID, seed=4)
    ref.run(2, "jacobi")
    snap = ref.snapshot()
    mp = {k: np.zeros(GRID) for k in oracle.MP_FIELDS}
    oracle.plant_microphysics(snap, mp, 9)
    with A.Integration(GRID) as a, A.Integration(GRID) as b:
        a.load(snap)
        for k, v in mp.items():
            a.set_mp(k, v)
        a.ice_scheme()
        data = a.save_state()
        _, arrays = A.read_restart(data)
        for name in L.RESTART_ARRAYS[25:]:
            assert arrays[name].any(), name
            assert np.array_equal(arrays[name], a.get_mp(name)), name
        b.init_state(1)
        b.load_state(data)
        for name in L.RESTART_ARRAYS[25:]:
            assert np.array_equal(b.get_mp(name), a.get_mp(name)), name
        assert b.save_state() == data
