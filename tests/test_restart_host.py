"""Restart files in the reference's ATM1 layout (atmosphere/FileIO_Atm.cpp:293-356), host side:
the header and array order written by write_restart, and read_restart's inverse."""
import numpy as np
import pytest

from athad_cond_amd import _lib as L
from athad_cond_amd.integration import read_restart, write_restart


def test_layout_matches_the_reference_order():
    im, jm, km = 3, 5, 7
    rng = np.random.default_rng(0)
    arrays = {name: rng.standard_normal((im, jm, km)) for name in L.RESTART_ARRAYS[:25]}
    data = write_restart({"im": im, "jm": jm, "km": km, "total_iter_count": 1234}, arrays)
    assert len(data) == 20 + 29 * im * jm * km * 8
    # header: magic "ATM1", dimensions, total_iter_count (FileIO_Atm.cpp:317)
    assert list(np.frombuffer(data[:20], "<i4")) == [0x41544D31, im, jm, km, 1234]
    # arrays follow restart_arrays() (FileIO_Atm.cpp:297-304), each [i][j][k] row-major
    n = im * jm * km
    body = np.frombuffer(data, "<f8", offset=20)
    for q, name in enumerate(L.RESTART_ARRAYS):
        want = arrays.get(name, np.zeros((im, jm, km)))
        assert np.array_equal(body[q * n:(q + 1) * n], want.ravel()), name
    assert L.RESTART_ARRAYS[:4] == ["t", "u", "v", "w"] and L.RESTART_ARRAYS[-4:] == ["P_rain", "P_snow", "P_rainn", "P_snown"]
    hdr, back = read_restart(data)
    assert hdr == {"im": im, "jm": jm, "km": km, "total_iter_count": 1234}
    for name in arrays:
        assert np.array_equal(back[name], arrays[name])


def test_reader_rejects_other_files():
    with pytest.raises(ValueError):
        read_restart(np.array([0, 1, 1, 1, 0], "<i4").tobytes() + b"\0" * 8 * 29)
