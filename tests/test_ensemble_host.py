"""Host-side member derivation of the batched ensemble (CPU; no device calls).

Known answers from test/cond_column_selftest.cpp for the reference's own configuration
(60 bar, 513.15 K sea, dry composition H2O 1 %, CO2 92 %, N2 7 %), and the oracle's
independent thermodynamics for drawn members.
"""
import math

import pytest

from athad_cond_amd import ensemble as E


def close(a, b, rel):
    return abs(a - b) <= rel * abs(b)


def test_reference_configuration():
    p = E.member_params(E.SurfaceState(60.0, 513.15, 0.010, 0.920))
    assert close(p.x_H2O, 0.5578, 5e-3)          # cond_column_selftest.cpp:92
    assert close(p.x_CO2, 0.4110, 1e-2)          # :99
    assert close(p.x_N2, 0.0313, 2e-2)           # :100
    assert close(p.r_air, 40.8, 2e-2)            # :131-132
    assert close(p.c_0, 0.3464, 2e-2)            # :123
    assert close(p.c_h2o_dry_top, 0.00423, 2e-2)  # :145
    assert p.t_surf_pole == p.t_surf_equator - 10.0
    # the surface anchor gives back the surface pressure (:134-136)
    R_mix = 8.314462618 / (p.x_H2O * 0.018015 + p.x_CO2 * 0.044010 + p.x_N2 * 0.028014)
    assert close(1e-2 * p.r_air * R_mix * 513.15, 60000.0, 1e-12)


def test_dew_point():
    assert close(E.dew_point(60000.0), 548.7, 2e-3)     # cond_column_selftest.cpp:74
    assert close(E.dew_point(1013.25), 373.124, 1e-3)   # saturation_selftest.cpp:43
    assert E.dew_point(0.0) == 0.0


def test_drawn_members_against_oracle(oracle):
    L = oracle.lib()
    states = E.draw_members(32, seed=7)
    assert len(states) == 32
    for s in states:
        assert 40.0 <= s.p_surf_bar <= 100.0 and 503.0 <= s.t_sea <= 523.0
        assert s.t_sea < E.dew_point(s.p_surf_bar * 1000.0)
        p = E.member_params(s)
        x = L.orc_p_sat(s.t_sea) / (s.p_surf_bar * 1000.0)
        assert close(p.x_H2O, x, 1e-14)
        assert math.isclose(p.x_H2O + p.x_CO2 + p.x_N2, 1.0, rel_tol=1e-14)
        assert close(p.x_CO2 / p.x_N2, s.x_co2_dry / (1.0 - s.x_h2o_dry - s.x_co2_dry), 1e-12)
        C = oracle.resolve(p.x_H2O, p.x_CO2, p.x_N2)
        assert close(p.c_0, C["q_H2O"], 1e-14)
        assert close(p.r_air, s.p_surf_bar * 1e5 / (C["R_mix"] * s.t_sea), 1e-14)


def test_boiling_sea_rejected():
    with pytest.raises(Exception):
        E.member_params(E.SurfaceState(40.0, 540.0, 0.01, 0.9))
