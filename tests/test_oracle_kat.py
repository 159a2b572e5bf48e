"""Known-answer tests of the oracle's thermodynamics (CPU).

Values and tolerances are the reference's own self-tests (test/saturation_selftest.cpp,
test/cond_column_selftest.cpp, cited per check), plus NIST/JANAF heat capacities and
closed-form limits that a transposed coefficient or a dropped term would fail.
"""
import math

import pytest

T_SURF, P_SURF = 513.15, 60000.0   # cond_column_selftest.cpp:40-42 (K, hPa)


def close(got, want, rel):
    return abs(got - want) <= rel * abs(want)


@pytest.fixture(scope="module")
def L(oracle):
    return oracle.lib()


@pytest.mark.parametrize("T,want,tol", [
    (273.16, 6.11657, 2e-3),      # saturation_selftest.cpp:41
    (373.124, 1013.25, 2e-3),     # :43
    (400.0, 2457.4, 5e-3),        # :45
    (500.0, 26392.0, 5e-3),       # :46
    (600.0, 123450.0, 5e-3),      # :47
    (647.095, 220640.0, 1e-3),    # :49
])
def test_p_sat_points(L, T, want, tol):
    assert close(L.orc_p_sat(T), want, tol)


def test_p_sat_sentinels(L):
    assert L.orc_p_sat(647.096) == -1.0          # :53-54
    assert L.orc_p_sat(1500.0) == -1.0           # :55-56
    assert L.orc_p_auto(700.0) == -1.0


@pytest.mark.parametrize("T,want,tol", [(273.16, 6.11657, 1e-4), (250.0, 0.76053, 5e-3),
                                        (200.0, 1.6362e-3, 1e-2)])   # :62-64
def test_p_subl_points(L, T, want, tol):
    assert close(L.orc_p_subl(T), want, tol)


def test_p_auto_monotone_and_continuous(L):
    prev = 0.0
    T = 150.0
    while T < 647.0:                                 # saturation_selftest.cpp:68-75
        p = L.orc_p_auto(T)
        assert math.isfinite(p) and p > prev
        prev = p
        T += 0.5
    # the two branches meet at the triple point to their stated accuracy
    assert close(L.orc_p_sat(273.16), L.orc_p_subl(273.16), 2e-3)


def test_latent_heat(L):
    assert close(L.orc_latent_heat(273.15), 2.501e6, 1e-6)    # :79
    assert close(L.orc_latent_heat(373.15), 2.257e6, 3e-2)    # :80
    assert L.orc_latent_heat(647.096) == 0.0                  # :81
    assert close(L.orc_latent_heat(513.0), 1.72e6, 5e-2)      # cond_column_selftest.cpp:172


def test_q_sat_limits(L):
    assert close(L.orc_q_sat(999.0, 1000.0, 0.034), 0.998114, 1e-4)   # saturation_selftest.cpp:117
    assert L.orc_q_sat(2000.0, 1000.0, 0.034) == 1.0                   # :118
    assert L.orc_q_sat(-1.0, 1000.0, 0.034) == 1.0                     # :119
    # dilute limit against the textbook ep*E/(p-(1-ep)E), :106-113
    E = L.orc_p_sat(288.0)
    ep = 0.62198
    assert close(L.orc_q_sat(E, 1013.25, 0.028964), ep * E / (1013.25 - (1.0 - ep) * E), 5e-3)
    # exact conversion: q = x*Mw/(x*Mw + (1-x)*Mo) at x = 1/2 is Mw/(Mw+Mo)
    assert close(L.orc_q_sat(500.0, 1000.0, 0.04), 0.018015 / (0.018015 + 0.04), 1e-15)


def test_surface_composition(oracle, L):
    C = oracle.resolve(0.5578, 0.4109, 0.0313)      # cond_column_selftest.cpp:118-126
    assert close(C["M_mean"] * 1000, 29.01, 1e-2)
    assert close(C["R_mix"], 286.6, 1e-2)
    assert close(C["q_H2O"], 0.3464, 2e-2)
    cp = L.orc_cp_of(C["q_H2O"], C["q_CO2"], T_SURF, C["M_bg"])
    assert close(cp, 1354.0, 3e-2)
    Mnw = L.orc_M_nonwater(C["q_H2O"], C["q_CO2"], C["M_bg"])
    assert close(Mnw * 1000, 42.888, 1e-2)          # :164-165
    # the saturated sea surface: p_sat(513.15) = 33.470 bar and x_H2O = 0.5578 (:88-92)
    assert close(L.orc_p_sat(T_SURF) / 1000.0, 33.470, 5e-3)
    assert close(L.orc_p_sat(T_SURF) / P_SURF, 0.5578, 5e-3)
    # boiling point at 60 bar is 548.7 K (:74): p_sat(548.7) is 60 bar
    assert close(L.orc_p_sat(548.7), P_SURF, 2e-3 * 12)
    D = oracle.resolve(0.010, 0.920, 0.070)         # :141-145
    assert close(D["M_mean"] * 1000, 42.63, 1e-3)
    assert close(D["R_mix"], 195.0, 1e-3)
    assert close(D["q_H2O"], 0.00423, 2e-2)


def test_heat_capacities_against_nist(L):
    # Pure-species limits of cp_of: NIST-JANAF molar cp [J/(mol K)] / M.
    assert close(L.orc_cp_of(1.0, 0.0, 500.0, 0.028) * 0.018015, 35.22, 2e-3)   # H2O(g), 500 K
    assert close(L.orc_cp_of(0.0, 1.0, 500.0, 0.028) * 0.044010, 44.63, 2e-3)   # CO2, 500 K
    assert close(L.orc_cp_of(0.0, 0.0, 300.0, 0.028014) * 0.028014, 29.12, 3e-3)  # N2, 300 K
    # clamped outside the fit range
    assert L.orc_cp_of(0.3, 0.5, 100.0, 0.028) == L.orc_cp_of(0.3, 0.5, 298.0, 0.028)


def test_lapse_rates(oracle, L):
    C = oracle.resolve(0.5578, 0.4109, 0.0313)
    cp = L.orc_cp_of(C["q_H2O"], C["q_CO2"], T_SURF, C["M_bg"])
    Mnw = L.orc_M_nonwater(C["q_H2O"], C["q_CO2"], C["M_bg"])
    assert close(9.81 / cp * 1000.0, 7.27, 1e-2)                                      # :196
    moist = L.orc_moist_lapse(T_SURF, P_SURF, Mnw, cp, C["R_mix"], 9.81)
    assert close(moist * 1000.0, 5.04, 2e-2)                                          # :197
    assert close(L.orc_moist_lapse(300.0, 1000.0, 0.02896, 1005.0, 287.0, 9.81) * 1000, 3.79, 3e-2)  # :209
    assert L.orc_moist_lapse(700.0, 60000.0, Mnw, cp, C["R_mix"], 9.81) == 9.81 / cp  # :215-216
    assert close(L.orc_moist_lapse(200.0, 60000.0, Mnw, cp, C["R_mix"], 9.81), 9.81 / cp, 1e-5)  # :222
    # dq_sat/dT exact / dilute ratio (:178-184)
    q = L.orc_q_sat(L.orc_p_sat(T_SURF), P_SURF, Mnw)
    Lh = L.orc_latent_heat(T_SURF)
    dilute = q * Lh / ((8.314462618 / 0.018015) * T_SURF * T_SURF)
    x = L.orc_p_sat(T_SURF) / P_SURF
    assert close(L.orc_dqsat_dT(T_SURF, P_SURF, Mnw) / dilute,
                 Mnw / (x * 0.018015 + (1 - x) * Mnw), 1e-2)
    assert close(L.orc_dqsat_dT(T_SURF, P_SURF, Mnw) / dilute, 1.478, 1e-2)
