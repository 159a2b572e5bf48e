"""Closed forms that pin the oracle's mode-2 radiation (CPU).

MultiLayerRadiation.h:59-384 with the 5 % blend of cAtmosphereModel.cpp:863-879 and
updateSkinTemperature (:276-311). A transparent column (no mass: every emissivity exactly 0)
takes every layer to the classical skin temperature T_surf 2^(-1/4) and carries the surface
emission to the top unchanged; an opaque column (every emissivity exactly 1) makes each layer's
sigma T^4 the mean of its neighbours'; the short wave is the pole-to-equator parabola; t_skin
takes one relaxation step toward the emission temperature of the absorbed short wave.
"""
import math

import numpy as np
import pytest

SIGMA, T0, OMEGA, C_H, GEO, ALB = 5.67028e-08, 273.15, 0.05, 15.0, 150.0, 0.08
GRID = (9, 21, 17)


def blank(oracle, t_profile, p_profile, c=0.0):
    im, jm, km = GRID
    st = oracle.OracleState(*GRID, blank=True)
    st.field("t")[...] = (np.asarray(t_profile) / T0)[:, None, None]
    st.field("p_stat")[...] = np.asarray(p_profile, dtype=float)[:, None, None]
    st.field("c")[...] = c
    return st


def short_wave(jm):
    jh = (jm - 1) // 2
    sw = np.zeros(jm)
    for j in range(jh + 1):
        x = j / jh
        sw[j] = -298.0 * (x * x - 2.0 * x)
    sw[jh + 1:] = sw[:jh][::-1][: jm - jh - 1]
    return sw


def test_short_wave_parabola(oracle):
    st = blank(oracle, np.full(GRID[0], 300.0), np.zeros(GRID[0]))
    st.radiation()
    sw = st.rad_outputs()["sw"]
    jh = (GRID[1] - 1) // 2
    assert sw[0] == 0.0 and sw[-1] == 0.0 and sw[jh] == 298.0
    np.testing.assert_array_equal(sw, sw[::-1])
    np.testing.assert_allclose(sw, short_wave(GRID[1]), rtol=1e-15, atol=1e-13)


def test_transparent_column_reaches_the_skin_temperature(oracle):
    im = GRID[0]
    T_in = np.linspace(320.0, 250.0, im)
    st = blank(oracle, T_in, np.zeros(im))   # no mass above any level: tau = 0, eps = 0
    st.radiation()
    out = st.rad_outputs()
    assert (out["epsilon"] == 0.0).all() and (out["albedo"] == ALB).all()
    t_skin = T_in[0] * 2.0 ** -0.25
    e0 = SIGMA * T_in[0] ** 4
    sw = out["sw"]
    d4 = 4.0 * SIGMA * T_in[0] ** 3
    Ts = ((1.0 - ALB) * sw + GEO - e0 + d4 * T_in[0] + C_H * t_skin) / (d4 + C_H)
    t = st.field("t")
    for i in range(1, im):
        np.testing.assert_allclose(t[i] * T0, (1 - OMEGA) * T_in[i] + OMEGA * t_skin, rtol=1e-14)
        np.testing.assert_allclose(out["radiation"][i], e0, rtol=1e-14)
    col = lambda v: np.broadcast_to(v[:, None], t[0].shape)   # noqa: E731  per latitude
    np.testing.assert_allclose(t[0] * T0, col((1 - OMEGA) * T_in[0] + OMEGA * Ts), rtol=1e-14)
    np.testing.assert_allclose(out["radiation"][0], col(SIGMA * Ts ** 4), rtol=1e-13)


def test_opaque_column_averages_neighbours_in_sigma_t4(oracle):
    im = GRID[0]
    T_in = np.array([330.0, 300.0, 290.0, 305.0, 280.0, 270.0, 265.0, 250.0, 240.0])
    p = 1.0e5 * (im - np.arange(im))          # 1e7 Pa per layer, broadened: tau >> 700
    st = blank(oracle, T_in, p, c=0.5)
    st.radiation()
    out = st.rad_outputs()
    assert (out["epsilon"] == 1.0).all()
    s = T_in ** 4
    for _ in range(4):                        # Lambda iterations: s_i <- (s_{i-1} + s_{i+1}) / 2, none above the top
        up = s.copy()
        dn = np.append(s, 0.0)
        s = np.concatenate([[s[0]], 0.5 * (up[:-1] + dn[2:])])
    T = s ** 0.25
    t = st.field("t")
    for i in range(1, im):
        np.testing.assert_allclose(t[i] * T0, (1 - OMEGA) * T_in[i] + OMEGA * T[i], rtol=1e-13)
        np.testing.assert_allclose(out["radiation"][i], SIGMA * T[i] ** 4, rtol=1e-13)
    # the surface sees the lowest layer's emission as downwelling long wave
    sw = out["sw"]
    d4 = 4.0 * SIGMA * T_in[0] ** 3
    Ts = ((1.0 - ALB) * sw + GEO + SIGMA * T[1] ** 4 - SIGMA * T_in[0] ** 4 + d4 * T_in[0] + C_H * T[1]) / (d4 + C_H)
    np.testing.assert_allclose(t[0] * T0, np.broadcast_to(((1 - OMEGA) * T_in[0] + OMEGA * Ts)[:, None], t[0].shape),
                               rtol=1e-13)


def test_skin_temperature_relaxes_toward_the_absorbed_flux(oracle):
    im, jm, km = GRID
    st = blank(oracle, np.full(im, 300.0), np.zeros(im))
    assert st.scalar("t_skin") == 254.0
    st.radiation()
    sw = st.rad_outputs()["sw"]
    w = np.cos((np.arange(jm) / (jm - 1) - 0.5) * math.pi)
    absorbed = np.sum(w * (1.0 - ALB) * sw) / np.sum(w)
    target = ((absorbed + GEO) / SIGMA) ** 0.25
    assert st.scalar("t_skin") == pytest.approx(254.0 + 0.25 * (target - 254.0), rel=1e-14)
    # a second call relaxes again from the new value
    t1 = st.scalar("t_skin")
    st.radiation()
    assert st.scalar("t_skin") == pytest.approx(t1 + 0.25 * (target - t1), rel=1e-14)
