"""Checkpoint diagnostics computed on the device: searchMinMax_3D (atmosphere/MinMax_Atm.cpp:51-108)
and findResiduumAtm (atmosphere/UtilsAtm.h:70-218)."""
import math

import numpy as np
import pytest

import athad_cond_amd as A

pytestmark = pytest.mark.gpu


def first_extremes(a):
    """numpy restatement of the row-major scan with strict comparisons over finite values."""
    flat = a.ravel()
    fin = np.isfinite(flat)
    if not fin.any():
        return 0.0, (0, 0, 0), 0.0, (0, 0, 0)
    lo, hi = flat[fin].min(), flat[fin].max()
    qlo = int(np.flatnonzero(fin & (flat == lo))[0])
    qhi = int(np.flatnonzero(fin & (flat == hi))[0])
    return float(lo), tuple(int(x) for x in np.unravel_index(qlo, a.shape)), float(hi), \
        tuple(int(x) for x in np.unravel_index(qhi, a.shape))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_minmax_matches_the_row_major_scan(precision):
    grid = (11, 21, 41)
    print(f"case minmax {grid} {precision}", flush=True)
    with A.Integration(grid, precision=precision) as integ:
        integ.init_state(4)
        integ.run(6)
        for name in ["t", "u", "v", "c", "tke", "p_stat", "p_dyn"]:
            assert integ.field_minmax(name) == first_extremes(integ.get(name)), name
        # non-finite values are skipped; ties keep the first position; nothing finite: 0 at 0
        x = integ.get("u")
        x[0, 0, 0] = np.nan
        x[3, 4, 5] = np.inf
        x[2, 2, 2] = -np.inf
        integ.set("u", x)
        assert integ.field_minmax("u") == first_extremes(integ.get("u"))
        integ.set("u", np.full(x.shape, 0.25))
        assert integ.field_minmax("u") == (0.25, (0, 0, 0), 0.25, (0, 0, 0))
        integ.set("u", np.full(x.shape, np.nan))
        assert integ.field_minmax("u") == (0.0, (0, 0, 0), 0.0, (0, 0, 0))
        # the minimum starts from DBL_MAX and the maximum from -DBL_MAX with strict comparisons,
        # so a field holding only that bound reports 0 at (0, 0, 0) for that side
        big = np.finfo(np.float64).max
        if precision == "fp64":
            integ.set("u", np.full(x.shape, big))
            assert integ.field_minmax("u") == (0.0, (0, 0, 0), big, (0, 0, 0))
            integ.set("u", np.full(x.shape, -big))
            assert integ.field_minmax("u") == (-big, (0, 0, 0), 0.0, (0, 0, 0))
            x = np.full(x.shape, big)
            x[1, 2, 3] = -big
            integ.set("u", x)
            assert integ.field_minmax("u") == (-big, (1, 2, 3), big, (0, 0, 0))


def test_residuum_closed_forms():
    im, jm, km = 33, 11, 21
    print(f"case residuum {(im, jm, km)}", flush=True)
    dr, dthe, dphi = 1.0 / (im - 1), math.pi / (jm - 1), 2 * math.pi / (km - 1)
    rad = [1.0]
    for _ in range(im - 1):
        rad.append(rad[-1] + dr)
    rad10 = rad[10]
    i, j, k = np.meshgrid(np.arange(im), np.arange(jm), np.arange(km), indexing="ij")
    s = 2.0 ** -9   # multiples of a power of two subtract exactly: equal differences tie exactly
    with A.Integration((im, jm, km)) as integ:
        integ.init_state(1)
        zero = np.zeros((im, jm, km))
        # u linear in i: du/dr = s / dr / (rm + 1), largest at the lowest scanned level
        integ.set("u", s * i); integ.set("v", zero); integ.set("w", zero)
        r = integ.residuum()
        assert r["pos_max"] == (10, 1, 1) and r["n_nonfinite"] == 0
        assert math.isclose(r["max"], s / dr / (rad10 + 1.0) / math.sqrt(3.0), rel_tol=1e-12)
        assert r["n_finite"] == 21 * (jm - 2) * (km - 2)
        rms = np.array([s / dr / (rad[ii] + 1.0) / math.sqrt(3.0) for ii in range(10, 31)])
        assert math.isclose(r["avg_abs"], rms.mean(), rel_tol=1e-12)
        # no velocity gradient: residual 0 everywhere, reported at (0, 0, 0)
        integ.set("u", np.full((im, jm, km), 2.0))
        r = integ.residuum()
        assert r["max"] == 0.0 and r["pos_max"] == (0, 0, 0) and r["avg_abs"] == 0.0
        assert r["n_finite"] == 21 * (jm - 2) * (km - 2) and r["pos_first_nonfinite"] == (-1, -1, -1)
        # v linear in j: dv/dtheta = s / (rm dtheta)
        integ.set("u", zero); integ.set("v", s * j)
        r = integ.residuum()
        assert r["pos_max"] == (10, 1, 1)
        assert math.isclose(r["max"], s / (rad10 * dthe) / math.sqrt(3.0), rel_tol=1e-12)
        # w linear in k: dw/dphi = s / (rm max(0.55, sin theta) dphi), floored rows tie: first j
        integ.set("v", zero); integ.set("w", s * k)
        r = integ.residuum()
        assert r["pos_max"] == (10, 1, 1)
        assert math.isclose(r["max"], s / (rad10 * 0.55 * dphi) / math.sqrt(3.0), rel_tol=1e-12)
        # a non-finite velocity makes its two radial neighbours' residuals non-finite
        u = s * i.astype(float)
        u[15, 5, 5] = np.nan
        integ.set("u", u); integ.set("w", zero)
        r = integ.residuum()
        assert r["n_nonfinite"] == 2 and r["pos_first_nonfinite"] == (14, 5, 5)
        assert r["n_finite"] == 21 * (jm - 2) * (km - 2) - 2


def residuum_numpy(u, v, w, rad, sint, dr, dthe, dphi):
    """Vectorised restatement of findResiduumAtm over i 10..30 and the interior (j, k)."""
    i = slice(10, 31)
    rm = np.asarray(rad)[i][:, None, None]
    st = np.asarray(sint)[1:-1][None, :, None]
    dudr = (u[11:32, 1:-1, 1:-1] - u[9:30, 1:-1, 1:-1]) / (2.0 * dr) * (1.0 / (rm + 1.0))
    dvdthe = (v[i, 2:, 1:-1] - v[i, :-2, 1:-1]) / (2.0 * rm * dthe)
    dwdphi = (w[i, 1:-1, 2:] - w[i, 1:-1, :-2]) / (2.0 * (rm * st) * dphi)
    return np.sqrt((dudr * dudr + dvdthe * dvdthe + dwdphi * dwdphi) / 3.0)


def test_residuum_of_a_developed_state():
    im, jm, km = 41, 21, 41
    print(f"case residuum developed {(im, jm, km)}", flush=True)
    dr, dthe, dphi = 1.0 / (im - 1), (180.0 / (jm - 1)) / (180.0 / math.pi), (360.0 / (km - 1)) / (180.0 / math.pi)
    rad, the = [1.0], [0.0]
    for _ in range(im - 1):
        rad.append(rad[-1] + dr)
    for _ in range(jm - 1):
        the.append(the[-1] + dthe)
    sint = [max(0.55, math.sin(t)) for t in the]
    with A.Integration((im, jm, km)) as integ:
        integ.init_state(7)
        integ.run(8)
        res = residuum_numpy(integ.get("u"), integ.get("v"), integ.get("w"), rad, sint, dr, dthe, dphi)
        r = integ.residuum()
        assert res.max() > 0.0
        assert math.isclose(r["max"], float(res.max()), rel_tol=1e-13)
        # the state is near mirror-symmetric, so maxima can tie to rounding: the reported cell
        # holds the maximum to rounding (exact tie-breaks are checked on the closed forms)
        pi, pj, pk = r["pos_max"]
        assert 10 <= pi <= 30 and 1 <= pj <= jm - 2 and 1 <= pk <= km - 2
        assert math.isclose(float(res[pi - 10, pj - 1, pk - 1]), float(res.max()), rel_tol=1e-13)
        assert r["n_finite"] == res.size and r["n_nonfinite"] == 0
        assert math.isclose(r["avg_abs"], float(res.mean()), rel_tol=1e-12)


def test_residuum_needs_32_levels():
    print("case residuum im < 32", flush=True)
    with A.Integration((11, 21, 41)) as integ:
        integ.init_state(1)
        with pytest.raises(A.AthadError, match="status 4"):
            integ.residuum()


TWO_PI_A = 2.0 * math.pi * (6370.001 * 1000.0)   # 2 pi r_Earth, r_Earth = 6370.001 km
U_0 = 8.0


def layer_heights(im):
    """init_layer_heights: (exp(zeta (r - r0)) - 1) L_atm, stored as float."""
    dr, rad = 1.0 / (im - 1), [1.0]
    for _ in range(im - 1):
        rad.append(rad[-1] + dr)
    return np.array([np.float32((math.exp(3.0 * (r - rad[0])) - 1.0) * 6287.5) for r in rad], dtype=np.float32)


def colatitudes(jm):
    dthe, the = (180.0 / (jm - 1)) / (180.0 / math.pi), [0.0]
    for _ in range(jm - 1):
        the.append(the[-1] + dthe)
    return the


def streamfunction_numpy(v, rho, r_air):
    """write_meridional_streamfunction's loops: sequential sums over k, integral down in i."""
    im, jm, _ = v.shape
    fin = np.isfinite(v)
    rho = np.where(np.isfinite(rho) & (rho > 0.0), rho, r_air)
    vv = np.where(fin, v, 0.0)
    s = np.cumsum(vv, axis=2)[..., -1]            # cumsum adds in order, as the reference's loop
    rs = np.cumsum(rho * vv, axis=2)[..., -1]
    n = fin.sum(axis=2)
    vbar = np.where(n > 0, (s / np.maximum(n, 1)) * U_0, 0.0)
    rvbar = np.where(n > 0, (rs / np.maximum(n, 1)) * U_0, 0.0)
    h = layer_heights(im)
    coeff = TWO_PI_A * np.sin(np.array(colatitudes(jm)))
    psi = np.zeros((im, jm))
    for i in range(im - 2, -1, -1):
        dz = np.float64(h[i + 1] - h[i])
        psi[i] = psi[i + 1] + coeff * (0.5 * (rvbar[i] + rvbar[i + 1])) * dz
    return vbar, rvbar, psi


def test_streamfunction_closed_form():
    im, jm, km = 21, 19, 41
    print(f"case streamfunction closed form {(im, jm, km)}", flush=True)
    r_air = A.default_params().r_air
    c, r = 2.0 ** -4, 0.5   # every partial sum of these is exact
    with A.Integration((im, jm, km)) as integ:
        integ.init_state(1)
        integ.set("v", np.full((im, jm, km), c))
        integ.set("r_humid", np.full((im, jm, km), r))
        out = integ.meridional_streamfunction()
        assert (out["vbar"] == c * U_0).all() and (out["rho_vbar"] == r * c * U_0).all()
        # constant rho v: psi(i) = 2 pi a sin(theta) rho v u_0 (sum of the layer thicknesses above
        # i), each thickness the float difference of the float heights (get_layer_height returns
        # float); the double difference of the same heights is off by ~5e-10 here
        h = layer_heights(im)
        dz = (h[1:] - h[:-1]).astype(np.float64)
        above = np.concatenate([np.cumsum(dz[::-1])[::-1], [0.0]])
        coeff = TWO_PI_A * np.sin(np.array(colatitudes(jm)))
        expect = coeff[None, :] * (r * c * U_0) * above[:, None]
        assert (out["psi"][-1] == 0.0).all()
        np.testing.assert_allclose(out["psi"], expect, rtol=1e-13, atol=1e-6)
        # a non-finite v leaves the mean; a non-positive density is replaced by r_air
        v = np.full((im, jm, km), c)
        v[3, 4, 5] = np.nan
        rho = np.full((im, jm, km), r)
        rho[6, 7, 8] = -1.0
        integ.set("v", v)
        integ.set("r_humid", rho)
        out = integ.meridional_streamfunction()
        assert out["vbar"][3, 4] == ((c * (km - 1)) / (km - 1)) * U_0
        assert out["rho_vbar"][3, 4] == ((r * c * (km - 1)) / (km - 1)) * U_0
        assert out["rho_vbar"][6, 7] == ((r * c * (km - 1) + r_air * c) / km) * U_0
        # no finite v in a row: 0
        v[2, 3, :] = np.inf
        integ.set("v", v)
        out = integ.meridional_streamfunction()
        assert out["vbar"][2, 3] == 0.0 and out["rho_vbar"][2, 3] == 0.0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_streamfunction_of_a_developed_state(precision):
    grid = (21, 37, 73)
    print(f"case streamfunction developed {grid} {precision}", flush=True)
    with A.Integration(grid, precision=precision) as integ:
        integ.init_state(3)
        integ.run(10)
        out = integ.meridional_streamfunction()
        vbar, rvbar, psi = streamfunction_numpy(integ.get("v"), integ.get("r_humid"), A.default_params().r_air)
        np.testing.assert_array_equal(out["vbar"], vbar)
        np.testing.assert_array_equal(out["rho_vbar"], rvbar)
        np.testing.assert_array_equal(out["psi"], psi)
        assert np.abs(psi).max() > 0.0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_slices_are_the_planes_of_the_field(precision):
    grid = (13, 21, 41)
    im, jm, km = grid
    print(f"case slices {grid} {precision}", flush=True)
    with A.Integration(grid, precision=precision) as integ:
        integ.init_state(5)
        integ.run(4)
        for name in ["t", "v", "w", "p_dyn", "p_stat", "c"]:
            x = integ.get(name)
            for i in (0, im // 2, im - 1):
                np.testing.assert_array_equal(integ.get_slice(name, 0, i), x[i])
            for j in (0, jm // 2, jm - 1):
                np.testing.assert_array_equal(integ.get_slice(name, 1, j), x[:, j, :])
            for k in (0, km // 3, km - 1):
                np.testing.assert_array_equal(integ.get_slice(name, 2, k), x[:, :, k])
        for axis, index in [(0, im), (1, -1), (2, km), (3, 0)]:
            with pytest.raises(A.AthadError, match="status 1"):
                integ.get_slice("t", axis, index)
