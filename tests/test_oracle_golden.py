"""The oracle against golden vectors written by the reference's own compiled code (CPU).

tests/golden/ref_vectors.npz holds, at the reference's built-in grid 61 x 181 x 361, the
reference's outputs (oracle/_ref/ref_harness, compiled from the reference's sources by
oracle/Makefile.ref and run by tests/golden/make_ref_vectors.py) for
  - the thermodynamic init chain (whole columns at five latitudes),
  - two whole scoped iterations in the reference's single-thread order,
  - each operator once, on a state spun up by the oracle.
Inputs are regenerated here from the same seed; a sample of 512 cells per field and every
field's sum and L2 norm are compared.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "ref_vectors.npz")

pytestmark = pytest.mark.skipif(not os.path.exists(GOLD), reason="golden vectors not generated")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2)) / den) if den > 0 else float(np.sqrt(np.sum((a - b) ** 2)))


def check_pack(prefix, snap, gold, tol, skip=()):
    idx = gold["index"]
    bad = {}
    for name in snap:
        if name in skip or f"{prefix}/{name}/sample" not in gold:
            continue
        a = snap[name].reshape(-1)
        e = rel(a[idx], gold[f"{prefix}/{name}/sample"])
        s_ref, s = float(gold[f"{prefix}/{name}/sum"]), float(np.sum(a))
        l_ref, l2 = float(gold[f"{prefix}/{name}/l2"]), float(np.sqrt(np.sum(a * a)))
        e = max(e, abs(s - s_ref) / max(abs(s_ref), l_ref, 1e-300), abs(l2 - l_ref) / max(l_ref, 1e-300))
        if e > tol:
            bad[name] = e
    assert not bad, f"{prefix}: {bad}"


def test_init_chain(oracle, gold):
    g = tuple(int(x) for x in gold["grid"])
    st = oracle.OracleState(*g, thermo_only=True)
    for name in ["t", "p_stat", "r_humid", "r_dry", "c", "co2"]:
        got = st.field(name)[:, [0, 45, 90, 135, 180], 0]
        assert rel(got, gold[f"init/{name}"]) < 1e-13, name
    assert rel(st.field2d("t_top_init")[:, 0], gold["init/t_top_init"]) < 1e-13


def test_two_steps_lex(oracle, gold):
    g = tuple(int(x) for x in gold["grid"])
    st = oracle.OracleState(*g, seed=int(gold["seed"]))
    st.run(2, "lex")
    # S_c_c is a diagnostic whose Newton break-off can flip at the 1e-16 level; it never
    # feeds back into the prognostic fields.
    check_pack("steps2", st.snapshot(), gold, 1e-11, skip=("S_c_c",))


TURB_GOLD = os.path.join(HERE, "golden", "ref_turb.npz")


@pytest.mark.parametrize("model", ["k_epsilon", "k_omega", "k_omega_SST"])
def test_turbulence_closures(oracle, model):
    """Each closure of TurbulenceAtm (ref_turb.npz): TurbulenceAtm::init on the nominal state
    (the boundary-layer profile, the closure's wall value, compute_sources), then two whole
    iterations in the reference's single-thread order with the closure's RHS terms."""
    gold = np.load(TURB_GOLD, allow_pickle=False)
    g = tuple(int(x) for x in gold["grid"])
    st = oracle.OracleState(*g, seed=int(gold["seed"]))
    st.set_turb_model(model)
    st.turbulence_init()
    check_pack(f"{model}/init", st.snapshot(), gold, 1e-13)
    # the plane-parallel sweep, bit for bit the sequential lex order
    # (test_lex_planes_equal_sequential_lex)
    st.run(int(gold["steps"]), "lex_planes")
    check_pack(f"{model}/steps", st.snapshot(), gold, 1e-11, skip=("S_c_c",))


def test_turbulence_closures_differ():
    """The golden outputs of the three closures are not one another's: dis is epsilon in the
    k-epsilon closure (orders of magnitude below omega) and the k-omega and SST sources differ."""
    gold = np.load(TURB_GOLD, allow_pickle=False)
    l2 = lambda m, f: float(gold[f"{m}/steps/{f}/l2"])
    assert l2("k_epsilon", "dis") < 1e-3 * l2("k_omega", "dis")
    assert abs(l2("k_omega", "dis_source") - l2("k_omega_SST", "dis_source")) > 0.1 * l2("k_omega_SST", "dis_source")
    assert abs(l2("k_omega", "nue") - l2("k_omega_SST", "nue")) > 1e-4 * l2("k_omega_SST", "nue")


def test_unknown_turbulence_model_is_refused(oracle):
    st = oracle.OracleState(11, 21, 41, seed=1)
    with pytest.raises(ValueError):
        st.set_turb_model(3)
    with pytest.raises(ValueError):
        st.set_turb_model("laminar")


OPS = ["poisson", "satadj", "damp_moist", "moist_clamp", "densities", "evaporation", "turbulence",
       "damp_turb", "bc", "rk4", "polar", "radial", "project"]


@pytest.fixture(scope="module")
def spun(oracle, gold):
    g = tuple(int(x) for x in gold["grid"])
    st = oracle.OracleState(*g, seed=int(gold["seed"]))
    st.run(int(gold["spin"]), "jacobi")
    return st


@pytest.mark.parametrize("op", OPS)
def test_operator(oracle, gold, spun, op):
    g = tuple(int(x) for x in gold["grid"])
    st = oracle.OracleState(*g, blank=True)
    st.load(spun.snapshot())
    st.set_iter(int(gold["spin"]), spun.total_iter)
    st.op(op, "lex")
    tol = 1e-10 if op == "satadj" else 1e-12
    check_pack(f"op/{op}", st.snapshot(), gold, tol, skip=("S_c_c",))


GOLD_RAD = os.path.join(HERE, "golden", "ref_radiation.npz")


def check_rad(prefix, rad, gold, tol):
    for name, ix in (("radiation", gold["index"]), ("epsilon", gold["index"]), ("albedo", gold["index2d"])):
        a = rad[name].reshape(-1)
        assert rel(a[ix], gold[f"{prefix}/{name}/sample"]) < tol, name
        s_ref, l_ref = float(gold[f"{prefix}/{name}/sum"]), float(gold[f"{prefix}/{name}/l2"])
        assert abs(float(np.sum(a)) - s_ref) <= tol * max(abs(s_ref), l_ref), name
        assert abs(float(np.sqrt(np.sum(a * a))) - float(gold[f"{prefix}/{name}/l2"])) <= tol * float(gold[f"{prefix}/{name}/l2"]), name
    assert rel(rad["sw"], gold[f"{prefix}/sw"]) < 1e-15
    assert abs(rad["t_skin"] - float(gold[f"{prefix}/t_skin"])) < 1e-11 * float(gold[f"{prefix}/t_skin"])


@pytest.mark.skipif(not os.path.exists(GOLD_RAD), reason="radiation vectors not generated")
def test_radiation_once(oracle, spun):
    gold = np.load(GOLD_RAD, allow_pickle=False)
    g = tuple(int(x) for x in gold["grid"])
    st = oracle.OracleState(*g, blank=True)
    st.load(spun.snapshot())
    st.set_iter(int(gold["spin"]), spun.total_iter)
    st.radiation()
    check_pack("op/radiation", st.snapshot(), gold, 1e-13, skip=("S_c_c",))
    check_rad("op/radiation", st.rad_outputs(), gold, 1e-13)


@pytest.mark.skipif(not os.path.exists(GOLD_RAD), reason="radiation vectors not generated")
def test_two_steps_with_radiation_lex(oracle):
    gold = np.load(GOLD_RAD, allow_pickle=False)
    g = tuple(int(x) for x in gold["grid"])
    st = oracle.OracleState(*g, seed=int(gold["seed"]))
    for _ in range(2):
        st.step("lex")
        st.radiation()
    check_pack("steps2_rad", st.snapshot(), gold, 1e-11, skip=("S_c_c",))
    # the albedo follows the cloud water, which two iterations carry to ~1e-11 (SatAdj)
    check_rad("steps2_rad", st.rad_outputs(), gold, 1e-10)


GOLD_ICE = os.path.join(HERE, "golden", "ref_ice.npz")
ICE_STATE = ["t", "c", "cloud", "ice", "gr", "S_c_c"]


@pytest.mark.skipif(not os.path.exists(GOLD_ICE), reason="ice-scheme vectors not generated")
def test_ice_scheme_twice(oracle):
    """TwoCatIceScheme (+ the damping of P_rain, P_snow) twice in a row on the spun-up state
    with planted condensate, fluxes and temperature spread (oracle.plant_microphysics), against
    the reference's compiled scheme on the same input. The reference is built with -ffast-math
    -O2, which reassociates the rate sums: 2e-14 measured, bar 1e-12."""
    gold = np.load(GOLD_ICE, allow_pickle=False)
    g = tuple(int(x) for x in gold["grid"])
    spin = int(gold["spin"])
    st = oracle.OracleState(*g, seed=int(gold["seed"]))
    st.run(spin, "jacobi")
    snap = st.snapshot()
    mp = {k: np.zeros(g) for k in oracle.MP_FIELDS}
    oracle.plant_microphysics(snap, mp, int(gold["plant"]))
    st.load(snap)
    for k, v in mp.items():
        st.mp(k)[...] = v
    idx = gold["index"]
    for call in (1, 2):
        st.ice_scheme()
        got = {k: st.field(k) for k in ICE_STATE}
        got.update(st.mp_snapshot())
        bad = {}
        for name, a in got.items():
            a = a.reshape(-1)
            l_ref = float(gold[f"call{call}/{name}/l2"])
            e = rel(a[idx], gold[f"call{call}/{name}/sample"])
            e = max(e, abs(float(np.sum(a)) - float(gold[f"call{call}/{name}/sum"])) / max(l_ref, 1e-300),
                    abs(float(np.sqrt(np.sum(a * a))) - l_ref) / max(l_ref, 1e-300))
            if e > 1e-12:
                bad[name] = e
        assert not bad, f"call {call}: {bad}"
        # the scheme did something: rain and snow fall, every source is live
        for name in ("P_rain", "P_snow", "S_v", "S_c", "S_i", "S_r", "S_s"):
            assert float(gold[f"call{call}/{name}/l2"]) > 0.0, name
        st.ice_store()


@pytest.mark.skipif(not os.path.exists(GOLD_ICE), reason="ice-scheme vectors not generated")
def test_two_steps_with_the_ice_scheme_lex(oracle):
    """Two whole iterations (the second a moist one) with the ice scheme in the moist block,
    its sources in the RHS, the source boundaries and the evaporation of falling precipitation
    in saturation adjustment, in the reference's single-thread order, from the planted state.
    S_c_c can flip at the Newton break-off as in the scoped step: 1.1e-10 on the sampled cells
    (4.7e-14 over the whole field after four iterations), bar 1e-8; every other array 1e-11."""
    gold = np.load(GOLD_ICE, allow_pickle=False)
    g = tuple(int(x) for x in gold["grid"])
    spin = int(gold["spin"])
    st = oracle.OracleState(*g, seed=int(gold["seed"]))
    st.run(spin, "jacobi")
    snap = st.snapshot()
    mp = {k: np.zeros(g) for k in oracle.MP_FIELDS}
    oracle.plant_microphysics(snap, mp, int(gold["plant"]))
    st = oracle.OracleState(*g, blank=True)
    st.load(snap)
    st.set_iter(spin, spin)
    for k, v in mp.items():
        st.mp(k)[...] = v
    st.set_ice_scheme(True)
    st.run(2, "lex")
    got = st.snapshot()
    got.update(st.mp_snapshot())
    idx = gold["index"]
    bad = {}
    for name in oracle.FIELDS + oracle.MP_FIELDS:
        a = got[name].reshape(-1)
        l_ref = float(gold[f"steps2_ice/{name}/l2"])
        e = rel(a[idx], gold[f"steps2_ice/{name}/sample"])
        e = max(e, abs(float(np.sum(a)) - float(gold[f"steps2_ice/{name}/sum"])) / max(l_ref, 1e-300),
                abs(float(np.sqrt(np.sum(a * a))) - l_ref) / max(l_ref, 1e-300))
        if e > (1e-8 if name == "S_c_c" else 1e-11):
            bad[name] = e
    assert not bad, bad


GOLD_CONV = os.path.join(HERE, "golden", "ref_conv.npz")


def conv_input(oracle, gold):
    """The convection vectors' input: the Jacobi-spun-up state with the ice-scheme vectors'
    planted condensate, and the seeded evaporation diagnostic (make_ref_vectors.conv_evaporation)."""
    g = tuple(int(x) for x in gold["grid"])
    spin = int(gold["spin"])
    st = oracle.OracleState(*g, seed=int(gold["seed"]))
    st.run(spin, "jacobi")
    snap = st.snapshot()
    oracle.plant_microphysics(snap, {k: np.zeros(g) for k in oracle.MP_FIELDS}, int(gold["plant"]))
    st = oracle.OracleState(*g, blank=True)
    st.load(snap)
    st.set_iter(spin, spin)
    evap = np.random.default_rng(int(gold["evap_seed"])).uniform(-1.0, 6.0, g[1:])
    return st, snap, evap


@pytest.mark.skipif(not os.path.exists(GOLD_CONV), reason="moist-convection vectors not generated")
def test_moist_convection_twice(oracle):
    """MoistConvection::run twice in a row from zeroed convection arrays, against the reference's
    compiled scheme (-ffast-math -O2) on the same input: all 32 arrays, 2.6e-13 measured (g_p,
    whose precipitation threshold sits on the cloud-water recurrence), bar 1e-11. The input
    makes every branch live: updrafts, downdrafts, parcel condensation, precipitation that
    survives the downdraft evaporation."""
    gold = np.load(GOLD_CONV, allow_pickle=False)
    st, _, evap = conv_input(oracle, gold)
    st.evap()[...] = evap
    idx = gold["index"]
    for call in (1, 2):
        st.moist_convection()
        bad = {}
        for name in oracle.MC_FIELDS:
            a = st.mc(name).reshape(-1)
            l_ref = float(gold[f"call{call}/{name}/l2"])
            den = max(l_ref, 1e-300)
            e = rel(a[idx], gold[f"call{call}/{name}/sample"])
            e = max(e, abs(float(np.sum(a)) - float(gold[f"call{call}/{name}/sum"])) / den,
                    abs(float(np.sqrt(np.sum(a * a))) - l_ref) / den)
            if e > 1e-11:
                bad[name] = e
        assert not bad, f"call {call}: {bad}"
        for name in ("M_u", "M_d", "q_c_u", "c_u", "g_p", "e_d", "e_p", "P_conv", "MC_t", "MC_q", "MC_v", "MC_w"):
            assert float(gold[f"call{call}/{name}/l2"]) > 0.0, name
        # computeCAPE's diagnostic: each level holds the last column (k outer, j inner) that
        # reached it; zero where none did
        cape, ref = st.cape(), gold[f"call{call}/cape"]
        assert np.count_nonzero(ref) > 10
        assert np.array_equal(cape == 0.0, ref == 0.0)
        assert rel(cape, ref) <= 1e-11, rel(cape, ref)


GOLD_CONV_STEPS = os.path.join(HERE, "golden", "ref_conv_steps.npz")


def planted_mc(oracle, grid, seed=33):
    """Convection arrays at the start of the coupled steps (tests/golden/make_ref_vectors.py,
    conv_planted_mc): uniform seeded MC_t, MC_q, MC_v, MC_w and P_conv, the rest zero."""
    out = np.zeros((len(oracle.MC_FIELDS),) + tuple(grid))
    rng = np.random.default_rng(seed)
    for name, lo, hi in (("MC_t", -1e-3, 1e-3), ("MC_q", -1e-6, 1e-6), ("MC_v", -1e-4, 1e-4),
                         ("MC_w", -1e-4, 1e-4), ("P_conv", 0.0, 1e-3)):
        out[oracle.MC_FIELDS.index(name)] = rng.uniform(lo, hi, grid)
    return out


@pytest.mark.skipif(not os.path.exists(GOLD_CONV_STEPS), reason="coupled convection vectors not generated")
def test_steps_with_convection_lex(oracle):
    """Four lex iterations of the configured moist block (the ice scheme, then moist convection
    with the damping of eleven of its arrays and P_conv in precipitationSum, MC_* in the RHS, the
    Evaporation diagnostic written by the evaporation step), against the reference's own loop
    (harness mode steps_conv) from the same planted state, with seeded MC_* tendencies that the
    scheme carries where no column convects: every state field, ice-scheme array and convection
    array, and the Evaporation diagnostic."""
    gold = np.load(GOLD_CONV_STEPS, allow_pickle=False)
    g = tuple(int(x) for x in gold["grid"])
    spin = int(gold["spin"])
    st = oracle.OracleState(*g, seed=int(gold["seed"]))
    st.run(spin, "jacobi")
    snap = st.snapshot()
    mp = {k: np.zeros(g) for k in oracle.MP_FIELDS}
    oracle.plant_microphysics(snap, mp, int(gold["plant"]))
    st = oracle.OracleState(*g, blank=True)
    st.load(snap)
    st.set_iter(spin, spin)
    for k, v in mp.items():
        st.mp(k)[...] = v
    st.evap()[...] = np.random.default_rng(int(gold["evap_seed"])).uniform(-1.0, 6.0, g[1:])
    planted = planted_mc(oracle, g)
    for q, name in enumerate(oracle.MC_FIELDS):
        st.mc(name)[...] = planted[q]
    st.set_moist_convection(True)
    st.run(int(gold["steps"]), "lex_planes")
    snap = st.snapshot()
    snap.update({k: st.mp(k) for k in oracle.MP_FIELDS})
    snap.update({"mc/" + k: st.mc(k) for k in oracle.MC_FIELDS})
    check_pack("steps", snap, gold, 1e-10, skip=("S_c_c",))
    ev = gold["steps/evap"]
    assert rel(st.evap().reshape(-1), ev) < 1e-12
    assert float(gold["steps/mc/MC_t/l2"]) > 0.0 and float(gold["steps/mc/MC_q/l2"]) > 0.0
