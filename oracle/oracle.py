"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of bench.py. The product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FIELDS = [
    "t", "u", "v", "w", "c", "cloud", "ice", "gr", "co2", "tke", "dis",
    "tn", "un", "vn", "wn", "cn", "cloudn", "icen", "grn", "co2n", "tken", "disn",
    "p_dyn", "p_stat", "r_humid", "r_dry", "t_eq", "nue", "prod", "tke_source",
    "dis_source", "aux_u", "aux_v", "aux_w", "S_c_c",
]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
FIELDS2D = ["vel_star", "t_top_init"]
OPS = ["poisson", "satadj", "damp_moist", "moist_clamp", "densities", "evaporation",
       "turbulence", "damp_turb", "bc", "rk4", "polar", "radial", "store", "level_mean", "project"]
ORDER = {"lex": 0, "jacobi": 1, "lex_planes": 2, "jacobi_reverse": 3}
# the turbulence closures in TurbulenceAtm.h:24 order
TURB_MODELS = ["k_epsilon", "k_omega", "k_omega_SST"]


class Member(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "t_surf_equator", "t_surf_pole", "r_air", "x_H2O", "x_CO2", "x_N2",
        "c_0", "co2_0", "c_h2o_dry_top")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, d, i, u64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_uint64
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i, i, i, u64, ctypes.POINTER(Member)]
        L.orc_create_thermo.restype = vp
        L.orc_create_thermo.argtypes = [i, i, i, ctypes.POINTER(Member)]
        L.orc_create_blank.restype = vp
        L.orc_create_blank.argtypes = [i, i, i, ctypes.POINTER(Member)]
        L.orc_destroy.argtypes = [vp]
        L.orc_field_ptr.restype = dp
        L.orc_field_ptr.argtypes = [vp, i]
        L.orc_field2d_ptr.restype = dp
        L.orc_field2d_ptr.argtypes = [vp, i]
        L.orc_total_iter.restype = i
        L.orc_total_iter.argtypes = [vp]
        L.orc_set_iter.argtypes = [vp, i, i]
        L.orc_scalar.restype = d
        L.orc_scalar.argtypes = [vp, ctypes.c_char_p]
        L.orc_level_mean.restype = d
        L.orc_level_mean.argtypes = [vp, i]
        L.orc_step.argtypes = [vp, i]
        L.orc_run.argtypes = [vp, i, i]
        L.orc_set_sst_scan.argtypes = [vp, i]
        L.orc_set_turb_model.restype = i
        L.orc_set_turb_model.argtypes = [vp, i]
        L.orc_turbulence_init.argtypes = [vp]
        L.orc_set_moist_convection.argtypes = [vp, i]
        L.orc_set_threads.restype = i
        L.orc_set_threads.argtypes = [i]
        L.orc_op.argtypes = [vp, i, i]
        for name in ("orc_p_sat", "orc_p_subl", "orc_p_auto", "orc_latent_heat"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [d]
        L.orc_q_sat.restype = d
        L.orc_q_sat.argtypes = [d, d, d]
        L.orc_dqsat_dT.restype = d
        L.orc_dqsat_dT.argtypes = [d, d, d]
        L.orc_moist_lapse.restype = d
        L.orc_moist_lapse.argtypes = [d, d, d, d, d, d]
        L.orc_cp_of.restype = d
        L.orc_cp_of.argtypes = [d, d, d, d]
        L.orc_R_of.restype = d
        L.orc_R_of.argtypes = [d, d, d]
        L.orc_M_nonwater.restype = d
        L.orc_M_nonwater.argtypes = [d, d, d]
        L.orc_resolve.argtypes = [d, d, d, dp]
        L.orc_satadj_cells.argtypes = [i, dp, dp, dp, dp, dp, dp, dp, dp, d, i]
        L.orc_radiation.argtypes = [vp]
        L.orc_rad_ptr.restype = dp
        L.orc_rad_ptr.argtypes = [vp, i]
        L.orc_ice_scheme.argtypes = [vp]
        L.orc_ice_store.argtypes = [vp]
        L.orc_set_ice_scheme.argtypes = [vp, i]
        L.orc_mp_ptr.restype = dp
        L.orc_mp_ptr.argtypes = [vp, i]
        L.orc_moist_convection.argtypes = [vp]
        L.orc_cape.restype = ctypes.c_int
        L.orc_cape.argtypes = [vp, dp]
        L.orc_mc_ptr.restype = dp
        L.orc_mc_ptr.argtypes = [vp, i]
        L.orc_evap_ptr.restype = dp
        L.orc_evap_ptr.argtypes = [vp]
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's parallel loops (n <= 0: the OpenMP default); returns the
    number in use."""
    return lib().orc_set_threads(int(n))


def resolve(x_h2o, x_co2, x_n2):
    out = (ctypes.c_double * 7)()
    lib().orc_resolve(x_h2o, x_co2, x_n2, out)
    keys = ["M_mean", "R_mix", "q_H2O", "q_CO2", "q_bg", "M_bg", "R_bg"]
    return dict(zip(keys, list(out)))


class OracleState:
    """One grid's state in the oracle."""

    def __init__(self, im, jm, km, seed=1234, member=None, blank=False, thermo_only=False):
        L = lib()
        m = ctypes.byref(member) if member is not None else None
        if thermo_only:
            self.h = L.orc_create_thermo(im, jm, km, m)
        elif blank:
            self.h = L.orc_create_blank(im, jm, km, m)
        else:
            self.h = L.orc_create(im, jm, km, seed, m)
        if not self.h:
            raise ValueError("oracle rejected the grid")
        self.shape = (im, jm, km)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def field(self, name) -> np.ndarray:
        """A live (writable) view of one field."""
        p = lib().orc_field_ptr(self.h, FIELD_ID[name])
        return np.ctypeslib.as_array(p, shape=self.shape)

    def field2d(self, name) -> np.ndarray:
        p = lib().orc_field2d_ptr(self.h, FIELDS2D.index(name))
        return np.ctypeslib.as_array(p, shape=self.shape[1:])

    def snapshot(self) -> dict:
        d = {n: self.field(n).copy() for n in FIELDS}
        for n in FIELDS2D:
            d[n] = self.field2d(n).copy()
        return d

    def load(self, d: dict):
        for n in FIELDS:
            if n in d:
                self.field(n)[...] = d[n]
        for n in FIELDS2D:
            if n in d:
                self.field2d(n)[...] = d[n]

    @property
    def total_iter(self):
        return lib().orc_total_iter(self.h)

    def set_iter(self, iter_n, total):
        lib().orc_set_iter(self.h, iter_n, total)

    def scalar(self, name):
        return lib().orc_scalar(self.h, name.encode())

    def set_moist_convection(self, on=True):
        """Moist convection in the moist block after the ice scheme (switched on with it)."""
        lib().orc_set_moist_convection(self.h, 1 if on else 0)

    def set_sst_scan(self, on=True):
        """The reference's distance-to-land scans in the SST closure (same results, its cost)."""
        lib().orc_set_sst_scan(self.h, 1 if on else 0)

    def set_turb_model(self, model):
        """The turbulence closure, by name or TurbulenceAtm.h:24 index."""
        m = TURB_MODELS.index(model) if isinstance(model, str) else int(model)
        if lib().orc_set_turb_model(self.h, m) != 0:
            raise ValueError(f"unknown turbulence model {model!r}")

    def turbulence_init(self):
        """TurbulenceAtm::init with the state's closure."""
        lib().orc_turbulence_init(self.h)

    def step(self, order="lex"):
        lib().orc_step(self.h, ORDER[order])

    def run(self, n, order="lex"):
        lib().orc_run(self.h, n, ORDER[order])

    def op(self, name, order="lex"):
        lib().orc_op(self.h, OPS.index(name), ORDER[order])

    def radiation(self):
        """Mode-2 radiation after a step: MLR, the 5 % blend of t, one t_skin relaxation."""
        lib().orc_radiation(self.h)

    def rad_outputs(self) -> dict:
        im, jm, km = self.shape
        shapes = [(im, jm, km), (im, jm, km), (jm, km), (jm,)]
        out = {}
        for w, (name, shp) in enumerate(zip(RAD_OUTPUTS, shapes)):
            p = lib().orc_rad_ptr(self.h, w)
            out[name] = np.ctypeslib.as_array(p, shape=shp).copy() if p else None
        out["t_skin"] = self.scalar("t_skin")
        return out

    def ice_scheme(self):
        """The two-category ice scheme and the damping of its two precipitation fluxes."""
        lib().orc_ice_scheme(self.h)

    def set_ice_scheme(self, on=True):
        """The ice scheme inside every later step's moist block (CategoryIceScheme = 2)."""
        lib().orc_set_ice_scheme(self.h, 1 if on else 0)

    def ice_store(self):
        """P_rainn <- P_rain, P_snown <- P_snow, as the store keeps them."""
        lib().orc_ice_store(self.h)

    def mp(self, name) -> np.ndarray:
        """A live (writable) view of one ice-scheme array (MP_FIELDS)."""
        p = lib().orc_mp_ptr(self.h, MP_FIELDS.index(name))
        return np.ctypeslib.as_array(p, shape=self.shape)

    def mp_snapshot(self) -> dict:
        return {n: self.mp(n).copy() for n in MP_FIELDS}

    def moist_convection(self):
        """One call of MoistConvection::run on the state's fields (MoistConvection.h:106-140)."""
        lib().orc_moist_convection(self.h)

    def cape(self) -> np.ndarray:
        """computeCAPE's im values (MoistConvection.h:1259-1299) after the last convection call."""
        out = np.zeros(self.shape[0])
        if lib().orc_cape(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) != 0:
            raise RuntimeError("no convection call yet")
        return out

    def mc(self, name) -> np.ndarray:
        """A live (writable) view of one moist-convection array (MC_FIELDS)."""
        p = lib().orc_mc_ptr(self.h, MC_FIELDS.index(name))
        return np.ctypeslib.as_array(p, shape=self.shape)

    def evap(self) -> np.ndarray:
        """A live view of the per-column evaporation diagnostic [mm/d] convection reads."""
        return np.ctypeslib.as_array(lib().orc_evap_ptr(self.h), shape=self.shape[1:])

    def mc_snapshot(self) -> dict:
        return {n: self.mc(n).copy() for n in MC_FIELDS}

    def level_mean(self):
        return np.array([lib().orc_level_mean(self.h, i) for i in range(self.shape[0])])

    # Raw state file shared with oracle/ref_harness.cpp
    def save_bin(self, path, iter_n=0):
        im, jm, km = self.shape
        with open(path, "wb") as f:
            f.write(struct.pack("<5i", im, jm, km, iter_n, self.total_iter))
            for n in FIELDS:
                f.write(np.ascontiguousarray(self.field(n), dtype=np.float64).tobytes())
            for n in FIELDS2D:
                f.write(np.ascontiguousarray(self.field2d(n), dtype=np.float64).tobytes())


RAD_OUTPUTS = ["radiation", "epsilon", "albedo", "sw"]
# Arrays of the two-category ice scheme, in the order of oracle/athad_oracle.h's orc_mp.
MC_FIELDS = ["s", "M_u", "M_d", "s_u", "q_v_u", "q_c_u", "c_u", "e_l", "e_d", "e_p", "g_p", "s_d", "q_v_d",
             "P_conv", "E_u", "D_u", "u_u", "v_u", "w_u", "E_d", "D_d", "u_d", "v_d", "w_d", "MC_t", "MC_q",
             "MC_v", "MC_w", "Deep_beg", "Deep_end", "CloudBase", "LevelFreeSinking"]
MP_FIELDS = ["P_rain", "P_snow", "P_rainn", "P_snown", "Precipitation", "S_v", "S_c", "S_i", "S_r", "S_s"]


def _unit(seed, stream, n):
    """n uniform doubles in [0, 1): SplitMix64 of (seed, stream, index), 53 high bits.
    Integer arithmetic only, so the values do not depend on the numpy version."""
    with np.errstate(over="ignore"):
        x = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(stream) * np.uint64(0xD1B54A32D192ED03)
             + np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def plant_microphysics(fields: dict, mp: dict, seed: int):
    """Deterministic condensate, falling precipitation and temperature spread that drive every
    branch of the two-category ice scheme: cloud water above and below the autoconversion
    threshold, ice absent in a third of the cells (nucleation), stored rain and snow fluxes up
    to the flux cap, vapour scattered around saturation, and the isothermal upper levels spread
    over 224-259 K (homogeneous freezing, nucleation and deposition bands). Modifies the
    arrays in place."""
    shape = fields["t"].shape
    n = int(np.prod(shape))
    u = lambda q: _unit(seed, q, n).reshape(shape)   # noqa: E731
    fields["cloud"][...] = 2.0e-3 * u(1)
    ice = 6.0e-4 * u(2)
    ice[u(3) < 0.3] = 0.0
    fields["ice"][...] = ice
    fields["c"][...] *= 0.9 + 0.2 * u(4)
    fields["S_c_c"][...] = 2.0e-6 * (u(5) - 0.5)
    t = fields["t"]
    cold = t * 273.15 < 255.0
    t[...] = np.where(cold, t * (0.88 + 0.14 * u(6)), t * (0.99 + 0.02 * u(7)))
    rain = 4.0e-4 * u(8)
    rain[u(9) < 0.1] = 0.0
    snow = 4.0e-4 * u(10)
    snow[u(11) < 0.1] = 0.0
    mp["P_rainn"][...] = rain
    mp["P_snown"][...] = snow


def save_mp(path, d: dict):
    """Ice-scheme arrays next to a state file, in MP_FIELDS order (oracle/ref_harness.cpp)."""
    with open(path, "wb") as f:
        for n in MP_FIELDS:
            f.write(np.ascontiguousarray(d[n], dtype=np.float64).tobytes())


def read_mp(path, shape):
    raw = np.fromfile(path, dtype=np.float64)
    n = int(np.prod(shape))
    assert raw.size == len(MP_FIELDS) * n, "ice-scheme file size"
    return {name: raw[q * n:(q + 1) * n].reshape(shape) for q, name in enumerate(MP_FIELDS)}


def read_rad(path, shape):
    """Radiation outputs written by oracle/ref_harness.cpp next to a state file."""
    im, jm, km = shape
    n = im * jm * km
    raw = np.fromfile(path, dtype=np.float64)
    assert raw.size == 2 * n + jm * km + jm + 1, "radiation file size"
    return {"radiation": raw[:n].reshape(shape), "epsilon": raw[n:2 * n].reshape(shape),
            "albedo": raw[2 * n:2 * n + jm * km].reshape(jm, km), "sw": raw[2 * n + jm * km:-1],
            "t_skin": float(raw[-1])}


def read_bin(path):
    with open(path, "rb") as f:
        im, jm, km, iter_n, total = struct.unpack("<5i", f.read(20))
        n = im * jm * km
        d = {}
        for name in FIELDS:
            d[name] = np.frombuffer(f.read(n * 8), dtype=np.float64).reshape(im, jm, km).copy()
        for name in FIELDS2D:
            d[name] = np.frombuffer(f.read(jm * km * 8), dtype=np.float64).reshape(jm, km).copy()
    d["_iter_n"], d["_total"] = iter_n, total
    return d


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    if den == 0.0:
        return float(num)
    return float(num / den)
