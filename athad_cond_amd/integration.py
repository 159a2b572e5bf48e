"""Host-side mirror of the reference's model object for the scoped time step.

``Integration`` owns one device context holding one or more members (grids). Its
methods are the operators that ``cAtmosphereModel::run_3D_loop``
(atmosphere/cAtmosphereModel.cpp:744-1728) calls in the scoped step, in the same order
and with the same names where the reference has them; ``step()`` is one iteration.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Sequence

import numpy as np

from . import _lib as L


class Integration:
    def __init__(self, grids: Sequence[tuple] | tuple, params: Iterable | None = None,
                 precision: str = "fp64", order: str = "jacobi"):
        if isinstance(grids, tuple) and len(grids) == 3 and isinstance(grids[0], int):
            grids = [grids]
        self.grids = [tuple(int(x) for x in g) for g in grids]
        n = len(self.grids)
        garr = (L.Grid * n)(*[L.Grid(*g) for g in self.grids])
        if params is None:
            parr = None
        else:
            params = list(params)
            parr = (L.MemberParams * n)(*params)
        self.precision = {"fp64": L.FP64, "mixed": L.MIXED}[precision]
        self.order = {"lex": L.LEX, "jacobi": L.JACOBI, "symmetric": L.SYMMETRIC}[order]
        h = ctypes.c_void_p()
        st = L.lib().athad_create(n, garr, parr, self.precision, self.order, ctypes.byref(h))
        L.check(st, None)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            L.lib().athad_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- state ------------------------------------------------------------------------
    def _chk(self, st):
        L.check(st, self.h)

    def init_state(self, seed: int = 1234):
        self._chk(L.lib().athad_init_state(self.h, seed))

    def get(self, name: str, member: int = 0) -> np.ndarray:
        g = self.grids[member]
        if name in L.FIELDS2D:
            out = np.empty(g[1:], dtype=np.float64)
            self._chk(L.lib().athad_get_field2d(self.h, member, L.FIELDS2D.index(name),
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
            return out
        out = np.empty(g, dtype=np.float64)
        self._chk(L.lib().athad_get_field(self.h, member, L.FIELD_ID[name],
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set(self, name: str, value: np.ndarray, member: int = 0):
        a = np.ascontiguousarray(value, dtype=np.float64)
        ptr = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if name in L.FIELDS2D:
            self._chk(L.lib().athad_set_field2d(self.h, member, L.FIELDS2D.index(name), ptr))
        else:
            self._chk(L.lib().athad_set_field(self.h, member, L.FIELD_ID[name], ptr))

    def load(self, state: dict, member: int = 0):
        for n in L.FIELDS + L.FIELDS2D:
            if n in state:
                self.set(n, state[n], member)

    def snapshot(self, member: int = 0, names=None) -> dict:
        return {n: self.get(n, member) for n in (names or (L.FIELDS + L.FIELDS2D))}

    def set_iter(self, iter_n: int, total: int, member: int | None = None):
        ms = range(len(self.grids)) if member is None else [member]
        for m in ms:
            self._chk(L.lib().athad_set_iter(self.h, m, iter_n, total))

    def get_iter(self, member: int = 0):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._chk(L.lib().athad_get_iter(self.h, member, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # ---- operators (run_3D_loop call sites) ------------------------------------------
    def project_initial_velocity(self, n_sweeps: int = 200):
        """PressureSolverAtm::project_initial_velocity (PressureSolverAtm.h:686-767)."""
        self._chk(L.lib().athad_project_initial_velocity(self.h, n_sweeps))

    def poisson_sweep(self):
        self._chk(L.lib().athad_poisson_sweep(self.h))

    def saturation_adjustment(self):
        self._chk(L.lib().athad_satadj(self.h))

    def damp_moist(self):
        self._chk(L.lib().athad_damp_moist(self.h))

    def densities(self, set_water_profile=False, with_evaporation=False):
        self._chk(L.lib().athad_densities(self.h, int(set_water_profile), int(with_evaporation)))

    def turbulence(self):
        self._chk(L.lib().athad_turbulence(self.h))

    def set_turbulence_model(self, model):
        """The turbulence closure of every member: "k_epsilon", "k_omega" or "k_omega_SST"
        (the default), or its TurbulenceAtm.h:24 index."""
        m = L.TURB_MODELS.index(model) if isinstance(model, str) else int(model)
        self._chk(L.lib().athad_set_turbulence_model(self.h, m))

    def turbulence_init(self):
        """TurbulenceAtm::init with the current closure (TurbulenceAtm.h:39-59)."""
        self._chk(L.lib().athad_turbulence_init(self.h))

    def boundary(self):
        self._chk(L.lib().athad_boundary(self.h))

    def rk4(self):
        self._chk(L.lib().athad_rk4(self.h))

    def filters(self):
        self._chk(L.lib().athad_filters(self.h))

    def store(self):
        self._chk(L.lib().athad_store(self.h))

    def step(self):
        self._chk(L.lib().athad_step(self.h))

    def run(self, n: int, graph: bool = False):
        self._chk(L.lib().athad_run(self.h, n, int(graph)))

    def field_stats(self, name: str, member: int = 0):
        """(sum, sum of squares) of a field, reduced on the device in a fixed order."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._chk(L.lib().athad_field_stats(self.h, member, L.FIELD_ID[name], ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_member_params(self, params, member: int = 0):
        """New surface state / composition for a member (reuse of a context); then init_state."""
        self._chk(L.lib().athad_set_member_params(self.h, member, ctypes.byref(params)))

    def damp_tile(self):
        """(tj, tk, levels) of the fused damping tile chosen for this context."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(L.lib().athad_damp_tile(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_damp_tile(self, shape: int, levels: int):
        self._chk(L.lib().athad_set_damp_tile(self.h, shape, levels))

    def busy(self) -> bool:
        """True while work of this context is still in flight (non-blocking)."""
        st = L.lib().athad_query(self.h)
        if st == L.NOT_READY:
            return True
        self._chk(st)
        return False

    def sync(self):
        self._chk(L.lib().athad_sync(self.h))

    # ---- restart files (the reference's ATM1 checkpoint, FileIO_Atm.cpp:293-356) --------
    def save_state(self, path: str | None = None, member: int = 0):
        """Writes member's restart to `path`, or returns it as bytes when path is None."""
        if path is not None:
            self._chk(L.lib().athad_save_state(self.h, member, os.fsencode(path)))
            return None
        buf = ctypes.create_string_buffer(restart_bytes(self.grids[member]))
        self._chk(L.lib().athad_save_state_mem(self.h, member, buf, len(buf)))
        return buf.raw

    def load_state(self, src, member: int = 0):
        """Loads a restart from a path or from bytes (total iteration count from its header)."""
        if isinstance(src, (bytes, bytearray)):
            buf = ctypes.create_string_buffer(bytes(src), len(src))
            self._chk(L.lib().athad_load_state_mem(self.h, member, buf, len(src)))
        else:
            self._chk(L.lib().athad_load_state(self.h, member, os.fsencode(src)))

    # ---- radiation (run_3D_loop's radiation_mode 2) ------------------------------------
    def radiation(self):
        """MultiLayerRadiation + the 5 % blend of t + one t_skin relaxation, every member."""
        self._chk(L.lib().athad_radiation(self.h))

    def set_radiation(self, on: bool = True):
        """Radiation (mode 2) and the skin-temperature step after every later step's store."""
        self._chk(L.lib().athad_set_radiation(self.h, 1 if on else 0))

    def radiation_outputs(self, member: int = 0) -> dict:
        im, jm, km = self.grids[member]
        out = {"radiation": np.empty((im, jm, km)), "epsilon": np.empty((im, jm, km)), "albedo": np.empty((jm, km))}
        ts = ctypes.c_double()
        ptr = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in out.values()]
        self._chk(L.lib().athad_get_radiation(self.h, member, *ptr, ctypes.byref(ts)))
        out["t_skin"] = ts.value
        return out

    # ---- two-category ice scheme (CategoryIceScheme = 2) -------------------------------
    def ice_scheme(self):
        """TwoCatIceScheme::run and the damping of P_rain, P_snow, every member."""
        self._chk(L.lib().athad_ice_scheme(self.h))

    def set_ice_scheme(self, on: bool = True):
        """CategoryIceScheme = 2 inside every later step's moist block."""
        self._chk(L.lib().athad_set_ice_scheme(self.h, 1 if on else 0))

    def set_moist_convection(self, on: bool = True):
        """Moist convection in the moist block after the ice scheme (switched on with it)."""
        self._chk(L.lib().athad_set_moist_convection(self.h, 1 if on else 0))

    def moist_convection(self):
        """MoistConvection::run on every member's current fields (MoistConvection.h:106-140)."""
        self._chk(L.lib().athad_moist_convection(self.h))

    def _conv_shape(self, name, member):
        return tuple(self.grids[member])[1:] if name == "Evaporation" else tuple(self.grids[member])

    def get_conv(self, name: str, member: int = 0) -> np.ndarray:
        """One convection array (CONV_FIELDS; "Evaporation" is the jm x km input) as FP64."""
        out = np.empty(self._conv_shape(name, member))
        self._chk(L.lib().athad_get_conv_field(self.h, member, L.CONV_FIELDS.index(name),
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def conv_cape(self, member: int = 0) -> np.ndarray:
        """computeCAPE's im values after the last convection call (MoistConvection.h:1259-1299)."""
        out = np.empty(self.grids[member][0])
        self._chk(L.lib().athad_get_conv_cape(self.h, member, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set_conv(self, name: str, arr, member: int = 0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self._conv_shape(name, member), "shape mismatch"
        self._chk(L.lib().athad_set_conv_field(self.h, member, L.CONV_FIELDS.index(name),
                                               a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def get_mp(self, name: str, member: int = 0) -> np.ndarray:
        """One ice-scheme array (MP_FIELDS) as FP64."""
        out = np.empty(self.grids[member])
        self._chk(L.lib().athad_get_mp_field(self.h, member, L.MP_FIELDS.index(name),
                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set_mp(self, name: str, arr, member: int = 0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == tuple(self.grids[member]), "shape mismatch"
        self._chk(L.lib().athad_set_mp_field(self.h, member, L.MP_FIELDS.index(name),
                                             a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    # ---- checkpoint diagnostics on the device (MinMax_Atm.cpp, UtilsAtm.h) --------------
    def field_minmax(self, name: str, member: int = 0):
        """searchMinMax_3D: (min, (i, j, k), max, (i, j, k)) of the finite values."""
        lo, hi = ctypes.c_double(), ctypes.c_double()
        plo, phi = (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
        self._chk(L.lib().athad_field_minmax(self.h, member, L.FIELD_ID[name], ctypes.byref(lo), plo,
                                             ctypes.byref(hi), phi))
        return lo.value, tuple(plo), hi.value, tuple(phi)

    def residuum(self, member: int = 0) -> dict:
        """findResiduumAtm: maximum residual and its (i, j, k), average, counts."""
        mx, avg = ctypes.c_double(), ctypes.c_double()
        pm, pn = (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
        nf, nn = ctypes.c_int64(), ctypes.c_int64()
        self._chk(L.lib().athad_residuum(self.h, member, ctypes.byref(mx), pm, ctypes.byref(avg),
                                         ctypes.byref(nf), ctypes.byref(nn), pn))
        return {"max": mx.value, "pos_max": tuple(pm), "avg_abs": avg.value, "n_finite": nf.value,
                "n_nonfinite": nn.value, "pos_first_nonfinite": tuple(pn)}

    def meridional_streamfunction(self, member: int = 0) -> dict:
        """write_meridional_streamfunction's arrays: vbar, rho_vbar, psi, each (im, jm)."""
        im, jm, _ = self.grids[member]
        out = {k: np.empty((im, jm)) for k in ("vbar", "rho_vbar", "psi")}
        ptr = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in out.values()]
        self._chk(L.lib().athad_meridional_streamfunction(self.h, member, *ptr))
        return out

    def get_slice(self, name: str, axis: int, index: int, member: int = 0) -> np.ndarray:
        """One plane of a field: axis 0 fixes i (jm, km), 1 fixes j (im, km), 2 fixes k (im, jm)."""
        im, jm, km = self.grids[member]
        shape = [(jm, km), (im, km), (im, jm)][axis] if axis in (0, 1, 2) else (1,)
        out = np.empty(shape)
        self._chk(L.lib().athad_get_slice(self.h, member, L.FIELD_ID[name], axis, index,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    # ---- measurement -------------------------------------------------------------------
    def profile(self, kernel: str, enable: bool = True):
        self._chk(L.lib().athad_profile(self.h, L.KERNEL_ID[kernel], int(enable)))

    def profile_result(self, kernel: str):
        ms, n = ctypes.c_double(), ctypes.c_int()
        self._chk(L.lib().athad_profile_result(self.h, L.KERNEL_ID[kernel], ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


def restart_bytes(grid: tuple) -> int:
    g = L.Grid(*grid)
    return int(L.lib().athad_restart_bytes(ctypes.byref(g)))


def read_restart(data: bytes):
    """Parses an ATM1 restart: (header dict, {array name: [im][jm][km] float64 array})."""
    hdr = np.frombuffer(data, dtype="<i4", count=5)
    if int(hdr[0]) != L.RESTART_MAGIC:
        raise ValueError("not an ATM1 restart")
    im, jm, km = (int(v) for v in hdr[1:4])
    n = im * jm * km
    arrs = np.frombuffer(data, dtype="<f8", offset=20, count=len(L.RESTART_ARRAYS) * n)
    return ({"im": im, "jm": jm, "km": km, "total_iter_count": int(hdr[4])},
            {name: arrs[q * n:(q + 1) * n].reshape(im, jm, km) for q, name in enumerate(L.RESTART_ARRAYS)})


def write_restart(header: dict, arrays: dict) -> bytes:
    """An ATM1 restart from a header dict and arrays by name (missing arrays are zeros)."""
    im, jm, km = header["im"], header["jm"], header["km"]
    out = [np.array([L.RESTART_MAGIC, im, jm, km, header["total_iter_count"]], dtype="<i4").tobytes()]
    for name in L.RESTART_ARRAYS:
        a = arrays.get(name)
        out.append(np.zeros((im, jm, km), "<f8").tobytes() if a is None else np.ascontiguousarray(a, "<f8").tobytes())
    return b"".join(out)


def footprint_bytes(grid: tuple, precision: str = "fp64") -> int:
    g = L.Grid(*grid)
    return int(L.lib().athad_footprint_bytes(ctypes.byref(g), {"fp64": L.FP64, "mixed": L.MIXED}[precision]))
