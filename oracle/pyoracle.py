"""Test infrastructure only: ctypes binding of the C oracle (librc_oracle.so).

Imported by tests/, ``__graft_entry__.smoke`` and bench.py's CPU-baseline leg,
never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
MAXSP = 8


class RcoMech(C.Structure):
    _fields_ = [("nsp", C.c_int), ("mw", C.c_double * MAXSP), ("rconstant", C.c_double),
                ("t_low", C.c_double), ("t_crit", C.c_double), ("t_high", C.c_double),
                ("low", (C.c_double * 7) * MAXSP), ("high", (C.c_double * 7) * MAXSP),
                ("low_b1", C.c_double * MAXSP), ("high_b1", C.c_double * MAXSP),
                ("model", C.c_int), ("coef_f", C.c_double * MAXSP), ("coef_r", C.c_double * MAXSP),
                ("a_f", C.c_double), ("e_act", C.c_double)]


class RcoParams(C.Structure):
    _fields_ = [("cfl", C.c_double), ("limit_id", C.c_int), ("combustion", C.c_int),
                ("p_out", C.c_double),
                ("ox_y", C.c_double * MAXSP), ("ox_t", C.c_double), ("ox_g", C.c_double),
                ("fu_y", C.c_double * MAXSP), ("fu_t", C.c_double), ("fu_g", C.c_double)]


class RcoTransport(C.Structure):
    _fields_ = [("t_low", C.c_double), ("t_crit", C.c_double), ("t_high", C.c_double),
                ("mu_low", (C.c_double * 4) * MAXSP), ("mu_high", (C.c_double * 4) * MAXSP),
                ("k_low", (C.c_double * 4) * MAXSP), ("k_high", (C.c_double * 4) * MAXSP),
                ("k_b", C.c_double), ("sigma", C.c_double * MAXSP), ("eps", C.c_double * MAXSP)]


def make_transport(tr) -> RcoTransport:
    t = RcoTransport()
    t.t_low, t.t_crit, t.t_high, t.k_b = tr.t_low, tr.t_crit, tr.t_high, tr.k_b
    for i in range(len(tr.sigma)):
        t.sigma[i], t.eps[i] = tr.sigma[i], tr.eps[i]
        for k in range(4):
            t.mu_low[i][k], t.mu_high[i][k] = tr.mu_low[i, k], tr.mu_high[i, k]
            t.k_low[i][k], t.k_high[i][k] = tr.k_low[i, k], tr.k_high[i, k]
    return t


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "librc_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        P = C.POINTER
        L.rco_create.argtypes = [P(RcoMech), P(RcoParams), C.c_int, P(C.c_double), C.c_int, P(C.c_int),
                                 C.c_int, P(C.c_int), P(C.c_int), P(C.c_void_p)]
        L.rco_create_periodic.argtypes = [P(RcoMech), P(RcoParams), C.c_int, P(C.c_double), C.c_int, P(C.c_int),
                                          C.c_int, P(C.c_int), P(C.c_int), C.c_int, P(C.c_int), P(C.c_double),
                                          P(C.c_void_p)]
        L.rco_destroy.argtypes = [C.c_void_p]
        L.rco_last_error.argtypes = [C.c_void_p]
        L.rco_last_error.restype = C.c_char_p
        L.rco_sizes.argtypes = [C.c_void_p, P(C.c_int)]
        L.rco_darray.argtypes = [C.c_void_p, C.c_char_p]
        L.rco_darray.restype = P(C.c_double)
        L.rco_iarray.argtypes = [C.c_void_p, C.c_char_p]
        L.rco_iarray.restype = P(C.c_int)
        for n in ("rco_timestep", "rco_gradient", "rco_limiter", "rco_flux_sourceterm",
                  "rco_lusgs_steady", "rco_stage"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.rco_set_pa.argtypes = [C.c_void_p, P(C.c_double)]
        L.rco_residual.argtypes = [C.c_void_p, P(C.c_double)]
        L.rco_gas_param.argtypes = [P(RcoMech), P(C.c_double), C.c_double, P(C.c_double)]
        L.rco_enthalpy.argtypes = [P(RcoMech), C.c_double, P(C.c_double)]
        L.rco_newton.argtypes = [P(RcoMech), C.c_double, P(C.c_double), C.c_double, P(C.c_double)]
        L.rco_production_rate.argtypes = [P(RcoMech), C.c_double, C.c_double, P(C.c_double), P(C.c_double)]
        L.rco_con_to_pri_jaco.argtypes = [P(RcoMech), P(C.c_double), P(C.c_double)]
        L.rco_dwdp.argtypes = [P(RcoMech), P(C.c_double), P(C.c_double)]
        L.rco_solve_dgesvx.argtypes = [C.c_int, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.rco_be_lusgs.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.rco_bdf2_lusgs.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double]
        L.rco_bdf2_begin_step.argtypes = [C.c_void_p, C.c_int]
        L.rco_bdf2_begin_step.restype = C.c_double
        L.rco_be_stage.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, P(C.c_int), P(C.c_double)]
        L.rco_set_p_out.argtypes = [C.c_void_p, C.c_double]
        L.rco_set_transport.argtypes = [C.c_void_p, P(RcoTransport), C.c_int]
        L.rco_transport_property.argtypes = [P(RcoMech), P(RcoTransport), C.c_double, P(C.c_double),
                                             P(C.c_double), P(C.c_double)]
        L.rco_diffusion_velocity.argtypes = [P(RcoMech), P(RcoTransport), C.c_double, C.c_double, P(C.c_double),
                                             P(C.c_double)]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def make_mech(mech) -> RcoMech:
    m = RcoMech()
    m.nsp = mech.nsp
    for k in range(mech.nsp):
        m.mw[k] = mech.mole_weight[k]
        m.low_b1[k] = mech.low_b1[k]
        m.high_b1[k] = mech.high_b1[k]
        m.coef_f[k] = mech.coef_f[k]
        m.coef_r[k] = mech.coef_r[k]
        for j in range(7):
            m.low[k][j] = mech.low_coef[k, j]
            m.high[k][j] = mech.high_coef[k, j]
    m.rconstant = mech.rconstant
    m.t_low, m.t_crit, m.t_high = mech.t_low, mech.t_crit, mech.t_high
    m.model = mech.model
    m.a_f, m.e_act = mech.a_f, mech.e_act
    return m


def make_params(case) -> RcoParams:
    p = RcoParams()
    p.cfl = case.info.cfl
    p.limit_id = case.info.limit_id
    p.combustion = 1 if case.info.combustion else 0
    p.p_out = case.p_out
    for k in range(case.mech.nsp):
        p.ox_y[k] = case.oxidizer.y[k]
        p.fu_y[k] = case.fuel.y[k]
    p.ox_t = case.oxidizer.temp
    p.ox_g = case.inlet_mass_flux
    p.fu_t = case.fuel.temp
    p.fu_g = case.fuel_mass_flux
    return p


class OracleCase:
    """One case held by the C oracle."""

    def __init__(self, case):
        L = lib()
        self.case = case
        self._mech = make_mech(case.mech)
        self._par = make_params(case)
        m = case.mesh
        coor = np.ascontiguousarray(m.coor, dtype=np.float64)
        cells = np.ascontiguousarray(m.cells, dtype=np.int32)
        be = np.ascontiguousarray(m.bedges, dtype=np.int32)
        bt = np.ascontiguousarray(m.btype, dtype=np.int32)
        h = C.c_void_p()
        periodic = getattr(m, "periodic", ())
        con = np.ascontiguousarray([[p[0], p[1]] for p in periodic], dtype=np.int32).reshape(-1, 2)
        dxyz = np.ascontiguousarray([[p[2], p[3]] for p in periodic], dtype=np.float64).reshape(-1, 2)
        rc = L.rco_create_periodic(C.byref(self._mech), C.byref(self._par), len(coor), _dp(coor), len(cells),
                                   _ip(cells), len(be), _ip(be), _ip(bt), len(con), _ip(con), _dp(dxyz), C.byref(h))
        self.h = h
        if rc:
            raise RuntimeError("oracle set-up failed: " + L.rco_last_error(h).decode())
        s = (C.c_int * 4)()
        L.rco_sizes(h, s)
        self.ne, self.nf, self.ng, self.nvar = s[0], s[1], s[2], s[3]
        if getattr(case.info, "viscous", False):
            if case.transport is None:
                raise ValueError("viscous case without transport tables")
            self._tr = make_transport(case.transport)
            L.rco_set_transport(h, C.byref(self._tr), 1)
        self.set_pa(case.pa0)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.rco_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc:
            raise RuntimeError(_LIB.rco_last_error(self.h).decode())

    _DSHAPE = {"PA": "nv", "WA": "nv", "hWA": "nv", "gradC": "g", "GPA": "gv", "alpaF": "f", "step": "e",
               "dd": "e", "fixed": "nv", "VECX": "f2", "vol": "e", "cellxy": "e2", "reflen": "e2", "wdis": "e3",
               "AMLS": "a", "vol_Ghostcell": "gg", "CellXY_Ghostcell": "gg2", "production_rate": "ps"}
    _ISHAPE = {"Neighbor": "f2", "FNode": "f2", "NeighF": "e3", "NeighC": "e4", "Neigh": "e",
               "Num_ghostcell": "e", "Ghostcell": "e3", "Hostcell": "gg", "BCtype_Ghostcell": "gg"}

    def _shape(self, code):
        ne, nf, ng, nv = self.ne, self.nf, self.ng, self.nvar
        return {"nv": (ne, nv), "g": (ne, nv, 2), "gv": (ng, nv), "f": (nf,), "e": (ne,), "f2": (nf, 2),
                "e2": (ne, 2), "e3": (ne, 3), "e4": (ne, 4), "a": (ne, 3, 2), "gg": (ng,), "gg2": (ng, 2),
                "ps": (ne, self.case.mech.nsp)}[code]

    def get(self, name: str) -> np.ndarray:
        """A copy of a reference array, cell-major (row i is cell i+1)."""
        if name in self._DSHAPE:
            shp = self._shape(self._DSHAPE[name])
            p = _LIB.rco_darray(self.h, name.encode())
            return np.ctypeslib.as_array(p, shape=(int(np.prod(shp)),)).reshape(shp).copy()
        if name in self._ISHAPE:
            shp = self._shape(self._ISHAPE[name])
            p = _LIB.rco_iarray(self.h, name.encode())
            return np.ctypeslib.as_array(p, shape=(int(np.prod(shp)),)).reshape(shp).copy()
        if name == "total_volume":
            return float(_LIB.rco_darray(self.h, b"total_volume")[0])
        raise KeyError(name)

    def set_pa(self, pa):
        pa = np.ascontiguousarray(pa, dtype=np.float64)
        self._check(_LIB.rco_set_pa(self.h, _dp(pa)))

    def timestep(self):
        self._check(_LIB.rco_timestep(self.h))

    def gradient(self):
        self._check(_LIB.rco_gradient(self.h))

    def limiter(self):
        self._check(_LIB.rco_limiter(self.h))

    def reconstruction(self):
        self.gradient()
        if self.case.info.limit_id > 0:
            self.limiter()

    def flux_sourceterm(self):
        self._check(_LIB.rco_flux_sourceterm(self.h))

    def lusgs_steady(self):
        self._check(_LIB.rco_lusgs_steady(self.h))

    def stage(self):
        self._check(_LIB.rco_stage(self.h))

    def residual(self) -> np.ndarray:
        out = np.zeros(self.nvar + 1)
        _LIB.rco_residual(self.h, _dp(out))
        return out

    def set_transport(self, transport, viscous: bool = True):
        """Switches the viscous path on (IfViscous) with these tables."""
        self._tr = make_transport(transport)
        _LIB.rco_set_transport(self.h, C.byref(self._tr), 1 if viscous else 0)

    # dual-time backward Euler (implicit_time_method = 3)
    def set_p_out(self, p_out: float):
        _LIB.rco_set_p_out(self.h, float(p_out))

    def be_lusgs(self, dt: float, inner: int):
        self._check(_LIB.rco_be_lusgs(self.h, float(dt), int(inner)))

    def bdf2_begin_step(self, first: bool) -> float:
        """The WP/WP2 prologue of BDF2_stage; returns phi."""
        return float(_LIB.rco_bdf2_begin_step(self.h, 1 if first else 0))

    def bdf2_lusgs(self, dt: float, inner: int, phi: float):
        self._check(_LIB.rco_bdf2_lusgs(self.h, float(dt), int(inner), float(phi)))

    def be_stage(self, dt: float, innerloop: int, reltol: float) -> np.ndarray:
        """One physical step's inner iterations; returns their RoRes."""
        n = C.c_int(0)
        ro = np.zeros(max(int(innerloop), 1))
        self._check(_LIB.rco_be_stage(self.h, float(dt), int(innerloop), float(reltol), C.byref(n), _dp(ro)))
        return ro[:n.value]


def gas_param(mech, y, t):
    m = make_mech(mech)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(3)
    rc = lib().rco_gas_param(C.byref(m), _dp(y), float(t), _dp(out))
    if rc:
        raise ValueError("temperature out of range")
    return out


def enthalpy(mech, t):
    m = make_mech(mech)
    out = np.zeros(mech.nsp)
    rc = lib().rco_enthalpy(C.byref(m), float(t), _dp(out))
    if rc:
        raise ValueError("temperature out of range")
    return out


def newton(mech, et, y, t0):
    m = make_mech(mech)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(4)
    rc = lib().rco_newton(C.byref(m), float(et), _dp(y), float(t0), _dp(out))
    if rc:
        raise ValueError("temperature out of range")
    return out


def production_rate(mech, t, rho, y):
    m = make_mech(mech)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(mech.nsp)
    lib().rco_production_rate(C.byref(m), float(t), float(rho), _dp(y), _dp(out))
    return out


def con_to_pri_jaco(mech, pri):
    """dU/dPA at one primitive state (nvar x nvar)."""
    m = make_mech(mech)
    pri = np.ascontiguousarray(pri, dtype=np.float64)
    out = np.zeros((len(pri), len(pri)))
    if lib().rco_con_to_pri_jaco(C.byref(m), _dp(pri), _dp(out)):
        raise ValueError("temperature out of range")
    return out


def dwdp(mech, pri):
    """d(omega_i)/dPA for the first nsp-1 species ((nsp-1) x nvar)."""
    m = make_mech(mech)
    pri = np.ascontiguousarray(pri, dtype=np.float64)
    out = np.zeros((mech.nsp - 1, len(pri)))
    if lib().rco_dwdp(C.byref(m), _dp(pri), _dp(out)):
        raise ValueError("temperature out of range")
    return out


def solve_dgesvx(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    if lib().rco_solve_dgesvx(len(b), _dp(a), _dp(b), _dp(x)):
        raise np.linalg.LinAlgError("singular matrix")
    return x


def transport_property(mech, tr, t, y):
    """(mu, K) of the mixture: compute_transport_property."""
    m, T = make_mech(mech), make_transport(tr)
    y = np.ascontiguousarray(y, dtype=np.float64)
    mu, k = C.c_double(), C.c_double()
    if lib().rco_transport_property(C.byref(m), C.byref(T), float(t), _dp(y), C.byref(mu), C.byref(k)):
        raise ValueError("temperature out of the transport range")
    return mu.value, k.value


def diffusion_velocity(mech, tr, p, t, y):
    """Mixture-averaged diffusion coefficients: compute_diffusion_velocity."""
    m, T = make_mech(mech), make_transport(tr)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(mech.nsp)
    lib().rco_diffusion_velocity(C.byref(m), C.byref(T), float(p), float(t), _dp(y), _dp(out))
    return out
