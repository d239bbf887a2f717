"""Input decks of the reference solver: mechanism, inlet and solver settings.

The reference reads positional, list-directed text files from its working
directory (one comment line before each value group):

* ``combustion/species.in``            -- ``setup_species_info``     (src/combustion_modeling/species.f90:14-51)
* ``combustion/thermo.in``             -- ``setup_thermo_property``  (src/combustion_modeling/thermo_property.f90:18-59)
* ``combustion/chemical_kinetics.in``  -- ``setup_chemical_kinetics`` (src/combustion_modeling/chemical_kinetics.f90:16-57)
* ``combustion/oxidizer.in``           -- ``setup_oxidizer_info``    (src/combustion_modeling/oxidizer.f90:13-41)
* ``combustion/fuel.in``               -- ``setup_fuel_info``        (src/combustion_modeling/fuel.f90:13-41)
* ``flow/flow_solver_info.in``         -- ``startup``                (src/startup.f90:55-128)

This module reads and writes those layouts so that the same deck drives the
reference and this library.  The values held by :func:`default_mechanism`
are the reference's shipped H2/O2/H2O/Ar mechanism (species masses, the
7-term cp/R polynomials and the one-step reaction), i.e. input data.
"""
from __future__ import annotations

import dataclasses
import os
import re
from typing import List, Optional

import numpy as np

NFLOW = 4  # src/module.f90:8


@dataclasses.dataclass
class Mechanism:
    """Species, thermodynamic and kinetic tables (module ``species``,
    ``thermo_property`` and ``chemical_kinetics`` of the reference)."""
    names: List[str]
    mole_weight: np.ndarray          # (Nsp,) kg/mol
    rconstant: float                 # J/mol/K
    tref: float
    t_low: float
    t_crit: float
    t_high: float
    low_coef: np.ndarray             # (Nsp, 7)  cp/R polynomial, basis T^(j-3)
    low_b1: np.ndarray               # (Nsp,)
    high_coef: np.ndarray            # (Nsp, 7)
    high_b1: np.ndarray              # (Nsp,)
    hof: np.ndarray                  # (Nsp,) J/mol (unused on the path)
    model: int                       # code_chemical_model: 0 Arrhenius, 1 threshold
    coef_f: np.ndarray               # (Nsp,) forward stoichiometric coefficients of reaction 1
    coef_r: np.ndarray               # (Nsp,) reverse stoichiometric coefficients of reaction 1
    a_f: float
    beta: float
    e_act: float
    dentropy: float = 0.0
    denthalpy: float = 0.0
    p_a: float = 1.0e5

    @property
    def nsp(self) -> int:
        return len(self.names)

    @property
    def nvar(self) -> int:
        return NFLOW + self.nsp - 1

    @property
    def r_species(self) -> np.ndarray:
        # R_species(i)= Rconstant/mole_weight(i)  (species.f90:42-44)
        return self.rconstant / self.mole_weight


@dataclasses.dataclass
class Inlet:
    """Mass-flux inlet (``oxidizer.in`` / ``fuel.in``).  ``mass_flow`` is the
    deck value; the reference divides it by the inlet height measure set in
    ``compute_fuel_oxidizer_diameter`` (src/initialization.f90:22-23,194-195)."""
    y: np.ndarray                    # (Nsp,)
    temp: float
    pres: float
    mass_flow: float


@dataclasses.dataclass
class Transport:
    """``transport.in`` and ``diffusion.in`` (modules ``transport_property``
    and ``diffusion_velocity``): per-species viscosity and conductivity fits
    mu, K = exp(A ln T + B/T + C/T^2 + D) (micropoise, microW/cm/K) on two
    temperature ranges, and Lennard-Jones collision data for the binary
    diffusion coefficients."""
    t_low: float
    t_crit: float
    t_high: float
    mu_low: np.ndarray               # (Nsp, 4) A, B, C, D
    mu_high: np.ndarray
    k_low: np.ndarray
    k_high: np.ndarray
    k_b: float                       # read, unused by the reference's formulas
    sigma: np.ndarray                # (Nsp,) collision diameter, Angstrom
    eps: np.ndarray                  # (Nsp,) Lennard-Jones energy / k_B, K


def default_transport() -> Transport:
    """The reference's shipped ``combustion/transport.in`` and
    ``combustion/diffusion.in`` tables (four species, in the deck's order)."""
    rows = np.array([
        [0.57388074E0, -0.98544160E2, 0.20012204E4, 0.17536015E1],
        [0.65074534E0, 0.23936771E2, -0.22020183E5, 0.11244058E1],
        [0.11770360E1, -0.17422121E3, 0.22865563E5, -0.55146852E0],
        [0.49214767E0, -0.91598343E3, 0.87265127E5, 0.48489412E1],
        [0.61936357E0, -0.44608607E2, -0.13460714E4, 0.19597562E1],
        [0.63839563E0, -0.12344438E1, -0.22885810E5, 0.18056937E1],
        [0.81595343E0, -0.34366856E2, 0.22785080E4, 0.10050999E1],
        [0.80805788E0, 0.11982181E3, -0.47335931E5, 0.95189193E0],
        [0.54330318E0, -0.18823898E3, 0.88726567E4, 0.24499362E1],
        [0.65318879E0, 0.51738759E2, -0.62834882E5, 0.15227045E1],
        [0.53726173E0, -0.49928331E3, 0.37397504E5, 0.32903619E1],
        [0.66068182E0, -0.12741845E3, -0.81580328E5, 0.21817907E1],
        [0.78387780E0, -0.38260408E3, 0.49040158E5, 0.85222785E0],
        [0.50714993E0, -0.68966913E3, 0.87454750E5, 0.30285155E1],
        [0.15541443E1, 0.66106305E2, 0.55969886E4, -0.39259598E1],
        [0.79349503E0, -0.13340063E4, 0.37864327E6, 0.23591474E1]]).reshape(4, 4, 4)
    return Transport(t_low=200.0, t_crit=1000.0, t_high=50000.0, mu_low=rows[:, 0].copy(),
                     mu_high=rows[:, 1].copy(), k_low=rows[:, 2].copy(), k_high=rows[:, 3].copy(),
                     k_b=1.38064852E-23, sigma=np.array([3.758, 3.467, 3.941, 2.641]),
                     eps=np.array([148.6, 106.7, 195.2, 809.1]))


def read_transport(case_dir: str, nsp: int) -> Transport:
    """The layouts ``setup_transport_property`` (transport_property.f90:18-45)
    and ``setup_diffusion_velocity`` (diffusion_velocity.f90:15-38) read."""
    comb = os.path.join(case_dir, "combustion")
    r = _Reader(os.path.join(comb, "transport.in"))
    r.skip(); tl, tc, th = [_f(t) for t in r.values()[:3]]
    r.skip()
    tab = np.zeros((nsp, 4, 4))
    for i in range(nsp):
        for k in range(4):
            tab[i, k] = [_f(t) for t in r.values()[:4]]
    r = _Reader(os.path.join(comb, "diffusion.in"))
    r.skip(); kb = _f(r.values()[0])
    r.skip()
    sig, eps = [], []
    for _ in range(nsp):
        v = r.values(); sig.append(_f(v[0])); eps.append(_f(v[1]))
    return Transport(t_low=tl, t_crit=tc, t_high=th, mu_low=tab[:, 0].copy(), mu_high=tab[:, 1].copy(),
                     k_low=tab[:, 2].copy(), k_high=tab[:, 3].copy(), k_b=kb, sigma=np.array(sig),
                     eps=np.array(eps))


def write_transport(case_dir: str, tr: Transport) -> None:
    comb = os.path.join(case_dir, "combustion")
    os.makedirs(comb, exist_ok=True)
    with open(os.path.join(comb, "transport.in"), "w") as f:
        f.write("# transport_lowest_tem,transport_critical_tem,transport_highest_tem\n")
        f.write("%.17g, %.17g, %.17g\n" % (tr.t_low, tr.t_crit, tr.t_high))
        f.write("# low_mu_A,low_mu_B,low_mu_C,low_mu_D\n")
        for i in range(len(tr.sigma)):
            for tab in (tr.mu_low, tr.mu_high, tr.k_low, tr.k_high):
                f.write(", ".join("%.17g" % x for x in tab[i]) + "\n")
    with open(os.path.join(comb, "diffusion.in"), "w") as f:
        f.write("# Boltzmann's constant k_B (m2*kg*s-2*K-1)\n%.17g\n" % tr.k_b)
        f.write("# hard-sphere collision diameter sigma (A), Lennard-Jones energy epsilon/k_B (K)\n")
        for sg, ep in zip(tr.sigma, tr.eps):
            f.write("%.17g, %.17g\n" % (sg, ep))


@dataclasses.dataclass
class SolverInfo:
    """The fields of ``flow/flow_solver_info.in`` (src/startup.f90:60-120)."""
    combustion: bool = True
    if_right_domain: bool = False
    n_previous_par: int = 1
    npar: int = 1
    unit: float = 1.0
    steady: bool = True
    viscous: bool = False
    read_backup: bool = True
    explicit: bool = False
    implicit_time_method: int = 5
    cfl: float = 5.0
    totaltime: float = 1.0e30
    nstep_max: int = 2000
    save_start_time: float = 0.0
    save_end_time: float = 0.0
    fix_dt: bool = False
    inloop: int = 1
    innerloop: int = 1
    reltol: float = 1.0e-3
    n_screen: int = 1
    n_file_output: int = 100000000
    n_file_backup: int = 100000000
    n_write_monitor: int = 100000000
    rc_detector: float = 3.0
    n_wbap: int = 10
    limit_id: int = 1
    p_out_ratio: float = 0.02
    a_p: float = 0.0
    f_p: float = 0.0


def default_mechanism(model: int = 1, b1_in_kelvin: bool = False) -> Mechanism:
    """The reference's shipped mechanism (src/combustion/species.in,
    thermo.in, chemical_kinetics.in).

    ``b1_in_kelvin`` divides the tabulated H2O formation constant (given in
    J/mol in the shipped file, but multiplied by R_k as if it were K,
    thermo_property.f90:131) by R; both codes read whatever the file holds.
    """
    names = ["H2", "O2", "H2O", "Ar"]
    mw = np.array([2.0e-3, 32.0e-3, 18.0e-3, 40.0e-3])
    rconst = 8.31446261815324
    coef = np.zeros((4, 7))
    coef[:, 2] = 3.6084
    b1 = np.array([0.0, 0.0, -241826.0, 0.0])
    if b1_in_kelvin:
        b1 = b1 / rconst
    return Mechanism(
        names=names, mole_weight=mw, rconstant=rconst, tref=298.15,
        t_low=200.0, t_crit=1000.0, t_high=60000.0,
        low_coef=coef.copy(), low_b1=b1.copy(), high_coef=coef.copy(), high_b1=b1.copy(),
        hof=np.array([0.0, 0.0, -241826.0, 0.0]),
        model=model,
        coef_f=np.array([2.0, 1.0, 0.0, 0.0]), coef_r=np.array([0.0, 0.0, 2.0, 0.0]),
        a_f=2.12e11, beta=0.0, e_act=2.025e5, dentropy=0.0, denthalpy=0.0, p_a=1.0e5)


# ----------------------------------------------------------------------------
# list-directed text helpers
# ----------------------------------------------------------------------------

class _Reader:
    """Reads value records of a list-directed file, skipping one comment line
    where the reference does ``read(iofile,*)``."""

    def __init__(self, path: str):
        with open(path) as f:
            self.lines = f.read().splitlines()
        self.pos = 0

    def skip(self):
        self.pos += 1

    def values(self) -> List[str]:
        line = self.lines[self.pos]
        self.pos += 1
        toks = [t for t in re.split(r"[,\s]+", line.strip()) if t]
        return toks


def _f(tok: str) -> float:
    return float(tok.lower().replace("d", "e"))


def _b(tok: str) -> bool:
    t = tok.strip().lower().strip(".")
    return t in ("true", "t")


def read_mechanism(case_dir: str, model: Optional[int] = None) -> Mechanism:
    comb = os.path.join(case_dir, "combustion")
    r = _Reader(os.path.join(comb, "species.in"))
    r.skip(); nsp = int(r.values()[0]); r.skip()
    names, mw = [], []
    for _ in range(nsp):
        v = r.values(); names.append(v[0]); mw.append(_f(v[1]))
    r.skip(); rconst = _f(r.values()[0])

    r = _Reader(os.path.join(comb, "thermo.in"))
    r.skip(); tref = _f(r.values()[0])
    r.skip(); order = int(r.values()[0])
    if order != 7:
        raise ValueError("only 7-term cp polynomials are supported")
    r.skip(); tl, tc, th = [_f(t) for t in r.values()[:3]]
    low, lb1, high, hb1 = [], [], [], []
    for _ in range(nsp):
        low.append([_f(t) for t in r.values()[:7]]); lb1.append(_f(r.values()[0]))
        high.append([_f(t) for t in r.values()[:7]]); hb1.append(_f(r.values()[0]))
    r.skip(); hof = [_f(r.values()[0]) for _ in range(nsp)]

    r = _Reader(os.path.join(comb, "chemical_kinetics.in"))
    r.skip(); v = r.values(); nreact, code = int(v[0]), int(v[1])
    if nreact != 1:
        raise ValueError("the reference's production rate handles one reaction")
    r.skip()
    cf, cr = [], []
    for _ in range(nsp):
        v = r.values(); cf.append(_f(v[0])); cr.append(_f(v[1]))
    r.skip(); a, beta, e = [_f(t) for t in r.values()[:3]]
    r.skip(); ds, dh = [_f(t) for t in r.values()[:2]]
    r.skip(); p_a = _f(r.values()[0])
    return Mechanism(names=names, mole_weight=np.array(mw), rconstant=rconst, tref=tref,
                     t_low=tl, t_crit=tc, t_high=th,
                     low_coef=np.array(low), low_b1=np.array(lb1),
                     high_coef=np.array(high), high_b1=np.array(hb1), hof=np.array(hof),
                     model=code if model is None else model,
                     coef_f=np.array(cf), coef_r=np.array(cr), a_f=a, beta=beta, e_act=e,
                     dentropy=ds, denthalpy=dh, p_a=p_a)


def read_inlet(path: str, nsp: int) -> Inlet:
    r = _Reader(path)
    r.skip(); y = [_f(r.values()[0]) for _ in range(nsp)]
    r.skip(); v = r.values(); t, p = _f(v[0]), _f(v[1])
    r.skip(); m = _f(r.values()[0])
    return Inlet(y=np.array(y), temp=t, pres=p, mass_flow=m)


def write_mechanism(case_dir: str, mech: Mechanism) -> None:
    comb = os.path.join(case_dir, "combustion")
    os.makedirs(comb, exist_ok=True)
    with open(os.path.join(comb, "species.in"), "w") as f:
        f.write("# number of species\n%d\n# species name, mole weight (kg/mol)\n" % mech.nsp)
        for n, m in zip(mech.names, mech.mole_weight):
            f.write("%s , %.17g\n" % (n, m))
        f.write("# gas constant R (J/mol/K)\n%.17g\n" % mech.rconstant)
    with open(os.path.join(comb, "thermo.in"), "w") as f:
        f.write("# reference temperature\n%.17g\n# order of Cp polynomial\n7\n" % mech.tref)
        f.write("# Cp polynomial coefficients\n%.17g,%.17g,%.17g\n" % (mech.t_low, mech.t_crit, mech.t_high))
        for k in range(mech.nsp):
            f.write(", ".join("%.17g" % c for c in mech.low_coef[k]) + "\n")
            f.write("%.17g\n" % mech.low_b1[k])
            f.write(", ".join("%.17g" % c for c in mech.high_coef[k]) + "\n")
            f.write("%.17g\n" % mech.high_b1[k])
        f.write("# heat of formation (J/mol)\n")
        for h in mech.hof:
            f.write("%.17g\n" % h)
    with open(os.path.join(comb, "chemical_kinetics.in"), "w") as f:
        f.write("# number of reactions, chemical model\n1, %d\n" % mech.model)
        f.write("# forward reaction coefficient, reverse reaction coefficient\n")
        for a, b in zip(mech.coef_f, mech.coef_r):
            f.write("%.17g, %.17g\n" % (a, b))
        f.write("# pre-exponential factor, temperature exponent, activation energy (J)\n")
        f.write("%.17g, %.17g, %.17g\n" % (mech.a_f, mech.beta, mech.e_act))
        f.write("# entropy change, enthalpy change\n%.17g, %.17g\n" % (mech.dentropy, mech.denthalpy))
        f.write("# p_a\n%.17g\n" % mech.p_a)


def write_inlet(path: str, inlet: Inlet, what: str) -> None:
    with open(path, "w") as f:
        f.write("# %s mass fraction\n" % what)
        for y in inlet.y:
            f.write("%.17g\n" % y)
        f.write("# temperature (K), pressure (Pa)\n%.17g, %.17g\n" % (inlet.temp, inlet.pres))
        f.write("# mass flow rate (kg/s)\n%.17g\n" % inlet.mass_flow)


def write_transport_placeholders(case_dir: str, nsp: int) -> None:
    """``transport.in`` and ``diffusion.in`` in the layouts that
    ``setup_transport_property`` (transport_property.f90:31-42) and
    ``setup_diffusion_velocity`` (diffusion_velocity.f90:30-35) read.  The
    reference reads both at start-up even when IfViscous is false; on the
    inviscid path their values are never used, so neutral numbers are written."""
    comb = os.path.join(case_dir, "combustion")
    os.makedirs(comb, exist_ok=True)
    with open(os.path.join(comb, "transport.in"), "w") as f:
        f.write("# transport_lowest_tem,transport_critical_tem,transport_highest_tem\n200.0, 1000.0, 60000.0\n")
        f.write("# mu/K polynomial coefficients A,B,C,D (unused without viscous fluxes)\n")
        for _ in range(4 * nsp):
            f.write("0.5, 0.0, 0.0, 1.0\n")
    with open(os.path.join(comb, "diffusion.in"), "w") as f:
        f.write("# Boltzmann's constant k_B\n1.38064852E-23\n")
        f.write("# collision diameter (A), LJ energy (K) (unused without viscous fluxes)\n")
        for _ in range(nsp):
            f.write("3.0, 100.0\n")


def write_solver_info(case_dir: str, s: SolverInfo) -> None:
    d = os.path.join(case_dir, "flow")
    os.makedirs(d, exist_ok=True)
    L = lambda b: ".true." if b else ".false."
    rows = [
        ("# If_turn_on_combustion", L(s.combustion)),
        ("# IfRightDomain", L(s.if_right_domain)),
        ("# NPreviousPar", str(s.n_previous_par)),
        ("# Npar, number of processors", str(s.npar)),
        ("# unit", "%.17g" % s.unit),
        ("# IfSteady, IfViscous, IfReadBackup, IfExplicit",
         ", ".join(L(b) for b in (s.steady, s.viscous, s.read_backup, s.explicit))),
        ("# implicit_time_method", str(s.implicit_time_method)),
        ("# cfl,totaltime,nstep_max,save_start_time, save_end_time",
         "%.17g, %.17g, %d, %.17g, %.17g" % (s.cfl, s.totaltime, s.nstep_max, s.save_start_time, s.save_end_time)),
        ("# IfFixDt", L(s.fix_dt)),
        ("# InLoop, InnerLoop, relTol", "%d, %d, %.17g" % (s.inloop, s.innerloop, s.reltol)),
        ("# NScreenOutput,NFileOutput,NFileBackup,NWriteMonitor",
         "%d, %d, %d, %d" % (s.n_screen, s.n_file_output, s.n_file_backup, s.n_write_monitor)),
        ("# rc_detector", "%.17g" % s.rc_detector),
        ("# n_wbap,limit_ID", "%d, %d" % (s.n_wbap, s.limit_id)),
        ("# number of monitor points", "0"),
        ("# x_monitor_point, y_monitor_point", None),
        ("# exit pressure ratio, A_p, f_p", "%.17g, %.17g, %.17g" % (s.p_out_ratio, s.a_p, s.f_p)),
    ]
    with open(os.path.join(d, "flow_solver_info.in"), "w") as f:
        for c, v in rows:
            f.write(c + "\n")
            if v is not None:
                f.write(v + "\n")


def write_grid(path: str, coor: np.ndarray, cells: np.ndarray, bedges: np.ndarray,
               btype: np.ndarray, nface: int, periodic=()) -> None:
    """``grid1.in`` for one rank (src/mesh.f90:10-65).  ``cells`` are 0-based
    counter-clockwise triangles; ``bedges`` 0-based node pairs with their BC
    type in ``btype`` (one boundary group per type; labels >= 100 are
    periodic boundaries, written with type 100).  ``periodic`` holds the
    connections (source label, target label, dx, dy), written as NCon lines
    of 1-based boundary indices and the translation (mesh.f90:60-64)."""
    ne, nn = len(cells), len(coor)
    types = sorted(set(int(t) for t in btype))
    groups = [(t, bedges[btype == t]) for t in types]
    imax = max(len(g) for _, g in groups) if groups else 0
    index = {t: k + 1 for k, t in enumerate(types)}
    with open(path, "w") as f:
        f.write("%d %d %d %d %d %d\n" % (ne, ne, nn, nn, nface, nface))
        np.savetxt(f, np.column_stack([np.full(ne, 3), cells + 1]), fmt="%d")
        np.savetxt(f, np.column_stack([coor, np.zeros((nn, 2))]), fmt=["%.17g", "%.17g", "%d", "%d"])
        f.write("0\n")
        f.write("%d %d\n" % (len(groups), imax))
        for t, g in groups:
            f.write("%d %d\n" % (min(t, 100), len(g)))
            np.savetxt(f, g + 1, fmt="%d")
        f.write("%d\n" % len(periodic))
        for src, dst, dx, dy in periodic:
            f.write("%d %d %.17g %.17g\n" % (index[int(src)], index[int(dst)], dx, dy))


def write_backup(path: str, cellxy: np.ndarray, pa: np.ndarray, ttime: float = 0.0) -> None:
    """``fvp1.sav`` (src/output.f90:256-273 writer, :316-327 reader): ``ttime``,
    ``NE``, then ``cx cy PA(1:Nvar)`` per cell."""
    with open(path, "w") as f:
        f.write("%.17g\n%d\n" % (ttime, len(pa)))
        np.savetxt(f, np.column_stack([cellxy, pa]), fmt="%.17g")


def read_backup(path: str, nvar: int):
    with open(path) as f:
        ttime = _f(f.readline().split()[0])
        ne = int(f.readline().split()[0])
        data = np.loadtxt(f, ndmin=2)
    data = data.reshape(ne, 2 + nvar)
    return ttime, data[:, :2], data[:, 2:]


def read_grid(path: str, unit: float = 1.0, with_periodic: bool = False):
    """``grid<N>.in`` as mesh.f90:10-65 reads it (free-format tokens):
    returns (coor, cells, bedges, btype, nface) with 0-based node numbers,
    and with ``with_periodic`` the periodic connections too: the faces of
    a boundary that a connection names carry the label 100 + its 1-based
    index, and each connection is (source label, target label, dx, dy) with
    dxyz as read (mesh.f90:60-64 does not scale it by ``unit``).  Multi-rank
    connectivity lists are not supported."""
    with open(path) as f:
        tok = f.read().split()
    pos = 0

    def take(n=1, conv=int):
        nonlocal pos
        out = [conv(_f(t)) if conv is float else int(t) for t in tok[pos:pos + n]]
        pos += n
        return out

    ne, ne_vir, nnt, nnt_vir, nf, _ = take(6)
    cells = np.empty((ne_vir, 3), dtype=np.int32)
    for i in range(ne_vir):
        k = take()[0]
        nodes = take(k)
        if k != 3:
            raise ValueError("cell %d has %d nodes: only triangles are supported" % (i + 1, k))
        cells[i] = nodes
    xy = np.array(take(4 * nnt_vir, float), dtype=np.float64).reshape(nnt_vir, 4)
    ncone = take()[0]
    if ncone:
        raise ValueError("partition connectivity lists (multi-rank grids) are not supported")
    nbound, _ = take(2)
    bedges, btype = [], []
    for _ in range(nbound):
        t, cnt = take(2)
        pairs = np.array(take(2 * cnt), dtype=np.int32).reshape(cnt, 2)
        bedges.append(pairs)
        btype.append(np.full(cnt, t, dtype=np.int32))
    ncon = take()[0] if pos < len(tok) else 0
    if ncon and not with_periodic:
        raise ValueError("periodic face pairs: read this grid with with_periodic=True")
    periodic = []
    for _ in range(ncon):
        gs, gt = take(2)
        dx, dy = take(2, float)
        if not (1 <= gs <= nbound and 1 <= gt <= nbound):
            raise ValueError("periodic connection names boundary %d or %d of %d" % (gs, gt, nbound))
        periodic.append((100 + gs, 100 + gt, dx, dy))
    for lab in {p[0] for p in periodic} | {p[1] for p in periodic}:
        btype[lab - 101] = np.full(len(btype[lab - 101]), lab, dtype=np.int32)
    be = np.concatenate(bedges) - 1 if bedges else np.zeros((0, 2), dtype=np.int32)
    bt = np.concatenate(btype) if btype else np.zeros(0, dtype=np.int32)
    out = (xy[:, :2] / unit, cells - 1, be.astype(np.int32), bt, nf)
    return out + (tuple(periodic),) if with_periodic else out


def read_solver_info(case_dir: str) -> SolverInfo:
    """``flow/flow_solver_info.in`` (startup.f90:55-128): the value lines in
    the order write_solver_info writes them (comment lines start with #)."""
    with open(os.path.join(case_dir, "flow", "flow_solver_info.in")) as f:
        vals = [ln.replace(",", " ").split() for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
    it = iter(vals)
    s = SolverInfo()
    s.combustion = _b(next(it)[0])
    s.if_right_domain = _b(next(it)[0])
    s.n_previous_par = int(next(it)[0])
    s.npar = int(next(it)[0])
    s.unit = _f(next(it)[0])
    v = next(it)
    s.steady, s.viscous, s.read_backup, s.explicit = (_b(x) for x in v[:4])
    s.implicit_time_method = int(next(it)[0])
    v = next(it)
    s.cfl, s.totaltime = _f(v[0]), _f(v[1])
    s.nstep_max = int(v[2])
    s.save_start_time, s.save_end_time = _f(v[3]), _f(v[4])
    s.fix_dt = _b(next(it)[0])
    v = next(it)
    s.inloop, s.innerloop, s.reltol = int(v[0]), int(v[1]), _f(v[2])
    v = next(it)
    s.n_screen, s.n_file_output, s.n_file_backup, s.n_write_monitor = (int(x) for x in v[:4])
    s.rc_detector = _f(next(it)[0])
    v = next(it)
    s.n_wbap, s.limit_id = int(v[0]), int(v[1])
    nmon = int(next(it)[0])
    for _ in range(nmon):
        next(it)
    v = next(it)
    s.p_out_ratio, s.a_p, s.f_p = _f(v[0]), _f(v[1]), _f(v[2])
    return s


def write_converge(path: str, rows, append: bool = False) -> None:
    """``converge.dat`` as step_residual writes it (residual.f90:83-93): a
    header line, then ``istep cputime Res(1:Nvar) Romax`` per report.
    ``rows`` holds (istep, cputime, res[Nvar+1]) with res[Nvar] = Romax."""
    with open(path, "a" if append else "w") as f:
        if not append:
            f.write(' VARIABLES = "Time step","CPU time (Sec)","L1(Res)","Romax"\n')
        for istep, cpu, res in rows:
            f.write(" %d %.17g %s\n" % (istep, cpu, " ".join("%.17g" % x for x in res)))


def read_converge(path: str) -> np.ndarray:
    """The rows of a ``converge.dat`` (reference's or ours): (n, 2+Nvar+1)."""
    with open(path) as f:
        lines = [ln for ln in f if ln.strip() and "VARIABLES" not in ln]
    return np.array([[_f(t) for t in ln.split()] for ln in lines], dtype=np.float64)


def write_vtk(path: str, coor: np.ndarray, cells: np.ndarray, pa: np.ndarray, rho: np.ndarray, mach: np.ndarray,
              prod: np.ndarray) -> None:
    """Legacy-VTK cell data in the layout of output_vtk (output.f90:59-223):
    points, triangles, then rho, u, v, p, T, Ma, Y1..Ynsp and
    Production_rate_1..nsp-1; values below 1e-15 in magnitude written as 0
    for Y and the rates as there (:45-53)."""
    ne, nn = len(cells), len(coor)
    nsp = pa.shape[1] - 3
    y = np.column_stack([pa[:, 4:], 1.0 - pa[:, 4:].sum(axis=1)])
    y = np.where(np.abs(y) < 1e-15, 0.0, y)
    prod = np.where(np.abs(prod[:, :1]) < 1e-15, 0.0, prod)
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 1.0\nUnstructured Grid Format\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write("POINTS %d double\n" % nn)
        np.savetxt(f, np.column_stack([coor, np.zeros(nn)]), fmt="%.8E")
        f.write("CELLS %d %d\n" % (ne, 4 * ne))
        np.savetxt(f, np.column_stack([np.full(ne, 3), cells]), fmt="%d")
        f.write("CELL_TYPES %d\n" % ne)
        np.savetxt(f, np.full(ne, 5), fmt="%d")
        f.write("CELL_DATA %d\n" % ne)
        fields = [("rho", rho, "%.8E"), ("u", pa[:, 1], "%.8E"), ("v", pa[:, 2], "%.8E"), ("p", pa[:, 0], "%.8E"),
                  ("T", pa[:, 3], "%.8E"), ("Ma", mach, "%.8E")]
        fields += [("Y%d" % (k + 1), y[:, k], "%.12E") for k in range(nsp)]
        fields += [("Production_rate_%d" % (k + 1), prod[:, k], "%.12E") for k in range(nsp - 1)]
        for name, vals, fmt in fields:
            f.write("SCALARS %s double\nLOOKUP_TABLE default\n" % name)
            np.savetxt(f, vals, fmt=fmt)


def read_vtk_cell_data(path: str) -> dict:
    """The CELL_DATA scalars of a legacy-VTK file written by write_vtk or
    output_vtk."""
    out, name, vals = {}, None, []
    in_data = False
    with open(path) as f:
        for ln in f:
            t = ln.split()
            if not t:
                continue
            if t[0] == "CELL_DATA":
                in_data = True
            elif in_data and t[0] == "SCALARS":
                if name:
                    out[name] = np.array(vals)
                name, vals = t[1], []
            elif in_data and t[0] == "LOOKUP_TABLE":
                continue
            elif in_data and name:
                vals.append(_f(t[0]))
    if name:
        out[name] = np.array(vals)
    return out
