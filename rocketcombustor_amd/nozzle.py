"""Seeded synthetic nozzle cases: geometry, triangle mesh, BCs, initial field.

The reference ships no mesh generator of its own for this geometry (its
meshes come from gmsh, inv-nozzle-no-BL-rocket/rocket.py).  This module
builds the 2-D planar half nozzle the workloads are defined on:

* chamber x in [0, Lc], half-height hc; cosine contraction over [Lc, Lc+Lv]
  to the throat half-height ht; cosine divergence over [Lc+Lv, Lc+Lv+Ld] to
  the exit half-height he.  The cosine blend is the profile of
  inv-nozzle-no-BL-rocket/rocket.py:145-160 (zero slope at both ends).
* a mapped nx x ny grid in (xi = x, eta = y/h(x)); interior nodes jittered by
  +-20 % of the local spacing from a SplitMix64 stream (seed 1); every quad
  split along the diagonal that satisfies the Delaunay empty-circle test in
  (xi/dxi, eta/deta); mapped back with y = eta*h(x).  Triangles are
  counter-clockwise, as ``ReadMesh`` requires (src/mesh.f90:164-165 makes
  VECX owner-outward only for CCW cells).
* BC types of the reference (src/flux_lax.f90:123-235): inlet x=0 -> 3
  (mass-flux inlet), outlet -> 5, upper wall -> 6 (slip), axis y=0 -> 2.
* initial field: quasi-1-D isentropic area-Mach state from (p0, T0), with each
  cell's mass fractions scaled by (1 + 0.01 U[-1,1]) (seed 2) and renormalised.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

from . import deck

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, idx: np.ndarray) -> np.ndarray:
    """Output number ``idx`` (0-based) of the SplitMix64 generator started at
    ``seed``: state_k = seed + (k+1)*golden, then the standard finaliser."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, idx: np.ndarray) -> np.ndarray:
    """U[0,1) with 53 random bits from :func:`splitmix64`."""
    return (splitmix64(seed, idx) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclasses.dataclass
class NozzleGeometry:
    l_chamber: float = 0.10
    h_chamber: float = 0.025
    l_converge: float = 0.02
    h_throat: float = 0.010
    l_diverge: float = 0.04
    h_exit: float = 0.040

    @property
    def length(self) -> float:
        return self.l_chamber + self.l_converge + self.l_diverge

    def height(self, x: np.ndarray) -> np.ndarray:
        x = np.asarray(x, dtype=np.float64)
        h = np.full_like(x, self.h_chamber)
        x1 = self.l_chamber
        x2 = x1 + self.l_converge
        s = np.clip((x - x1) / self.l_converge, 0.0, 1.0)
        hc = self.h_chamber + (self.h_throat - self.h_chamber) * 0.5 * (1.0 - np.cos(np.pi * s))
        s2 = np.clip((x - x2) / self.l_diverge, 0.0, 1.0)
        hd = self.h_throat + (self.h_exit - self.h_throat) * 0.5 * (1.0 - np.cos(np.pi * s2))
        h = np.where(x <= x1, h, np.where(x <= x2, hc, hd))
        return h

    def dheight(self, x: np.ndarray) -> np.ndarray:
        x = np.asarray(x, dtype=np.float64)
        x1 = self.l_chamber
        x2 = x1 + self.l_converge
        s = np.clip((x - x1) / self.l_converge, 0.0, 1.0)
        dc = (self.h_throat - self.h_chamber) * 0.5 * np.pi * np.sin(np.pi * s) / self.l_converge
        s2 = np.clip((x - x2) / self.l_diverge, 0.0, 1.0)
        dd = (self.h_exit - self.h_throat) * 0.5 * np.pi * np.sin(np.pi * s2) / self.l_diverge
        return np.where(x <= x1, 0.0, np.where(x <= x2, dc, dd))


@dataclasses.dataclass
class Mesh:
    coor: np.ndarray      # (NNT, 2) float64
    cells: np.ndarray     # (NE, 3) int32, 0-based, counter-clockwise
    bedges: np.ndarray    # (NB, 2) int32, 0-based node pairs
    btype: np.ndarray     # (NB,) int32, reference BC type (>= 100: label of a periodic boundary)
    nface: int
    # periodic connections (mesh.f90:60-64, period.f90): (source label,
    # target label, dx, dy), the translation taking source to target faces
    periodic: tuple = ()

    @property
    def ne(self) -> int:
        return len(self.cells)

    def centroids(self) -> np.ndarray:
        return self.coor[self.cells].mean(axis=1)

    def permuted(self, perm: np.ndarray) -> "Mesh":
        """Cells reordered so that new cell k is old cell ``perm[k]``."""
        return dataclasses.replace(self, cells=np.ascontiguousarray(self.cells[perm]))


def _incircle(ax, ay, bx, by, cx, cy, dx, dy):
    adx, ady = ax - dx, ay - dy
    bdx, bdy = bx - dx, by - dy
    cdx, cdy = cx - dx, cy - dy
    return ((adx * adx + ady * ady) * (bdx * cdy - cdx * bdy)
            - (bdx * bdx + bdy * bdy) * (adx * cdy - cdx * ady)
            + (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady))


def nozzle_mesh(nx: int, ny: int, geom: Optional[NozzleGeometry] = None, seed: int = 1,
                jitter: float = 0.2, wall_type: int = 6, fuel_faces: int = 0) -> Mesh:
    """Triangulated nozzle with ``2*nx*ny`` cells in natural (x-major) order.
    ``wall_type`` 6 (slip) or 1 (viscous wall); the top ``fuel_faces`` inlet
    faces become a fuel inlet (type 4)."""
    geom = geom or NozzleGeometry()
    L = geom.length
    dxi, deta = L / nx, 1.0 / ny
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    nid = (ii * (ny + 1) + jj).ravel()
    xi = (ii * dxi).ravel().astype(np.float64)
    eta = (jj * deta).ravel().astype(np.float64)
    interior = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)).ravel()
    u1 = uniform01(seed, 2 * nid)
    u2 = uniform01(seed, 2 * nid + 1)
    xi = np.where(interior, xi + jitter * dxi * (2.0 * u1 - 1.0), xi)
    eta = np.where(interior, eta + jitter * deta * (2.0 * u2 - 1.0), eta)
    # the exact end values for boundary nodes
    xi = np.where(ii.ravel() == nx, L, xi)
    eta = np.where(jj.ravel() == ny, 1.0, eta)
    coor = np.column_stack([xi, eta * geom.height(xi)])

    # quads, x-major
    qi, qj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    qi, qj = qi.ravel(), qj.ravel()
    a = qi * (ny + 1) + qj
    b = (qi + 1) * (ny + 1) + qj
    c = b + 1
    d = a + 1
    sx, sy = xi / dxi, eta / deta
    inc = _incircle(sx[a], sy[a], sx[b], sy[b], sx[c], sy[c], sx[d], sy[d])
    use_bd = inc > 0.0
    t1 = np.where(use_bd[:, None], np.column_stack([a, b, d]), np.column_stack([a, b, c]))
    t2 = np.where(use_bd[:, None], np.column_stack([b, c, d]), np.column_stack([a, c, d]))
    cells = np.empty((2 * nx * ny, 3), dtype=np.int32)
    cells[0::2] = t1
    cells[1::2] = t2

    j = np.arange(ny)
    i = np.arange(nx)
    inlet = np.column_stack([j, j + 1])
    outlet = np.column_stack([nx * (ny + 1) + j, nx * (ny + 1) + j + 1])
    axis = np.column_stack([i * (ny + 1), (i + 1) * (ny + 1)])
    wall = np.column_stack([i * (ny + 1) + ny, (i + 1) * (ny + 1) + ny])
    bedges = np.vstack([inlet, outlet, axis, wall]).astype(np.int32)
    in_type = np.full(ny, 3)
    if fuel_faces:
        in_type[ny - fuel_faces:] = 4
    btype = np.concatenate([in_type, np.full(ny, 5), np.full(nx, 2), np.full(nx, wall_type)]).astype(np.int32)
    nface = nx * (ny + 1) + (nx + 1) * ny + nx * ny
    return Mesh(coor, cells, bedges, btype, nface)


# ----------------------------------------------------------------------------
# gas state
# ----------------------------------------------------------------------------

def mixture_r_gamma(mech: deck.Mechanism, y: np.ndarray, t: float):
    """R and gamma of a mixture at temperature t from the cp/R polynomials."""
    basis = np.array([t ** (j - 3) for j in range(1, 8)])
    coef = mech.low_coef if mech.t_low <= t <= mech.t_crit else mech.high_coef
    cp = float(np.sum((coef @ basis) * mech.r_species * y))
    r = float(np.sum(mech.r_species * y))
    return r, cp / (cp - r)


def area_mach(ratio: np.ndarray, gamma: float, supersonic: np.ndarray) -> np.ndarray:
    """Mach number for A/A* = ratio on the chosen branch (bisection)."""
    ratio = np.maximum(np.asarray(ratio, dtype=np.float64), 1.0)
    g = gamma
    e = (g + 1.0) / (2.0 * (g - 1.0))

    def f(m):
        return (1.0 / m) * ((2.0 / (g + 1.0)) * (1.0 + 0.5 * (g - 1.0) * m * m)) ** e - ratio

    lo = np.where(supersonic, 1.0, 1e-6)
    hi = np.where(supersonic, 50.0, 1.0)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        fm = f(mid)
        # subsonic branch: f decreasing in m; supersonic: increasing
        go_right = np.where(supersonic, fm < 0.0, fm > 0.0)
        lo = np.where(go_right, mid, lo)
        hi = np.where(go_right, hi, mid)
    return 0.5 * (lo + hi)


@dataclasses.dataclass
class Case:
    """One nozzle case: mesh, mechanism, inlet/outlet data, initial PA."""
    mesh: Mesh
    mech: deck.Mechanism
    oxidizer: deck.Inlet
    fuel: deck.Inlet
    info: deck.SolverInfo
    pa0: np.ndarray                 # (NE, Nvar)
    geom: NozzleGeometry
    p0: float
    t0: float
    transport: Optional[deck.Transport] = None   # used when info.viscous

    @property
    def inlet_mass_flux(self) -> float:
        """G of the type-3 inlet as ``initialization`` sets it:
        mass_flow / (2*(ymax-ymin)) over the inlet faces
        (src/initialization.f90:22, :194)."""
        ins = self.mesh.bedges[self.mesh.btype == 3]
        if len(ins) == 0:
            return 0.0
        y = self.mesh.coor[ins.ravel(), 1]
        return self.oxidizer.mass_flow / (2.0 * (y.max() - y.min()))

    @property
    def fuel_mass_flux(self) -> float:
        """G of type-4 faces: 0.5*mass_flow/(ymax-ymin) (src/initialization.f90:23,195)."""
        ins = self.mesh.bedges[self.mesh.btype == 4]
        if len(ins) == 0:
            return 0.0
        y = self.mesh.coor[ins.ravel(), 1]
        return 0.5 * self.fuel.mass_flow / (y.max() - y.min())

    @property
    def p_out(self) -> float:
        # p_static= p_out_ratio*oxidizer_pre_in  (src/initialization.f90:32)
        return self.info.p_out_ratio * self.oxidizer.pres

    def permuted(self, perm: np.ndarray) -> "Case":
        return dataclasses.replace(self, mesh=self.mesh.permuted(perm), pa0=np.ascontiguousarray(self.pa0[perm]))


def nozzle_case(nx: int, ny: int, p0: float = 5.0e6, t0: float = 3000.0, phi: float = 1.0,
                geom: Optional[NozzleGeometry] = None, mech: Optional[deck.Mechanism] = None,
                cfl: float = 5.0, mesh_seed: int = 1, y_seed: int = 2, y_perturb: float = 0.01,
                limit_id: int = 1, combustion: bool = True, p_out_ratio: float = 0.02,
                wall_type: int = 6, fuel_faces: int = 0, subsonic_exit: bool = False,
                viscous: bool = False) -> Case:
    """A nozzle case.  ``subsonic_exit`` starts from the subsonic branch of
    the area-Mach relation in the divergent part too (the outlet then takes
    the fixed-pressure branch); the default mechanism is the reference's with
    kinetics model 1 and the formation constant in K.  ``viscous`` switches
    on the viscous path (IfViscous) with the shipped transport tables."""
    geom = geom or NozzleGeometry()
    mech = mech or deck.default_mechanism(model=1, b1_in_kelvin=True)
    mesh = nozzle_mesh(nx, ny, geom, seed=mesh_seed, wall_type=wall_type, fuel_faces=fuel_faces)
    nsp = mech.nsp
    # H2/O2 at equivalence ratio phi: mass ratio H2:O2 = phi*4:32
    y_in = np.zeros(nsp)
    y_in[0] = phi * 4.0 / (phi * 4.0 + 32.0)
    y_in[1] = 1.0 - y_in[0]
    r_mix, gam = mixture_r_gamma(mech, y_in, t0)
    area_c = geom.h_chamber / geom.h_throat
    m_c = float(area_mach(np.array([area_c]), gam, np.array([False]))[0])
    t_in = t0 / (1.0 + 0.5 * (gam - 1.0) * m_c * m_c)
    p_in = p0 * (t_in / t0) ** (gam / (gam - 1.0))
    rho_in = p_in / (r_mix * t_in)
    g_in = rho_in * m_c * math.sqrt(gam * r_mix * t_in)
    oxid = deck.Inlet(y=y_in, temp=t_in, pres=p0, mass_flow=g_in * 2.0 * geom.h_chamber)
    # fuel stream (type-4 faces only): pure H2 at 300 K with a tenth of the oxidizer mass flux
    fuel = deck.Inlet(y=np.eye(nsp)[0], temp=300.0, pres=p0, mass_flow=0.1 * g_in * 2.0 * geom.h_chamber)

    # initial field
    cxy = mesh.centroids()
    x, y = cxy[:, 0], cxy[:, 1]
    h = geom.height(x)
    throat_x = geom.l_chamber + geom.l_converge
    mach = area_mach(h / geom.h_throat, gam, (x >= throat_x) & (not subsonic_exit))
    t = t0 / (1.0 + 0.5 * (gam - 1.0) * mach * mach)
    p = p0 * (t / t0) ** (gam / (gam - 1.0))
    c = np.sqrt(gam * r_mix * t)
    u = mach * c
    v = u * (y / h) * geom.dheight(x)
    ne = mesh.ne
    k = np.arange(ne, dtype=np.int64)
    yy = np.empty((ne, nsp))
    for s in range(nsp):
        yy[:, s] = y_in[s] * (1.0 + y_perturb * (2.0 * uniform01(y_seed, nsp * k + s) - 1.0))
    yy /= yy.sum(axis=1, keepdims=True)
    pa0 = np.column_stack([p, u, v, t, yy[:, :nsp - 1]])
    info = deck.SolverInfo(cfl=cfl, limit_id=limit_id, combustion=combustion, p_out_ratio=p_out_ratio,
                           viscous=viscous)
    return Case(mesh=mesh, mech=mech, oxidizer=oxid, fuel=fuel, info=info, pa0=pa0,
                geom=geom, p0=p0, t0=t0, transport=deck.default_transport() if viscous else None)


def periodic_channel_case(nx: int, ny: int, direction: str = "x", **kw) -> Case:
    """A straight channel (the nozzle mesh at constant height) whose
    inlet and outlet (``direction`` "x") or axis and wall ("y") are joined
    as periodic boundary pairs: labels 101 -> 102 and 102 -> 101 with the
    channel's length or height as translation."""
    g = NozzleGeometry(h_throat=NozzleGeometry.h_chamber, h_exit=NozzleGeometry.h_chamber)
    case = nozzle_case(nx, ny, geom=g, **kw)
    m = case.mesh
    bt = m.btype.copy()
    if direction == "x":
        src, dst = bt == 3, bt == 5
        d = (g.length, 0.0)
    elif direction == "y":
        src, dst = bt == 2, np.isin(bt, (1, 6))
        d = (0.0, g.h_chamber)
    else:
        raise ValueError("direction is 'x' or 'y'")
    bt[src] = 101
    bt[dst] = 102
    mesh = dataclasses.replace(m, btype=bt, periodic=((101, 102, d[0], d[1]), (102, 101, -d[0], -d[1])))
    return dataclasses.replace(case, mesh=mesh)


def write_reference_case(case_dir: str, case: Case, nstep_max: Optional[int] = None,
                         extra_decks: Optional[str] = None) -> None:
    """Writes the full input deck the reference reads for a one-rank steady
    LU-SGS run restarted from ``fvp1.sav`` (``IfReadBackup``)."""
    os.makedirs(case_dir, exist_ok=True)
    deck.write_mechanism(case_dir, case.mech)
    deck.write_inlet(os.path.join(case_dir, "combustion", "oxidizer.in"), case.oxidizer, "oxidizer")
    deck.write_inlet(os.path.join(case_dir, "combustion", "fuel.in"), case.fuel, "fuel")
    info = case.info if nstep_max is None else dataclasses.replace(case.info, nstep_max=nstep_max)
    deck.write_solver_info(case_dir, info)
    deck.write_grid(os.path.join(case_dir, "grid1.in"), case.mesh.coor, case.mesh.cells,
                    case.mesh.bedges, case.mesh.btype, case.mesh.nface, periodic=case.mesh.periodic)
    deck.write_backup(os.path.join(case_dir, "fvp1.sav"), case.mesh.centroids(), case.pa0)
    if case.transport is not None:
        deck.write_transport(case_dir, case.transport)
    elif extra_decks:
        for name in ("transport.in", "diffusion.in"):
            src = os.path.join(extra_decks, name)
            with open(src) as f, open(os.path.join(case_dir, "combustion", name), "w") as g:
                g.write(f.read())
    else:
        deck.write_transport_placeholders(case_dir, case.mech.nsp)


@dataclasses.dataclass
class CaseSpec:
    """Parameters of one sampled nozzle/chamber case."""
    nx: int
    ny: int
    p0: float
    t0: float
    phi: float
    throat_ratio: float

    @property
    def cells(self) -> int:
        return 2 * self.nx * self.ny

    def build(self, **kw) -> Case:
        g = NozzleGeometry(h_throat=self.throat_ratio * NozzleGeometry.h_chamber)
        return nozzle_case(self.nx, self.ny, p0=self.p0, t0=self.t0, phi=self.phi, geom=g, **kw)


def sample_cases(n: int, seed: int = 3, cells_lo: float = 5e3, cells_hi: float = 8e4, first: int = 0) -> list:
    """Cases ``first .. first+n-1`` of the seeded family of independent
    nozzle/chamber cases: cell count log-uniform in [cells_lo, cells_hi]
    (mapped grid with nx = 5 ny), chamber pressure U[1, 10] MPa, T0
    U[2500, 3500] K, equivalence ratio U[0.7, 1.3], throat-to-chamber
    height ratio U[0.3, 0.5]; five SplitMix64 draws per case."""
    out = []
    for k in range(first, first + n):
        u = uniform01(seed, 5 * k + np.arange(5))
        cells = math.exp(math.log(cells_lo) + u[0] * (math.log(cells_hi) - math.log(cells_lo)))
        ny = max(2, int(round(math.sqrt(cells / 10.0))))
        out.append(CaseSpec(nx=5 * ny, ny=ny, p0=float(1e6 + 9e6 * u[1]), t0=float(2500.0 + 1000.0 * u[2]),
                            phi=float(0.7 + 0.6 * u[3]), throat_ratio=float(0.3 + 0.2 * u[4])))
    return out


def case_from_dir(case_dir: str) -> Case:
    """A case from a reference case directory (grid1.in, fvp1.sav,
    combustion/*.in, flow/flow_solver_info.in), as the reference's start-up
    reads it (startup.f90, mesh.f90:10-65, output.f90:276-338)."""
    info = deck.read_solver_info(case_dir)
    mech = deck.read_mechanism(case_dir)
    oxid = deck.read_inlet(os.path.join(case_dir, "combustion", "oxidizer.in"), mech.nsp)
    fuel = deck.read_inlet(os.path.join(case_dir, "combustion", "fuel.in"), mech.nsp)
    coor, cells, bedges, btype, nface, periodic = deck.read_grid(os.path.join(case_dir, "grid1.in"), info.unit,
                                                                 with_periodic=True)
    mesh = Mesh(coor=coor, cells=cells, bedges=bedges, btype=btype, nface=nface, periodic=periodic)
    _, _, pa = deck.read_backup(os.path.join(case_dir, "fvp1.sav"), 3 + mech.nsp)
    return Case(mesh=mesh, mech=mech, oxidizer=oxid, fuel=fuel, info=info, pa0=pa, geom=NozzleGeometry(),
                p0=oxid.pres, t0=oxid.temp)
