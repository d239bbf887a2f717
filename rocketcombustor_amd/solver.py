"""Host-side mirror of the reference's operator interface for the steady
path and the dual-time backward-Euler path, driving the HIP library
through its C-ABI.

The reference's operators are argument-less subroutines over module state;
here that state lives in a device-resident :class:`SteadyCase`, and the
methods keep the reference's names and call sequence:

    timestep()          TimeStep            src/TimeStep.f90:4
    gradient()          gradient            src/gradient.f90:4
    limiter_barth()     Limiter_Barth       src/limiter.f90:5
    reconstruction()    reconstruction      src/fvm2d.f90:183
    flux_sourceterm()   flux_sourceterm     src/fvm2d.f90:196
    lu_sgs_steady()     LU_SGS_steady       src/Time_LUSGS_steady.f90:32
    lu_sgs_stage()      lu_sgs_stage        src/Time_LUSGS_steady.f90:4
    run_steady(n)       steady main loop    src/fvm2d.f90:48-150
    residual()          Residual/step_residual  src/residual.f90:4-21,61-79

    set_transport(tables)         IfViscous + transport/diffusion decks  src/startup.f90:76
    set_outlet_pressure(p)        compute_p_out         src/fvm2d.f90:229-239
    backward_euler_lusgs(dt, k)   backwardEuler_LU_SGS  src/Time_BackwardEuler.f90:58
    backward_euler_stage(dt, ..)  backwardEuler_stage   src/Time_BackwardEuler.f90:4
    bdf2_begin_step(first)        BDF2_stage prologue   src/Time_BDF2.f90:12-20
    bdf2_lusgs(dt, k, phi)        BDF2_LU_SGS           src/Time_BDF2.f90:71
    bdf2_stage(dt, .., first)     BDF2_stage            src/Time_BDF2.f90:4
    run_unsteady(n)     unsteady main loop  src/fvm2d.f90:48-120 (implicit_time_method=3 or 5)

Errors where the reference stops raise :class:`native.RcError`.

:class:`SteadyBatch` steps many resident cases with the same launches (no
reference counterpart); each member computes exactly what it computes alone.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np

from . import native
from .native import NVAR, check, dptr, iptr, lib


def outlet_pressure(p_static: float, a_p: float, f_p: float, ttime: float) -> float:
    """compute_p_out (src/fvm2d.f90:229-239): p_static*(1 + A_p sin(2 pi f_p t))."""
    return p_static * (1. + a_p * math.sin(2 * math.pi * f_p * ttime))


class SteadyCase:
    """One nozzle/chamber case resident on a GPU (steady LU-SGS steps and
    dual-time backward-Euler steps act on the same resident state)."""

    def __init__(self, case, device: int = 0):
        L = lib()
        self.case = case
        m = case.mesh
        self._mech = native.make_mechanism(case.mech)
        self._par = native.make_params(case.info.cfl, case.info.limit_id, case.info.combustion, case.p_out,
                                       case.oxidizer, case.inlet_mass_flux, case.fuel, case.fuel_mass_flux)
        coor = np.ascontiguousarray(m.coor, dtype=np.float64)
        cells1 = np.ascontiguousarray(m.cells, dtype=np.int32) + 1
        bn1 = np.ascontiguousarray(m.bedges, dtype=np.int32) + 1
        bt = np.ascontiguousarray(m.btype, dtype=np.int32)
        h = C.c_void_p()
        periodic = getattr(m, "periodic", ())
        if periodic:
            con = np.ascontiguousarray([[p[0], p[1]] for p in periodic], dtype=np.int32)
            dxyz = np.ascontiguousarray([[p[2], p[3]] for p in periodic], dtype=np.float64)
            rc = L.rc_case_create_periodic(C.byref(self._mech), C.byref(self._par), len(coor), dptr(coor),
                                           len(cells1), iptr(cells1), len(bn1), iptr(bn1), iptr(bt), len(con),
                                           iptr(con), dptr(dxyz), int(device), C.byref(h))
        else:
            rc = L.rc_case_create(C.byref(self._mech), C.byref(self._par), len(coor), dptr(coor), len(cells1),
                                  iptr(cells1), len(bn1), iptr(bn1), iptr(bt), int(device), C.byref(h))
        check(rc, None)
        self.h = h
        info = (C.c_longlong * 6)()
        check(L.rc_case_info(h, info), h)
        self.ne, self.nf, self.ng, self.nvar, self.nlev, self.device_bytes = [int(v) for v in info]
        if getattr(case.info, "viscous", False):
            self.set_transport(case.transport)
        self.set_primitive(case.pa0)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().rc_case_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _call(self, name: str, *args) -> None:
        check(getattr(lib(), name)(self.h, *args), self.h)

    # -- state -------------------------------------------------------------
    def set_primitive(self, pa: np.ndarray) -> None:
        pa = np.ascontiguousarray(pa, dtype=np.float64)
        if pa.shape != (self.ne, NVAR):
            raise ValueError("PA must be (NE, Nvar)")
        self._call("rc_set_primitive", dptr(pa))

    def _get(self, name: str, shape) -> np.ndarray:
        out = np.empty(shape, dtype=np.float64)
        self._call(name, dptr(out))
        return out

    def primitive(self) -> np.ndarray:
        return self._get("rc_get_primitive", (self.ne, NVAR))

    def conserved(self) -> np.ndarray:
        return self._get("rc_get_conserved", (self.ne, NVAR))

    def increment(self) -> np.ndarray:
        """hWA: the residual after flux_sourceterm, dW after LU-SGS."""
        return self._get("rc_get_increment", (self.ne, NVAR))

    def gradient_array(self) -> np.ndarray:
        return self._get("rc_get_gradient", (self.ne, NVAR, 2))

    def ghost(self) -> np.ndarray:
        out = np.empty((self.ng, NVAR))
        if self.ng:
            self._call("rc_get_ghost", dptr(out))
        return out

    def level_cells(self) -> np.ndarray:
        """Cells in each LU-SGS wavefront level."""
        out = np.zeros(max(self.nlev, 1), dtype=np.int32)
        check(lib().rc_case_levels(self.h, iptr(out), len(out)), self.h)
        return out[:self.nlev]

    def timestep_arrays(self):
        step = np.empty(self.ne)
        alpaf = np.empty(self.nf)
        self._call("rc_get_timestep", dptr(step), dptr(alpaf))
        return step, alpaf

    # -- operators -----------------------------------------------------------
    def timestep(self) -> None:
        self._call("rc_timestep")

    def gradient(self) -> None:
        self._call("rc_gradient")

    def limiter_barth(self) -> None:
        self._call("rc_limiter_barth")

    def reconstruction(self) -> None:
        self._call("rc_reconstruction")

    def flux_sourceterm(self) -> None:
        self._call("rc_flux_sourceterm")

    def lu_sgs_steady(self) -> None:
        self._call("rc_lusgs_steady")

    def lu_sgs_stage(self) -> None:
        self._call("rc_lusgs_stage")

    def residual(self) -> np.ndarray:
        """Res(1:Nvar)/total_volume and max(rho), as step_residual reports."""
        return self._get("rc_residual", (NVAR + 1,))

    def run_steady(self, nsteps: int, every: int = 0, history: bool = True) -> Optional[np.ndarray]:
        """``nsteps`` steady LU-SGS steps; with ``history`` the step_residual
        norms at step 1 and every ``every`` steps, shape (nreports, Nvar+1)."""
        every = every if every > 0 else nsteps + 1
        nrep = 1 + nsteps // every if nsteps > 0 else 0
        buf = np.zeros((max(nrep, 1), NVAR + 1)) if history else None
        done = C.c_int(0)
        rc = lib().rc_run_steady(self.h, int(nsteps), int(every), dptr(buf) if history else None, C.byref(done))
        check(rc, self.h)
        if not history:
            return None
        nrep_done = (1 if nsteps >= 1 else 0) + sum(1 for s in range(2, nsteps + 1) if s % every == 0)
        return buf[:nrep_done]

    def set_transport(self, transport) -> None:
        """The viscous path (IfViscous) with these transport.in /
        diffusion.in tables (setup_transport_property,
        setup_diffusion_velocity)."""
        if transport is None:
            raise ValueError("the viscous path needs transport tables")
        self._tr = native.make_transport(transport)
        check(lib().rc_set_transport(self.h, C.byref(self._tr)), self.h)

    # -- learned chemistry (no reference counterpart) ----------------------
    def set_surrogate(self, mlp, transform=None, tol: float = 0.05, rtol: float = 0.0, trust=None) -> None:
        """The source term from the network ``mlp`` (surrogate.MLP) with the
        per-cell gate (training range, mass balance within ``tol``; with a
        ``trust`` region (surrogate.TrustRegion) only its calibrated grid
        cells; with ``rtol`` > 0 also within ``rtol`` of the exact FP64
        rates) and the exact FP64 rate as fallback; ``mlp=None`` restores the
        exact source."""
        if mlp is None:
            check(lib().rc_set_surrogate(self.h, None, None, 0.0, 0.0), self.h)
            self._sur = None
            return
        xf = np.ascontiguousarray(transform.packed(), dtype=np.float64)
        check(lib().rc_set_surrogate(self.h, mlp.h, dptr(xf), float(tol), float(rtol)), self.h)
        if trust is not None:
            dims = (C.c_int * 3)(*trust.dims)
            lo = np.ascontiguousarray(trust.lo, dtype=np.float64)
            hi = np.ascontiguousarray(trust.hi, dtype=np.float64)
            acc = np.ascontiguousarray(trust.accept, dtype=np.uint8)
            check(lib().rc_set_surrogate_trust(self.h, dims, dptr(lo), dptr(hi),
                                               acc.ctypes.data_as(C.POINTER(C.c_ubyte))), self.h)
        self._sur = (mlp, xf)

    def surrogate_stats(self):
        """(accepted, evaluated) cell source evaluations since set_surrogate."""
        out = (C.c_longlong * 2)()
        check(lib().rc_get_surrogate_stats(self.h, out), self.h)
        return int(out[0]), int(out[1])

    # -- dual-time backward Euler (implicit_time_method = 3) ----------------
    def set_outlet_pressure(self, p_out: float) -> None:
        self._call("rc_set_outlet_pressure", float(p_out))

    def backward_euler_lusgs(self, dt: float, inner: int) -> None:
        """backwardEuler_LU_SGS(inner) after timestep, reconstruction and
        flux_sourceterm; increment() is then the primitive increment."""
        self._call("rc_backward_euler_lusgs", float(dt), int(inner))

    def backward_euler_stage(self, dt: float, innerloop: int, reltol: float) -> np.ndarray:
        """backwardEuler_stage: the inner iterations of one physical step;
        returns their inner_iteration_residual rows, shape (ninner, Nvar+1)."""
        buf = np.zeros((max(int(innerloop), 1), NVAR + 1))
        n = C.c_int(0)
        rc = lib().rc_backward_euler_stage(self.h, float(dt), int(innerloop), float(reltol), C.byref(n), dptr(buf))
        if rc:
            try:
                check(rc, self.h)
            except native.RcError as e:
                e.inner = n.value
                raise
        return buf[:n.value]

    def bdf2_begin_step(self, first: bool) -> float:
        """The WP/WP2 prologue of BDF2_stage; returns phi (0 at the first
        step, 0.5 after)."""
        phi = C.c_double(0.0)
        self._call("rc_bdf2_begin_step", 1 if first else 0, C.byref(phi))
        return phi.value

    def bdf2_lusgs(self, dt: float, inner: int, phi: float) -> None:
        """BDF2_LU_SGS(inner, phi) after timestep, reconstruction and
        flux_sourceterm (bdf2_begin_step at the start of the step)."""
        self._call("rc_bdf2_lusgs", float(dt), int(inner), float(phi))

    def bdf2_stage(self, dt: float, innerloop: int, reltol: float, first: bool) -> np.ndarray:
        """BDF2_stage: the WP/WP2 prologue and the inner iterations of one
        physical step; returns their inner_iteration_residual rows."""
        buf = np.zeros((max(int(innerloop), 1), NVAR + 1))
        n = C.c_int(0)
        rc = lib().rc_bdf2_stage(self.h, float(dt), int(innerloop), float(reltol), 1 if first else 0, C.byref(n),
                                 dptr(buf))
        if rc:
            try:
                check(rc, self.h)
            except native.RcError as e:
                e.inner = n.value
                raise
        return buf[:n.value]

    def run_unsteady(self, nsteps: int, dt: Optional[float] = None, innerloop: Optional[int] = None,
                     reltol: Optional[float] = None, a_p: Optional[float] = None, f_p: Optional[float] = None,
                     ttime: float = 0.0, step_rows: Optional[list] = None) -> np.ndarray:
        """The unsteady branch of the main loop with backward Euler
        (src/fvm2d.f90:48-120): per step ttime += dt, compute_p_out, the
        stage, and compute_p_out at the step's end time.  Parameters default
        to the case's solver settings (dt = totaltime/nstep_max, InnerLoop,
        relTol, A_p, f_p); p_static is the case's exit pressure.  Returns
        rows (istep, inner, Res/total_volume, max rho) of every inner
        iteration; self.ttime holds the time reached.  With ``step_rows``
        the step_residual norms after each step (fvm2d.f90:137, of the
        residual the stage's closing flux_sourceterm leaves) are appended
        to it as (istep, Res/total_volume, max rho)."""
        info = self.case.info
        dt = info.totaltime / info.nstep_max if dt is None else dt
        innerloop = info.innerloop if innerloop is None else innerloop
        reltol = info.reltol if reltol is None else reltol
        a_p = info.a_p if a_p is None else a_p
        f_p = info.f_p if f_p is None else f_p
        p_static = self.case.p_out
        rows = []
        for istep in range(1, int(nsteps) + 1):
            tn = ttime
            ttime = tn + 1. * dt
            self.set_outlet_pressure(outlet_pressure(p_static, a_p, f_p, ttime))
            if info.implicit_time_method == 5:
                h = self.bdf2_stage(dt, innerloop, reltol, istep == 1)
            else:
                h = self.backward_euler_stage(dt, innerloop, reltol)
            rows += [np.r_[istep, k + 1, h[k]] for k in range(len(h))]
            ttime = tn + dt
            if step_rows is not None:
                step_rows.append(np.r_[istep, self.residual()])
        self.set_outlet_pressure(outlet_pressure(p_static, a_p, f_p, ttime))
        self.ttime = ttime
        return np.array(rows).reshape(-1, NVAR + 3)

    # -- output and restart in the reference's formats ---------------------
    def derived(self) -> dict:
        """rho, Ma and the production rates of the current state, as
        output_vtk writes them (computed on the device)."""
        rho, ma = np.empty(self.ne), np.empty(self.ne)
        prod = np.empty((self.ne, NVAR - 4))
        check(lib().rc_get_derived(self.h, dptr(rho), dptr(ma), dptr(prod)), self.h)
        return {"rho": rho, "Ma": ma, "production_rate": prod}

    def write_vtk(self, path: str) -> None:
        """The current field as output_vtk writes it (output.f90:5-227)."""
        from . import deck
        d = self.derived()
        deck.write_vtk(path, self.case.mesh.coor, self.case.mesh.cells, self.primitive(), d["rho"], d["Ma"],
                       d["production_rate"])

    def backup(self, path: str, ttime: float = 0.0) -> None:
        """``fvp1.sav`` of the current PA, as BackupData writes it
        (output.f90:256-273); the reference restarts from it."""
        from . import deck
        deck.write_backup(path, self.case.mesh.centroids(), self.primitive(), ttime)

    def restart(self, path: str) -> float:
        """PA from a backup file, then WA from PA, as Read_Backup and
        primitve_to_conserved_variables do; returns the stored time."""
        from . import deck
        ttime, _, pa = deck.read_backup(path, NVAR)
        if len(pa) != self.ne:
            raise ValueError("backup holds %d cells, the case has %d" % (len(pa), self.ne))
        self.set_primitive(pa)
        return ttime

    def set_timing(self, on: bool) -> None:
        self._call("rc_set_timing", 1 if on else 0)

    TIMING_KEYS = ("gradient_limiter", "flux_face", "residual_source", "lusgs_forward", "lusgs_top", "lusgs_backward",
                   "update")

    def timing(self) -> dict:
        """Mean device ms per step of each kernel since set_timing(True)."""
        t = self._get("rc_get_timing", (8,))
        n = max(t[7], 1.0)
        out = {k: t[i] / n for i, k in enumerate(self.TIMING_KEYS)}
        out["steps"] = int(t[7])
        return out


class SteadyBatch:
    """Resident :class:`SteadyCase` objects of one device stepped together:
    every kernel launch of a step covers all members.  Members can be added
    and removed between calls (continuous batching); a member's results are
    bitwise the same alone and in any batch."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().rc_batch_create(int(device), C.byref(h)), None)
        self.h = h
        self.members = []
        self.launch = (256, 64, -1)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().rc_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, case: SteadyCase) -> None:
        check(lib().rc_batch_add(self.h, case.h), None)
        self.members.append(case)

    def remove(self, case: SteadyCase) -> None:
        check(lib().rc_batch_remove(self.h, case.h), None)
        self.members.remove(case)

    def set_launch(self, cell_threads: int = 256, run_threads: int = 64, xcd: int = -1) -> None:
        """Launch shape of the step (rc_batch_set_launch): threads per block
        of the cell/face/ghost kernels and of the LU-SGS level kernels, and
        XCD banding of the work lists (-1: as the members were laid out).
        Results do not depend on it."""
        check(lib().rc_batch_set_launch(self.h, int(cell_threads), int(run_threads), int(xcd)), None)
        self.launch = (int(cell_threads), int(run_threads), int(xcd))

    def info(self) -> dict:
        v = (C.c_longlong * 3)()
        check(lib().rc_batch_info(self.h, v), None)
        return {"members": int(v[0]), "cells": int(v[1]), "levels": int(v[2])}

    def run_steady(self, nsteps: int, every: int = 0, history: bool = True) -> Optional[np.ndarray]:
        """``nsteps`` steps of every member; with ``history`` the step_residual
        norms at step 1 and every ``every`` steps, shape (nreports, members,
        Nvar+1).  A member's device error raises RcError naming the member."""
        n = len(self.members)
        every = every if every > 0 else nsteps + 1
        nrep = (1 if nsteps >= 1 else 0) + sum(1 for s in range(2, nsteps + 1) if s % every == 0)
        buf = np.zeros((max(nrep, 1), max(n, 1), NVAR + 1)) if history else None
        done, bad = C.c_int(0), C.c_int(-1)
        rc = lib().rc_batch_run_steady(self.h, int(nsteps), int(every), dptr(buf) if history else None,
                                       C.byref(done), C.byref(bad))
        if rc != 0:
            err = native.RcError(rc, lib().rc_last_error(None).decode())
            err.done, err.member = int(done.value), int(bad.value)
            raise err
        return buf[:nrep, :n] if history else None
