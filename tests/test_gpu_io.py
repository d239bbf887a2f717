"""Output and restart through the device path: the derived fields of the
VTK output against the oracle's thermo, and a backup/restart in the
reference's fvp1.sav format against the oracle doing the same."""
import sys

import numpy as np
import pytest

from helpers import relerr
from rocketcombustor_amd import deck, native, nozzle
from rocketcombustor_amd.solver import SteadyCase

oracle = pytest.importorskip("oracle.pyoracle")

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    sys.__stderr__.write("[gpu-test] " + msg + "\n")
    sys.__stderr__.flush()


def test_derived_fields_and_vtk(tmp_path, xcd_layout):
    case = nozzle.nozzle_case(60, 12)
    announce("derived fields NE=%d" % case.mesh.ne)
    with SteadyCase(case) as g:
        g.run_steady(20, history=False)
        pa = g.primitive()
        d = g.derived()
        g.write_vtk(str(tmp_path / "out.vtk"))
    mech = case.mech
    rho, ma, prod = [], [], []
    for row in pa:
        y = np.append(row[4:], 1.0 - row[4:].sum())
        cp, gam, r = oracle.gas_param(mech, y, row[3])
        rho.append(row[0] / (r * row[3]))
        ma.append(np.sqrt(row[1] ** 2 + row[2] ** 2) / np.sqrt(gam * r * row[3]))
        prod.append(oracle.production_rate(mech, row[3], row[0] / (row[3] * r), y)[:3])
    assert relerr(d["rho"], np.array(rho)) < 1e-14
    assert relerr(d["Ma"], np.array(ma)) < 1e-13
    assert relerr(d["production_rate"], np.array(prod), -1) < 1e-13
    v = deck.read_vtk_cell_data(str(tmp_path / "out.vtk"))
    assert np.allclose(v["rho"], d["rho"], rtol=1e-8) and np.allclose(v["T"], pa[:, 3], rtol=1e-8)


def test_backup_restart_vs_oracle(tmp_path, xcd_layout):
    case = nozzle.nozzle_case(30, 8)
    perm, _ = native.order_multicolor(case.mesh.cells)
    case = case.permuted(perm)
    path = str(tmp_path / "fvp1.sav")
    announce("backup/restart NE=%d" % case.mesh.ne)
    with SteadyCase(case) as g:
        g.run_steady(40, history=False)
        g.backup(path)
        pa40 = g.primitive()
    _, _, pa_file = deck.read_backup(path, 7)
    assert np.array_equal(pa_file, pa40)            # the text format round-trips exactly
    with SteadyCase(case) as g2:
        g2.restart(path)
        g2.run_steady(20, history=False)
        pa_gpu = g2.primitive()
    o = oracle.OracleCase(case)
    for _ in range(40):
        o.stage()
    o2 = oracle.OracleCase(case)
    o2.set_pa(pa_file)
    for _ in range(20):
        o2.stage()
    assert relerr(pa_gpu, o2.get("PA"), -1) < 1e-10


def test_boundary_type_outside_flux_lax_is_rejected():
    """Boundary types other than 1..6 are refused at set-up: the reference's
    flux_lax stops on any such face every step (flux_lax.f90:232-235), before
    flux_viscous could reach its type 7 ("induced") branch."""
    case = nozzle.nozzle_case(6, 2)
    case.mesh.btype = np.array(case.mesh.btype, copy=True)
    case.mesh.btype[0] = 7
    with pytest.raises(native.RcError, match="flux_lax handles types 1-6") as ei:
        SteadyCase(case)
    assert ei.value.code == native.RC_ERR_MESH
