"""The drop-in: the reference's own Fortran host (start-up, mesh, restart,
residual monitor, backup) with its steady step -- or, for unsteady decks,
its backward-Euler or BDF2 stage -- replaced by the MI355X library through
include/rocketcombustor_iface.f90 (oracle/ref_host_gpu.f90, linked from the
reference's objects by oracle/ref_build.sh).  Its residual history and final
field must match the reference's own run (golden vectors).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from helpers import case_from_fixture, load, relerr
from rocketcombustor_amd import nozzle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "rc_ref_host_gpu")


@pytest.mark.parametrize("name", ["n24x6_natural", "n24x6_bc_variants", "n24x6_be_bc_variants", "n24x6_be_model1",
                                  "n24x6_viscous_bc_variants", "n24x6_be_viscous", "n24x6_be_bdf2_bc_variants",
                                  "n24x6_be_bdf2_viscous"])
def test_reference_host_with_gpu_step(name):
    if not os.path.exists(BIN):
        pytest.skip("reference host binary not built (oracle/ref_build.sh needs the reference sources)")
    d = load(name)
    case = case_from_fixture(d)
    nsteps = int(d["nsteps"])
    sys.__stderr__.write("[gpu-test] fortran host %s steps=%d\n" % (name, nsteps))
    sys.__stderr__.flush()
    with tempfile.TemporaryDirectory() as tmp:
        nozzle.write_reference_case(tmp, case, nstep_max=nsteps)
        os.makedirs(os.path.join(tmp, "out"))
        r = subprocess.run([BIN, "out"], cwd=tmp, capture_output=True, text=True, timeout=120)
        assert "rc_ref_host_gpu finished" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
        unsteady = not case.info.steady
        res = np.fromfile(os.path.join(tmp, "out", "resid.bin")).reshape(-1, 10 if unsteady else 8)
        pa = np.fromfile(os.path.join(tmp, "out", "final_PA.bin")).reshape(case.mesh.ne, 7)
        assert os.path.exists(os.path.join(tmp, "converge.dat"))
        if unsteady:
            assert os.path.exists(os.path.join(tmp, "converge_uns.dat"))
            # step_residual after every stage (NScreenOutput = 1)
            sr = np.fromfile(os.path.join(tmp, "out", "resid_step.bin")).reshape(-1, 9)
        assert os.path.exists(os.path.join(tmp, "fvp1.sav"))
    ref = d["ref_resid"]
    assert res.shape == ref.shape
    if unsteady:
        assert np.array_equal(res[:, :2], ref[:, :2])   # same inner-iteration counts
        res, ref = res[:, 2:], ref[:, 2:]
        assert sr.shape == d["ref_step_resid"].shape
        assert relerr(sr[:, 1:], d["ref_step_resid"][:, 1:], -1) < 1e-10
    assert relerr(res[:, :7], ref[:, :7], -1) < 1e-10
    assert relerr(pa, d["ref_final_PA"], -1) < 1e-10
