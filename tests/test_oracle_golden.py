"""The C oracle against the reference's own outputs (golden vectors written
by tests/golden/make_golden.py from the compiled reference).

Tolerances: topology bit-exact; geometry bit-exact except AMLS, which the
reference forms with LAPACK dgesvd and the oracle with a QR (<=1e-14
relative); per-call arrays <=1e-12 relative per variable (they differ only by
the AMLS rounding and libm); the 8000-step converged field <=1e-10.
"""
import numpy as np
import pytest

from helpers import case_from_fixture, fixture_names, load, relerr

oracle = pytest.importorskip("oracle.pyoracle")

NAMES = fixture_names()


@pytest.fixture(scope="module", params=NAMES)
def fx(request):
    d = load(request.param)
    return request.param, d, case_from_fixture(d)


def test_setup_arrays(fx):
    name, d, case = fx
    o = oracle.OracleCase(case)
    for n in ("Neighbor", "FNode", "NeighF", "NeighC", "Neigh", "Num_ghostcell", "Ghostcell", "Hostcell",
              "BCtype_Ghostcell"):
        assert np.array_equal(o.get(n), d["ref_" + n]), n
    for n in ("VECX", "vol", "cellxy", "reflen", "wdis", "vol_Ghostcell", "CellXY_Ghostcell"):
        assert np.array_equal(o.get(n), d["ref_" + n]), n
    assert relerr(o.get("AMLS"), d["ref_AMLS"]) < 1e-14
    assert o.get("total_volume") == float(d["ref_total_volume"])
    assert case.p_out == float(d["ref_p_out"])
    assert np.array_equal(o.get("WA"), d["ref_init_WA"])


def test_per_call_arrays(fx):
    name, d, case = fx
    o = oracle.OracleCase(case)
    s = 1
    while "ref_s%d_step" % s in d:
        o.timestep()
        assert relerr(o.get("step"), d["ref_s%d_step" % s]) < 1e-13
        assert relerr(o.get("alpaF"), d["ref_s%d_alpaF" % s]) < 1e-13
        o.gradient()
        assert relerr(o.get("GPA"), d["ref_s%d_GPA" % s], -1) < 1e-13
        assert relerr(o.get("gradC"), d["ref_s%d_gradC0" % s], -1) < 1e-12
        if case.info.limit_id > 0:
            o.limiter()
        assert relerr(o.get("gradC"), d["ref_s%d_gradC" % s], -1) < 1e-12
        o.flux_sourceterm()
        assert relerr(o.get("hWA"), d["ref_s%d_res" % s], -1) < 1e-12
        o.lusgs_steady()
        assert relerr(o.get("hWA"), d["ref_s%d_dW" % s], -1) < 1e-12
        assert relerr(o.get("PA"), d["ref_s%d_PA" % s], -1) < 1e-13
        assert relerr(o.get("WA"), d["ref_s%d_WA" % s], -1) < 1e-13
        s += 1


def test_trajectory(fx):
    """Whole steps: every step_residual row and the final state."""
    name, d, case = fx
    o = oracle.OracleCase(case)
    nsteps = int(d["nsteps"])
    every = int(d["resid_every"])
    ref = d["ref_resid"]
    if "ref_error_step" in d:
        err_step = int(d["ref_error_step"])
        for s in range(1, err_step):
            o.stage()
        with pytest.raises(RuntimeError, match="temperature"):
            o.stage()
        return
    got = []
    for s in range(1, nsteps + 1):
        o.stage()
        if (s - 1) % every == 0:
            got.append(o.residual())
    got = np.array(got)
    assert got.shape == ref.shape
    # dW norms, each row against its own size, while the increments are
    # above rounding noise (below ~1e-8 of the first row they are noise once
    # the field has converged)
    rel = np.abs(got[:, :7] - ref[:, :7]) / np.maximum(np.abs(ref[:, :7]), 1e-300)
    tol = 1e-8 if nsteps > 100 else 1e-10
    big = np.abs(ref[:, :7]) > 1e-6 * np.abs(ref[0, :7])
    assert np.max(rel[big]) < tol
    assert relerr(o.get("PA"), d["ref_final_PA"], -1) < (1e-10 if nsteps > 100 else 1e-12)
    if "ref_final_WA" in d:
        assert relerr(o.get("WA"), d["ref_final_WA"], -1) < (1e-10 if nsteps > 100 else 1e-12)


@pytest.mark.parametrize("name", [n for n in fixture_names("trajectory") if n not in NAMES])
def test_long_run_trajectory(name):
    """Whole runs at workload size (the 12k-triangle nozzle, 2000 steps): every
    step_residual row and the final PA of the reference."""
    d = load(name)
    test_trajectory((name, d, case_from_fixture(d)))
