"""The served request that stops (rid 19 of the serving stream, 73,960
cells; tests/golden/serving_rid19_stop.npz) run on the device to its stop:
it stops at the oracle's step and face with the reference's error code,
and its step_residual rows follow the reference's while the reference and
the oracle agree (to step 2501) and the oracle's to 1e-12 up to the stop.  The
reference itself runs on past step 5000: the two codes' rounding
differences grow until the fields part (test_serving_stop_fixture.py).
"""
import sys

import numpy as np
import pytest

from helpers import load
from rocketcombustor_amd import native
from rocketcombustor_amd.solver import SteadyCase
from test_serving_stop_fixture import stop_case

pytestmark = pytest.mark.gpu


def test_served_stop_matches_the_oracle():
    d = load("serving_rid19_stop")
    case = stop_case(d)
    sys.__stderr__.write("[gpu-test] serving stop rid 19 NE=%d, up to 5000 steps\n" % case.mesh.ne)
    sys.__stderr__.flush()
    rows = []
    with SteadyCase(case) as g:
        with pytest.raises(native.RcError) as ei:
            for _ in range(int(d["nsteps"]) // 10):
                # one report per chunk: the step_residual of steps 1, 11, 21, ...
                rows.append(g.run_steady(10, every=11)[0])
    e = ei.value
    assert e.code == native.RC_ERR_TEMPERATURE
    assert "(index %d, step %d)" % (int(d["oracle_stop_face"]), int(d["oracle_stop_step"])) in str(e), str(e)
    rows = np.array(rows)
    ref, orc = d["ref_resid"], d["oracle_resid"]
    # the chunk that holds the stop reports nothing
    assert len(rows) == (int(d["oracle_stop_step"]) - 1) // 10
    orc = orc[:len(rows)]
    steps = 1 + 10 * np.arange(len(rows))

    def rel(a, b):
        return np.max(np.abs(a[:, :7] - b[:, :7]) / np.abs(b[:, :7]), axis=1)

    r_ref, r_orc = rel(rows, ref[:len(rows)]), rel(rows, orc)
    sys.__stderr__.write("[gpu-test] max rel diff vs reference to 2501: %.2e; vs oracle to 2501 / 4001 / stop: "
                         "%.2e / %.2e / %.2e\n" % (r_ref[steps <= 2501].max(), r_orc[steps <= 2501].max(),
                                                   r_orc[steps <= 4001].max(), r_orc.max()))
    assert r_ref[steps <= 2501].max() < 1e-9
    # the device and the oracle share the set-up's arithmetic (AMLS by the
    # same thin QR) and part only by the last bits of exp/log: they stay
    # together to the stop
    assert r_orc.max() < 1e-12
