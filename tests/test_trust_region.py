"""The calibrated trust region of the learned source's gate (host side):
grid cells are accepted exactly where the held-out states in them stayed
within the bound and were numerous enough; states outside the grid are
never accepted; the host gate restatement applies it."""
import numpy as np

from rocketcombustor_amd import surrogate


def synthetic(n=20000, seed=0):
    rng = np.random.default_rng(seed)
    t = rng.uniform(300.0, 3300.0, n)
    p = np.exp(rng.uniform(np.log(1e5), np.log(5e6), n))
    y1 = rng.uniform(0.0, 0.11, n)
    y2 = rng.uniform(0.0, 0.89, n)
    y3 = 1.0 - y1 - y2
    raw = np.column_stack([t, p, y1, y2, y3, np.zeros(n)])
    w_exact = np.column_stack([-y1 * 1e3, -y1 * 8e3, y1 * 9e3])
    return raw, w_exact


def test_calibration_marks_exactly_the_good_cells():
    raw, w_exact = synthetic()
    tf = surrogate.Transform.fit(raw, w_exact)
    # the network is exact except for T in [1000, 1400) K, where it is 1% off
    w_net = w_exact.copy()
    bad = (raw[:, 0] >= 1000.0) & (raw[:, 0] < 1400.0)
    w_net[bad] *= 1.01
    tr = surrogate.TrustRegion.calibrate(tf, raw, w_net, w_exact, dims=(30, 4, 5), tau=1e-4, nmin=3)
    idx = tr.index(raw)
    assert np.all(idx >= 0)
    # every grid cell holding a bad state is rejected; the others with
    # enough states are accepted
    n = np.bincount(idx, minlength=tr.accept.size)
    has_bad = np.bincount(idx[bad], minlength=tr.accept.size) > 0
    assert np.array_equal(tr.accept.astype(bool), (n >= 3) & ~has_bad)
    ok = surrogate.gate(tf, raw, w_net, w_exact, 1e9, 0.0, tr)
    assert not np.any(ok & bad)
    assert ok.mean() > 0.5
    # the accepted states' errors are within the bound
    err = surrogate.rel_error(tf, w_net, w_exact)
    assert err[ok].max() <= 1e-4


def test_states_outside_the_grid_are_rejected():
    raw, w_exact = synthetic()
    tf = surrogate.Transform.fit(raw, w_exact)
    tr = surrogate.TrustRegion.calibrate(tf, raw, w_exact, w_exact, dims=(8, 4, 4), tau=1e-4, nmin=1)
    assert tr.accept.all()
    out = raw[:4].copy()
    out[0, 0] = raw[:, 0].max() * 1.5        # hotter than the grid
    out[1, 1] = raw[:, 1].min() * 0.5        # lower pressure
    out[2, 2] = -0.01                        # Y_1 below
    assert np.all(tr.index(out[:3]) == -1)
    assert tr.index(out[3:])[0] >= 0
