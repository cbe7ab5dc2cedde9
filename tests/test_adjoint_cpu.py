"""The FP64 fixed-spike adjoint of the oracle (oracle/adjoint_oracle.c) on the CPU: its gradient
against FP64 central finite differences, and its forward at the default parameters against the
reference step (the oracle's membrane)."""
import numpy as np
import pytest

import oracle
from snncuda_amd import workloads as W

DEFAULT = np.array([1.0, 1.0, 1.0, 1.0, 5.0, 100.0, 10.0, 150.0])


def setup(n=300, fan_in=30, ticks=200, seed=0):
    net = W.config4(seed, n=n, fan_in=fan_in, ticks=ticks)
    o = oracle.Oracle(net, log_events=True)
    o.run_commands(net.commands)
    adj = oracle.AdjointOracle(net, ticks, o.event_log())
    w = net.weight.astype(np.float64)
    rng = np.random.default_rng(seed)
    w_target = w + rng.normal(0.0, 0.05, w.shape)
    target = adj.forward(w_target, DEFAULT[:4], DEFAULT[4:])
    return net, o, adj, w, target


def test_default_forward_reduces_to_reference_membrane():
    net, o, adj, w, _ = setup()
    v = adj.forward(w, DEFAULT[:4], DEFAULT[4:])
    log = o.event_log()
    last = np.full(net.n_neurons, -1, np.int64)
    np.maximum.at(last, log["target"].astype(np.int64), log["tick"].astype(np.int64))
    mem = o.neurons()["membrane"].astype(np.float64)
    has = last >= 0
    assert has.sum() > 0.9 * net.n_neurons
    got = v[last[has], np.nonzero(has)[0]]
    assert np.allclose(got, mem[has], rtol=1e-5, atol=1e-6)


def test_fired_events_are_the_resets_and_the_raster():
    net, o, adj, w, _ = setup()
    log = o.event_log()
    ticks, neurons = o.fires()
    assert np.array_equal(log["target"][log["fired"] == 1], neurons)
    assert np.array_equal(log["tick"][log["fired"] == 1], ticks)


@pytest.mark.parametrize("seed", [0, 1])
def test_adjoint_matches_central_differences(seed):
    net, o, adj, w, target = setup(seed=seed)
    alpha, tau = DEFAULT[:4].copy(), DEFAULT[4:].copy()
    alpha[1] = 0.9
    tau[1] = 80.0
    loss, gw, ga, gt = adj.grad(w, alpha, tau, target)
    assert loss > 0
    h = 1e-3
    rng = np.random.default_rng(10 + seed)
    for j in rng.choice(net.n_synapses, 60, replace=False):
        i = int(net.tgt[j])
        wp, wm = w.copy(), w.copy()
        wp[j] += h
        wm[j] -= h
        fd = (adj.loss_target(i, wp, alpha, tau, target) - adj.loss_target(i, wm, alpha, tau, target)) / (2 * h)
        assert abs(fd - gw[j]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(gw).max()), (j, fd, gw[j])
    for k in range(4):
        ap, am = alpha.copy(), alpha.copy()
        ap[k] += h
        am[k] -= h
        fd = (adj.loss(w, ap, tau, target) - adj.loss(w, am, tau, target)) / (2 * h)
        assert abs(fd - ga[k]) <= 1e-6 * max(abs(fd), 1e-9), ("alpha", k, fd, ga[k])
        tp, tm = tau.copy(), tau.copy()
        tp[k] += h
        tm[k] -= h
        fd = (adj.loss(w, alpha, tp, target) - adj.loss(w, alpha, tm, target)) / (2 * h)
        assert abs(fd - gt[k]) <= 1e-4 * max(abs(fd), 1e-9), ("tau", k, fd, gt[k])


def test_planted_mistake_is_caught_by_finite_differences():
    """A gradient with the reset ignored differs from the finite differences: the check is sharp."""
    net, o, adj, w, target = setup()
    log = o.event_log()
    log_noreset = dict(log)
    log_noreset["fired"] = np.zeros_like(log["fired"])
    wrong = oracle.AdjointOracle(net, 200, log_noreset)
    _, gw_wrong, _, _ = wrong.grad(w, DEFAULT[:4], DEFAULT[4:], target)
    _, gw, _, _ = adj.grad(w, DEFAULT[:4], DEFAULT[4:], target)
    assert np.abs(gw_wrong - gw).max() > 1e-3 * np.abs(gw).max()
