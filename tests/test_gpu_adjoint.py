"""The fixed-spike adjoint on the MI355X (snn_adjoint_* through snncuda_amd.adjoint): its event
lists equal the oracle's processing log, its FP32 forward and gradient match the FP64 CPU adjoint
(oracle/adjoint_oracle.c), its gradient matches FP64 central finite differences (h = 1e-3) on
1,000 random parameters of the 10,000-neuron workload, and the torch.autograd.Function returns the
same gradient."""
import numpy as np
import pytest

import oracle
import snncuda_amd as sa
from snncuda_amd import workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

DEFAULT = np.array([1.0, 1.0, 1.0, 1.0, 5.0, 100.0, 10.0, 150.0])
PARAMS = np.array([1.0, 0.9, 1.1, 1.0, 5.0, 80.0, 12.0, 150.0])


@pytest.fixture(scope="module", autouse=True)
def _needs_device():
    if sa.device_count() == 0:
        pytest.fail("no HIP device visible: the GPU tests must run on an MI355X")


def _record(net):
    run = sa.Network(net, max_events_per_tick=1 << 24, fire_log_capacity=1 << 22, raster_capacity=1 << 24,
                     track_pre_count=False)
    run.record_fire_events()
    run.run_commands(net.commands)
    return run


def _case(net, ticks):
    import time
    from snncuda_amd.adjoint import Adjoint
    t0 = time.time()
    o = oracle.Oracle(net, log_events=True)
    o.run_commands(net.commands)
    print(f"  oracle run {time.time() - t0:.1f} s", flush=True)
    run = _record(net)
    adj = Adjoint(run, net, ticks)
    print(f"  device run + adjoint event lists {time.time() - t0:.1f} s", flush=True)
    ref = oracle.AdjointOracle(net, ticks, o.event_log())
    print(f"  FP64 adjoint built {time.time() - t0:.1f} s", flush=True)
    return o, run, adj, ref


def _cases():
    return {
        "config4_small": (lambda: W.config4(0, n=300, fan_in=30, ticks=200), 200),
        "feature1": (lambda: W.feature_net(1), 300),
        "feature3": (lambda: W.feature_net(3), 300),
    }


@pytest.mark.parametrize("name", sorted(_cases()))
def test_event_lists_equal_oracle_log(name):
    make, ticks = _cases()[name]
    net = make()
    print(f"case {name}", flush=True)
    o, run, adj, _ = _case(net, ticks)
    log = o.event_log()
    keep = log["tick"] < ticks
    order = np.argsort(log["target"][keep], kind="stable")
    ev = adj.events()
    counts = np.bincount(log["target"][keep], minlength=net.n_neurons)
    assert np.array_equal(np.diff(ev["ptr"]), counts)
    assert np.array_equal(ev["tick"], log["tick"][keep][order])
    assert np.array_equal(ev["synapse"], log["synapse"][keep][order])
    assert np.array_equal(ev["reset"], log["fired"][keep][order])
    ext = ev["synapse"] == 0xFFFFFFFF
    assert np.array_equal(ev["weight"][ext], log["weight"][keep][order][ext])


def _torch_run(adj, w, params, target):
    wt = torch.tensor(w, dtype=torch.float32, device="cuda")
    tr = torch.empty((adj.T, adj.n), dtype=torch.float32, device="cuda")
    adj.forward(wt, params, tr)
    tg = torch.tensor(target, dtype=torch.float32, device="cuda")
    g = (2.0 / (adj.n * adj.T)) * (tr - tg)
    gw = torch.empty_like(wt)
    gp = adj.backward(g.contiguous(), gw)
    return tr.cpu().numpy().astype(np.float64), gw.cpu().numpy().astype(np.float64), gp


@pytest.mark.parametrize("name", sorted(_cases()))
def test_forward_and_gradient_match_fp64_adjoint(name):
    make, ticks = _cases()[name]
    net = make()
    print(f"case {name}", flush=True)
    o, run, adj, ref = _case(net, ticks)
    w = net.weight.astype(np.float64)
    target = ref.forward(w + np.random.default_rng(1).normal(0, 0.05, w.shape), DEFAULT[:4], DEFAULT[4:])
    v64 = ref.forward(w, PARAMS[:4], PARAMS[4:])
    _, gw64, ga64, gt64 = ref.grad(w, PARAMS[:4], PARAMS[4:], target)
    v, gw, gp = _torch_run(adj, w, PARAMS, target)
    assert np.abs(v - v64).max() <= 1e-5 * np.abs(v64).max()
    assert np.abs(gw - gw64).max() <= 1e-4 * np.abs(gw64).max()
    g64 = np.concatenate([ga64, gt64])
    for k in range(8):
        assert abs(gp[k] - g64[k]) <= 1e-4 * max(abs(g64[k]), 1e-6 * np.abs(g64).max()), (k, gp[k], g64[k])


def test_gradient_matches_finite_differences_on_1000_parameters():
    """The 10,000-neuron workload (fan-in 500, 5e6 synapses, 500 ticks): GPU FP32 adjoint against
    FP64 central differences (h = 1e-3) on 992 random weights and the 8 coupling / time-constant
    parameters, and against the FP64 CPU adjoint on every parameter."""
    net = W.config4(0)
    T = 500
    print(f"config4: {net.n_neurons} neurons, {net.n_synapses} synapses", flush=True)
    o, run, adj, ref = _case(net, T)
    print(f"events {adj.n_events}, fires {run.fired}", flush=True)
    assert run.fired == o.fired and adj.n_events > 10_000_000
    w = net.weight.astype(np.float64)
    target = ref.forward(w + np.random.default_rng(2).normal(0, 0.05, w.shape), DEFAULT[:4], DEFAULT[4:])
    alpha, tau = PARAMS[:4], PARAMS[4:]
    _, gw64, ga64, gt64 = ref.grad(w, alpha, tau, target)
    _, gw, gp = _torch_run(adj, w, PARAMS, target)
    # against the FP64 adjoint, every parameter
    assert np.abs(gw - gw64).max() <= 1e-4 * np.abs(gw64).max()
    g64 = np.concatenate([ga64, gt64])
    for k in range(8):
        assert abs(gp[k] - g64[k]) <= 1e-4 * max(abs(g64[k]), 1e-6 * np.abs(g64).max()), (k, gp[k], g64[k])
    # against FP64 central differences on 1,000 random parameters
    # The FP32 trace rounds each residual v - target to ~2e-6 absolute (|v| <= 40), so a component
    # whose residual contributions cancel carries that absolute error: the bar is 1e-4 of the
    # vector norm (norm-wise relative error) and, per component, 1e-4 of max(|fd|, rms of the fds).
    h = 1e-3
    rng = np.random.default_rng(3)
    js = rng.choice(net.n_synapses, 992, replace=False)
    fds = np.zeros(js.shape[0])
    for q, j in enumerate(js):
        i = int(net.tgt[j])
        wp, wm = w.copy(), w.copy()
        wp[j] += h
        wm[j] -= h
        fds[q] = (ref.loss_target(i, wp, alpha, tau, target) - ref.loss_target(i, wm, alpha, tau, target)) / (2 * h)
    err = np.abs(gw[js] - fds)
    rms = np.sqrt(np.mean(fds ** 2))
    normwise = np.linalg.norm(gw[js] - fds) / np.linalg.norm(fds)
    percomp = (err / np.maximum(np.abs(fds), rms)).max()
    print(f"weights vs FD: norm-wise relative error {normwise:.3e}; max per-component error / max(|fd|, rms) "
          f"{percomp:.3e}; max relative error where |fd| >= rms "
          f"{(err / np.abs(fds))[np.abs(fds) >= rms].max():.3e}", flush=True)
    assert normwise <= 1e-4
    assert percomp <= 1e-4
    for k in range(4):
        for arr, gi in ((alpha, k), (tau, 4 + k)):
            p, m = arr.copy(), arr.copy()
            p[k] += h
            m[k] -= h
            if gi < 4:
                fd = (ref.loss(w, p, tau, target) - ref.loss(w, m, tau, target)) / (2 * h)
            else:
                fd = (ref.loss(w, alpha, p, target) - ref.loss(w, alpha, m, target)) / (2 * h)
            print(f"param {gi}: gpu {gp[gi]:.6e} fd {fd:.6e}", flush=True)
            assert abs(gp[gi] - fd) <= 1e-4 * max(abs(fd), 1e-6 * np.abs(g64).max()), (gi, gp[gi], fd)


def test_autograd_function_gives_the_adjoint_gradient():
    from snncuda_amd.adjoint import somatic_trace
    net = W.config4(0, n=300, fan_in=30, ticks=200)
    o, run, adj, ref = _case(net, 200)
    w = net.weight.astype(np.float64)
    wt = torch.tensor(net.weight, dtype=torch.float32, device="cuda", requires_grad=True)
    pt = torch.tensor(PARAMS, dtype=torch.float32, device="cuda", requires_grad=True)
    with torch.no_grad():
        target = somatic_trace(adj, wt * 1.05, pt)
    trace = somatic_trace(adj, wt, pt)
    with torch.no_grad():
        somatic_trace(adj, wt * 0.5, pt)  # a later forward replaces the checkpoints
    loss = ((trace - target) ** 2).mean()
    loss.backward()
    _, gw64, ga64, gt64 = ref.grad(w, PARAMS[:4], PARAMS[4:], target.cpu().numpy().astype(np.float64))
    assert np.abs(wt.grad.cpu().numpy() - gw64).max() <= 1e-4 * np.abs(gw64).max()
    g64 = np.concatenate([ga64, gt64])
    gp = pt.grad.cpu().numpy().astype(np.float64)
    assert np.abs(gp - g64).max() <= 1e-4 * np.abs(g64).max()
