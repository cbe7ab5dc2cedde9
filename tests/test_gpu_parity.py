"""The HIP step on the MI355X against the reference fixtures and the oracle, through the C ABI."""
import numpy as np
import pytest

import oracle
import snncuda_amd as sa
from snncuda_amd import workloads as W
from tests import golden_util as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _needs_device():
    if sa.device_count() == 0:
        pytest.fail("no HIP device visible: the GPU tests must run on an MI355X")


def run_device(net, **kw):
    dev = sa.Network(net, **kw)
    dev.run_commands(net.commands)
    return dev


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_device_matches_reference_fixture(name):
    print(f"case {name}", flush=True)
    net = G.CASES[name]()
    g = G.load(name)
    with run_device(net, max_events_per_tick=1 << 22, fire_log_capacity=1 << 20) as dev:
        d, f = dev.tick_counts()
        bad = G.mismatches(g, d, f, dev.raster(), dev.neurons(), dev.synapses())
        assert dev.delivered == int(g["delivered"]) and dev.fired == int(g["fired"])
    assert not bad, f"{name}: device differs from the reference in {bad}"


@pytest.mark.parametrize("seed", [1, 3, 5])
def test_device_matches_oracle_other_seeds(seed):
    print(f"seed {seed}", flush=True)
    for net in (W.config0(seed, threshold=1.1, ticks=400), W.feature_net(100 + seed),
                W.config1_reduced(seed, ticks=150)):
        ref = oracle.Oracle(net)
        ref.run_commands(net.commands)
        with run_device(net) as dev:
            assert np.array_equal(dev.tick_counts()[0], ref.tick_counts()[0]), net.name
            assert np.array_equal(dev.raster(), ref.fires()[1]), net.name
            dn, rn = dev.neurons(), ref.neurons()
            for k in ("membrane", "current", "spike_count", "t_match", "t_learned", "last_fired"):
                assert np.array_equal(dn[k], rn[k]), (net.name, k)
            ds, rs = dev.synapses(), ref.synapses()
            for k in ("weight", "pre_count", "upd_count", "last_pre", "last_post", "c_match", "c_learned"):
                assert np.array_equal(ds[k], rs[k].astype(ds[k].dtype)), (net.name, k)
            pl = dev.plasticity()
            assert np.array_equal(pl["ltp"], rs["ltp"]), (net.name, "ltp")
            assert np.array_equal(pl["ltp"] + pl["ltd"], ds["upd_count"]), (net.name, "ltp + ltd")


@pytest.mark.parametrize("knob", ["SNNCUDA_WIDE_G", "SNNCUDA_NO_GROUP_STARTS"])
def test_large_network_paths_equal_the_oracle(monkeypatch, knob):
    """Paths that only larger networks take, forced on small ones: SNNCUDA_WIDE_G=1 sorts each
    target's events by full entries instead of 32-bit keys (ticks of more than 2^27 events), and
    SNNCUDA_NO_GROUP_STARTS=1 finds each segment from the row start and the group offsets (networks
    whose table of group starts would take more than 1/16 of free memory). Small networks (sparse
    and rounds fold paths) and the 1M-neuron law at 8,000 neurons (chunked path, STDP off and on)
    still equal the oracle field by field."""
    print(f"{knob}=1", flush=True)
    monkeypatch.setenv(knob, "1")
    test_device_matches_oracle_other_seeds(3)
    test_bench_law_matches_oracle(False, 2.5)
    test_bench_law_matches_oracle(True, 2.75)


@pytest.mark.parametrize("delay_max", [130, 300])
def test_long_delays_plan_in_global_memory_and_equal_the_oracle(delay_max):
    """Delays and code offsets past 127 ticks make the tick's group table (2 * lmax + 1 groups) too
    large for LDS: one wave plans it in global memory (k_plan) and the segment kernel reads it
    there. Every feature of the step with delays up to delay_max (code offsets added), field by
    field against the oracle."""
    net = W.feature_net(40 + delay_max, n=80, fanout=8, ticks=2 * delay_max + 60, delay_max=delay_max)
    print(f"{net.name}: delays up to {delay_max}", flush=True)
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    with run_device(net) as dev:
        assert np.array_equal(dev.tick_counts()[0], ref.tick_counts()[0])
        assert np.array_equal(dev.raster(), ref.fires()[1])
        dn, rn = dev.neurons(), ref.neurons()
        for k in ("membrane", "current", "spike_count", "t_match", "t_learned", "last_fired"):
            assert np.array_equal(dn[k], rn[k]), k
        ds, rs = dev.synapses(), ref.synapses()
        for k in ("weight", "pre_count", "upd_count", "last_pre", "last_post", "c_match", "c_learned"):
            assert np.array_equal(ds[k], rs[k].astype(ds[k].dtype)), k
    assert ref.delivered > 0


def test_capacity_overflow_raises():
    net = G.CASES["config0_burst"]()
    with pytest.raises(sa.CapacityError):
        run_device(net, fire_log_capacity=1 << 12)
    with pytest.raises(sa.CapacityError):
        run_device(net, max_events_per_tick=1 << 12, fire_log_capacity=1 << 20)


def test_split_runs_equal_one_run():
    net = W.config0(5, threshold=1.1, ticks=300)
    injects = [c for c in net.commands if c[0] == 0]
    a = run_device(net)
    b = sa.Network(net)
    b.run_commands(injects)
    for t in (0, 1, 57, 58, 200, 299):
        b.run_until(t)
    assert np.array_equal(a.raster(), b.raster())
    assert np.array_equal(a.synapses()["weight"], b.synapses()["weight"])
    a.close()
    b.close()


@pytest.mark.parametrize("seed", [2, 9])
def test_injects_out_of_tick_order_from_several_calls_equal_the_oracle(seed):
    """Injects issued in shuffled tick order, split over calls made at different clock positions, so
    one tick receives injects of several calls (ordered by the call's clock, then by call order) and
    late ones are clamped to the clock; the run is cut into segments of uneven length. Threshold
    1.2 keeps the clamped bursts from igniting the network (at 1.1 they do, and activity grows
    without bound)."""
    print(f"seed {seed}", flush=True)
    rng = np.random.default_rng(seed)
    base = W.config0(seed, threshold=1.2, ticks=260)
    inj = [c for c in base.commands if c[0] == 0]
    inj = [inj[i] for i in rng.permutation(len(inj))]
    extra = [(0, int(rng.integers(0, base.n_neurons)), int(rng.integers(0, 260)),
              float(np.float32(rng.uniform(0.05, 1.0)))) for _ in range(200)]
    groups = np.array_split(np.arange(len(inj) + len(extra)), 4)
    allc = inj + extra
    cmds = []
    for g, end in zip(groups, (37, 38, 150, 259)):
        cmds += [allc[i] for i in g]
        cmds.append((1, 0, end, 0.0))
    ref = oracle.Oracle(base)
    ref.run_commands(cmds)
    with sa.Network(base) as dev:
        dev.run_commands(cmds)
        assert np.array_equal(dev.tick_counts()[0], ref.tick_counts()[0])
        assert np.array_equal(dev.raster(), ref.fires()[1])
        dn, rn = dev.neurons(), ref.neurons()
        for k in ("membrane", "current", "spike_count", "last_fired"):
            assert np.array_equal(dn[k], rn[k]), k
        assert np.array_equal(dev.synapses()["weight"], ref.synapses()["weight"].astype(np.float32))
    assert ref.fired > 250


@pytest.mark.parametrize("plastic", [False, True, "exc"])
def test_device_generator_equals_host_build(plastic):
    # snn_net_create_ei must build the same network as snn_net_create on workloads.config2's arrays
    net = W.ei_network(3000, W.lognormal_fanout(3, 3000), 3, n_sources=60, source_fanout=200, p_source=0.05,
                       ticks=120, threshold=2.5, delay_max=32, compartments=(W.BASAL, W.APICAL), plastic=plastic,
                       w_inh=5.0)
    a = run_device(net)
    b = sa.Network(spec=net.info["spec"])
    b.run_commands(net.commands)
    assert a.delivered == b.delivered and a.fired == b.fired
    assert np.array_equal(a.raster(), b.raster())
    sa_, sb = a.synapses(), b.synapses()
    for k in ("weight", "c_match", "c_learned", "pre_count", "last_pre", "last_post", "upd_count"):
        assert np.array_equal(sa_[k], sb[k]), k
    na, nb = a.neurons(), b.neurons()
    for k in ("membrane", "current", "spike_count", "t_match"):
        assert np.array_equal(na[k], nb[k]), k
    a.close()
    b.close()


@pytest.mark.parametrize("plastic", [False, True])
def test_per_source_code_patterns_equal_per_synapse_records(plastic, monkeypatch):
    """With one code offset per synapse and uniform target parameters the code patterns are kept per
    source neuron; every synapse's readback must equal the per-synapse records (and the oracle),
    also when the run ends with fires still in flight on their longest lags."""
    net = W.config0(7, threshold=1.1, ticks=333, plastic=plastic)
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("SNNCUDA_CODE_PER_SYNAPSE", mode)
        with run_device(net) as dev:
            out[mode] = dev.synapses()
    rs = ref.synapses()
    for k in ("c_match", "c_learned", "last_pre", "last_post", "weight", "pre_count"):
        assert np.array_equal(out["0"][k], out["1"][k]), k
        assert np.array_equal(out["0"][k], rs[k].astype(out["0"][k].dtype)), k


@pytest.mark.parametrize("stdp,theta", [(False, 2.5), (True, 2.75), (False, 2.4)])
def test_bench_law_matches_oracle(stdp, theta):
    """The 1M-neuron workload's law at 8,000 neurons (lognormal fan-out, delays U{1..32}, 2%
    ignition inject, Poisson inputs; STDP on for every synapse but those of inhibitory neurons at
    threshold 2.75, as the bench runs it; the denser regime's threshold 2.4): device = oracle, field
    by field."""
    n = 8000
    net = W.ei_network(n, W.lognormal_fanout(0, n), 0, n_sources=n // 100, source_fanout=1000, p_source=0.005,
                       ticks=100, threshold=theta, delay_max=32, compartments=(W.BASAL, W.APICAL),
                       plastic="exc" if stdp else False, w_inh=5.0)
    hit = np.nonzero(W.unif24(W.hash3(0, 40, np.arange(n, dtype=np.uint64))) < np.float32(0.02))[0]
    net.commands = [(0, int(i), 0, theta) for i in hit] + net.commands
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    print(f"  stdp={stdp} theta={theta}: oracle {ref.delivered} delivered, {ref.fired} fired", flush=True)
    assert ref.fired > 10 * n // 100  # recurrent activity, not only the inputs
    with run_device(net) as dev:
        assert np.array_equal(dev.tick_counts()[0], ref.tick_counts()[0])
        assert np.array_equal(dev.raster(), ref.fires()[1])
        dn, rn = dev.neurons(), ref.neurons()
        for k in ("membrane", "current", "spike_count", "t_match", "t_learned", "last_fired"):
            assert np.array_equal(dn[k], rn[k]), k
        ds, rs = dev.synapses(), ref.synapses()
        for k in ("weight", "upd_count", "last_pre", "last_post", "c_match", "c_learned"):
            assert np.array_equal(ds[k], rs[k].astype(ds[k].dtype)), k


def test_a_clamp_bound_per_plastic_synapse_matches_the_oracle():
    net = W.feature_net(21, n=2000, ticks=600, plastic_frac=0.9, threshold=0.6, max_weight_range=(0.05, 0.6))
    print("feature_net(21, n=2000) with a max_weight per synapse", flush=True)
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    with run_device(net) as dev:
        assert np.array_equal(dev.synapses()["weight"], ref.synapses()["weight"])
        assert np.array_equal(dev.raster(), ref.fires()[1])


@pytest.mark.parametrize("case,limit,under,name", [
    ("candidate_runs", {}, {"distinct_ticks": 8}, "neuron-temporal-pattern"),
    ("candidate_runs_wide", {"pattern_runs": 32}, {"distinct_ticks": 32}, "neuron-temporal-pattern"),
    ("learned_patterns", {"learned_pattern_slots": 4}, {"max_refs": 4}, "neuron-temporal-pattern"),
    ("code_patterns", {}, {"fires": 5}, "synapse-code-pattern"),
])
def test_pattern_capacity_overflow_raises_and_just_under_it_equals_the_oracle(case, limit, under, name):
    from tests import capacity_nets as C
    print(f"capacity case {case}", flush=True)
    net = getattr(C, case)()
    with pytest.raises(sa.CapacityError, match=name):
        run_device(net, **limit)
    net = getattr(C, case)(**under)
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    with run_device(net, **limit) as dev:
        for k in ("t_match", "t_learned", "membrane"):
            assert np.array_equal(dev.neurons()[k], ref.neurons()[k]), k
        assert np.array_equal(dev.synapses()["c_learned"], ref.synapses()["c_learned"].astype(np.uint64))


@pytest.mark.parametrize("seed", [0, 1])
def test_neurons_holding_many_learned_patterns_match_the_oracle(seed):
    from tests import capacity_nets as C
    print(f"many_learned({seed})", flush=True)
    net = C.many_learned(seed)
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    with run_device(net) as dev:
        dn, rn = dev.neurons(), ref.neurons()
        for k in ("membrane", "current", "spike_count", "t_match", "t_learned", "last_fired"):
            assert np.array_equal(dn[k], rn[k]), k
        assert np.array_equal(dev.raster(), ref.fires()[1])


@pytest.mark.parametrize("device_pages,one_call", [(1, False), (2, False), (1, True)])
def test_paged_pattern_pool_equals_the_device_pool_and_the_oracle(device_pages, one_call):
    # 3,000 neurons (three 1,024-neuron pages) learning up to 500 patterns each; the paged runs hold
    # one or two pages in the device cache and move pages at every run_until call and every 64 ticks
    from tests import capacity_nets as C
    net = C.many_learned(3, n=3000, ticks=600, rate=600)
    if one_call:
        net.commands = [c for c in net.commands if c[0] == 0] + [(1, 0, 599, 0.0)]
    print(f"paged pool: device_pages {device_pages} one_call {one_call}", flush=True)
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    with run_device(net) as a, run_device(net, pattern_device_pages=device_pages) as b:
        pg = b.pattern_paging()
        assert pg["pages"] == 3 and pg["device_pages"] == device_pages and pg["moves"] > 0, pg
        assert a.pattern_paging()["device_pages"] == 0
        na, nb, nr = a.neurons(), b.neurons(), ref.neurons()
        for k in ("membrane", "current", "spike_count", "t_match", "t_learned", "last_fired"):
            assert np.array_equal(na[k], nb[k]), k
            assert np.array_equal(nb[k], nr[k]), k
        assert np.array_equal(a.raster(), b.raster())
        assert int(nr["t_learned"].max()) > 8

