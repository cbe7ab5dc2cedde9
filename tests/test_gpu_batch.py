"""The batched form on the MI355X through the C ABI (snn_batch_*): every network of a batch equals
that network run alone in the oracle, and schedules (partition size, scatter workgroups, HIP graph
replay) and the online autotuner leave every result unchanged."""
import numpy as np
import pytest

import oracle
import snncuda_amd as sa
from tests import batch_nets as B

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _needs_device():
    if sa.device_count() == 0:
        pytest.fail("no HIP device visible: the GPU tests must run on an MI355X")


def _run_batch(nets, schedule=None, autotune=0):
    b = sa.Batch(nets, max_events_per_tick=1 << 22, fire_log_capacity=1 << 20)
    if schedule:
        b.union.set_schedule(**schedule)
    if autotune:
        b.union.autotune(autotune)
    b.run_commands([n.commands for n in nets])
    return b


def _snapshot(b):
    return dict(raster=b.union.raster(), neurons=b.union.neurons(), synapses=b.union.synapses(),
                ticks=b.union.tick_counts(), counts=b.net_counts())


def _assert_same(a, b):
    assert np.array_equal(a["raster"], b["raster"])
    for k in a["neurons"]:
        assert np.array_equal(a["neurons"][k], b["neurons"][k]), k
    for k in a["synapses"]:
        assert np.array_equal(a["synapses"][k], b["synapses"][k]), k
    assert np.array_equal(a["ticks"][0], b["ticks"][0]) and np.array_equal(a["ticks"][1], b["ticks"][1])


@pytest.mark.parametrize("maker", ["mixed_batch", "code3_batch"])
def test_batch_matches_oracle_per_network(maker):
    print(f"batch {maker}", flush=True)
    nets, _ = getattr(B, maker)()
    with _run_batch(nets) as b:
        raster, un, us = b.union.raster(), b.union.neurons(), b.union.synapses()
        dl, fd = b.net_counts()
        for k, net in enumerate(nets):
            o = oracle.Oracle(net)
            o.run_commands(net.commands)
            assert np.array_equal(b.net_raster(k, raster), o.fires()[1]), net.name
            dn, rn = b.net_neurons(k, un), o.neurons()
            for key in ("membrane", "current", "spike_count", "t_match", "t_learned", "last_fired"):
                assert np.array_equal(dn[key], rn[key]), (net.name, key)
            ds, rs = b.net_synapses(k, us), o.synapses()
            for key in ("weight", "pre_count", "upd_count", "last_pre", "last_post", "c_match", "c_learned"):
                assert np.array_equal(ds[key], rs[key].astype(ds[key].dtype)), (net.name, key)
            assert int(dl[k]) == o.delivered and int(fd[k]) == o.fired, net.name


def test_schedules_and_autotune_leave_results_unchanged():
    nets, _ = B.code3_batch()
    with _run_batch(nets) as b:
        ref = _snapshot(b)
        base = b.union.schedule()
    for sc in (dict(graph_ticks=16), dict(partition_shift=12, scatter_workgroups=64),
               dict(partition_shift=10, scatter_workgroups=1024, graph_ticks=7), dict(fold_threads=256),
               dict(partition_shift=11, fold_threads=256, graph_ticks=5)):
        print(f"schedule {sc}", flush=True)
        with _run_batch(nets, schedule=sc) as b:
            got = b.union.schedule()
            assert all(got[k] == v for k, v in sc.items())
            _assert_same(ref, _snapshot(b))
    print("autotune", flush=True)
    with _run_batch(nets, autotune=4) as b:
        log = b.union.tune_log()
        assert log.count("\n") >= 5, log
        _assert_same(ref, _snapshot(b))
    assert base["partition_shift"] in (10, 11)


def test_plain_step_graph_replay_equals_stream_launches():
    from snncuda_amd import workloads as W
    net = W.config0(2, threshold=1.1, ticks=400)
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    with sa.Network(net) as dev:
        dev.set_schedule(graph_ticks=32)
        dev.run_commands(net.commands)
        assert np.array_equal(dev.raster(), ref.fires()[1])
        assert np.array_equal(dev.synapses()["weight"], ref.synapses()["weight"])
        assert np.array_equal(dev.neurons()["membrane"], ref.neurons()["membrane"])


def test_batch_rejects_bad_input():
    from snncuda_amd import workloads as W
    net = W.feature_net(1)
    with pytest.raises(sa.SNNError):
        bad = W.feature_net(2)
        bad.tgt = bad.tgt.copy()
        bad.tgt[0] = bad.n_neurons  # out of range
        sa.Batch([net, bad])
    with sa.Batch([net, W.feature_net(2)]) as b:
        with pytest.raises(sa.SNNError):
            b.inject(2, 0, 0, 1.0)  # no network 2
        with pytest.raises(sa.SNNError):
            b.inject(0, net.n_neurons, 0, 1.0)  # neuron beyond network 0
        with pytest.raises(sa.SNNError):
            b.union.set_schedule(partition_shift=13)


def _gen_cases():
    from snncuda_amd import workloads as W
    nets = [W.config3_net(s, ticks=200, n_range=(500, 3000), fan_range=(50, 150)) for s in range(4)]
    # one code {0}; three codes (non-uniform code-pattern length: general code records)
    nets.append(W.ei_network(800, 40, 11, n_sources=50, source_fanout=30, p_source=0.02, ticks=200, threshold=2.0,
                             delay_max=9, compartments=(W.SOMA, W.BASAL, W.APICAL), plastic=False))
    nets.append(W.ei_network(600, 30, 12, n_sources=40, source_fanout=20, p_source=0.02, ticks=200, threshold=2.0,
                             delay_max=5, compartments=(W.BASAL,), n_codes=3, code_max=4, w_inh=2.0))
    nets, _ = B.equalize(nets)
    return nets


def test_device_generated_batch_equals_host_build():
    nets = _gen_cases()
    specs = [n.info["spec"] for n in nets]
    assert all(s is not None for s in specs)
    with _run_batch(nets) as host:
        want = _snapshot(host)
        assert want["counts"][1].sum() > 0
    with sa.Batch(specs=specs, max_events_per_tick=1 << 22, fire_log_capacity=1 << 20) as dev:
        assert np.array_equal(dev.neuron_off, host.neuron_off) and np.array_equal(dev.synapse_off, host.synapse_off)
        dev.run_commands([n.commands for n in nets])
        _assert_same(want, _snapshot(dev))


def test_device_generated_network_with_codes_equals_host_build():
    nets = _gen_cases()
    net = nets[5]
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    with sa.Network(spec=net.info["spec"]) as dev:
        dev.run_commands(net.commands)
        assert np.array_equal(dev.raster(), ref.fires()[1])
        ds, rs = dev.synapses(), ref.synapses()
        for key in ("weight", "pre_count", "c_match", "c_learned", "last_pre", "last_post"):
            assert np.array_equal(ds[key], rs[key].astype(ds[key].dtype)), key


def test_full_size_batch_equals_each_network_run_alone():
    # The 512-network workload at full size (neurons log-uniform in [500, 50,000], fan-out U{50..500},
    # 2 code offsets, STDP on; about 5.5e6 neurons and 1.5e9 synapses in all), generated on the
    # device and stepped together under the online autotuner: a network's raster, fire count, every
    # neuron field and its STDP weights (the union's synapse fields would be tens of GB on the host)
    # must equal that network run alone, for the smallest, the largest and a spread of the others.
    from snncuda_amd import workloads as W
    T = 150
    specs, cmds = zip(*[W.config3_spec(s, ticks=T) for s in range(512)])

    def weights(net):
        w = np.zeros(net.e, np.float32)
        sa._check(sa.lib().snn_net_read_synapses(net._h, sa._p(w), None, None, None, None, None, None))
        return w

    with sa.Batch(specs=list(specs), track_pre_count=False, max_events_per_tick=1 << 24,
                  fire_log_capacity=1 << 22, raster_capacity=1 << 22) as b:
        b.union.autotune(8)
        b.run_commands(list(cmds))
        raster = b.union.raster()
        _, fd = b.net_counts()
        sizes = np.array([sp["n"] for sp in specs])
        pick = sorted({int(sizes.argmin()), int(sizes.argmax()), 0, 1, 137, 300, 511})
        print(f"batch: {int(b.neuron_off[-1])} neurons, {int(b.synapse_off[-1])} synapses; "
              f"{int(fd.sum())} fires; schedule {b.union.schedule()}; checking networks {pick}", flush=True)
        assert int(fd.sum()) > 0
        nrn = b.union.neurons()
        w = weights(b.union)
        got = {k: (b.net_raster(k, raster), b.net_neurons(k, nrn), b.net_synapses(k, {"weight": w})["weight"],
                   int(fd[k])) for k in pick}
        del nrn, w
    for k in pick:
        with sa.Network(spec=specs[k], track_pre_count=False, max_events_per_tick=1 << 22,
                        fire_log_capacity=1 << 20, raster_capacity=1 << 20) as net:
            net.run_commands(cmds[k])
            r, nn, ww, f = got[k]
            assert net.fired == f, k
            assert np.array_equal(net.raster(), r), k
            alone = net.neurons()
            for key in nn:
                assert np.array_equal(nn[key], alone[key]), (k, key)
            assert np.array_equal(ww, weights(net)), k
