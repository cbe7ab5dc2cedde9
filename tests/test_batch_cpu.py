"""The batched form's principle on the CPU: the disjoint union of independent networks, stepped as
one network by the CPU model of the kernels (tests/emu), equals every network run alone in the
oracle; and the command merge keeps each network's inject poll ticks."""
import numpy as np

import oracle
from snncuda_amd import workloads as W
from tests import batch_nets as B
from tests.emu import Emu


def _oracle(net):
    o = oracle.Oracle(net)
    o.run_commands(net.commands)
    return o


def test_merge_commands_keeps_poll_ticks():
    a = [(0, 0, 0, 1.0), (1, 0, 5, 0.0), (0, 1, 3, 1.0), (1, 0, 9, 0.0)]
    b = [(0, 0, 2, 1.0), (1, 0, 7, 0.0), (0, 0, 8, 1.0), (1, 0, 9, 0.0)]
    cmds, owner = W.merge_commands([a, b], [0, 10])
    # a's second inject is issued after tick 5 and before tick 7; b's after tick 7
    assert cmds == [(0, 0, 0, 1.0), (0, 10, 2, 1.0), (1, 0, 5, 0.0), (0, 1, 3, 1.0), (1, 0, 7, 0.0),
                    (0, 10, 8, 1.0), (1, 0, 9, 0.0)]
    assert owner == [0, 1, 0, 1]


def test_union_model_equals_each_network_alone():
    nets, _ = B.mixed_batch()
    u = W.union(nets)
    m = Emu(u, ev_cap=1 << 22, fire_cap=1 << 20)
    m.run_commands(u.commands)
    assert m.flags == 0
    noff, eoff = u.info["neuron_off"], u.info["synapse_off"]
    raster, un, us = m.raster(), m.neurons(), m.synapses()
    for k, net in enumerate(nets):
        o = _oracle(net)
        lo, hi = int(noff[k]), int(noff[k + 1])
        mine = raster[(raster >= lo) & (raster < hi)] - lo
        assert np.array_equal(mine, o.fires()[1]), net.name
        rn, rs = o.neurons(), o.synapses()
        for key in ("membrane", "spike_count", "t_match", "t_learned"):
            assert np.array_equal(un[key][lo:hi], rn[key]), (net.name, key)
        assert np.array_equal(un["current"][:, lo:hi], rn["current"]), net.name
        es = slice(int(eoff[k]), int(eoff[k + 1]))
        for key in ("weight", "pre_count", "c_match", "c_learned"):
            assert np.array_equal(us[key][es], rs[key].astype(us[key].dtype)), (net.name, key)


def test_plasticity_modes_of_the_ei_family():
    # plastic="exc": every synapse is plastic except those of inhibitory network neurons
    n = 300
    net = W.ei_network(n, 20, 1, n_sources=10, source_fanout=5, p_source=0.01, ticks=10, threshold=2.0,
                       delay_max=5, compartments=(W.BASAL,), plastic="exc", w_inh=5.0)
    n_exc = net.info["spec"]["n_exc"]
    inh = (net.src >= n_exc) & (net.src < n)
    assert inh.any() and (~inh).any()
    assert np.array_equal(net.plastic, (~inh).astype(np.uint8))
    assert np.array_equal(net.receptor[net.src < n] == W.GABA_A, inh[net.src < n])
    assert net.info["spec"]["plastic"] == 2
    for mode, flag in ((False, 0), (True, 1)):
        m = W.ei_network(n, 20, 1, n_sources=10, source_fanout=5, p_source=0.01, ticks=10, threshold=2.0,
                         delay_max=5, compartments=(W.BASAL,), plastic=mode)
        assert (m.plastic == flag).all() and m.info["spec"]["plastic"] == flag
