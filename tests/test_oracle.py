"""The oracle (oracle/snn_oracle.c) pinned against the reference: golden fixtures written by the
reference's own runtime, the reference tests' known answers, and closed forms."""
import math

import numpy as np
import pytest

import oracle
from tests import golden_util as G
from tests import kat_nets as K


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_oracle_matches_reference_fixture(name):
    net = G.CASES[name]()
    g = G.load(name)
    assert str(g["digest"]) == G.digest(net), "generator output changed since the fixture was written"
    res = oracle.run_oracle(net, patterns=False)
    n, s = res.neuron, res.synapse
    neurons = dict(membrane=n["membrane"], current=n["current"], tau=n["tau"], spike_count=n["spike_count"],
                   last_fired=n["last_fired"], t_match=n["t_match"], t_learned=n["t_learned"])
    syn = {k: s[k] for k in ("weight", "pre_count", "upd_count", "last_pre", "last_post", "c_match", "c_learned")}
    bad = G.mismatches(g, res.delivered_per_tick, res.fired_per_tick, None, neurons, syn)
    if not np.array_equal(g["raster"], res.raster):
        bad.append("raster")
    assert res.delivered == int(g["delivered"]) and res.fired == int(g["fired"])
    assert not bad, f"{name}: oracle differs from the reference in {bad}"


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_oracle_learned_pattern_contents_match_reference_fixture(name):
    # Every neuron's candidate offsets and learned patterns and every synapse's code offsets and
    # learned code patterns, word for word as the reference runtime left them (the fixture holds the
    # synapse words of the large cases as a SHA-256).
    import hashlib
    net = G.CASES[name]()
    g = G.load(name)
    res = oracle.run_oracle(net, patterns=True)
    assert np.array_equal(res.neuron_words, g["neuron_words"]), f"{name}: neuron pattern contents differ"
    if "synapse_words" in g:
        assert np.array_equal(res.synapse_words, g["synapse_words"]), f"{name}: synapse code patterns differ"
    else:
        assert hashlib.sha256(res.synapse_words.tobytes()).hexdigest() == str(g["synapse_words_sha256"]), \
            f"{name}: synapse code patterns differ"
    # the decoded counts agree with the count fields
    assert [len(p) for _, p in res.neuron_patterns] == list(res.neuron["t_learned"])
    assert [len(o) for o, _ in res.neuron_patterns] == list(res.neuron["t_n_offsets"])


def _run(net):
    o = oracle.Oracle(net)
    o.run_commands(net.commands)
    return o


def test_kat_three_stage_chain_fires_at_0_2_5():
    # propagation_tests.cpp:144-174
    o = _run(K.three_stage())
    t, n = o.fires()
    assert list(zip(t.tolist(), n.tolist())) == [(0, 0), (2, 1), (5, 2)]
    assert np.all(o.neurons()["membrane"] == 0.0)


def test_kat_subthreshold_persists():
    # propagation_tests.cpp:176-193
    o = _run(K.ALL["subthreshold"]())
    assert o.fired == 0 and o.neurons()["membrane"][0] == np.float32(0.5)


def test_kat_spike_code_02():
    # propagation_tests.cpp:92-142: delivered 6 (2 external + 4 coded), pre_spike_count 4
    o = _run(K.spike_code_02())
    s = o.synapses()
    assert o.delivered == 6 and s["pre_count"][0] == 4
    assert s["c_learned"][0] >= 1 and s["c_match"][0] >= 1


def test_kat_apical_and_inhibitory():
    # synapse_dendrite_tests.cpp:137-187
    o = _run(K.kat_apical())
    n = o.neurons()
    assert n["current"][0, 1] == 0.0 and abs(n["current"][2, 1] - 0.6) < 1e-4 and abs(n["membrane"][1] - 0.6) < 1e-4
    assert o.synapses()["pre_count"][0] == 1
    net = K.kat_inhibitory()
    o1 = oracle.Oracle(net)
    o1.run_commands(net.commands[:2])
    inh = o1.neurons()["current"][3, 1]
    o1.run_commands(net.commands[2:])
    dec = o1.neurons()["current"][3, 1]
    assert inh < -0.79 and inh < dec < 0.0
    # closed form: one GABA_A event of weight 0.8, then 10 ticks of decay with tau 10
    assert dec == np.float32(inh * np.float32(math.exp(-1.0)))


def test_kat_ltp_closed_form():
    # synapse_dendrite_tests.cpp:189-210 and the closed form 0.5 + 0.01*exp(-1/20)
    o = _run(K.kat_ltp())
    s = o.synapses()
    assert s["upd_count"][0] == 1
    assert abs(float(s["weight"][0]) - (0.5 + 0.01 * math.exp(-1 / 20))) < 1e-7
    assert s["ltp"][0] == 1  # post after pre: the one change is potentiation


def test_kat_ltd_direction():
    # synapse_dendrite_tests.cpp:212-236: pre at 5 after post at 0 -> depression by 0.012*exp(-5/20)
    o = _run(K.kat_ltd())
    s = o.synapses()
    w = float(s["weight"][0])
    assert w < 0.5 and abs(w - (0.5 - 0.012 * math.exp(-5 / 20))) < 1e-7
    assert s["upd_count"][0] == 1 and s["ltp"][0] == 0  # pre after post: depression only
    assert o.neurons()["spike_count"][1] == 1


def test_kat_temporal_learn_then_match():
    # synapse_dendrite_tests.cpp:238-259
    net = K.kat_temporal()
    o = oracle.Oracle(net)
    o.run_commands(net.commands[:3])
    n = o.neurons()
    assert n["t_learned"][0] == 1 and n["t_match"][0] == 0
    o.run_commands(net.commands[3:])
    assert o.neurons()["t_match"][0] == 1


def test_kat_coded_retrigger_delivers_8():
    # cuda_synapse_dendrite_tests.cpp:200-237: CPU delivered 8
    o = _run(K.coded_retrigger())
    assert o.delivered == 8 and o.synapses()["c_match"][0] >= 1


def test_kat_learning_experiments():
    # learning_temporal_experiment.cpp:125-194 and cuda_learning_consistency_experiment.cpp:289-297
    net = K.learning_temporal()
    o = _run(net)
    assert o.synapses()["weight"][0] >= 0.5
    assert o.neurons()["t_match"][net.info["probe"]] >= 1
    net = K.learning_consistency()
    o = _run(net)
    s = o.synapses()
    assert s["weight"][1] > 0.5 and o.neurons()["t_match"][net.info["probe"]] >= 1
    # cuda_learning_consistency_experiment.cpp:297: net causal LTP
    ltp = int(s["ltp"].sum())
    assert ltp > int(s["upd_count"].sum()) - ltp


def test_delivery_ticks_follow_fire_tick_plus_delay_plus_offset():
    # invariant: every synaptic delivery at tick t comes from a fire at t - max(1, delay + offset)
    net = G.CASES["feature1"]()
    o = _run(net)
    t, nrn = o.fires()
    d, _ = o.tick_counts()
    inj = np.zeros_like(d)
    pending = {}
    next_tick = 0
    for k, n_, tk, w in net.commands:
        if k == 0:
            pending.setdefault(max(tk, next_tick), 0)
            pending[max(tk, next_tick)] += 1
        else:
            next_tick = tk + 1
    for tk, c in pending.items():
        if tk < inj.shape[0]:
            inj[tk] += c
    exp = inj.astype(np.int64).copy()
    cp = net.code_ptr.astype(np.int64)
    for ft, fn in zip(t.tolist(), nrn.tolist()):
        for j in np.nonzero(net.src == fn)[0]:
            offs = net.codes[cp[j]:cp[j + 1]].tolist() or [0]
            for off in offs:
                lag = max(1, int(net.delay[j]) + int(off))
                if ft + lag < exp.shape[0]:
                    exp[ft + lag] += 1
    assert np.array_equal(exp, d.astype(np.int64))
