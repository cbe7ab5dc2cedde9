"""The reference test suite's networks and drive sequences, restated in connectome order.

Each builder returns a ``workloads.Net`` whose neuron / synapse order is what ``ConnectomeBuilder``
produces for the reference test's NetworkIR (populations in declaration order, projections in
declaration order) and whose commands replay the test's inject / run_until calls. ``EXPECT`` holds
the assertions the reference test makes, by name.
"""
from __future__ import annotations

import numpy as np

from snncuda_amd import workloads as W

BASAL, APICAL, SOMA, INH = W.BASAL, W.APICAL, W.SOMA, W.INHIBITORY
AMPA, GABA_A = W.AMPA, W.GABA_A


def _net(params, synapses, commands, name):
    """params: list of (threshold, window, sim_thr, max_refs) per neuron.
    synapses: list of dicts with src, tgt, w, mw, delay, comp, rec, plastic, codes."""
    thr = np.array([p[0] for p in params], np.float32)
    win = np.array([p[1] for p in params], np.uint64)
    sim = np.array([p[2] for p in params], np.float32)
    mref = np.array([p[3] for p in params], np.uint64)
    s = synapses
    return W.make_net(
        (thr, win, sim, mref), [x["src"] for x in s], [x["tgt"] for x in s], [x["w"] for x in s],
        [x.get("mw", 2.0) for x in s], [x.get("delay", 1) for x in s], [x.get("comp", BASAL) for x in s],
        [x.get("rec", AMPA) for x in s], [1 if x.get("plastic", True) else 0 for x in s],
        codes_per_syn=[list(x.get("codes", [0])) for x in s], commands=commands, name=name)


DEFAULT = (1.0, 500, 0.93, 500)  # NeuronParamsIR defaults (reference NetworkIR.h:14-19)


def three_stage(inject_weight=1.0, runs=(0, 2, 5)):
    """propagation_tests.cpp:16-90 / :144-174 (and :176-193 with weight 0.5)."""
    p = (1.0, 10, 0.9, 4)
    syn = [dict(src=0, tgt=1, w=1.0, delay=2), dict(src=1, tgt=2, w=1.0, delay=3)]
    cmds = [(0, 0, 0, inject_weight)] + [(1, 0, t, 0.0) for t in runs]
    return _net([p, p, p], syn, cmds, "three_stage")


def spike_code_02():
    """propagation_tests.cpp:92-142: source -> sink with spike_code_offsets {0, 2}."""
    src = (1.0, 10, 0.9, 4)
    sink = (100.0, 8, 0.99, 4)
    syn = [dict(src=0, tgt=1, w=0.1, delay=1, codes=[0, 2])]
    cmds = [(0, 0, 0, 1.0), (0, 0, 20, 1.0), (1, 0, 25, 0.0)]
    return _net([src, sink], syn, cmds, "spike_code_02")


def two_neuron(target_threshold, weight, comp=BASAL, rec=AMPA, plastic=False, commands=()):
    """synapse_dendrite_tests.cpp:23-90 (input: defaults with threshold 1.0; target: window 10,
    similarity 0.99, max 4)."""
    inp = (1.0, 500, 0.93, 500)
    tgt = (target_threshold, 10, 0.99, 4)
    syn = [dict(src=0, tgt=1, w=weight, delay=1, comp=comp, rec=rec, plastic=plastic)]
    return _net([inp, tgt], syn, list(commands), "two_neuron")


def kat_apical():
    """synapse_dendrite_tests.cpp:137-160"""
    return two_neuron(1.0, 0.6, APICAL, commands=[(0, 0, 0, 1.0), (1, 0, 1, 0.0)])


def kat_inhibitory():
    """synapse_dendrite_tests.cpp:162-187"""
    return two_neuron(0.2, 0.8, INH, GABA_A, commands=[(0, 0, 0, 1.0), (1, 0, 1, 0.0), (0, 1, 11, 0.0), (1, 0, 11, 0.0)])


def kat_ltp():
    """synapse_dendrite_tests.cpp:189-210"""
    return two_neuron(0.4, 0.5, plastic=True, commands=[(0, 0, 0, 1.0), (1, 0, 1, 0.0)])


def kat_ltd():
    """synapse_dendrite_tests.cpp:212-236"""
    return two_neuron(1.0, 0.5, plastic=True, commands=[(0, 1, 0, 1.0), (1, 0, 0, 0.0), (0, 0, 5, 1.0), (1, 0, 6, 0.0)])


def kat_temporal():
    """synapse_dendrite_tests.cpp:238-259: single neuron, window 10, similarity 0.99, max 2."""
    p = (10.0, 10, 0.99, 2)
    cmds = [(0, 0, 10, 0.1), (0, 0, 12, 0.1), (1, 0, 12, 0.0), (0, 0, 30, 0.1), (0, 0, 32, 0.1), (1, 0, 32, 0.0)]
    return _net([p], [], cmds, "kat_temporal")


def coded_retrigger():
    """cuda_synapse_dendrite_tests.cpp:200-237: codes {0,2,3} plus a delay-20 self retrigger."""
    inp = (1.0, 500, 0.93, 500)
    tgt = (100.0, 8, 0.99, 4)
    syn = [dict(src=0, tgt=1, w=0.1, delay=1, plastic=True, codes=[0, 2, 3]),
           dict(src=0, tgt=0, w=1.0, delay=20, plastic=False)]
    return _net([inp, tgt], syn, [(0, 0, 0, 1.0), (1, 0, 25, 0.0)], "coded_retrigger")


def learning_consistency():
    """cuda_learning_consistency_experiment.cpp:28-182 with its CPU drive (:221-235): every
    initially-active neuron injected at tick 0 with weight 1.0, run to tick 3543."""
    epochs, spacing = 35, 100
    first_t, second_t = 3520, 3540
    inp = (1.0, 500, 0.93, 500)
    det = (0.5, 500, 0.93, 500)
    rd = (0.5, 500, 0.93, 500)
    tmp = (10.0, 8, 0.99, 1)
    # populations: cue, detector, readout, temporal_probe, temporal_source_0..3, teacher_0..34
    params = [inp, det, rd, tmp] + [inp] * 4 + [inp] * epochs
    cue, detector, readout, probe = 0, 1, 2, 3
    tsrc = [4 + i for i in range(4)]
    teach = [8 + i for i in range(epochs)]
    syn = [dict(src=cue, tgt=cue, w=1.0, mw=1.0, delay=spacing, plastic=False),
           dict(src=cue, tgt=detector, w=0.25, mw=1.5, delay=1, plastic=True),
           dict(src=detector, tgt=readout, w=0.75, mw=1.0, delay=1, plastic=False)]
    delays = [first_t, first_t + 2, second_t, second_t + 2]
    for i in range(4):
        syn.append(dict(src=tsrc[i], tgt=probe, w=0.1, mw=0.1, delay=delays[i], comp=SOMA, plastic=False))
    for ep in range(epochs):
        syn.append(dict(src=teach[ep], tgt=detector, w=1.0, mw=1.0, delay=2 + ep * spacing, comp=SOMA, plastic=False))
    active = [cue] + teach + tsrc
    cmds = [(0, a, 0, 1.0) for a in active] + [(1, 0, second_t + 3, 0.0)]
    net = _net(params, syn, cmds, "learning_consistency")
    net.info.update(cue=cue, detector=detector, readout=readout, probe=probe, plastic_index=1)
    return net


def learning_temporal():
    """learning_temporal_experiment.cpp:19-128 driven as :130-194: 35 epochs of cue + teacher
    injects with run_until between them, then a cue-only test and temporal probe injects."""
    inp = (1.0, 500, 0.93, 500)
    det = (0.5, 12, 0.99, 8)
    rd = (0.5, 500, 0.93, 500)
    tmp = (10.0, 8, 0.99, 4)
    cue, detector, readout, probe = 0, 1, 2, 3
    syn = [dict(src=cue, tgt=detector, w=0.25, mw=1.5, delay=1, plastic=True),
           dict(src=detector, tgt=readout, w=0.75, mw=1.0, delay=1, plastic=False)]
    cmds = [(0, cue, 0, 1.0), (1, 0, 3, 0.0)]
    for ep in range(35):
        base = 20 + 100 * ep
        cmds += [(0, cue, base, 1.0), (0, detector, base + 2, 1.0), (1, 0, base + 4, 0.0)]
    test = 20 + 100 * 35 + 20
    cmds += [(0, cue, test, 1.0), (1, 0, test + 3, 0.0)]
    cmds += [(0, probe, test + 20, 0.1), (0, probe, test + 22, 0.1), (1, 0, test + 22, 0.0)]
    cmds += [(0, probe, test + 40, 0.1), (0, probe, test + 42, 0.1), (1, 0, test + 42, 0.0)]
    net = _net([inp, det, rd, tmp], syn, cmds, "learning_temporal")
    net.info.update(cue=cue, detector=detector, readout=readout, probe=probe, test_tick=test)
    return net


ALL = {
    "three_stage": three_stage,
    "subthreshold": lambda: three_stage(0.5, runs=(10,)),
    "spike_code_02": spike_code_02,
    "kat_apical": kat_apical,
    "kat_inhibitory": kat_inhibitory,
    "kat_ltp": kat_ltp,
    "kat_ltd": kat_ltd,
    "kat_temporal": kat_temporal,
    "coded_retrigger": coded_retrigger,
    "learning_consistency": learning_consistency,
    "learning_temporal": learning_temporal,
}
