"""The readout oracle (oracle/readout_oracle.c) pinned on the CPU: the reference's own known answers
(tests/cuda_synapse_dendrite_tests.cpp:306-383), float64 dense restatements (one matrix per delay) on
random graphs, and hand cases for the top-k and threshold gates."""
import numpy as np
import pytest

import oracle
from snncuda_amd import merge_inputs
from tests import readout_nets as RN

NONE = RN.NONE


def test_merge_keeps_first_appearance_order_and_sums_repeats():
    ptr, x, w = merge_inputs([[(3, 0.25), (1, 1.0), (3, 0.5), (9, 2.0)], [], [(-1, 1.0), (0, 0.125)]], 5)
    assert ptr.tolist() == [0, 2, 2, 3]
    assert x.tolist() == [3, 1, 0]
    assert w.tolist() == [0.75, 1.0, 0.125]


def test_oracle_feedforward_matches_the_reference_known_answers():
    # cuda_synapse_dendrite_tests.cpp:306-331: input -> target weight 0.5, target threshold 0.4
    net = RN.ltp_net()
    ptr, x, w = merge_inputs([[(0, 0.5)], [(0, 1.0)], [(0, 0.25), (0, 0.75)]], 2)
    scores, spikes = oracle.readout_feedforward(net, ptr, x, w, [1])
    assert np.allclose(scores[:, 0], [0.25, 0.5, 0.5], atol=1e-4)
    assert spikes[:, 0].tolist() == [0, 1, 1]
    assert (ptr[3] - ptr[2]) == 1  # the two inputs of sample 2 merge into one delivered input


def test_oracle_recurrent_matches_the_reference_known_answers():
    # :333-383: target -> input feedback weight 1.0 delay 1; decay 1, top-1 readout, scores as gates
    net = RN.ltp_net(feedback=True)
    ptr, x, w = merge_inputs([[(0, 1.0)]], 2)
    kw = dict(feedback_decay=1.0, top_k_feedback_readouts=1, top_k_hidden=0, use_scores_not_spikes=True)
    direct, _, T0 = oracle.readout_recurrent(net, ptr, x, w, [1], [0], iterations=0, **kw)
    assert abs(direct[0, 0] - 0.5) < 1e-4 and T0 == 2
    scores, spikes, T1 = oracle.readout_recurrent(net, ptr, x, w, [1], [0], iterations=1, **kw)
    assert abs(scores[0, 0] - 0.75) < 1e-4 and spikes[0, 0] == 1 and T1 == 4


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_feedforward_equals_the_dense_product(seed):
    net = RN.random_graph(seed, n=60, e=900, multapses=40)
    ptr, x, w = merge_inputs(RN.random_samples(seed, 60, n_samples=9, max_inputs=25), 60)
    readout = np.array([5, 17, NONE, 5, 33, 59, 0], np.uint32)
    scores, spikes = oracle.readout_feedforward(net, ptr, x, w, readout)
    ref, mag = RN.dense_feedforward(net, ptr, x, w, readout)
    assert np.all(np.abs(scores - ref) <= 1e-5 * mag + 1e-30)
    thr = np.array([net.threshold[r] if r != NONE else 1.0 for r in readout])
    assert np.array_equal(spikes, (scores >= thr).astype(np.uint64))
    assert np.all(scores[:, 2] == 0) and np.all(scores[:, 0] == 0)  # unknown readout; 5's first listing


CASES = [
    dict(iterations=0, feedback_decay=0.5, top_k_feedback_readouts=1, top_k_hidden=0, use_scores_not_spikes=True),
    dict(iterations=1, feedback_decay=0.5, top_k_feedback_readouts=2, top_k_hidden=0, use_scores_not_spikes=True),
    dict(iterations=2, feedback_decay=0.75, top_k_feedback_readouts=0, top_k_hidden=0, use_scores_not_spikes=True),
    dict(iterations=2, feedback_decay=0.5, top_k_feedback_readouts=3, top_k_hidden=6, use_scores_not_spikes=False),
    dict(iterations=1, feedback_decay=1.0, top_k_feedback_readouts=-1, top_k_hidden=4, use_scores_not_spikes=False),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_recurrent_equals_the_dense_timeline(case):
    kw = CASES[case]
    net = RN.random_graph(10 + case, n=40, e=500, max_delay=3, multapses=20)
    ptr, x, w = merge_inputs(RN.random_samples(case, 40, n_samples=6, max_inputs=12), 40)
    readout = np.array([3, 7, 11, NONE, 19], np.uint32)
    hidden = np.array([h for h in range(40) if h not in (3, 7, 11, 19)] + [NONE, 2], np.uint32)
    scores, spikes, T = oracle.readout_recurrent(net, ptr, x, w, readout, hidden, **kw)
    ref, Tr = RN.dense_recurrent(net, ptr, x, w, readout, hidden, **kw)
    assert T == Tr
    scale = 1e-5 * (1.0 + np.abs(ref))
    assert np.all(np.abs(scores - ref) <= scale), np.abs(scores - ref).max()


def test_dense_check_catches_a_planted_delay_mistake():
    # a timeline one step late (every delay + 1) must not pass the comparison above
    kw = CASES[1]
    net = RN.random_graph(11, n=40, e=500, max_delay=3)
    ptr, x, w = merge_inputs(RN.random_samples(1, 40, n_samples=6, max_inputs=12), 40)
    readout = np.array([3, 7, 11], np.uint32)
    hidden = np.array([h for h in range(40) if h not in (3, 7, 11)], np.uint32)
    scores, _, _ = oracle.readout_recurrent(net, ptr, x, w, readout, hidden, **kw)
    late = RN.random_graph(11, n=40, e=500, max_delay=3)
    late.delay = np.maximum(1, late.delay) + 1
    wrong, _ = RN.dense_recurrent(late, ptr, x, w, readout, hidden, **kw)
    assert np.abs(scores - wrong).max() > 1e-3


def test_top_k_ties_go_to_the_lower_index_and_threshold_gates_every_readout():
    # hidden 0 -> readouts 2, 3 (0.5 each), hidden 1 -> readouts 2, 3 (0.25 each); readout 2 feeds
    # hidden 1 (1.0), readout 3 feeds hidden 0 (2.0); every delay 1, so T = 1 + 2 * iterations + 1
    thr = np.array([1.0, 1.0, 0.5, 0.5], np.float32)
    from snncuda_amd import workloads as W
    params = (thr, np.full(4, 32, np.uint64), np.full(4, 0.9, np.float32), np.ones(4, np.uint64))
    src, tgt, wt = [0, 0, 1, 1, 2, 3], [2, 3, 2, 3, 1, 0], [0.5, 0.5, 0.25, 0.25, 1.0, 2.0]
    net = W.make_net(params, src, tgt, wt, [2.0] * 6, [1] * 6, [W.BASAL] * 6, [W.AMPA] * 6, [0] * 6,
                     codes_per_syn=None, name="ties")
    ptr, x, w = merge_inputs([[(0, 1.0), (1, 1.0)]], 4)
    # top-1 readout with gate 1.0: every step picks readout 2 (rows 0 and 1 tie; row 2 ties again), so
    # hidden 1 gets 1.0 at steps 1, 2, 3: readouts 0.75 + 0.25 + 0.25 (readout 3 would add 1.0 + 1.0)
    s1, _, T = oracle.readout_recurrent(net, ptr, x, w, [2, 3], [0, 1], iterations=1, feedback_decay=1.0,
                                        top_k_feedback_readouts=1, top_k_hidden=0, use_scores_not_spikes=False)
    assert T == 4
    assert s1[0].tolist() == [1.25, 1.25]
    # top-1 hidden at step 0 keeps hidden 0 (1.0 / 1.0 tie to the lower index): readouts 0.5 each
    s2, _, _ = oracle.readout_recurrent(net, ptr, x, w, [2, 3], [0, 1], iterations=0, feedback_decay=1.0,
                                        top_k_feedback_readouts=1, top_k_hidden=1, use_scores_not_spikes=True)
    assert s2[0].tolist() == [0.5, 0.5]
    # threshold gates: both readouts reach 0.5 at step 1 and both feed back (hidden 0: 2.0, hidden 1:
    # 1.0 at step 2), giving 2 * 0.5 + 1 * 0.25 more at step 3
    s3, _, _ = oracle.readout_recurrent(net, ptr, x, w, [2, 3], [0, 1], iterations=1, feedback_decay=1.0,
                                        top_k_feedback_readouts=0, top_k_hidden=0, use_scores_not_spikes=False)
    assert s3[0].tolist() == [2.0, 2.0]
