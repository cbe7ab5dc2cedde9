"""The device readouts (snn_graph_feedforward / snn_graph_recurrent, through the C ABI) against the
oracle (oracle/readout_oracle.c) on random graphs with multapses, delays 0..4, repeated and unknown
input ids, unknown and repeated readouts, every gate mode and the hidden top-k.

Tolerance: the device adds in the oracle's order (inputs / sources in list order, each source's
edges in connectome order, steps ascending); a cell may still differ by rounding where one wave
instruction adds to it twice, so scores are compared within 1e-5 of the sum of |terms| (feedforward)
or of 1 + |score| (recurrent), and spikes where the score is not within that bound of its threshold.
"""
import numpy as np
import pytest

import oracle
import snncuda_amd as sa
from snncuda_amd import merge_inputs
from tests import readout_nets as RN

pytestmark = pytest.mark.gpu
NONE = RN.NONE


def _spikes_agree(scores, spikes, ref_spikes, thr, tol):
    near = np.abs(scores - thr[None, :]) <= tol
    return np.array_equal(spikes[~near], ref_spikes[~near])


@pytest.mark.parametrize("seed,n,e,R,S", [(1, 300, 6000, 9, 37), (2, 5000, 40000, 3000, 11), (3, 64, 0, 4, 5)])
def test_feedforward_matches_the_oracle(seed, n, e, R, S):
    print(f"feedforward seed {seed} n {n} e {e} readouts {R} samples {S}", flush=True)
    net = RN.random_graph(seed, n=n, e=e, multapses=e // 20)
    ptr, x, w = merge_inputs(RN.random_samples(seed, n, n_samples=S, max_inputs=40), n)
    rng = np.random.default_rng(seed)
    readout = rng.integers(0, n, R).astype(np.uint32)
    readout[0] = NONE
    if R > 3:
        readout[3] = readout[1]
    with sa.Graph(net) as g:
        scores, spikes = g.feedforward(ptr, x, w, readout)
    ref, ref_spikes = oracle.readout_feedforward(net, ptr, x, w, readout)
    _, mag = RN.dense_feedforward(net, ptr, x, w, readout) if n <= 300 else (None, np.abs(ref) + 1.0)
    tol = 1e-5 * mag + 1e-30
    diff = np.abs(scores - ref)
    print(f"max |device - oracle| = {diff.max() if diff.size else 0.0:.3e}, exact cells "
          f"{int((scores == ref).sum())}/{scores.size}", flush=True)
    assert np.all(diff <= tol)
    thr = np.array([net.threshold[r] if r != NONE else 1.0 for r in readout], np.float32)
    assert _spikes_agree(scores, spikes, ref_spikes, thr, tol)


CASES = [
    dict(iterations=0, feedback_decay=0.5, top_k_feedback_readouts=1, top_k_hidden=0, use_scores_not_spikes=True),
    dict(iterations=1, feedback_decay=0.5, top_k_feedback_readouts=2, top_k_hidden=0, use_scores_not_spikes=True),
    dict(iterations=2, feedback_decay=0.75, top_k_feedback_readouts=0, top_k_hidden=0, use_scores_not_spikes=True),
    dict(iterations=2, feedback_decay=0.5, top_k_feedback_readouts=3, top_k_hidden=16, use_scores_not_spikes=False),
    dict(iterations=1, feedback_decay=1.0, top_k_feedback_readouts=-1, top_k_hidden=5, use_scores_not_spikes=False),
    dict(iterations=3, feedback_decay=0.25, top_k_feedback_readouts=4, top_k_hidden=1, use_scores_not_spikes=True),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_recurrent_matches_the_oracle(case):
    kw = CASES[case]
    print(f"recurrent case {case} {kw}", flush=True)
    n = 200
    net = RN.random_graph(20 + case, n=n, e=3000, max_delay=4, multapses=100)
    ptr, x, w = merge_inputs(RN.random_samples(case, n, n_samples=23, max_inputs=30), n)
    rng = np.random.default_rng(case)
    readout = rng.choice(n, 12, replace=False).astype(np.uint32)
    readout[5] = NONE
    hidden = np.array([h for h in range(n) if h not in set(readout.tolist())] + [NONE, int(readout[2])], np.uint32)
    if case % 2:
        hidden = hidden[: n // 3]
    with sa.Graph(net) as g:
        scores, spikes, T = g.recurrent(ptr, x, w, readout, hidden, **kw)
    ref, ref_spikes, Tr = oracle.readout_recurrent(net, ptr, x, w, readout, hidden, **kw)
    assert T == Tr
    tol = 1e-5 * (1.0 + np.abs(ref))
    diff = np.abs(scores - ref)
    print(f"T {T} max |device - oracle| = {diff.max():.3e}, exact cells {int((scores == ref).sum())}/{scores.size}",
          flush=True)
    assert np.all(diff <= tol)
    thr = np.array([net.threshold[r] if r != NONE else 1.0 for r in readout], np.float32)
    assert _spikes_agree(scores, spikes, ref_spikes, thr, tol)


def test_reference_known_answers_on_the_device():
    # tests/cuda_synapse_dendrite_tests.cpp:306-383
    with sa.Graph(RN.ltp_net()) as g:
        ptr, x, w = merge_inputs([[(0, 0.5)], [(0, 1.0)], [(0, 0.25), (0, 0.75)]], 2)
        scores, spikes = g.feedforward(ptr, x, w, [1])
    assert scores[:, 0].tolist() == [0.25, 0.5, 0.5] and spikes[:, 0].tolist() == [0, 1, 1]
    with sa.Graph(RN.ltp_net(feedback=True)) as g:
        ptr, x, w = merge_inputs([[(0, 1.0)]], 2)
        kw = dict(feedback_decay=1.0, top_k_feedback_readouts=1, top_k_hidden=0, use_scores_not_spikes=True)
        d, _, T0 = g.recurrent(ptr, x, w, [1], [0], iterations=0, **kw)
        s, sp, T1 = g.recurrent(ptr, x, w, [1], [0], iterations=1, **kw)
    assert d[0, 0] == 0.5 and T0 == 2
    assert s[0, 0] == 0.75 and sp[0, 0] == 1 and T1 == 4


def test_empty_batches_and_bad_indices():
    with sa.Graph(RN.ltp_net()) as g:
        ptr, x, w = merge_inputs([], 2)
        scores, spikes = g.feedforward(ptr, x, w, [1])
        assert scores.shape == (0, 1)
        ptr, x, w = merge_inputs([[(0, 1.0)]], 2)
        scores, spikes = g.feedforward(ptr, x, w, np.zeros(0, np.uint32))
        assert scores.shape == (1, 0)
        with pytest.raises(sa.SNNError):
            g.feedforward(ptr, x, w, [7])
        with pytest.raises(sa.SNNError):
            g.feedforward(ptr, np.array([5], np.uint32), w, [1])
