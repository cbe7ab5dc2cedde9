"""Networks and sample sets for the non-spiking readout tests, plus a float64 dense restatement of
the readouts (matrices per delay) used to pin the oracle.

``ltp_net`` is the reference's CUDA test network (tests/cuda_synapse_dendrite_tests.cpp:21-80: input
threshold 1.0, target threshold 0.4, input -> target weight 0.5 delay 1); ``feedback=True`` adds the
target -> input projection of :334-345 (weight 1.0, delay 1).
"""
from __future__ import annotations

import numpy as np

from snncuda_amd import workloads as W

NONE = 0xFFFFFFFF


def ltp_net(feedback: bool = False):
    thr = np.array([1.0, 0.4], np.float32)
    params = (thr, np.full(2, 500, np.uint64), np.full(2, 0.93, np.float32), np.full(2, 500, np.uint64))
    src, tgt, w, d = [0], [1], [0.5], [1]
    if feedback:
        src, tgt, w, d = [0, 1], [1, 0], [0.5, 1.0], [1, 1]
    m = len(src)
    return W.make_net(params, src, tgt, w, [2.0] * m, d, [W.BASAL] * m, [W.AMPA] * m, [1, 0][:m],
                      codes_per_syn=[[0]] * m, name="ltp")


def random_graph(seed: int, n: int, e: int, max_delay: int = 4, multapses: int = 0):
    """n neurons with thresholds U[0.2, 1.5), e synapses with weights U[-1, 1), delays U{0..max_delay}
    (0 is read as 1), plus `multapses` repeated (source, target) pairs."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, e)
    tgt = rng.integers(0, n, e)
    if multapses:
        pick = rng.integers(0, e, multapses)
        src = np.concatenate([src, src[pick]])
        tgt = np.concatenate([tgt, tgt[pick]])
    m = len(src)
    w = rng.uniform(-1.0, 1.0, m).astype(np.float32)
    d = rng.integers(0, max_delay + 1, m)
    thr = rng.uniform(0.2, 1.5, n).astype(np.float32)
    params = (thr, np.full(n, 32, np.uint64), np.full(n, 0.9, np.float32), np.ones(n, np.uint64))
    return W.make_net(params, src, tgt, w, np.full(m, 2.0, np.float32), d, np.full(m, W.BASAL), np.full(m, W.AMPA),
                      np.zeros(m, np.uint8), codes_per_syn=None, name=f"graph{seed}")


def random_samples(seed: int, n: int, n_samples: int, max_inputs: int):
    """Lists of (neuron, weight) with repeats and ids outside the network (skipped by the merge)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_samples):
        k = int(rng.integers(0, max_inputs + 1))
        xs = rng.integers(-2, n + 2, k)
        ws = rng.uniform(0.0, 1.0, k).astype(np.float32)
        if k > 2:
            xs[-1] = xs[0]  # a repeated neuron
        out.append(list(zip(xs.tolist(), ws.tolist())))
    return out


def dense_feedforward(net, ptr, neuron, weight, readout):
    """float64: X[s, x] @ M[x, r] with M the summed weights of the edges x -> readout r (a readout
    listed twice maps to its last listing). Also returns sum |terms| per cell for tolerances."""
    n = net.n_neurons
    S, R = len(ptr) - 1, len(readout)
    X = np.zeros((S, n))
    for s in range(S):
        X[s, neuron[ptr[s]:ptr[s + 1]]] = weight[ptr[s]:ptr[s + 1]]
    col = np.full(n, -1)
    for r, x in enumerate(readout):
        if x != NONE:
            col[x] = r
    M = np.zeros((n, R))
    A = np.zeros((n, R))
    for a, b, w in zip(net.src, net.tgt, net.weight):
        if col[b] >= 0:
            M[a, col[b]] += float(w)
            A[a, col[b]] += abs(float(w))
    return X @ M, np.abs(X) @ A


def dense_recurrent(net, ptr, neuron, weight, readout, hidden, *, iterations, feedback_decay, top_k_feedback_readouts,
                    top_k_hidden, use_scores_not_spikes):
    """float64 restatement with one dense hidden -> readout and readout -> hidden matrix per delay.
    Returns (scores [S, R], time_count)."""
    n, S, R, H = net.n_neurons, len(ptr) - 1, len(readout), len(hidden)
    d = np.maximum(1, net.delay.astype(np.int64))
    T = int(d.max(initial=1)) * (1 + 2 * iterations) + 1
    hcol = np.full(n, -1)
    for h, x in enumerate(hidden):
        if x != NONE:
            hcol[x] = h
    rcol = np.full(n, -1)
    for r, x in enumerate(readout):
        if x != NONE:
            rcol[x] = r
    hrow = np.full(n, -1)  # neuron -> the hidden entries that read its edges (all listings)
    HR = {}  # delay -> [H, R]
    RH = {}  # delay -> [R, H]
    for a, b, w, dl in zip(net.src, net.tgt, net.weight, d):
        for h in np.nonzero(np.asarray(hidden) == a)[0]:
            if rcol[b] >= 0:
                HR.setdefault(int(dl), np.zeros((H, R)))[h, rcol[b]] += float(w)
        for r in np.nonzero(np.asarray(readout) == a)[0]:
            if hcol[b] >= 0:
                RH.setdefault(int(dl), np.zeros((R, H)))[r, hcol[b]] += float(w)
    del hrow
    thr = np.array([net.threshold[x] if x != NONE else 1.0 for x in readout], np.float64)
    out = np.zeros((S, R))
    for s in range(S):
        hid = np.zeros((T, H))
        ro = np.zeros((T, R))
        for x, w in zip(neuron[ptr[s]:ptr[s + 1]], weight[ptr[s]:ptr[s + 1]]):
            if hcol[x] >= 0:
                hid[0, hcol[x]] += float(w)
        for t in range(T):
            if 0 < top_k_hidden < H:
                order = sorted(range(H), key=lambda h: (-hid[t, h], h))
                keep = [h for h in order[:top_k_hidden] if hid[t, h] > -np.finfo(np.float32).max]
                mask = np.zeros(H, bool)
                mask[keep] = True
                hid[t, ~mask] = 0.0
            for dl, M in HR.items():
                if t + dl < T:
                    ro[t + dl] += hid[t] @ M
            gate = np.zeros(R)
            if top_k_feedback_readouts <= 0:
                on = ro[t] >= thr
                gate[on] = ro[t][on] if use_scores_not_spikes else 1.0
            else:
                order = sorted(range(R), key=lambda r: (-ro[t, r], r))
                for r in order[:min(top_k_feedback_readouts, R)]:
                    if ro[t, r] > -np.finfo(np.float32).max:
                        gate[r] = ro[t, r] if use_scores_not_spikes else 1.0
            for dl, M in RH.items():
                if t + dl < T:
                    hid[t + dl] += (feedback_decay * gate) @ M
        out[s] = ro.sum(axis=0)
    return out, T
