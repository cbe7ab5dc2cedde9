"""Small networks that fill one capacity of the device's per-neuron and per-synapse pattern state.

With a similarity threshold above 1 no cosine ever matches, so every observation from the second on
learns a pattern (SynapseDendriteProcessor.cpp:14-41, :318-365) and each network reaches its limit
within a few ticks."""
from __future__ import annotations

from tests.kat_nets import _net

NEVER = 1.5  # similarity threshold no cosine reaches


def candidate_runs(distinct_ticks=10):
    """Neuron 0 receives injects at `distinct_ticks` ticks of one 100-tick window: its candidate
    pattern holds one run per distinct offset (8 are stored)."""
    p = (100.0, 100, NEVER, 1)
    syn = [dict(src=1, tgt=0, w=0.0, plastic=False)]
    cmds = [(0, 0, t, 0.1) for t in range(distinct_ticks)] + [(1, 0, distinct_ticks + 2, 0.0)]
    return _net([p, p], syn, cmds, "capacity_candidate_runs")


def candidate_runs_wide(distinct_ticks=34):
    """candidate_runs for a network created with pattern_runs=32 (32 runs stored)."""
    return candidate_runs(distinct_ticks)


def learned_patterns(max_refs=5, observations=6):
    """Neuron 0 learns observations - 1 patterns in one window and keeps the last max_refs."""
    p = (100.0, 100, NEVER, max_refs)
    syn = [dict(src=1, tgt=0, w=0.0, plastic=False)]
    cmds = [(0, 0, t, 0.1) for t in range(observations)] + [(1, 0, observations + 2, 0.0)]
    return _net([p, p], syn, cmds, "capacity_learned_patterns")


def code_patterns(max_refs=5, fires=7):
    """Neuron 0 fires `fires` ticks in a row into synapse 0 -> 1, whose code pattern learns fires - 1
    patterns (the general per-synapse record holds 4)."""
    src = (1.0, 100, 0.93, 1)
    tgt = (100.0, 100, NEVER, max_refs)
    syn = [dict(src=0, tgt=1, w=0.01, plastic=False)]
    cmds = [(0, 0, t, 2.0) for t in range(fires)] + [(1, 0, fires + 3, 0.0)]
    return _net([src, tgt], syn, cmds, "capacity_code_patterns")


def many_learned(seed, n=300, ticks=2000, max_refs=(5, 12, 40, 500), rate=60):
    """Inject-driven neurons that learn far more than four temporal patterns: windows of at most 7
    ticks (so a candidate never holds more than 8 distinct offsets), several injects to one neuron in
    a tick, and max_reference_patterns up to the reference default 500. The only synapses come from
    a silent neuron, so no code pattern is observed. Up to `rate` injects per tick."""
    import numpy as np
    rng = np.random.default_rng(seed)
    params = [(float(rng.uniform(0.8, 1.5)), int(rng.choice([3, 5, 7])), float(rng.choice([0.9, 0.97, 0.9999])),
               int(rng.choice(np.asarray(max_refs)))) for _ in range(n)]
    params.append((1e9, 10, 0.9, 1))
    syn = [dict(src=n, tgt=int(t), w=0.1, plastic=False) for t in rng.integers(0, n, 50)]
    cmds = []
    for t in range(ticks):
        for x in rng.integers(0, n, int(rng.integers(0, rate))):
            cmds.append((0, int(x), t, float(rng.uniform(0.1, 1.2))))
        if t % 50 == 49:
            cmds.append((1, 0, t, 0.0))
    return _net(params, syn, cmds, f"many_learned{seed}")
