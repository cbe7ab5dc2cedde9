"""The device step's logic, run on the CPU through the same item functions as the kernels
(tests/emu), against the reference fixtures — including the per-target sort (events arrive in a
shuffled order) and every code path the templates select."""
import numpy as np
import pytest

from tests import golden_util as G
from tests.emu import Emu


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_step_model_matches_reference_fixture(name):
    net = G.CASES[name]()
    g = G.load(name)
    m = Emu(net, ev_cap=1 << 22, fire_cap=1 << 20)
    m.run_commands(net.commands)
    assert m.flags == 0
    d, f = m.tick_counts()
    bad = G.mismatches(g, d, f, m.raster(), m.neurons(), m.synapses())
    assert not bad, f"{name}: step model differs from the reference in {bad}"


def test_capacity_overflow_is_an_error_not_a_drop():
    net = G.CASES["config0_burst"]()
    m = Emu(net, ev_cap=1 << 22, fire_cap=1 << 12)
    m.run_commands(net.commands)
    assert m.flags & 4  # fire-log capacity
    m = Emu(net, ev_cap=1 << 12, fire_cap=1 << 20)
    m.run_commands(net.commands)
    assert m.flags & 1  # events per tick


@pytest.mark.parametrize("small", [False, True])
@pytest.mark.parametrize("width", [2, 3, 4])
@pytest.mark.parametrize("narrow_g", [False, True])
def test_per_target_sort_orders_by_g_at_every_size(narrow_g, width, small):
    # 8-, 12- and 16-byte entries at sizes on both sides of the sorting networks (8, 16, 24) and the
    # shell sort past them; g is unique within a tick, the low 4 bits carry compartment and
    # receptor, and every other word must travel with its event. The largest g a tick can hold is
    # included (2^27 - 1 when the tick is narrow).
    from tests.emu import sort_by_g
    rng = np.random.default_rng(5 + width)
    gmax = (1 << 27) if narrow_g else (1 << 28) - 1
    for k in list(range(0, 41)) + [100, 1000]:
        for _ in range(20 if k <= 40 else 2):
            g = rng.choice(gmax, size=k, replace=False).astype(np.uint32)
            if k:
                g[0] = gmax - 1
            rows = rng.integers(0, 1 << 32, size=(k, width), dtype=np.uint64).astype(np.uint32)
            rows[:, 0] = (g << 4) | rng.integers(0, 16, size=k, dtype=np.uint32)
            out = sort_by_g(rows, narrow_g, small)
            assert np.array_equal(out, rows[np.argsort(g, kind="stable")]), k


@pytest.mark.parametrize("runs", [8, 32])
@pytest.mark.parametrize("sim", [0.9, 1.0, 0.5, 0.0, -0.5])
@pytest.mark.parametrize("max_refs,cap", [(1, 1), (1, 4), (2, 4), (4, 4), (5, 4), (3, 7), (9, 9), (40, 64),
                                          (500, 500), (500, 37)])
def test_per_tick_temporal_update_equals_single_observations(sim, max_refs, cap, runs):
    # The fold applies one neuron's k observations of a tick together (a window restart in closed
    # form, repeated matches in bulk); it must leave the header, candidate, learned ring and error
    # flags exactly as k calls of the one-observation update. Sparse and dense ticks, long gaps
    # (window restarts), thresholds at and below 0 (every cosine of a restart matches), and
    # max_reference_patterns below, at and above the ring's slots (above: the error path).
    from tests.emu import temporal_check
    rng = np.random.default_rng(max_refs * 1000 + cap * 10 + int(sim * 10))
    for trial in range(40):
        window = int(rng.choice([1, 3, 8, 32]))
        m = 300
        gaps = rng.choice([1, 1, 1, 2, 5, window + 1, window + 7], size=m)
        ticks = np.cumsum(gaps).astype(np.uint32)
        times = rng.choice([1, 1, 2, 3, 7, 12, 30], size=m).astype(np.uint32)
        assert temporal_check(ticks, times, window, sim, max_refs, cap, runs) == m, (trial, window)


def test_pattern_runs_above_the_maximum_is_refused():
    from tests import capacity_nets as C
    with pytest.raises(ValueError, match="pattern_runs"):
        Emu(C.candidate_runs(), pattern_runs=65)
    Emu(C.candidate_runs(), pattern_runs=64)


def test_network_too_large_for_one_device_is_refused_at_create():
    # The scatter keeps 8 B per target partition in 64 KB of LDS: at most 8,192 partitions of 4,096
    # neurons. Past that, create raises a clear error instead of a launch failure at the first tick.
    from tests.emu import partition_shift
    assert partition_shift(1000)[0] == 10
    assert partition_shift(100_000)[0] == 10
    assert partition_shift(300_000)[0] == 11
    assert partition_shift(1_010_000)[0] == 12
    assert partition_shift(8192 * 4096)[0] == 12
    r, msg = partition_shift(8192 * 4096 + 1)
    assert r == -1 and "8192" in msg and "target partitions" in msg


def test_a_clamp_bound_per_plastic_synapse_matches_the_oracle():
    """StdpRule::apply clamps each synapse to min(2, its own max_weight) (src/learning/STDP.cpp:37-40):
    thousands of distinct bounds, most of them below the initial weights so the clamp acts."""
    import oracle
    from snncuda_amd import workloads as W
    net = W.feature_net(21, n=2000, ticks=600, plastic_frac=0.9, threshold=0.6, max_weight_range=(0.05, 0.6))
    assert np.unique(net.max_weight[net.plastic != 0]).size > 1000
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    m = Emu(net, ev_cap=1 << 22, fire_cap=1 << 20)
    m.run_commands(net.commands)
    assert m.flags == 0
    rs, ms = ref.synapses(), m.synapses()
    assert int(rs["upd_count"].sum()) > 1000
    at_bound = (net.plastic != 0) & (rs["upd_count"] > 0) & (rs["weight"] == np.minimum(net.max_weight, 2.0))
    assert np.unique(net.max_weight[at_bound]).size > 200
    assert np.array_equal(ms["weight"], rs["weight"])
    assert np.array_equal(m.raster(), ref.fires()[1])


@pytest.mark.parametrize("case,limit,under,flag", [
    ("candidate_runs", {}, {"distinct_ticks": 8}, 8),      # 8 stored runs of a candidate
    ("candidate_runs_wide", {"pattern_runs": 32}, {"distinct_ticks": 32}, 8),  # 32 with pattern_runs=32
    ("learned_patterns", {"ref_cap": 4}, {"max_refs": 4}, 8),  # the neuron's learned-pattern slots
    ("code_patterns", {}, {"fires": 5}, 16),                # 4 patterns of a synapse's code record
])
def test_pattern_capacity_overflow_is_an_error_and_just_under_it_equals_the_oracle(case, limit, under, flag):
    import oracle
    from tests import capacity_nets as C
    net = getattr(C, case)()
    m = Emu(net, **limit)
    m.run_commands(net.commands)
    assert m.flags & flag, m.flags
    net = getattr(C, case)(**under)
    m = Emu(net, **limit)
    m.run_commands(net.commands)
    assert m.flags == 0
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    for k in ("t_match", "t_learned", "membrane"):
        assert np.array_equal(m.neurons()[k], ref.neurons()[k]), k
    assert np.array_equal(m.synapses()["c_learned"], ref.synapses()["c_learned"])


@pytest.mark.parametrize("seed", [0, 1])
def test_neurons_holding_many_learned_patterns_match_the_oracle(seed):
    # learn_pattern keeps up to max_reference_patterns copies (default 500, NetworkIR.h:18); the ring
    # holds as many slots as the largest of them
    import oracle
    from tests import capacity_nets as C
    net = C.many_learned(seed)
    ref = oracle.Oracle(net)
    ref.run_commands(net.commands)
    rn = ref.neurons()
    assert rn["t_learned"].max() > 16
    m = Emu(net)
    m.run_commands(net.commands)
    assert m.flags == 0
    mn = m.neurons()
    for k in ("membrane", "current", "spike_count", "t_match", "t_learned"):
        assert np.array_equal(mn[k], rn[k]), k
    assert np.array_equal(m.raster(), ref.fires()[1])
