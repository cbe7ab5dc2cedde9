"""Synthetic networks for the per-tick step, generated from seeds.

Every random draw comes from a counter-based hash (``mix64`` of seed, stream and index), so the
same network can be rebuilt anywhere (host numpy here, or a device kernel) without sharing a
generator state. Networks are returned in *connectome order*: neuron ``i`` is NeuronId ``i+1``,
synapse ``j`` is SynapseId ``j+1``, and each source's out-edges are contiguous in the order they
were drawn — the order ``ConnectomeBuilder`` gives an explicit, source-major projection list
(reference ``src/declarative/Connectome.cpp:141-182``).

Pinned free parameters (the workload descriptions leave them open):

* Inhibitory synapses use the GABA_A receptor with a *positive* weight magnitude ``|U[-1,0]|``, so
  the receptor's sign (reference ``src/runtime/SynapseDendriteProcessor.cpp:259-269``) makes their
  current negative and STDP's ``[0, min(2, max_w)]`` clamp (``src/learning/STDP.cpp:37-40``) keeps
  them meaningful. A negative weight would turn GABA_A into an excitatory current.
* Input sources are neurons of their own (ids after the network neurons) with threshold 1.0 that
  are driven by ``inject(source, tick, 1.0)`` — each inject fires the source once.
* Source synapses: AMPA, Basal, weight U[0, 0.5), delays as the network's.
* Neuron temporal parameters: window 32 ticks, similarity threshold 0.9, ``max_refs`` 1.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

SOMA, BASAL, APICAL, INHIBITORY = 0, 1, 2, 3
AMPA, NMDA, GABA_A, GABA_B = 0, 1, 2, 3

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def mix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser (Steele, Lea, Flood 2014), vectorised over uint64."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
    return x ^ (x >> np.uint64(31))


def hash3(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """64-bit hash of (seed, stream, idx); idx is any integer array."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        k = mix64(np.uint64(seed) * _GOLD + np.uint64(stream)) + idx * _GOLD
    return mix64(k)


def unif24(h: np.ndarray) -> np.ndarray:
    """float32 uniform in [0, 1) with 24 random bits (exactly representable)."""
    return ((h >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))).astype(np.float32)


def below(h: np.ndarray, n) -> np.ndarray:
    """integer uniform in [0, n) by multiply-shift of the high 32 bits (n < 2**32)."""
    hi = (h >> np.uint64(32)).astype(np.uint64)
    with np.errstate(over="ignore"):
        return ((hi * np.asarray(n, dtype=np.uint64)) >> np.uint64(32)).astype(np.int64)


@dataclass
class Net:
    """A network in connectome order plus the commands that drive it."""

    threshold: np.ndarray  # f32 [N]
    window: np.ndarray  # u64 [N]
    sim_thr: np.ndarray  # f32 [N]
    max_refs: np.ndarray  # u64 [N]
    src: np.ndarray  # u32 [E]
    tgt: np.ndarray  # u32 [E]
    weight: np.ndarray  # f32 [E]
    max_weight: np.ndarray  # f32 [E]
    delay: np.ndarray  # u32 [E]
    comp: np.ndarray  # u8 [E]
    receptor: np.ndarray  # u8 [E]
    plastic: np.ndarray  # u8 [E]
    code_ptr: np.ndarray  # u64 [E+1]
    codes: np.ndarray  # u32 [code_ptr[-1]]
    # commands in call order: (kind 0 inject / 1 run_until, neuron, tick, weight)
    commands: List[Tuple[int, int, int, float]] = field(default_factory=list)
    name: str = ""
    info: dict = field(default_factory=dict)

    @property
    def n_neurons(self) -> int:
        return int(self.threshold.shape[0])

    @property
    def n_synapses(self) -> int:
        return int(self.src.shape[0])

    def injects(self) -> np.ndarray:
        return np.array([(n, t, w) for k, n, t, w in self.commands if k == 0], dtype=np.float64).reshape(-1, 3)

    def last_tick(self) -> int:
        ticks = [t for k, _, t, _ in self.commands if k == 1]
        return max(ticks) if ticks else -1


def neurons(n: int, threshold=1.0, window=32, sim_thr=0.9, max_refs=1):
    return (
        np.full(n, threshold, dtype=np.float32),
        np.full(n, window, dtype=np.uint64),
        np.full(n, sim_thr, dtype=np.float32),
        np.full(n, max_refs, dtype=np.uint64),
    )


def make_net(neuron_arrays, src, tgt, weight, max_weight, delay, comp, receptor, plastic,
             codes_per_syn: Optional[List[List[int]]] = None, code_ptr=None, codes=None,
             commands=None, name="", info=None) -> Net:
    e = len(src)
    if code_ptr is None:
        if codes_per_syn is None:
            code_ptr = np.arange(e + 1, dtype=np.uint64)
            codes = np.zeros(e, dtype=np.uint32)
        else:
            lens = np.array([len(c) for c in codes_per_syn], dtype=np.uint64)
            code_ptr = np.zeros(e + 1, dtype=np.uint64)
            code_ptr[1:] = np.cumsum(lens)
            codes = np.array([x for c in codes_per_syn for x in c], dtype=np.uint32)
    thr, win, sim, mref = neuron_arrays
    return Net(
        threshold=np.asarray(thr, np.float32), window=np.asarray(win, np.uint64),
        sim_thr=np.asarray(sim, np.float32), max_refs=np.asarray(mref, np.uint64),
        src=np.asarray(src, np.uint32), tgt=np.asarray(tgt, np.uint32),
        weight=np.asarray(weight, np.float32), max_weight=np.asarray(max_weight, np.float32),
        delay=np.asarray(delay, np.uint32), comp=np.asarray(comp, np.uint8),
        receptor=np.asarray(receptor, np.uint8), plastic=np.asarray(plastic, np.uint8),
        code_ptr=np.asarray(code_ptr, np.uint64), codes=np.asarray(codes, np.uint32),
        commands=list(commands or []), name=name, info=dict(info or {}),
    )


# --------------------------------------------------------------------------- random E/I networks

def _fanout_edges(seed: int, stream: int, sources: np.ndarray, fanout: np.ndarray, n_targets: int,
                  target_base: int, exclude_self: bool):
    """Targets for each (source, j): uniform over [target_base, target_base+n_targets), without
    self-loops when exclude_self. Returns (src, tgt, edge_counter) in source-major order."""
    fanout = np.asarray(fanout, dtype=np.int64)
    src = np.repeat(sources.astype(np.int64), fanout)
    starts = np.zeros(len(fanout), dtype=np.int64)
    np.cumsum(fanout[:-1], out=starts[1:]) if len(fanout) > 1 else None
    j = np.arange(src.shape[0], dtype=np.int64) - np.repeat(starts, fanout)
    ctr = (src.astype(np.uint64) << np.uint64(24)) + j.astype(np.uint64)
    h = hash3(seed, stream, ctr)
    if exclude_self:
        t = below(h, n_targets - 1)
        t = t + (t >= (src - target_base)).astype(np.int64)
    else:
        t = below(h, n_targets)
    return src, t + target_base, ctr


PLASTIC_MODES = {False: 0, True: 1, "exc": 2}  # snn_ei_desc::plastic


def plastic_flags(mode: int, src: np.ndarray, n: int, n_exc: int) -> np.ndarray:
    """Per-synapse plastic flags of ei_network's plasticity mode (0 none, 1 all, 2 all but the
    synapses of inhibitory neurons [n_exc, n))."""
    if mode == 2:
        return (~((src >= n_exc) & (src < n))).astype(np.uint8)
    return np.full(src.shape[0], 1 if mode == 1 else 0, np.uint8)


def ei_network(n: int, fanout, seed: int, *, n_sources: int, source_fanout: int, p_source: float,
               ticks: int, threshold: float, delay_max: int, compartments: Tuple[int, ...],
               n_codes: int = 1, code_max: int = 0, plastic=True, exc_frac: float = 0.8,
               w_exc: float = 0.5, w_inh: float = 1.0, window: int = 32, sim_thr: float = 0.9,
               max_refs: int = 1, name: str = "") -> Net:
    """E/I network of ``n`` neurons plus ``n_sources`` Poisson input neurons.

    Neurons [0, n*exc_frac) are excitatory (AMPA, weight U[0, w_exc)), the rest inhibitory
    (GABA_A, weight magnitude U[0, w_inh)). Compartments are drawn uniformly from
    ``compartments``; delays from U{1..delay_max}; code offsets (``n_codes`` per synapse) from
    U{0..code_max} in draw order. Source s fires at tick t when hash < p_source. ``plastic``:
    False (no STDP), True (every synapse) or "exc" (every synapse but those of inhibitory neurons).
    """
    plastic_mode = PLASTIC_MODES[plastic]
    n_exc = int(round(n * exc_frac))
    fan = np.broadcast_to(np.asarray(fanout, dtype=np.int64), (n,)).copy()
    spec = dict(n=n, n_sources=n_sources, fanout=fan, source_fanout=source_fanout, seed=seed, n_exc=n_exc,
                w_exc=w_exc, w_inh=w_inh, delay_max=delay_max, compartments=tuple(compartments),
                plastic=plastic_mode, threshold=threshold, window=window, sim_thr=sim_thr, max_refs=max_refs,
                n_codes=max(1, n_codes), code_max=code_max)
    src, tgt, ctr = _fanout_edges(seed, 1, np.arange(n), fan, n, 0, True)
    ssrc, stgt, sctr = _fanout_edges(seed, 2, np.arange(n, n + n_sources), np.full(n_sources, source_fanout), n, 0, False)
    e_net = src.shape[0]
    exc = src < n_exc
    u = unif24(hash3(seed, 3, ctr))
    weight = np.where(exc, u * np.float32(w_exc), u * np.float32(w_inh)).astype(np.float32)
    receptor = np.where(exc, AMPA, GABA_A).astype(np.uint8)
    comps = np.asarray(compartments, dtype=np.uint8)
    comp = comps[below(hash3(seed, 4, ctr), len(comps))]
    delay = (1 + below(hash3(seed, 5, ctr), delay_max)).astype(np.uint32)
    sw = unif24(hash3(seed, 6, sctr)) * np.float32(0.5)
    sdelay = (1 + below(hash3(seed, 7, sctr), delay_max)).astype(np.uint32)

    all_src = np.concatenate([src, ssrc]).astype(np.uint32)
    all_tgt = np.concatenate([tgt, stgt]).astype(np.uint32)
    e = all_src.shape[0]
    all_ctr = np.concatenate([ctr, sctr])
    if n_codes <= 1 and code_max == 0:
        code_ptr = np.arange(e + 1, dtype=np.uint64)
        codes = np.zeros(e, dtype=np.uint32)
    else:
        k = np.arange(n_codes, dtype=np.uint64)
        hc = hash3(seed, 8, (all_ctr[:, None] << np.uint64(4)) + k[None, :])
        codes = below(hc, code_max + 1).astype(np.uint32).reshape(-1)
        code_ptr = np.arange(e + 1, dtype=np.uint64) * np.uint64(n_codes)

    thr, win, sim, mref = neurons(n + n_sources, threshold, window, sim_thr, max_refs)
    thr[n:] = 1.0
    net = make_net(
        (thr, win, sim, mref), all_src, all_tgt,
        np.concatenate([weight, sw]), np.full(e, 2.0, np.float32),
        np.concatenate([delay, sdelay]),
        np.concatenate([comp, np.full(ssrc.shape[0], BASAL, np.uint8)]),
        np.concatenate([receptor, np.full(ssrc.shape[0], AMPA, np.uint8)]),
        plastic_flags(plastic_mode, all_src, n, n_exc), code_ptr=code_ptr, codes=codes, name=name,
        info=dict(n=n, n_sources=n_sources, e_net=int(e_net), seed=seed, ticks=ticks,
                  threshold=threshold, p_source=p_source, spec=spec if n_codes >= 1 or code_max == 0 else None),
    )
    net.commands = poisson_commands(seed, n, n_sources, p_source, ticks)
    return net


def poisson_commands(seed: int, first_source: int, n_sources: int, p: float, ticks: int):
    """inject(source, t, 1.0) for every (source, t) with hash < p, then run_until(ticks-1)."""
    s = np.arange(n_sources, dtype=np.uint64)
    t = np.arange(ticks, dtype=np.uint64)
    h = hash3(seed, 9, (s[None, :] << np.uint64(32)) + t[:, None])
    fire = unif24(h) < np.float32(p)
    tt, ss = np.nonzero(fire)  # tick-major, then source order
    cmds = [(0, int(first_source + s_), int(t_), 1.0) for t_, s_ in zip(tt, ss)]
    cmds.append((1, 0, ticks - 1, 0.0))
    return cmds


def config0(seed: int = 0, threshold: float = 1.0, ticks: int = 1000, plastic: bool = True) -> Net:
    """1,000 neurons (80/20 E/I), fan-out 100 without self-loops, delays U{1..20}, Basal/Apical
    compartments, codes {0}; 100 Poisson sources (p=0.01, 50 targets each); STDP on; 1,000 ticks."""
    return ei_network(1000, 100, seed, n_sources=100, source_fanout=50, p_source=0.01, ticks=ticks,
                      threshold=threshold, delay_max=20, compartments=(BASAL, APICAL),
                      plastic=plastic, name="config0")


def feature_net(seed: int, n: int = 60, fanout: int = 6, ticks: int = 300, n_codes: int = 3, code_max: int = 4,
                delay_max: int = 6, plastic_frac: float = 0.7, threshold: float = 0.8, max_refs=(1, 2, 4),
                window: int = 10, sim_thr: float = 0.9, interleave: bool = True, max_weight_range=None) -> Net:
    """Small network exercising every feature of the step: 0..n_codes code offsets per synapse
    (including empty lists and duplicates), delay 0, all four compartments and receptors, mixed
    plasticity and max_refs, repeated fires per tick, and injects issued between run_until calls
    (each inject's stage is the scheduler's next poll tick at the time of the call).
    max_weight_range=(lo, hi) draws every synapse's max_weight uniformly from [lo, hi) (one clamp
    bound per synapse) from a generator of its own, so the other arrays are those of the default."""
    rng = np.random.default_rng(seed)
    e = n * fanout
    src = np.repeat(np.arange(n), fanout)
    tgt = rng.integers(0, n, e)
    weight = rng.uniform(0.0, 0.5, e).astype(np.float32)
    receptor = rng.integers(0, 4, e).astype(np.uint8)
    comp = rng.integers(0, 4, e).astype(np.uint8)
    delay = rng.integers(0, delay_max + 1, e).astype(np.uint32)
    plastic = (rng.random(e) < plastic_frac).astype(np.uint8)
    max_w = rng.choice(np.array([0.5, 1.0, 2.0, 3.0], np.float32), e)
    if max_weight_range is not None:
        max_w = np.random.default_rng(seed + 7919).uniform(*max_weight_range, e).astype(np.float32)
    codes = [list(rng.integers(0, code_max + 1, rng.integers(0, n_codes + 1))) for _ in range(e)]
    nrn = neurons(n, threshold, window, sim_thr, 1)
    nrn[3][:] = rng.choice(np.asarray(max_refs, np.uint64), n)
    nrn[0][:] = rng.uniform(threshold * 0.8, threshold * 1.2, n).astype(np.float32)
    net = make_net(nrn, src, tgt, weight, max_w, delay, comp, receptor, plastic, codes_per_syn=codes,
                   name=f"feature{seed}")
    cmds = []
    t = 0
    while t < ticks:
        step = int(rng.integers(5, 40)) if interleave else ticks
        for _ in range(int(rng.integers(4, 30))):
            cmds.append((0, int(rng.integers(0, n)), int(t + rng.integers(-3, step + 5)) if t else int(rng.integers(0, step)),
                         float(rng.uniform(0.2, 1.5))))
        t = min(ticks, t + step)
        cmds.append((1, 0, t - 1, 0.0))
    net.commands = [(k, a, max(b, 0), w) for k, a, b, w in cmds]
    return net


def pattern_commands(seed: int, first_source: int, n_sources: int, ticks: int, n_patterns: int = 10,
                     spikes: int = 8, span: int = 32, jitter: int = 1, p_onset: float = 0.01,
                     p_background: float = 0.005):
    """Input of the spike-code workload: ``n_patterns`` fixed patterns of ``spikes`` (source, offset)
    pairs inside a ``span``-tick window, replayed at random onsets (probability ``p_onset`` per
    pattern per tick) with +-``jitter`` ticks, over Bernoulli(``p_background``) background spikes.
    Injects are issued in (tick, source) order before one run_until(ticks-1)."""
    k = np.arange(n_patterns * spikes, dtype=np.uint64)
    psrc = below(hash3(seed, 20, k), n_sources).reshape(n_patterns, spikes)
    poff = below(hash3(seed, 21, k), span).reshape(n_patterns, spikes)
    events = []
    t = np.arange(ticks, dtype=np.uint64)
    for p in range(n_patterns):
        on = np.nonzero(unif24(hash3(seed, 22, (np.uint64(p) << np.uint64(32)) + t)) < np.float32(p_onset))[0]
        for o in on:
            jit = below(hash3(seed, 23, (np.uint64(o) << np.uint64(16)) + np.uint64(p) * np.uint64(64) + np.arange(spikes, dtype=np.uint64)),
                        2 * jitter + 1) - jitter
            tt = o + poff[p] + jit
            for s_, t_ in zip(psrc[p], tt):
                if 0 <= t_ < ticks:
                    events.append((int(t_), int(s_)))
    s = np.arange(n_sources, dtype=np.uint64)
    h = hash3(seed, 24, (s[None, :] << np.uint64(32)) + t[:, None])
    bt, bs = np.nonzero(unif24(h) < np.float32(p_background))
    events += list(zip(bt.tolist(), bs.tolist()))
    events.sort()
    cmds = [(0, first_source + s_, t_, 1.0) for t_, s_ in events]
    cmds.append((1, 0, ticks - 1, 0.0))
    return cmds


def config1(seed: int = 0, n: int = 100_000, fanout: int = 1000, n_sources: int = 1000, source_fanout: int = 1000,
            ticks: int = 10_000, threshold: float = 1.1, plastic: bool = True, name: str = "config1") -> Net:
    """Spike-code network: 80/20 E/I, delays U{1..20}, 4 compartments, 4 code offsets per synapse
    from U{0..7} in draw order; input neurons replay 10 fixed 8-spike / 32-tick patterns with +-1
    jitter over 5 Hz background; STDP on."""
    net = ei_network(n, fanout, seed, n_sources=n_sources, source_fanout=source_fanout, p_source=0.0,
                     ticks=ticks, threshold=threshold, delay_max=20, compartments=(SOMA, BASAL, APICAL, INHIBITORY),
                     n_codes=4, code_max=7, plastic=plastic, name=name)
    net.commands = pattern_commands(seed, n, n_sources, ticks)
    return net


def config1_spec(seed: int = 0, n: int = 100_000, fanout: int = 1000, n_sources: int = 1000, source_fanout: int = 1000,
                 ticks: int = 10_000, threshold: float = 1.1, plastic=True, w_exc: float = 0.5, w_inh: float = 1.0):
    """The config-1 network as a device-generator spec plus its pattern-replay commands, without
    building the synapse arrays on the host (equal to config1(...) built there). Excitatory weights
    are U[0, w_exc), inhibitory magnitudes U[0, w_inh)."""
    spec = dict(n=n, n_sources=n_sources, fanout=np.full(n, fanout, np.int64), source_fanout=source_fanout, seed=seed,
                n_exc=int(round(n * 0.8)), w_exc=w_exc, w_inh=w_inh, delay_max=20,
                compartments=(SOMA, BASAL, APICAL, INHIBITORY), plastic=PLASTIC_MODES[plastic], threshold=threshold,
                window=32, sim_thr=0.9, max_refs=1, n_codes=4, code_max=7)
    return spec, pattern_commands(seed, n, n_sources, ticks)


def config1_reduced(seed: int = 0, ticks: int = 300, threshold: float = 4.5, plastic: bool = True) -> Net:
    """Config 1's law at 2,000 neurons (fan-out 100, 100 input neurons)."""
    return config1(seed, n=2000, fanout=100, n_sources=100, source_fanout=100, ticks=ticks, threshold=threshold,
                   plastic=plastic, name="config1_reduced")


def lognormal_fanout(seed: int, n: int, median: float = 1500.0, sigma: float = 0.8, cap: int = 20_000) -> np.ndarray:
    """ceil(median * exp(sigma * z)) capped, z standard normal by Box-Muller from two hashed
    uniforms (computed on the host; the device generator takes these counts as input)."""
    i = np.arange(n, dtype=np.uint64)
    u1 = (hash3(seed, 30, i) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    u2 = (hash3(seed, 31, i) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    f = np.ceil(median * np.exp(sigma * z))
    return np.clip(f, 1, cap).astype(np.int64)


def config2(seed: int = 0, n: int = 1_000_000, n_sources: int = 10_000, source_fanout: int = 1000,
            ticks: int = 2000, threshold: float = 8.0, plastic: bool = False, name: str = "config2") -> Net:
    """Large plain network: lognormal fan-out (median 1,500, sigma 0.8, cap 20,000), 80/20 E/I,
    delays U{1..32}, Basal/Apical compartments, codes {0}; 10,000 Poisson input neurons at
    p=0.005 with fan-out 1,000. STDP off unless ``plastic``. Host generation: use for reduced n."""
    fan = lognormal_fanout(seed, n)
    net = ei_network(n, fan, seed, n_sources=n_sources, source_fanout=source_fanout, p_source=0.005,
                     ticks=ticks, threshold=threshold, delay_max=32, compartments=(BASAL, APICAL),
                     plastic=plastic, name=name)
    net.info["fanout_mean"] = float(fan.mean())
    return net


def config2_spec(seed: int = 0, n: int = 1_000_000, n_sources: int = 10_000, source_fanout: int = 1000,
                 ticks: int = 2000, threshold: float = 8.0, plastic=False, w_exc: float = 0.5,
                 w_inh: float = 1.0, kick: float = 0.0):
    """The config-2 network as a device-generator spec plus its input commands, without building the
    synapse arrays on the host (equal to config2(...) built there). ``kick`` > 0 adds, before the
    Poisson input, an inject of weight ``threshold`` at tick 0 into each network neuron with
    probability ``kick``. ``plastic``: False, True or "exc" (as ei_network)."""
    fan = lognormal_fanout(seed, n)
    spec = dict(n=n, n_sources=n_sources, fanout=fan, source_fanout=source_fanout, seed=seed, n_exc=int(round(n * 0.8)),
                w_exc=w_exc, w_inh=w_inh, delay_max=32, compartments=(BASAL, APICAL),
                plastic=PLASTIC_MODES[plastic], threshold=threshold, window=32, sim_thr=0.9, max_refs=1)
    cmds = poisson_commands(seed, n, n_sources, 0.005, ticks)
    if kick > 0:
        hit = np.nonzero(unif24(hash3(seed, 40, np.arange(n, dtype=np.uint64))) < np.float32(kick))[0]
        cmds = [(0, int(i), 0, float(threshold)) for i in hit] + cmds
    return spec, cmds


# --------------------------------------------------------------------------- batches of networks

def config3_shape(seed: int, n_range=(500, 50_000), fan_range=(50, 500)) -> Tuple[int, int]:
    """Neuron count (log-uniform over n_range) and fan-out (uniform integer over fan_range) of the
    batched workload's network ``seed``."""
    u = float(unif24(hash3(seed, 50, np.zeros(1, np.uint64)))[0])
    lo, hi = np.log(n_range[0]), np.log(n_range[1])
    n = int(np.floor(np.exp(lo + (hi - lo) * u)))
    f = int(fan_range[0] + below(hash3(seed, 51, np.zeros(1, np.uint64)), fan_range[1] - fan_range[0] + 1)[0])
    return max(n_range[0], min(n_range[1], n)), f


def config3_net(seed: int, ticks: int = 2000, threshold: float = 3.0, n_range=(500, 50_000), fan_range=(50, 500),
                plastic: bool = True, w_inh: float = 1.0) -> Net:
    """Network ``seed`` of the batched workload: log-uniform size, uniform fan-out, 80/20 E/I,
    delays U{1..20}, Basal/Apical compartments, 2 code offsets per synapse from U{0..7}, 100 Poisson
    sources at p=0.01 with fan-out 50, STDP on."""
    n, f = config3_shape(seed, n_range, fan_range)
    return ei_network(n, f, seed, n_sources=100, source_fanout=50, p_source=0.01, ticks=ticks,
                      threshold=threshold, delay_max=20, compartments=(BASAL, APICAL), n_codes=2, code_max=7,
                      plastic=plastic, w_inh=w_inh, name=f"config3_{seed}")


def config3_spec(seed: int, ticks: int = 2000, threshold: float = 3.0, n_range=(500, 50_000), fan_range=(50, 500),
                 plastic: bool = True, w_inh: float = 1.0):
    """config3_net(seed, ...) as a device-generator spec plus its commands, without host arrays."""
    n, f = config3_shape(seed, n_range, fan_range)
    spec = dict(n=n, n_sources=100, fanout=np.full(n, f, np.int64), source_fanout=50, seed=seed,
                n_exc=int(round(n * 0.8)), w_exc=0.5, w_inh=w_inh, delay_max=20, compartments=(BASAL, APICAL),
                plastic=PLASTIC_MODES[plastic], threshold=threshold, window=32, sim_thr=0.9, max_refs=1, n_codes=2,
                code_max=7)
    return spec, poisson_commands(seed, n, 100, 0.01, ticks)


def merge_commands(commands_per_net: Sequence[Sequence], neuron_off: Sequence[int]):
    """One command list for the union of networks (network k's neuron i is neuron_off[k] + i) that
    gives every network its injects at the same scheduler poll tick as its own command list: each
    network's injects are issued when the shared clock reaches that network's previous run_until,
    and the clock runs to the smallest pending run_until of any network. Returns (command list,
    network of each inject in order)."""
    cmds = [list(c) for c in commands_per_net]
    pos = [0] * len(cmds)
    out, owner = [], []
    while True:
        for k, c in enumerate(cmds):
            while pos[k] < len(c) and c[pos[k]][0] == 0:
                _, n, t, w = c[pos[k]]
                out.append((0, int(neuron_off[k]) + int(n), int(t), float(w)))
                owner.append(k)
                pos[k] += 1
        pending = [c[pos[k]][2] for k, c in enumerate(cmds) if pos[k] < len(c)]
        if not pending:
            return out, owner
        m = min(pending)
        out.append((1, 0, int(m), 0.0))
        for k, c in enumerate(cmds):
            if pos[k] < len(c) and c[pos[k]][0] == 1 and c[pos[k]][2] == m:
                pos[k] += 1


def union(nets: Sequence[Net], name: str = "union") -> Net:
    """The disjoint union of independent networks as one network, with merged commands; info holds
    neuron_off / synapse_off."""
    noff = np.zeros(len(nets) + 1, np.int64)
    eoff = np.zeros(len(nets) + 1, np.int64)
    for k, n in enumerate(nets):
        noff[k + 1] = noff[k] + n.n_neurons
        eoff[k + 1] = eoff[k] + n.n_synapses
    cat = lambda f: np.concatenate([getattr(n, f) for n in nets])
    code_ptr = [np.zeros(1, np.uint64)]
    base = np.uint64(0)
    for n in nets:
        code_ptr.append(n.code_ptr[1:] - n.code_ptr[0] + base)
        base = code_ptr[-1][-1] if n.n_synapses else base
    cmds, _ = merge_commands([n.commands for n in nets], noff)
    return Net(threshold=cat("threshold"), window=cat("window"), sim_thr=cat("sim_thr"), max_refs=cat("max_refs"),
               src=np.concatenate([n.src.astype(np.int64) + noff[k] for k, n in enumerate(nets)]).astype(np.uint32),
               tgt=np.concatenate([n.tgt.astype(np.int64) + noff[k] for k, n in enumerate(nets)]).astype(np.uint32),
               weight=cat("weight"), max_weight=cat("max_weight"), delay=cat("delay"), comp=cat("comp"),
               receptor=cat("receptor"), plastic=cat("plastic"), code_ptr=np.concatenate(code_ptr).astype(np.uint64),
               codes=cat("codes"), commands=cmds, name=name,
               info=dict(neuron_off=noff, synapse_off=eoff))


# --------------------------------------------------------------------------- adjoint workload

def config4(seed: int = 0, n: int = 10_000, fan_in: int = 500, ticks: int = 500, threshold: float = 40.0,
            p_in: float = 0.01, w_in: float = 40.0, delay_max: int = 20) -> Net:
    """Network of the adjoint workload: n neurons with fan-in ``fan_in`` from uniform presynaptic
    neurons (no self-loops), weight magnitudes U[0, 0.5); 20% of the synapses are inhibitory (GABA_A
    or GABA_B, whose negative receptor sign negates the current, as in the other workloads), the rest
    AMPA or NMDA; Basal/Apical compartments, delays U{1..delay_max}, codes {0}, STDP off. Drive: every
    neuron receives an external inject of weight ``w_in`` with probability ``p_in`` per tick (soma,
    AMPA). Connectome order is target-major. Pinned regime: threshold 40 and inject weight 40 make
    the injects (1% per neuron per tick) the firing events and keep the recurrent input (fan-in 500,
    about 16 on average with NMDA's 100-tick time constant) subthreshold, so every neuron resets
    about five times in 500 ticks; threshold 10 runs away within 60 ticks at 4,000 neurons."""
    e = n * fan_in
    tgt = np.repeat(np.arange(n, dtype=np.int64), fan_in)
    ctr = np.arange(e, dtype=np.uint64)
    s_ = below(hash3(seed, 60, ctr), n - 1).astype(np.int64)
    src = s_ + (s_ >= tgt)
    w = (unif24(hash3(seed, 61, ctr)) * np.float32(0.5)).astype(np.float32)
    inh = unif24(hash3(seed, 62, ctr)) < np.float32(0.2)
    receptor = (below(hash3(seed, 63, ctr), 2) + np.where(inh, 2, 0)).astype(np.uint8)
    comp = np.where(below(hash3(seed, 64, ctr), 2) == 0, BASAL, APICAL).astype(np.uint8)
    delay = (1 + below(hash3(seed, 65, ctr), delay_max)).astype(np.uint32)
    thr, win, sim, mref = neurons(n, threshold, 32, 0.9, 1)
    net = make_net((thr, win, sim, mref), src.astype(np.uint32), tgt.astype(np.uint32), w, np.full(e, 2.0, np.float32),
                   delay, comp, receptor, np.zeros(e, np.uint8), name="config4",
                   info=dict(n=n, fan_in=fan_in, seed=seed, ticks=ticks, threshold=threshold, p_in=p_in, w_in=w_in))
    i = np.arange(n, dtype=np.uint64)
    t = np.arange(ticks, dtype=np.uint64)
    fire = unif24(hash3(seed, 66, (i[None, :] << np.uint64(32)) + t[:, None])) < np.float32(p_in)
    tt, nn = np.nonzero(fire)
    net.commands = [(0, int(a), int(b), float(w_in)) for b, a in zip(tt, nn)] + [(1, 0, ticks - 1, 0.0)]
    return net
