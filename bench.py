#!/usr/bin/env python3
"""Benchmark of the per-tick spiking step on one MI355X.

Workload of the headline line: the 1,000,000-neuron network with lognormal fan-out (median 1,500,
sigma 0.8, cap 20,000; ~2.07e9 synapses), 80/20 E/I, delays U{1..32}, Basal/Apical compartments, one
code offset per synapse, 10,000 Poisson input neurons at p=0.005 with fan-out 1,000, STDP off.
Pinned free parameters: threshold 2.5 (2.75 with STDP), excitatory weights U[0,0.5), inhibitory
GABA_A weights U[0,5), a 2% ignition inject at tick 0. The network is generated on the device and
is resident in HBM before timing.

A step is one tick. Every run first steps SETTLE_TICKS ticks untimed (the activity after the
ignition inject ramps up for ~100-200 ticks, then holds ~0.5% of the neurons firing per tick,
~1e7 delivered events per tick), then --warmup more ticks, then times exactly --steps ticks in one
synchronous run_until call. `value` is delivered synaptic events per second over those ticks
(input injects excluded). Every PROFILE_EVERY-th of the same timed ticks carries HIP events around
the scatter and the fold on the step's stream (events around every launch would slow the step by
~5%), so `roofline` and `value` describe the same ticks.

The JSON line carries:
  roofline       the dominant kernel (the partition fold): algorithmic bytes per launch = the
                 per-event byte model of the plain step (50 B per delivered event; with STDP the
                 pre- and post-synaptic terms) x the events of one launch, / the fold's mean launch
                 duration; traffic = FETCH_SIZE + WRITE_SIZE per dispatch from rocprofv3 --pmc
                 (profiles/), scaled by the copy-kernel calibration stated there; kernel_model: the
                 fold's own byte model.
  scatter_roofline, step_roofline   the partition scatter, and the whole tick (same byte model / tick time).
  stdp_on        the same measurement with STDP on (excitatory synapses plastic, threshold 2.75).
  dense_regime   the same measurement (STDP off) at threshold 2.4, after 1,300 settle ticks: the
                 activity grows to ~0.9% of the neurons firing per tick (~1.9e7 events per tick) and
                 holds there, so the per-tick fixed costs spread over about twice the events.
  forms          the batched form (512 uneven networks stepped together vs looping the plain step
                 over a sample of them), the fixed-spike adjoint, and the spike-code network (100,000
                 neurons, 4 code offsets per synapse, STDP on; its whole 10,000-tick run), measured in
                 this run.
  cpu_baseline   the oracle (plain-C restatement of the reference CPU step), one thread on this host,
                 on the same law at 20,000 neurons; cpu_baselines: the oracle and the reference's own
                 runtime (oracle/_ref/ref_harness, when built) on the other samples.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import pathlib
import sys
import tempfile
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
SETTLE_TICKS = 300
PROFILE_EVERY = 8  # timed ticks with kernel events: one in this many
# bytes per delivered event of the plain step (STDP off), by the per-event byte model: static edge
# fields 10 B (target, weight, delay, compartment/receptor) + event queue 8 B + synapse code-pattern
# record read and write 32 B
STEP_BYTES_PER_EVENT = 50.0
# STDP on adds 12 B per plastic event (pre-side record) and 20 B per (fire, incoming plastic synapse)
STDP_EVENT_BYTES = 12.0
STDP_POST_BYTES = 20.0
# partition scatter: reads the edge's target u32, weight f32 and compartment/receptor u8, writes the
# bucket record (8 B event + 2 B target index)
SCATTER_BYTES_PER_EVENT = 9.0 + 10.0
# partition fold: reads each bucket record (10 B) once, and per touched neuron its threshold (4 B),
# dynamic state (32 B) and temporal header (32 B), writing back state and header, and the
# candidate's last run (4 B read, 4 B written)
FOLD_BYTES_PER_EVENT = 10.0
FOLD_BYTES_PER_NEURON = 4.0 + 2 * 32.0 + 2 * 32.0 + 8.0
# with STDP the bucket record is 12 + 2 B and a plastic event reads and writes its 32-B STDP record
FOLD_STDP_BYTES_PER_EVENT = 14.0
FOLD_STDP_RECORD_BYTES = 2 * 32.0
# spike-code network: several code offsets per synapse keep one 96-B code record per synapse (offsets
# observed and learned), read and written by every event of the synapse, in place of the 32 B of the
# per-source record; STDP on every synapse
SPIKECODE_BYTES_PER_EVENT = 10.0 + 8.0 + 2 * 96.0 + STDP_EVENT_BYTES
SPIKECODE_THRESHOLD = 4.6
# the denser regime of the 1M-neuron workload: from threshold 2.4 the activity grows for ~1,200 ticks
# to about 0.9% of the neurons firing per tick (~1.9e7 events) and holds there (2.3 runs away)
DENSE_THRESHOLD = 2.4
DENSE_SETTLE_TICKS = 1300


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed ticks")
    ap.add_argument("--warmup", type=int, default=50, help="untimed ticks after the settle phase")
    ap.add_argument("--neurons", type=int, default=1_000_000)
    ap.add_argument("--form", choices=("all", "plain", "stdp", "dense", "batched", "adjoint", "spikecode"),
                    default="all",
                    help="all (default): the plain step as the headline line with STDP on, the denser regime, the "
                         "batched form, the adjoint, the spike-code network and the CPU baselines measured in the same "
                         "run; plain / stdp / dense: one plain-step line; batched / adjoint / spikecode: one line of "
                         "that form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--breakdown", action="store_true",
                    help="one more pass of --steps ticks with HIP events around every kernel (kernel_ms_per_tick)")
    ap.add_argument("--schedule", default="",
                    help="plain form: partition_shift,scatter_workgroups,graph_ticks[,fold_threads]")
    ap.add_argument("--cpu-neurons", type=int, default=20_000)
    ap.add_argument("--cpu-ticks", type=int, default=200)
    ap.add_argument("--nets", type=int, default=512, help="batched form: networks")
    ap.add_argument("--batch-ticks", type=int, default=2000, help="batched form: ticks per network")
    ap.add_argument("--spikecode-ticks", type=int, default=10_000, help="spike-code form: ticks of the run")
    ap.add_argument("--loop-nets", type=int, default=16,
                    help="batched form: networks in the looped baseline, every (nets/loop-nets)-th (0: all)")
    return ap.parse_args()


def threshold_of(stdp: bool) -> float:
    # STDP on potentiates the recurrent excitation: at 2.5 the 1M-neuron network runs away within
    # 500 ticks, at 2.75 it holds ~0.37% of the neurons firing per tick for 2,200 ticks
    return 2.75 if stdp else 2.5


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# --------------------------------------------------------------------------- plain step

def plain(args, stdp: bool, threshold: float = 0.0, settle: int = SETTLE_TICKS) -> dict:
    import numpy as np
    import snncuda_amd as sa
    from snncuda_amd import workloads as W

    n = args.neurons
    theta = threshold or threshold_of(stdp)
    total = settle + args.warmup + args.steps + (args.steps if args.breakdown else 0)
    # STDP on: every synapse but those of inhibitory neurons is plastic (the rule caps weights at
    # 2.0, below the pinned inhibitory magnitudes U[0,5), so plastic inhibition would collapse)
    spec, cmds = W.config2_spec(0, n=n, ticks=total, threshold=theta, w_inh=5.0, kick=0.02,
                                plastic="exc" if stdp else False)
    injects = [c for c in cmds if c[0] == 0]
    inj_ticks = np.array([t for _, _, t, _ in injects], dtype=np.int64)
    t_gen = time.perf_counter()
    with sa.Network(spec=spec, max_events_per_tick=1 << 27, fire_log_capacity=1 << 26,
                    raster_capacity=1 << 20, track_pre_count=False) as dev:
        t_gen = time.perf_counter() - t_gen
        dev.run_commands(injects)
        if args.schedule:
            v = [int(x) for x in args.schedule.split(",")]
            kw = dict(partition_shift=v[0], scatter_workgroups=v[1], graph_ticks=v[2])
            if len(v) > 3:
                kw["fold_threads"] = v[3]
            dev.set_schedule(**kw)
        first = settle + args.warmup
        dev.run_until(first - 1)
        # timed region: one synchronous call covering exactly --steps ticks, with HIP events around
        # the scatter and the fold on every PROFILE_EVERY-th tick
        dev.set_profiling(True, kernels=("scatter", "fold"), every=PROFILE_EVERY)
        d0 = dev.delivered
        t0 = time.perf_counter()
        dev.run_until(first + args.steps - 1)
        elapsed = time.perf_counter() - t0
        d1 = dev.delivered
        kt = dev.kernel_times()
        n_inj = int(((inj_ticks >= first) & (inj_ticks < first + args.steps)).sum())
        events = d1 - d0 - n_inj
        dpt, fpt = dev.tick_counts(first, args.steps)
        breakdown = None
        if args.breakdown:
            dev.set_profiling(True)
            dev.run_until(first + 2 * args.steps - 1)
            breakdown = {k: v[0] / max(v[1], 1) for k, v in dev.kernel_times().items() if v[1]}
        dev.set_profiling(False)
        E = dev.e
        schedule = dev.schedule()
    ev_tick = float(dpt.mean())
    inj_tick = n_inj / args.steps
    syn_tick = ev_tick - inj_tick  # synaptic events per launch
    # synaptic targets are uniform over the network's neurons: expected number touched per tick
    touched = n * (1.0 - math.exp(-syn_tick / n))
    fan = np.asarray(spec["fanout"], np.int64)
    plastic_syn = int(fan[:spec["n_exc"]].sum()) + spec["n_sources"] * spec["source_fanout"] if stdp else 0
    pfrac = plastic_syn / E  # plastic share of synapses, and of delivered events (targets uniform)
    fires_tick = float(fpt.mean())
    # the step's byte model (per tick), credited to the dominant kernel and to the whole tick
    step_bytes = STEP_BYTES_PER_EVENT * syn_tick
    if stdp:
        step_bytes += STDP_EVENT_BYTES * pfrac * syn_tick + STDP_POST_BYTES * fires_tick * plastic_syn / n
    if stdp:
        fold_model = (FOLD_STDP_BYTES_PER_EVENT + FOLD_STDP_RECORD_BYTES * pfrac) * ev_tick + FOLD_BYTES_PER_NEURON * touched
    else:
        fold_model = FOLD_BYTES_PER_EVENT * ev_tick + FOLD_BYTES_PER_NEURON * touched

    def traffic_of(name):
        # rocprofv3 counters (per delivered event, copy-kernel calibration applied) x this run's
        # events per launch; files *_stdp.json for STDP on
        f = ROOT / "profiles" / (name.replace(".json", "_stdp.json") if stdp else name)
        if not f.exists():
            return None
        tr = json.loads(f.read_text())
        return tr["bytes_per_event"] * ev_tick if tr.get("stdp", False) == stdp else None

    def kernel_roofline(key, kernel, bytes_per_launch, model_bytes, traffic_file):
        ms, launches = kt[key]
        avg_s = ms / max(launches, 1) / 1e3
        achieved = bytes_per_launch / avg_s / 1e9
        return {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_of(traffic_file),
                "bytes_per_launch": bytes_per_launch, "mean_launch_ms": avg_s * 1e3, "launches": launches,
                "kernel_model": {"bytes_per_launch": model_bytes, "achieved": model_bytes / avg_s / 1e9,
                                 "frac": model_bytes / avg_s / 1e9 / HBM_PEAK_GBS}}

    fold_rl = kernel_roofline("fold", "k_pfold (partition fold)", step_bytes, fold_model, "fold_traffic.json")
    fold_rl["kernel_model"].update(per_event=FOLD_STDP_BYTES_PER_EVENT + FOLD_STDP_RECORD_BYTES * pfrac if stdp
                                   else FOLD_BYTES_PER_EVENT, per_touched_neuron=FOLD_BYTES_PER_NEURON,
                                   touched_neurons=touched)
    fold_rl["bytes_per_event"] = step_bytes / syn_tick
    scat_rl = kernel_roofline("scatter", "k_pscatter (partition scatter)", SCATTER_BYTES_PER_EVENT * ev_tick,
                              SCATTER_BYTES_PER_EVENT * ev_tick, "scatter_traffic.json")
    step_gbs = step_bytes * args.steps / elapsed / 1e9
    return {
        "value": events / elapsed,
        "ms_per_step": elapsed / args.steps * 1e3,
        "events_per_tick": syn_tick,
        "fires_per_tick": fires_tick,
        "fire_fraction_per_tick": fires_tick / n,
        "threshold": theta,
        "settle_ticks": settle,
        "synapses": int(E),
        "neurons": n + spec["n_sources"],
        "plastic_fraction": pfrac,
        "roofline": fold_rl,
        "scatter_roofline": scat_rl,
        "step_roofline": {"achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                          "bytes_per_event": step_bytes / syn_tick},
        "kernel_ms_per_tick": breakdown,
        "schedule": schedule,
        "ticks_per_s": args.steps / elapsed,
        "generate_seconds": t_gen,
    }


def plain_line(args, r: dict, stdp: bool) -> dict:
    line = {
        "metric": "synaptic events/s (1M-neuron net, % HBM roofline)",
        "value": r["value"],
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator, built on the device)",
        "config": {
            "workload": f"{args.neurons}-neuron lognormal fan-out net (median 1500, sigma 0.8, cap 20000), "
                        f"delays U{{1..32}}, 2 compartments, codes {{0}}, "
                        f"STDP {'on (excitatory synapses)' if stdp else 'off'}",
            "neurons": r["neurons"],
            "synapses": r["synapses"],
            "settle_ticks": r["settle_ticks"],
            "timed_ticks": [r["settle_ticks"] + args.warmup, r["settle_ticks"] + args.warmup + args.steps - 1],
            "events_per_tick": r["events_per_tick"],
            "fires_per_tick": r["fires_per_tick"],
            "fire_fraction_per_tick": r["fire_fraction_per_tick"],
            "threshold": r["threshold"],
            "pre_spike_count": "not tracked in the timed run (a per-synapse runtime counter outside "
                               "CudaPropagationResult; track_pre_count=False)",
            "parallelism": "1 GPU",
        },
    }
    for k in ("roofline", "scatter_roofline", "step_roofline", "kernel_ms_per_tick", "schedule", "ticks_per_s",
              "generate_seconds"):
        if r.get(k) is not None:
            line[k] = r[k]
    return line


# --------------------------------------------------------------------------- batched form

def batched(args) -> dict:
    """512 independent networks of the batched workload (log-uniform sizes 500..50,000, fan-out
    U{50..500}, 2 code offsets, 100 Poisson sources each, STDP on, threshold 3.0), generated on the
    device, stepped together for --batch-ticks ticks under the online autotuner; against looping the
    plain step over a sample of the same networks (every k-th, stream launches), scaled to all of
    them. value = networks/s of the batched run (networks x their full run / wall time)."""
    import numpy as np
    import snncuda_amd as sa
    from snncuda_amd import workloads as W

    T = args.batch_ticks
    specs, cmds = zip(*[W.config3_spec(s, ticks=T) for s in range(args.nets)])
    n_total = sum(sp["n"] + sp["n_sources"] for sp in specs)
    e_total = sum(int(sp["fanout"].sum()) + sp["n_sources"] * sp["source_fanout"] for sp in specs)
    n_inj = sum(sum(1 for c in cm if c[0] == 0) for cm in cmds)
    log(f"batched: {args.nets} networks, {n_total} neurons, {e_total} synapses, {T} ticks")
    t_gen = time.perf_counter()
    with sa.Batch(specs=list(specs), track_pre_count=False, max_events_per_tick=1 << 24,
                  fire_log_capacity=1 << 22, raster_capacity=1 << 22) as b:
        t_gen = time.perf_counter() - t_gen
        u = b.union
        merged, owner = W.merge_commands(cmds, b.neuron_off)
        inj = np.array([c for c in merged if c[0] == 0], dtype=np.float64)
        own = np.asarray(owner, np.uint32)
        b.inject(own, inj[:, 1].astype(np.int64) - b.neuron_off[own].astype(np.int64), inj[:, 2].astype(np.uint64),
                 inj[:, 3].astype(np.float32))
        u.autotune(16)
        t0 = time.perf_counter()
        b.run_until(T - 1)
        t_batch = time.perf_counter() - t0
        delivered, fired, _ = u.counts()
        schedule = u.schedule()
        tune = u.tune_log()
    log(f"batched run {t_batch:.3f} s ({args.nets / t_batch:.1f} networks/s); schedule {schedule}")
    step = max(1, args.nets // args.loop_nets) if args.loop_nets else 1
    sample = list(range(0, args.nets, step))
    total = 0.0
    for k in sample:
        with sa.Network(spec=specs[k], track_pre_count=False, max_events_per_tick=1 << 22,
                        fire_log_capacity=1 << 20, raster_capacity=1 << 16) as net:
            net.run_commands([c for c in cmds[k] if c[0] == 0])
            t0 = time.perf_counter()
            net.run_until(T - 1)
            total += time.perf_counter() - t0
    loop_s = total * args.nets / len(sample)
    log(f"loop {loop_s:.3f} s for {args.nets} networks (sampled {len(sample)})")
    return {
        "metric": "networks/s (512 uneven networks, batched vs looped plain step)",
        "value": args.nets / t_batch,
        "unit": "networks/s",
        "ms_per_step": t_batch / T * 1e3,
        "config": {
            "workload": f"{args.nets} E/I networks, sizes log-uniform in [500, 50000], fan-out U{{50..500}}, "
                        "80/20 E/I, delays U{1..20}, 2 compartments, 2 code offsets U{0..7}, 100 Poisson "
                        "sources p=0.01 per network, STDP on, threshold 3.0",
            "neurons": n_total, "synapses": e_total, "ticks": T,
            "events_per_tick": (delivered - n_inj) / T, "fires_per_tick": fired / T,
        },
        "speedup_vs_loop": loop_s / t_batch,
        "loop_seconds": loop_s,
        "loop": f"plain step looped over {len(sample)} of the {args.nets} networks (every {step}-th, stream "
                f"launches), {total:.3f} s, scaled by {args.nets}/{len(sample)}",
        "batched_seconds": t_batch,
        "generate_seconds": t_gen,
        "events_per_s": (delivered - n_inj) / t_batch,
        "schedule": schedule,
        "autotune": tune.strip().split("\n"),
    }


# --------------------------------------------------------------------------- spike-code network

def spikecode(args) -> dict:
    """The spike-code network at full size: 100,000 neurons, fan-out 1,000 (1.01e8 synapses with the
    1,000 input neurons'), 80/20 E/I, delays U{1..20}, 4 compartments, 4 code offsets per synapse from
    U{0..7}, STDP on every synapse; the input neurons replay 10 fixed 8-spike / 32-tick patterns with
    +-1 jitter over 5 Hz background. Pinned threshold 4.6: from 4.6 to 4.9 the network holds
    ~0.5-1 recurrent fires per tick over the ~6 input fires for all 10,000 ticks; at 4.5 and below a
    fluctuation ignites every neuron within ~1,000 ticks, and from 5.0 only the input neurons fire.
    The whole run is timed (one run_until after --warmup untimed ticks, at most 100); value =
    synaptic events/s. The fold (per-event code patterns, STDP) is the dominant kernel."""
    import numpy as np
    import snncuda_amd as sa
    from snncuda_amd import workloads as W

    T = args.spikecode_ticks
    first = min(args.warmup, 100)
    spec, cmds = W.config1_spec(0, ticks=T, threshold=SPIKECODE_THRESHOLD)
    injects = [c for c in cmds if c[0] == 0]
    inj_ticks = np.array([t for _, _, t, _ in injects], dtype=np.int64)
    t_gen = time.perf_counter()
    with sa.Network(spec=spec, max_events_per_tick=1 << 24, fire_log_capacity=1 << 22, raster_capacity=1 << 22,
                    track_pre_count=False) as dev:
        t_gen = time.perf_counter() - t_gen
        dev.run_commands(injects)
        if first:
            dev.run_until(first - 1)
        dev.set_profiling(True, kernels=("scatter", "fold"), every=PROFILE_EVERY)
        d0 = dev.delivered
        t0 = time.perf_counter()
        dev.run_until(T - 1)
        elapsed = time.perf_counter() - t0
        events = dev.delivered - d0 - int((inj_ticks >= first).sum())
        kt = dev.kernel_times()
        dpt, fpt = dev.tick_counts(first, T - first)
        dev.set_profiling(False)
        E = dev.e
        schedule = dev.schedule()
    n = spec["n"]
    steps = T - first
    syn_tick = events / steps
    fires_tick = float(fpt.mean())
    # every synapse plastic: the post-synaptic term covers each fire's fan-in (E / neurons on average)
    step_bytes = SPIKECODE_BYTES_PER_EVENT * syn_tick + STDP_POST_BYTES * fires_tick * E / n
    ms, launches = kt["fold"]
    fold_s = ms / max(launches, 1) / 1e3
    sms, slaunches = kt["scatter"]
    achieved = step_bytes / fold_s / 1e9
    return {
        "metric": "synaptic events/s (spike-code net, 100k neurons, 4 code offsets, STDP on)",
        "value": events / elapsed,
        "unit": "synaptic events/s",
        "steps": steps,
        "warmup": first,
        "ms_per_step": elapsed / steps * 1e3,
        "ticks_per_s": steps / elapsed,
        "config": {
            "workload": "100,000 neurons, fan-out 1,000, 80/20 E/I, delays U{1..20}, 4 compartments, 4 code "
                        "offsets U{0..7} per synapse, STDP on, 1,000 input neurons replaying 10 fixed 8-spike / "
                        "32-tick patterns (+-1 jitter) over 5 Hz background, threshold 4.6",
            "neurons": n + spec["n_sources"], "synapses": int(E), "ticks": T,
            "events_per_tick": syn_tick, "fires_per_tick": fires_tick, "max_fires_per_tick": int(fpt.max()),
        },
        "roofline": {"kernel": "k_pfold (partition fold: per-event code patterns and STDP)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_launch": step_bytes, "bytes_per_event": step_bytes / syn_tick,
                     "mean_launch_ms": fold_s * 1e3, "launches": launches},
        "scatter_mean_launch_ms": sms / max(slaunches, 1),
        "schedule": schedule,
        "generate_seconds": t_gen,
    }


# --------------------------------------------------------------------------- adjoint

def adjoint_form(args, cpu: bool) -> dict:
    """Fixed-spike adjoint on the 10,000-neuron workload (fan-in 500, 5e6 synapses, 500 ticks): one
    step = forward (trace + checkpoints) + backward (segment recompute + adjoint recursion) of the
    MSE loss against a perturbed-weight target. value = steps/s; the FP64 CPU adjoint of the oracle
    (one thread) is timed on the same event lists as its CPU baseline."""
    import numpy as np
    import torch
    import snncuda_amd as sa
    from snncuda_amd import workloads as W
    from snncuda_amd.adjoint import Adjoint, DEFAULT_PARAMS

    net = W.config4(0)
    T = 500
    run = sa.Network(net, max_events_per_tick=1 << 24, fire_log_capacity=1 << 22, raster_capacity=1 << 24,
                     track_pre_count=False)
    run.record_fire_events()
    run.run_commands(net.commands)
    t0 = time.perf_counter()
    adj = Adjoint(run, net, T)
    t_build = time.perf_counter() - t0
    dev = torch.device("cuda")
    w = torch.tensor(net.weight, dtype=torch.float32, device=dev)
    params = np.asarray(DEFAULT_PARAMS, np.float32)
    target = torch.empty((T, adj.n), dtype=torch.float32, device=dev)
    adj.forward(w + 0.05 * torch.randn(w.shape, device=dev, generator=torch.Generator(device=dev).manual_seed(0)),
                params, target)
    trace = torch.empty_like(target)
    gw = torch.empty_like(w)

    def step():
        adj.forward(w, params, trace)
        g = (2.0 / trace.numel()) * (trace - target)
        return adj.backward(g, gw)

    for _ in range(3):
        step()
    K = max(1, min(args.steps, 50))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / K
    ev = adj.n_events
    # bytes per step: events (8 B) read once by the forward and twice by the backward (window
    # recompute, weight-gradient terms), the synapse weight gathered twice (4 B), each event's
    # gradient term written and read back through the synapse's event index (4 + 4 + 4 B); per
    # neuron-tick the trace written, the loss gradient formed from trace and target and read by the
    # backward (4 + 12 + 4 B); per synapse its event range read and its gradient written (8 + 4 B)
    bytes_step = ev * (3 * 8 + 2 * 4 + 3 * 4) + T * adj.n * (4 + 12 + 4) + adj.e * (8 + 4)
    line = {
        "metric": "adjoint steps/s (10k-neuron fixed-spike adjoint, forward + backward)",
        "value": 1.0 / dt,
        "unit": "steps/s",
        "steps": K,
        "ms_per_step": dt * 1e3,
        "config": {"workload": "10,000 neurons, fan-in 500 (5e6 synapses), 500 ticks, 2 compartments, checkpoints "
                               "every 64 ticks, MSE against a target from weights + N(0, 0.05)",
                   "events": ev, "fires": int(run.fired)},
        "events_per_s": 3 * ev / dt,
        "roofline": {"bound": "hbm", "achieved": bytes_step / dt / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes_step / dt / 1e9 / HBM_PEAK_GBS, "traffic": None, "bytes_per_step": bytes_step},
        "build_seconds": t_build,
    }
    if cpu:
        import oracle
        o = oracle.Oracle(net, log_events=True)
        o.run_commands(net.commands)
        ref = oracle.AdjointOracle(net, T, o.event_log())
        tg = target.cpu().numpy().astype(np.float64)
        wd = net.weight.astype(np.float64)
        pd = np.asarray(DEFAULT_PARAMS, np.float64)
        t0 = time.perf_counter()
        ref.grad(wd, pd[:4], pd[4:], tg)
        cdt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": 1.0 / cdt, "unit": "steps/s", "cores": 1, "kind": "port",
                                "sample": f"oracle/adjoint_oracle.c FP64 forward + backward on the same {ev} events, "
                                          f"one step in {cdt:.2f} s, 1 thread"}
    run.close()
    return line


def adjoint_child(args) -> dict:
    """The adjoint form in a child process. It works on torch tensors, and torch must load its HIP
    runtime before this library does, which this process (already stepping networks) cannot undo."""
    import subprocess
    cmd = [sys.executable, str(ROOT / "bench.py"), "--form", "adjoint", "--steps", str(args.steps)]
    if args.no_cpu_baseline:
        cmd.append("--no-cpu-baseline")
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if res.returncode != 0:
        raise RuntimeError("adjoint form failed:\n" + res.stderr[-4000:])
    return json.loads(res.stdout.strip().splitlines()[-1])


# --------------------------------------------------------------------------- CPU baselines

def scaled_config2(n: int, ticks: int, stdp: bool = False):
    """The 1M-neuron workload's law at n neurons (host-built): n/100 input neurons, the ignition
    inject, the same thresholds and weights."""
    import numpy as np
    from snncuda_amd import workloads as W
    net = W.ei_network(n, W.lognormal_fanout(0, n), 0, n_sources=max(1, n // 100), source_fanout=1000,
                       p_source=0.005, ticks=ticks, threshold=threshold_of(stdp), delay_max=32,
                       compartments=(W.BASAL, W.APICAL), plastic="exc" if stdp else False, w_inh=5.0)
    hit = np.nonzero(W.unif24(W.hash3(0, 40, np.arange(n, dtype=np.uint64))) < np.float32(0.02))[0]
    net.commands = [(0, int(i), 0, threshold_of(stdp)) for i in hit] + net.commands
    return net


def time_cpu(net, kind: str, label: str) -> dict:
    """One thread on this host: the oracle (kind 'port') or the reference runtime (kind
    'reference', oracle/_ref/ref_harness), the whole command list from a network file; synaptic
    events/s = (delivered - injects) / time inside run_until."""
    import oracle
    n_inj = sum(1 for c in net.commands if c[0] == 0)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "net.bin")
        oracle.write_net(path, net)
        if kind == "port":
            # the command list as the reference harness drives it, timing the run_until calls
            o = oracle.Oracle(net)
            sec = 0.0
            for k, a, t, w in net.commands:
                if k == 0:
                    o.inject(a, t, w)
                else:
                    t0 = time.perf_counter()
                    o.run_until(t)
                    sec += time.perf_counter() - t0
            delivered = o.delivered
        else:
            out = oracle.run_reference_file(path, os.path.join(d, "out.bin"), time_only=True)
            kv = dict(x.split("=") for x in out.split()[1:])
            delivered, sec = int(kv["delivered"]), float(kv["run_seconds"])
    ev = delivered - n_inj
    return {"value": ev / sec, "unit": "synaptic events/s", "cores": 1, "kind": kind,
            "sample": f"{label}: {net.n_neurons} neurons, {net.n_synapses} synapses, {ev} synaptic events in "
                      f"{sec:.2f} s, 1 thread ({'oracle/snn_oracle.c' if kind == 'port' else 'the reference runtime, oracle/_ref/ref_harness'})"}


def cpu_baselines(args):
    """(headline, others): the oracle on the 1M-neuron law at --cpu-neurons; the oracle and the
    reference runtime on config 0, config 1 reduced and a 5,000-neuron sample of the 1M-neuron law."""
    import oracle
    from snncuda_amd import workloads as W
    head = time_cpu(scaled_config2(args.cpu_neurons, args.cpu_ticks), "port",
                    f"the 1M-neuron workload's law at {args.cpu_neurons} neurons, {args.cpu_ticks} ticks incl. ignition")
    samples = [
        ("config 0 (1,000 neurons, fan-out 100, STDP on, threshold 1.1), 1,000 ticks", W.config0(0, threshold=1.1)),
        ("config 1 reduced (2,000 neurons, fan-out 100, 4 code offsets, STDP on), 300 ticks", W.config1_reduced(0)),
        ("the 1M-neuron workload's law at 5,000 neurons, 100 ticks incl. ignition", scaled_config2(5000, 100)),
    ]
    others = []
    have_ref = oracle.REF_HARNESS.exists()
    for label, net in samples:
        others.append(time_cpu(net, "port", label))
        if have_ref:
            others.append(time_cpu(net, "reference", label))
    if not have_ref:
        others.append({"kind": "reference", "sample": "oracle/_ref/ref_harness not built on this host"})
    return head, others


# --------------------------------------------------------------------------- main

def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the step shards by target neuron "
              "across GPUs by design, but no multi-GPU path is built or measured", file=sys.stderr)
        sys.exit(2)
    if args.form == "batched":
        line = batched(args)
        line.update(n_gpus=1, steps=args.batch_ticks, warmup=0, higher_is_better=True, scaling="weak",
                    vs_baseline=None, dtype="fp32", data="synthetic (seeded generator, built on the device)")
        print(json.dumps(line), flush=True)
        return
    if args.form == "spikecode":
        line = spikecode(args)
        line.update(n_gpus=1, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="fp32",
                    data="synthetic (seeded generator, built on the device)")
        print(json.dumps(line), flush=True)
        return
    if args.form == "adjoint":
        import torch  # noqa: F401  (torch's HIP runtime first; the library then uses it as well)
        line = adjoint_form(args, not args.no_cpu_baseline)
        line.update(n_gpus=1, warmup=3, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="fp32",
                    data="synthetic (seeded generator)")
        print(json.dumps(line), flush=True)
        return
    stdp = args.form == "stdp"
    t_all = time.perf_counter()
    if args.form == "dense":
        r = plain(args, False, DENSE_THRESHOLD, DENSE_SETTLE_TICKS)
    else:
        r = plain(args, stdp)
    log(f"plain{' stdp' if stdp else ''}: {r['value'] / 1e9:.2f} G events/s, {r['ms_per_step']:.4f} ms/tick, "
        f"{r['events_per_tick']:.3e} events/tick, fold {r['roofline']['mean_launch_ms'] * 1e3:.1f} us")
    line = plain_line(args, r, stdp)
    if args.form == "all":
        s = plain(args, True)
        log(f"stdp: {s['value'] / 1e9:.2f} G events/s, {s['ms_per_step']:.4f} ms/tick, "
            f"fold {s['roofline']['mean_launch_ms'] * 1e3:.1f} us")
        line["stdp_on"] = {k: s[k] for k in ("value", "ms_per_step", "events_per_tick", "fires_per_tick",
                                             "fire_fraction_per_tick", "threshold", "plastic_fraction", "roofline",
                                             "scatter_roofline", "step_roofline", "ticks_per_s")}
        line["stdp_on"]["unit"] = "synaptic events/s"
        line["stdp_on"]["workload"] = ("the same network with every synapse of excitatory and input neurons plastic, "
                                       "threshold 2.75, same settle/warmup/timed ticks")
        if s.get("kernel_ms_per_tick"):
            line["stdp_on"]["kernel_ms_per_tick"] = s["kernel_ms_per_tick"]
        d = plain(args, False, DENSE_THRESHOLD, DENSE_SETTLE_TICKS)
        log(f"dense: {d['value'] / 1e9:.2f} G events/s, {d['ms_per_step']:.4f} ms/tick, "
            f"{d['events_per_tick']:.3e} events/tick, fold {d['roofline']['mean_launch_ms'] * 1e3:.1f} us")
        line["dense_regime"] = {k: d[k] for k in ("value", "ms_per_step", "events_per_tick", "fires_per_tick",
                                                  "fire_fraction_per_tick", "threshold", "settle_ticks",
                                                  "step_roofline", "ticks_per_s")}
        # the traffic files were measured in the headline regime: no traffic figure here
        line["dense_regime"]["roofline"] = dict(d["roofline"], traffic=None)
        line["dense_regime"]["unit"] = "synaptic events/s"
        line["dense_regime"]["workload"] = (
            f"the same network (STDP off) at threshold {DENSE_THRESHOLD}: about 0.9% of the neurons fire per "
            f"tick once settled, {DENSE_SETTLE_TICKS} settle ticks")
        line["forms"] = {"batched": batched(args), "adjoint": adjoint_child(args), "spikecode": spikecode(args)}
    if not args.no_cpu_baseline:
        head, others = cpu_baselines(args)
        line["cpu_baseline"] = head
        line["cpu_baselines"] = others
    line["bench_seconds"] = time.perf_counter() - t_all
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
