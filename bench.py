#!/usr/bin/env python3
"""Benchmark of the per-tick spiking step on one MI355X.

Workload (default): the 1,000,000-neuron network with lognormal fan-out (median 1,500, sigma 0.8,
cap 20,000; ~2.07e9 synapses), 80/20 E/I, delays U{1..32}, Basal/Apical compartments, one code
offset per synapse, 10,000 Poisson input neurons at p=0.005 with fan-out 1,000, STDP off
(``--stdp`` for on, on every synapse but those of inhibitory neurons). Pinned free parameters:
threshold 2.5 (2.75 with STDP), excitatory weights U[0,0.5),
inhibitory GABA_A weights U[0,5), and a 2% ignition inject at tick 0; this keeps ~0.5% of the
neurons firing per tick (~1e7 delivered events per tick) for the whole run.

A step is one tick. ``value`` is delivered synaptic events per second over the timed ticks (input
injects excluded), with the network generated on the device and resident in HBM before timing.

The JSON line also carries:
  roofline      the dominant kernel (the partition fold) — algorithmic bytes per launch / its mean
                duration from HIP events on the step's stream during a second pass of the same
                length; traffic = FETCH_SIZE + WRITE_SIZE per dispatch from rocprofv3 --pmc
                (profiles/), scaled by the copy-kernel calibration factor stated there.
  scatter_roofline  the same for the second kernel (the partition scatter).
  step_roofline the whole tick: 50 B per delivered event (the plain-step byte model) / tick time.
  cpu_baseline  the oracle (plain-C restatement of the reference CPU step), one thread on this
                host, on a bounded sample of the same workload law (20,000 neurons).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# bytes per delivered event, plain step (STDP off), by the per-event byte model:
#   static edge fields 10 B (target, weight, delay, compartment/receptor) + event queue 8 B
#   + synapse code-pattern record read and write 32 B
STEP_BYTES_PER_EVENT = 50.0
# partition scatter: reads the edge's target u32, weight f32 and compartment/receptor u8, writes the
# bucket record (8 B event + 2 B target index); code patterns are advanced per source neuron
SCATTER_BYTES_PER_EVENT = 9.0 + 10.0
# partition fold: reads the bucket record (10 B) per event, and per touched neuron its parameters
# (16 B), dynamic state (32 B) and temporal header (32 B), writing back state and header, and the
# candidate's last run (4 B read, 4 B written)
FOLD_BYTES_PER_EVENT = 10.0
FOLD_BYTES_PER_NEURON = 16.0 + 2 * 32.0 + 2 * 32.0 + 8.0
# with STDP the bucket record is 12 + 2 B and a plastic event reads and writes its 32-B STDP record
FOLD_STDP_BYTES_PER_EVENT = 14.0
FOLD_STDP_RECORD_BYTES = 2 * 32.0
# STDP on adds 12 B per plastic event (pre-side record) and 20 B per (fire, incoming plastic synapse)
STDP_EVENT_BYTES = 12.0
STDP_POST_BYTES = 20.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed ticks")
    ap.add_argument("--warmup", type=int, default=200, help="untimed ticks before timing")
    ap.add_argument("--neurons", type=int, default=1_000_000)
    ap.add_argument("--stdp", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", type=int, default=0,
                    help="plain form: tune the launch schedule on the warm-up ticks with trials of this many ticks")
    ap.add_argument("--schedule", default="",
                    help="plain form: partition_shift,scatter_workgroups,graph_ticks[,fold_threads]")
    ap.add_argument("--cpu-neurons", type=int, default=20_000)
    ap.add_argument("--cpu-ticks", type=int, default=200)
    ap.add_argument("--form", choices=("plain", "batched", "adjoint"), default="plain",
                    help="plain: the 1M-neuron step (default); batched: 512 uneven networks in one launch "
                         "sequence vs looping the plain step over the same networks; adjoint: forward + "
                         "backward of the fixed-spike adjoint on the 10,000-neuron workload")
    ap.add_argument("--nets", type=int, default=512, help="batched form: networks")
    ap.add_argument("--batch-ticks", type=int, default=2000, help="batched form: ticks per network")
    ap.add_argument("--loop-nets", type=int, default=0,
                    help="batched form: networks in the looped baseline (0: all of them)")
    return ap.parse_args()


def threshold_of(args) -> float:
    # STDP on potentiates the recurrent excitation: at 2.5 the 1M-neuron network runs away within
    # 500 ticks, at 2.75 it holds ~0.37% of the neurons firing per tick for 2,200 ticks
    return 2.75 if args.stdp else 2.5


def cpu_baseline(args):
    """Oracle, one thread, on the same law at --cpu-neurons neurons."""
    import numpy as np
    import oracle
    from snncuda_amd import workloads as W
    n = args.cpu_neurons
    net = W.ei_network(n, W.lognormal_fanout(0, n), 0, n_sources=max(1, n // 100), source_fanout=1000,
                       p_source=0.005, ticks=args.cpu_ticks, threshold=threshold_of(args), delay_max=32,
                       compartments=(W.BASAL, W.APICAL), plastic="exc" if args.stdp else False, w_inh=5.0)
    hit = np.nonzero(W.unif24(W.hash3(0, 40, np.arange(n, dtype=np.uint64))) < np.float32(0.02))[0]
    cmds = [(0, int(i), 0, threshold_of(args)) for i in hit] + net.commands
    o = oracle.Oracle(net)
    n_inj = 0
    for k, a, t, w in cmds:
        if k == 0:
            o.inject(a, t, w)
            n_inj += 1
    t0 = time.perf_counter()
    o.run_until(args.cpu_ticks - 1)
    dt = time.perf_counter() - t0
    ev = o.delivered - n_inj
    return {"value": ev / dt, "unit": "synaptic events/s", "cores": 1, "kind": "port",
            "sample": f"oracle/snn_oracle.c on the same law at {n} neurons ({net.n_synapses} synapses), "
                      f"{args.cpu_ticks} ticks incl. ignition, {ev} events in {dt:.1f} s, 1 thread"}


def batched(args):
    """512 independent networks of the batched workload (log-uniform sizes 500..50,000, fan-out
    U{50..500}, 2 code offsets, 100 Poisson sources each, STDP on, threshold 3.0), generated on the
    device, stepped together for --batch-ticks ticks under the online autotuner; against looping the
    plain step over the same networks, once with stream launches and once with HIP graph replay.
    value = networks/s of the batched run (networks x their full run / wall time)."""
    import numpy as np
    import snncuda_amd as sa
    from snncuda_amd import workloads as W

    T = args.batch_ticks
    specs, cmds = zip(*[W.config3_spec(s, ticks=T) for s in range(args.nets)])
    n_total = sum(sp["n"] + sp["n_sources"] for sp in specs)
    e_total = sum(int(sp["fanout"].sum()) + sp["n_sources"] * sp["source_fanout"] for sp in specs)
    n_inj = sum(sum(1 for c in cm if c[0] == 0) for cm in cmds)
    print(f"batched: {args.nets} networks, {n_total} neurons, {e_total} synapses, {T} ticks", file=sys.stderr, flush=True)
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
    print(f"batched run {t_batch:.3f} s ({args.nets / t_batch:.1f} networks/s); schedule {schedule}",
          file=sys.stderr, flush=True)

    loop_nets = args.loop_nets or args.nets
    loops = {}
    for mode, graph in (("stream", 0), ("graph", 32)):
        total = 0.0
        for k in range(loop_nets):
            with sa.Network(spec=specs[k], track_pre_count=False, max_events_per_tick=1 << 22,
                            fire_log_capacity=1 << 20, raster_capacity=1 << 16) as net:
                net.set_schedule(graph_ticks=graph)
                net.run_commands([c for c in cmds[k] if c[0] == 0])
                t0 = time.perf_counter()
                net.run_until(T - 1)
                total += time.perf_counter() - t0
        loops[mode] = total * args.nets / loop_nets
        print(f"loop ({mode}) {loops[mode]:.3f} s for {args.nets} networks", file=sys.stderr, flush=True)
    best_loop = min(loops.values())
    value = args.nets / t_batch
    return {
        "metric": "networks/s (512 uneven networks, batched vs looped plain step)",
        "value": value,
        "unit": "networks/s",
        "n_gpus": 1,
        "steps": T,
        "warmup": 0,
        "ms_per_step": t_batch / T * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator, built on the device)",
        "config": {
            "workload": f"{args.nets} E/I networks, sizes log-uniform in [500, 50000], fan-out U{{50..500}}, "
                        "80/20 E/I, delays U{1..20}, 2 compartments, 2 code offsets U{0..7}, 100 Poisson "
                        "sources p=0.01 per network, STDP on, threshold 3.0",
            "neurons": n_total,
            "synapses": e_total,
            "ticks": T,
            "events_per_tick": (delivered - n_inj) / T,
            "fires_per_tick": fired / T,
            "parallelism": "1 GPU",
        },
        "speedup_vs_loop": best_loop / t_batch,
        "loop_seconds": loops,
        "loop_networks_timed": loop_nets,
        "batched_seconds": t_batch,
        "generate_seconds": t_gen,
        "events_per_s": (delivered - n_inj) / t_batch,
        "schedule": schedule,
        "autotune": tune.strip().split("\n"),
    }


def adjoint_form(args):
    """Fixed-spike adjoint on the 10,000-neuron workload (fan-in 500, 5e6 synapses, 500 ticks):
    one step = forward (trace + checkpoints) + backward (segment recompute + adjoint recursion) of
    the MSE loss against a perturbed-weight target. value = steps/s; the FP64 CPU adjoint of the
    oracle (one thread) is timed on the same event lists as the CPU baseline."""
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

    for _ in range(args.warmup if args.warmup < 50 else 3):
        step()
    K = min(args.steps, 50)
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
        "n_gpus": 1,
        "steps": K,
        "warmup": 3,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "10,000 neurons, fan-in 500 (5e6 synapses), 500 ticks, 2 compartments, checkpoints "
                               "every 64 ticks, MSE against a target from weights + N(0, 0.05)",
                   "events": ev, "fires": int(run.fired), "parallelism": "1 GPU"},
        "events_per_s": 3 * ev / dt,
        "roofline": {"bound": "hbm", "achieved": bytes_step / dt / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes_step / dt / 1e9 / HBM_PEAK_GBS, "traffic": None, "bytes_per_step": bytes_step},
        "build_seconds": t_build,
    }
    if not args.no_cpu_baseline:
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
    return line


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the step shards by target neuron "
              "across GPUs by design, but no multi-GPU path is built or measured", file=sys.stderr)
        sys.exit(2)
    if args.form == "batched":
        print(json.dumps(batched(args)), flush=True)
        return
    if args.form == "adjoint":
        print(json.dumps(adjoint_form(args)), flush=True)
        return
    import numpy as np
    import snncuda_amd as sa
    from snncuda_amd import workloads as W

    total = args.warmup + 2 * args.steps
    # STDP on: every synapse but those of inhibitory neurons is plastic (the rule caps weights at
    # 2.0, below the pinned inhibitory magnitudes U[0,5), so plastic inhibition would collapse)
    spec, cmds = W.config2_spec(0, n=args.neurons, ticks=total, threshold=threshold_of(args), w_inh=5.0, kick=0.02,
                                plastic="exc" if args.stdp else False)
    injects = [c for c in cmds if c[0] == 0]
    inj_ticks = np.array([t for _, _, t, _ in injects], dtype=np.int64)
    with sa.Network(spec=spec, max_events_per_tick=1 << 27, fire_log_capacity=1 << 26,
                    raster_capacity=1 << 20, track_pre_count=False) as dev:
        dev.run_commands(injects)
        if args.schedule:
            v = [int(x) for x in args.schedule.split(",")]
            kw = dict(partition_shift=v[0], scatter_workgroups=v[1], graph_ticks=v[2])
            if len(v) > 3:
                kw["fold_threads"] = v[3]
            dev.set_schedule(**kw)
        if args.autotune:
            dev.run_until(min(99, args.warmup - 1))  # past the ignition ramp first
            dev.autotune(args.autotune)
        dev.run_until(args.warmup - 1)
        if args.autotune:
            print(dev.tune_log(), file=sys.stderr, flush=True)
        # timed region: one synchronous call covering K ticks
        d0 = dev.delivered
        t0 = time.perf_counter()
        dev.run_until(args.warmup + args.steps - 1)
        t1 = time.perf_counter()
        d1 = dev.delivered
        n_inj = int(((inj_ticks >= args.warmup) & (inj_ticks < args.warmup + args.steps)).sum())
        events = d1 - d0 - n_inj
        elapsed = t1 - t0
        dpt, fpt = dev.tick_counts(args.warmup, args.steps)
        # second pass of the same length with HIP events around every kernel
        dev.set_profiling(True)
        dev.run_until(args.warmup + 2 * args.steps - 1)
        kt = dev.kernel_times()
        dpt2, _ = dev.tick_counts(args.warmup + args.steps, args.steps)
        E = dev.e
        schedule = dev.schedule()
    ev_per_launch = float(dpt2.mean())
    # synaptic targets are uniform over the network's neurons (the input neurons only receive
    # injects): the expected number touched by E events
    touched = args.neurons * (1.0 - math.exp(-ev_per_launch / args.neurons))

    def traffic_of(name):
        # counters scaled by the copy-kernel calibration (profiles/copy_calibration.txt), per
        # delivered event, times this run's events per launch (files *_stdp.json for STDP on)
        f = ROOT / "profiles" / (name.replace(".json", "_stdp.json") if args.stdp else name)
        if not f.exists():
            return None
        tr = json.loads(f.read_text())
        return tr["bytes_per_event"] * ev_per_launch if tr.get("stdp", False) == args.stdp else None

    def kernel_roofline(key, kernel, bytes_per_launch, traffic_file):
        ms, n = kt[key]
        avg_s = ms / max(n, 1) / 1e3
        achieved = bytes_per_launch / avg_s / 1e9
        return {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_of(traffic_file),
                "bytes_per_launch": bytes_per_launch, "mean_launch_ms": avg_s * 1e3}

    # plastic share of the synapses (excitatory and input neurons' synapses; targets uniform, so it
    # is also the share of the delivered events and the mean plastic in-degree over n)
    fan = np.asarray(spec["fanout"], np.int64)
    plastic_syn = int(fan[:spec["n_exc"]].sum()) + spec["n_sources"] * spec["source_fanout"] if args.stdp else 0
    pfrac = plastic_syn / E
    if args.stdp:
        fold_per_event = FOLD_STDP_BYTES_PER_EVENT + FOLD_STDP_RECORD_BYTES * pfrac
    else:
        fold_per_event = FOLD_BYTES_PER_EVENT
    fold_rl = kernel_roofline("fold", "k_pfold (partition fold)",
                              fold_per_event * ev_per_launch + FOLD_BYTES_PER_NEURON * touched,
                              "fold_traffic.json")
    fold_rl["bytes_model"] = {"per_event": fold_per_event, "per_touched_neuron": FOLD_BYTES_PER_NEURON,
                              "touched_neurons": touched, "plastic_fraction": pfrac}
    scat_rl = kernel_roofline("scatter", "k_pscatter (partition scatter)", SCATTER_BYTES_PER_EVENT * ev_per_launch,
                              "scatter_traffic.json")
    scat_rl["bytes_per_event"] = SCATTER_BYTES_PER_EVENT
    ms_per_step = elapsed / args.steps * 1e3
    # the survey's step model; with STDP the pre-side bytes per plastic event and the post-side bytes
    # per (fire, incoming plastic synapse), fires x mean plastic in-degree
    step_bytes = STEP_BYTES_PER_EVENT * events
    if args.stdp:
        fires = float(fpt.mean()) * args.steps
        step_bytes += STDP_EVENT_BYTES * pfrac * events + STDP_POST_BYTES * fires * plastic_syn / args.neurons
    step_gbs = step_bytes / elapsed / 1e9
    line = {
        "metric": "synaptic events/s (1M-neuron net, % HBM roofline)",
        "value": events / elapsed,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator, built on the device)",
        "config": {
            "workload": f"{args.neurons}-neuron lognormal fan-out net (median 1500, sigma 0.8, cap 20000), "
                        f"delays U{{1..32}}, 2 compartments, codes {{0}}, "
                        f"STDP {'on (excitatory synapses)' if args.stdp else 'off'}",
            "neurons": args.neurons + spec["n_sources"],
            "synapses": int(E),
            "events_per_tick": float(dpt.mean()),
            "fires_per_tick": float(fpt.mean()),
            "fire_fraction_per_tick": float(fpt.mean()) / args.neurons,
            "threshold": threshold_of(args),
            "parallelism": "1 GPU",
        },
        "roofline": fold_rl,
        "scatter_roofline": scat_rl,
        "step_roofline": {
            "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
            "bytes_per_event": step_bytes / max(events, 1),
        },
        "kernel_ms_per_tick": {k: v[0] / max(v[1], 1) for k, v in kt.items() if v[1]},
        "schedule": schedule,
        "ticks_per_s": args.steps / elapsed,
    }
    # the other two forms are measured by their own invocations (bench.py --form batched / adjoint,
    # a few minutes each); their last committed lines are attached for reference
    forms = {}
    for form, fname in (("batched", "batched_config3.json"), ("adjoint", "adjoint_config4.json")):
        f = ROOT / "profiles" / fname
        if f.exists():
            d = json.loads(f.read_text())
            forms[form] = {k: d.get(k) for k in ("metric", "value", "unit", "ms_per_step", "speedup_vs_loop",
                                                 "events_per_s") if k in d}
            forms[form]["source"] = f"profiles/{fname} (python bench.py --form {form})"
    if forms:
        line["forms"] = forms
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
