#!/usr/bin/env python3
"""Benchmark of the per-tick spiking step on one MI355X.

Workload (default): the 1,000,000-neuron network with lognormal fan-out (median 1,500, sigma 0.8,
cap 20,000; ~2.07e9 synapses), 80/20 E/I, delays U{1..32}, Basal/Apical compartments, one code
offset per synapse, 10,000 Poisson input neurons at p=0.005 with fan-out 1,000, STDP off
(``--stdp`` for on). Pinned free parameters: threshold 2.5, excitatory weights U[0,0.5),
inhibitory GABA_A weights U[0,5), and a 2% ignition inject at tick 0; this keeps ~0.5% of the
neurons firing per tick (~1e7 delivered events per tick) for the whole run.

A step is one tick. ``value`` is delivered synaptic events per second over the timed ticks (input
injects excluded), with the network generated on the device and resident in HBM before timing.

The JSON line also carries:
  roofline      the dominant kernel (partition scatter) — algorithmic bytes per launch / its mean
                duration from HIP events on the step's stream during a second pass of the same
                length; traffic = FETCH_SIZE + WRITE_SIZE per dispatch from rocprofv3 --pmc
                (profiles/), scaled by the copy-kernel calibration factor stated there.
  step_roofline the whole tick: 50 B per delivered event (the plain-step byte model) / tick time.
  cpu_baseline  the oracle (plain-C restatement of the reference CPU step), one thread on this
                host, on a bounded sample of the same workload law (20,000 neurons).
"""
from __future__ import annotations

import argparse
import json
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
# the partition-scatter kernel performs the static edge reads and the code-pattern RMW
SCATTER_BYTES_PER_EVENT = 10.0 + 32.0
# STDP on adds 12 B per event (pre-side record) and 20 B per (fire, incoming plastic synapse)
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
    ap.add_argument("--cpu-neurons", type=int, default=20_000)
    ap.add_argument("--cpu-ticks", type=int, default=200)
    return ap.parse_args()


def cpu_baseline(args):
    """Oracle, one thread, on the same law at --cpu-neurons neurons."""
    import numpy as np
    import oracle
    from snncuda_amd import workloads as W
    n = args.cpu_neurons
    net = W.ei_network(n, W.lognormal_fanout(0, n), 0, n_sources=max(1, n // 100), source_fanout=1000,
                       p_source=0.005, ticks=args.cpu_ticks, threshold=2.5, delay_max=32,
                       compartments=(W.BASAL, W.APICAL), plastic=args.stdp, w_inh=5.0)
    hit = np.nonzero(W.unif24(W.hash3(0, 40, np.arange(n, dtype=np.uint64))) < np.float32(0.02))[0]
    cmds = [(0, int(i), 0, 2.5) for i in hit] + net.commands
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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the step shards by target neuron "
              "across GPUs by design, but no multi-GPU path is built or measured", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import snncuda_amd as sa
    from snncuda_amd import workloads as W

    total = args.warmup + 2 * args.steps
    spec, cmds = W.config2_spec(0, n=args.neurons, ticks=total, threshold=2.5, w_inh=5.0, kick=0.02,
                                plastic=args.stdp)
    injects = [c for c in cmds if c[0] == 0]
    inj_ticks = np.array([t for _, _, t, _ in injects], dtype=np.int64)
    with sa.Network(spec=spec, max_events_per_tick=1 << 27, fire_log_capacity=1 << 26,
                    raster_capacity=1 << 20, track_pre_count=False) as dev:
        dev.run_commands(injects)
        dev.run_until(args.warmup - 1)
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
    scat_ms, scat_n = kt["scatter"]
    scat_avg_s = scat_ms / max(scat_n, 1) / 1e3
    ev_per_launch = float(dpt2.mean())
    achieved = SCATTER_BYTES_PER_EVENT * ev_per_launch / scat_avg_s / 1e9
    traffic = None
    tr_file = ROOT / "profiles" / "scatter_traffic.json"
    if tr_file.exists():
        tr = json.loads(tr_file.read_text())
        if tr.get("stdp", False) == args.stdp:
            # counters scaled by the copy-kernel calibration (profiles/copy_calibration.txt), per
            # delivered event, times this run's events per launch
            traffic = tr["bytes_per_event"] * ev_per_launch
    ms_per_step = elapsed / args.steps * 1e3
    step_gbs = STEP_BYTES_PER_EVENT * events / elapsed / 1e9
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
                        f"delays U{{1..32}}, 2 compartments, codes {{0}}, STDP {'on' if args.stdp else 'off'}",
            "neurons": args.neurons + spec["n_sources"],
            "synapses": int(E),
            "events_per_tick": float(dpt.mean()),
            "fires_per_tick": float(fpt.mean()),
            "fire_fraction_per_tick": float(fpt.mean()) / args.neurons,
            "threshold": 2.5,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "kernel": "k_pscatter (partition scatter)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_event": SCATTER_BYTES_PER_EVENT,
            "mean_launch_ms": scat_avg_s * 1e3,
        },
        "step_roofline": {
            "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
            "bytes_per_event": STEP_BYTES_PER_EVENT,
        },
        "kernel_ms_per_tick": {k: v[0] / max(v[1], 1) for k, v in kt.items() if v[1]},
        "ticks_per_s": args.steps / elapsed,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
