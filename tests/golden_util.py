"""Loading the golden fixtures and comparing a run against them."""
from __future__ import annotations

import hashlib
import os
import pathlib
import tempfile

import numpy as np

import oracle
from snncuda_amd import workloads as W
from tests import kat_nets as K

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"

CASES = {name: fn for name, fn in K.ALL.items()}
CASES["config0"] = lambda: W.config0(0, threshold=1.1)
CASES["config0_stdp_off"] = lambda: W.config0(0, threshold=1.1, plastic=False)
CASES["config0_burst"] = lambda: W.config0(0, threshold=1.0, ticks=262)
CASES["config1_reduced"] = lambda: W.config1_reduced(0)
for _s in range(4):
    CASES[f"feature{_s}"] = (lambda s=_s: W.feature_net(s))


def load(name: str):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


def digest(net) -> str:
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "net.bin")
        oracle.write_net(p, net)
        return hashlib.sha256(pathlib.Path(p).read_bytes()).hexdigest()


def raster_counts(tick_fired: np.ndarray, fires_in_order: np.ndarray) -> np.ndarray:
    """(tick, neuron, count) rows sorted by tick then neuron, from fires in processing order."""
    ticks = np.repeat(np.arange(tick_fired.shape[0], dtype=np.int64), tick_fired.astype(np.int64))
    if ticks.size == 0:
        return np.zeros((0, 3), np.uint32)
    key = ticks * (1 << 32) + fires_in_order.astype(np.int64)
    u, c = np.unique(key, return_counts=True)
    return np.stack([(u >> 32).astype(np.uint32), (u & 0xFFFFFFFF).astype(np.uint32), c.astype(np.uint32)], axis=1)


def mismatches(g: dict, ticks_d, ticks_f, fires, neurons: dict, synapses: dict, pre_count=True):
    """Names of the golden fields the run does not reproduce bit-exactly."""
    bad = []
    if not np.array_equal(g["delivered_per_tick"], ticks_d):
        bad.append("delivered_per_tick")
    if not np.array_equal(g["fired_per_tick"], ticks_f):
        bad.append("fired_per_tick")
    if fires is not None and not np.array_equal(g["raster"], raster_counts(np.asarray(ticks_f), np.asarray(fires))):
        bad.append("raster")
    for k in ("membrane", "current", "spike_count", "t_match", "t_learned"):
        if k in neurons and not np.array_equal(g[k].view(np.uint32) if g[k].dtype == np.float32 else g[k],
                                               neurons[k].view(np.uint32) if neurons[k].dtype == np.float32 else neurons[k].astype(g[k].dtype)):
            bad.append("neuron." + k)
    if "last_fired" in neurons and not np.array_equal(g["last_fired"], neurons["last_fired"]):
        bad.append("neuron.last_fired")
    if "tau" in neurons and not np.array_equal(g["tau"], neurons["tau"]):
        bad.append("neuron.tau")
    keys = ["weight", "c_match", "c_learned"] + (["pre_count"] if pre_count else [])
    keys += [k for k in ("upd_count", "last_pre", "last_post") if k in synapses]
    for k in keys:
        a, b = g[k], synapses[k]
        if a.dtype == np.float32:
            same = np.array_equal(a.view(np.uint32), b.view(np.uint32))
        else:
            same = np.array_equal(a, b.astype(a.dtype))
        if not same:
            bad.append("synapse." + k)
    return bad
