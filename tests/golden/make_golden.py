"""Write the golden fixtures: the reference's own CPU runtime (oracle/_ref/ref_harness, compiled
from the reference sources by oracle/Makefile) run on each case, reduced to compressed .npz files.

Run in the build container (where /root/reference exists):  python tests/golden/make_golden.py
Each fixture stores a SHA-256 of the exact SNNNET01 input, so a change in a generator is caught.
"""
from __future__ import annotations

import hashlib
import os
import pathlib
import sys
import tempfile

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from snncuda_amd import workloads as W  # noqa: E402
from tests import kat_nets as K  # noqa: E402

HERE = pathlib.Path(__file__).resolve().parent
SYNAPSE_WORDS_KEPT = 100_000


def cases():
    out = {name: fn for name, fn in K.ALL.items()}
    out["config0"] = lambda: W.config0(0, threshold=1.1)
    out["config0_stdp_off"] = lambda: W.config0(0, threshold=1.1, plastic=False)
    out["config0_burst"] = lambda: W.config0(0, threshold=1.0, ticks=262)
    out["config1_reduced"] = lambda: W.config1_reduced(0)
    for s in range(4):
        out[f"feature{s}"] = (lambda s=s: W.feature_net(s))
    return out


def net_digest(net) -> str:
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "net.bin")
        oracle.write_net(p, net)
        return hashlib.sha256(pathlib.Path(p).read_bytes()).hexdigest()


def save(name, net, res: "oracle.Results") -> None:
    n = res.neuron
    s = res.synapse
    np.savez_compressed(
        HERE / f"{name}.npz",
        digest=np.array(net_digest(net)),
        delivered_per_tick=res.delivered_per_tick, fired_per_tick=res.fired_per_tick, raster=res.raster,
        membrane=n["membrane"], current=n["current"], tau=n["tau"], spike_count=n["spike_count"],
        last_fired=n["last_fired"], t_match=n["t_match"], t_learned=n["t_learned"],
        t_window_start=n["t_window_start"], t_n_offsets=n["t_n_offsets"],
        weight=s["weight"], pre_count=s["pre_count"], upd_count=s["upd_count"], last_pre=s["last_pre"],
        last_post=s["last_post"], c_match=s["c_match"], c_learned=s["c_learned"],
        delivered=np.array(res.delivered), fired=np.array(res.fired),
        # learned-pattern contents: every neuron's candidate offsets and learned patterns, and every
        # synapse's code offsets and learned code patterns, in the harness's word encoding (the
        # synapse words of large cases as a SHA-256, to keep the fixtures small)
        neuron_words=res.neuron_words,
        **({"synapse_words": res.synapse_words} if res.synapse_words.size <= SYNAPSE_WORDS_KEPT else
           {"synapse_words_sha256": np.array(hashlib.sha256(res.synapse_words.tobytes()).hexdigest())}))


def main(names=None):
    oracle.build()
    for name, fn in cases().items():
        if names and name not in names:
            continue
        net = fn()
        res = oracle.run_reference(net, patterns=True)
        save(name, net, res)
        print(f"{name}: delivered={res.delivered} fired={res.fired}")


if __name__ == "__main__":
    main(sys.argv[1:])
