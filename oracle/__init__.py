"""ORACLE — test infrastructure, not product code.

Python access to the CPU checker: the plain-C restatement of the reference step
(``oracle/liboracle.so``, built from ``snn_oracle.c``) and, where it was built, the reference's own
runtime behind ``oracle/_ref/ref_harness``. Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this package.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import struct
import subprocess
import tempfile
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_HARNESS = HERE / "_ref" / "ref_harness"

_lib = None


def build() -> None:
    """Compile liboracle.so (and _ref/ when the reference sources are present)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        L.snno_create.restype = P
        L.snno_create.argtypes = [ctypes.c_uint64] + [P] * 4 + [ctypes.c_uint64] + [P] * 10
        L.snno_destroy.argtypes = [P]
        L.snno_inject.argtypes = [P, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float]
        L.snno_run_until.argtypes = [P, ctypes.c_uint64]
        for f in ("snno_delivered", "snno_fired", "snno_ticks_recorded"):
            getattr(L, f).restype = ctypes.c_uint64
            getattr(L, f).argtypes = [P]
        L.snno_read_tick_counts.argtypes = [P, P, P]
        L.snno_read_fires.argtypes = [P, P, P]
        L.snno_read_neurons.argtypes = [P] * 8
        L.snno_read_synapses.argtypes = [P] * 8
        L.snno_read_ltp.argtypes = [P, P]
        L.snno_set_event_log.argtypes = [P, ctypes.c_int]
        L.snno_event_log_size.argtypes = [P]
        L.snno_event_log_size.restype = ctypes.c_uint64
        L.snno_read_event_log.argtypes = [P] * 6
        L.adjo_create.restype = P
        L.adjo_create.argtypes = [ctypes.c_uint64] * 4 + [P] * 8
        L.adjo_destroy.argtypes = [P]
        L.adjo_event_count.argtypes = [P]
        L.adjo_event_count.restype = ctypes.c_uint64
        L.adjo_forward.argtypes = [P] * 5
        L.adjo_loss.argtypes = [P] * 5
        L.adjo_loss.restype = ctypes.c_double
        L.adjo_loss_target.argtypes = [P, ctypes.c_uint64] + [P] * 4
        L.adjo_loss_target.restype = ctypes.c_double
        L.adjo_grad.argtypes = [P] * 8
        L.adjo_grad.restype = ctypes.c_double
        U64 = ctypes.c_uint64
        L.roo_feedforward.argtypes = [U64, P, U64, P, P, P, U64, P, P, P, U64, P, P, P]
        L.roo_recurrent.argtypes = ([U64, P, U64, P, P, P, P, U64, P, P, P, U64, P, U64, P, ctypes.c_uint32,
                                     ctypes.c_float, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P, P, P])
        L.roo_feedforward.restype = ctypes.c_int
        L.roo_recurrent.restype = ctypes.c_int
        L.snno_run_file.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
        L.snno_run_file.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------- file formats

def write_net(path, net) -> None:
    """Write a workloads.Net as a SNNNET01 file (layout in oracle/netfile.h)."""
    cmds = np.zeros(len(net.commands), dtype=[("kind", "<u4"), ("neuron", "<u4"), ("tick", "<u8"),
                                               ("weight", "<f4"), ("pad", "<u4")])
    for i, (k, n, t, w) in enumerate(net.commands):
        cmds[i] = (k, n, t, w, 0)
    with open(path, "wb") as f:
        f.write(b"SNNNET01")
        f.write(struct.pack("<4Q", net.n_neurons, net.n_synapses, int(net.codes.shape[0]), len(net.commands)))
        for a, dt in ((net.threshold, "<f4"), (net.window, "<u8"), (net.sim_thr, "<f4"), (net.max_refs, "<u8"),
                      (net.src, "<u4"), (net.tgt, "<u4"), (net.weight, "<f4"), (net.max_weight, "<f4"),
                      (net.delay, "<u4"), (net.comp, "u1"), (net.receptor, "u1"), (net.plastic, "u1"),
                      (net.code_ptr, "<u8"), (net.codes, "<u4")):
            f.write(np.ascontiguousarray(a, dtype=dt).tobytes())
        f.write(cmds.tobytes())


@dataclass
class Results:
    """Contents of a SNNOUT01 file."""

    delivered_per_tick: np.ndarray
    fired_per_tick: np.ndarray
    raster: np.ndarray  # (tick, neuron, count) rows sorted by tick then neuron
    neuron: Dict[str, np.ndarray]
    synapse: Dict[str, np.ndarray]
    neuron_patterns: List[tuple]  # per neuron: (offsets, [learned patterns])
    synapse_patterns: List[tuple]
    delivered: int
    fired: int
    neuron_words: np.ndarray = None   # the encoded pattern blobs as written (see ref_harness.cpp Blob)
    synapse_words: np.ndarray = None


def _decode_blob(words: np.ndarray, count: int) -> List[tuple]:
    out, i = [], 0
    w = words.tolist()
    for _ in range(count):
        n = w[i]; offs = w[i + 1:i + 1 + n]; i += 1 + n
        k = w[i]; i += 1
        pats = []
        for _ in range(k):
            m = w[i]; pats.append(w[i + 1:i + 1 + m]); i += 1 + m
        out.append((offs, pats))
    return out


def read_results(path, patterns: bool = True) -> Results:
    data = pathlib.Path(path).read_bytes()
    assert data[:8] == b"SNNOUT01", "bad result file"
    n, e, nt, nrec = struct.unpack_from("<4Q", data, 8)
    off = 40

    def take(dt, count):
        nonlocal off
        a = np.frombuffer(data, dtype=dt, count=count, offset=off).copy()
        off += a.nbytes
        return a

    dpt = take("<u8", nt)
    fpt = take("<u8", nt)
    raster = take("<u4", 4 * nrec).reshape(-1, 4)[:, :3]
    neuron = {"membrane": take("<f4", n)}
    neuron["current"] = take("<f4", 4 * n).reshape(4, n)
    neuron["tau"] = take("<f4", 4 * n).reshape(4, n)
    neuron["last_update"] = take("<u8", 4 * n).reshape(4, n)
    for k in ("last_active", "last_fired", "spike_count", "t_window_start", "t_n_offsets", "t_match", "t_learned"):
        neuron[k] = take("<i8" if k == "last_fired" else "<u8", n)
    neuron["t_last_match"] = take("u1", n)
    synapse = {"weight": take("<f4", e), "last_delta": take("<f4", e)}
    for k in ("last_pre", "last_post"):
        synapse[k] = take("<i8", e)
    for k in ("pre_count", "upd_count", "c_window_start", "c_n_offsets", "c_match", "c_learned"):
        synapse[k] = take("<u8", e)
    synapse["c_last_match"] = take("u1", e)
    nw = struct.unpack_from("<Q", data, off)[0]; off += 8
    nwords = take("<u4", nw)
    sw = struct.unpack_from("<Q", data, off)[0]; off += 8
    swords = take("<u4", sw)
    delivered, fired = struct.unpack_from("<2Q", data, off)
    return Results(dpt, fpt, raster, neuron, synapse,
                   _decode_blob(nwords, n) if patterns else [],
                   _decode_blob(swords, e) if patterns else [], delivered, fired, nwords, swords)


def run_oracle_file(net_path, out_path) -> float:
    sec = ctypes.c_double(0.0)
    rc = lib().snno_run_file(str(net_path).encode(), str(out_path).encode() if out_path else None, ctypes.byref(sec))
    if rc != 0:
        raise RuntimeError(f"oracle run failed: {rc}")
    return sec.value


def run_reference_file(net_path, out_path, time_only: bool = False) -> str:
    if not REF_HARNESS.exists():
        raise FileNotFoundError("oracle/_ref/ref_harness is not built (reference sources absent)")
    args = [str(REF_HARNESS), str(net_path), str(out_path)] + (["--time"] if time_only else [])
    return subprocess.run(args, check=True, capture_output=True, text=True).stdout


def run_oracle(net, patterns: bool = True) -> Results:
    with tempfile.TemporaryDirectory() as d:
        write_net(os.path.join(d, "net.bin"), net)
        run_oracle_file(os.path.join(d, "net.bin"), os.path.join(d, "out.bin"))
        return read_results(os.path.join(d, "out.bin"), patterns)


def run_reference(net, patterns: bool = True) -> Results:
    with tempfile.TemporaryDirectory() as d:
        write_net(os.path.join(d, "net.bin"), net)
        run_reference_file(os.path.join(d, "net.bin"), os.path.join(d, "out.bin"))
        return read_results(os.path.join(d, "out.bin"), patterns)


class Oracle:
    """In-process oracle mirroring NetworkPropagator's inject / run_until / readers."""

    def __init__(self, net, log_events: bool = False):
        L = lib()
        self._keep = [np.ascontiguousarray(a) for a in (
            net.threshold, net.window, net.sim_thr, net.max_refs, net.src, net.tgt, net.weight,
            net.max_weight, net.delay, net.comp, net.receptor, net.plastic, net.code_ptr, net.codes)]
        k = self._keep
        self.n, self.e = net.n_neurons, net.n_synapses
        self.h = L.snno_create(self.n, _p(k[0]), _p(k[1]), _p(k[2]), _p(k[3]), self.e, *[_p(a) for a in k[4:]])
        if not self.h:
            raise ValueError("oracle rejected the network")
        if log_events:
            L.snno_set_event_log(self.h, 1)

    def event_log(self) -> Dict[str, np.ndarray]:
        """Every processed event in processing order: tick, target, synapse (2**32-1: external),
        weight, fired (the target fired on this event)."""
        m = int(lib().snno_event_log_size(self.h))
        o = dict(tick=np.zeros(m, np.uint64), target=np.zeros(m, np.uint32), synapse=np.zeros(m, np.uint32),
                 weight=np.zeros(m, np.float32), fired=np.zeros(m, np.uint8))
        lib().snno_read_event_log(self.h, *[_p(o[k]) for k in ("tick", "target", "synapse", "weight", "fired")])
        return o

    def __del__(self):
        if getattr(self, "h", None):
            lib().snno_destroy(self.h)
            self.h = None

    def inject(self, neuron: int, tick: int, weight: float) -> None:
        lib().snno_inject(self.h, neuron, tick, weight)

    def run_until(self, tick: int) -> None:
        lib().snno_run_until(self.h, tick)

    def run_commands(self, commands) -> None:
        for k, n, t, w in commands:
            if k == 0:
                self.inject(n, t, w)
            else:
                self.run_until(t)

    @property
    def delivered(self) -> int:
        return int(lib().snno_delivered(self.h))

    @property
    def fired(self) -> int:
        return int(lib().snno_fired(self.h))

    def tick_counts(self):
        nt = int(lib().snno_ticks_recorded(self.h))
        d = np.zeros(nt, np.uint64); f = np.zeros(nt, np.uint64)
        lib().snno_read_tick_counts(self.h, _p(d), _p(f))
        return d, f

    def fires(self):
        nf = self.fired
        t = np.zeros(nf, np.uint64); n = np.zeros(nf, np.uint32)
        lib().snno_read_fires(self.h, _p(t), _p(n))
        return t, n

    def neurons(self) -> Dict[str, np.ndarray]:
        n = self.n
        out = dict(membrane=np.zeros(n, np.float32), current=np.zeros((4, n), np.float32),
                   tau=np.zeros((4, n), np.float32), spike_count=np.zeros(n, np.uint64),
                   last_fired=np.zeros(n, np.int64), t_match=np.zeros(n, np.uint64),
                   t_learned=np.zeros(n, np.uint64))
        lib().snno_read_neurons(self.h, *[_p(out[k]) for k in ("membrane", "current", "tau", "spike_count",
                                                                 "last_fired", "t_match", "t_learned")])
        return out

    def synapses(self) -> Dict[str, np.ndarray]:
        e = self.e
        out = dict(weight=np.zeros(e, np.float32), pre_count=np.zeros(e, np.uint64),
                   upd_count=np.zeros(e, np.uint64), last_pre=np.zeros(e, np.int64),
                   last_post=np.zeros(e, np.int64), c_match=np.zeros(e, np.uint64),
                   c_learned=np.zeros(e, np.uint64))
        lib().snno_read_synapses(self.h, *[_p(out[k]) for k in ("weight", "pre_count", "upd_count", "last_pre",
                                                                  "last_post", "c_match", "c_learned")])
        out["ltp"] = np.zeros(e, np.uint64)
        lib().snno_read_ltp(self.h, _p(out["ltp"]))
        return out


class AdjointOracle:
    """FP64 fixed-spike adjoint (adjoint_oracle.c) over an Oracle run's event log."""

    def __init__(self, net, ticks: int, log: Dict[str, np.ndarray]):
        self.n, self.e, self.T = net.n_neurons, net.n_synapses, int(ticks)
        self._keep = [np.ascontiguousarray(x) for x in (log["tick"], log["target"], log["synapse"], log["weight"],
                                                          log["fired"], net.comp, net.receptor, net.tgt)]
        k = self._keep
        self.h = lib().adjo_create(self.n, self.e, self.T, int(k[0].shape[0]), *[_p(a) for a in k])

    def __del__(self):
        if getattr(self, "h", None):
            lib().adjo_destroy(self.h)
            self.h = None

    @staticmethod
    def _d(x):
        return np.ascontiguousarray(x, dtype=np.float64)

    def forward(self, w, alpha, tau) -> np.ndarray:
        """Somatic traces [T, N]."""
        out = np.zeros((self.T, self.n), np.float64)
        w, alpha, tau = self._d(w), self._d(alpha), self._d(tau)
        lib().adjo_forward(self.h, _p(w), _p(alpha), _p(tau), _p(out))
        return out

    def loss(self, w, alpha, tau, target) -> float:
        w, alpha, tau, target = self._d(w), self._d(alpha), self._d(tau), self._d(target)
        return float(lib().adjo_loss(self.h, _p(w), _p(alpha), _p(tau), _p(target)))

    def loss_target(self, i, w, alpha, tau, target) -> float:
        w, alpha, tau, target = self._d(w), self._d(alpha), self._d(tau), self._d(target)
        return float(lib().adjo_loss_target(self.h, int(i), _p(w), _p(alpha), _p(tau), _p(target)))

    def grad(self, w, alpha, tau, target):
        """(loss, d/dw [E], d/dalpha [4], d/dtau [4])."""
        w, alpha, tau, target = self._d(w), self._d(alpha), self._d(tau), self._d(target)
        gw = np.zeros(self.e, np.float64)
        ga = np.zeros(4, np.float64)
        gt = np.zeros(4, np.float64)
        loss = lib().adjo_grad(self.h, _p(w), _p(alpha), _p(tau), _p(target), _p(gw), _p(ga), _p(gt))
        return float(loss), gw, ga, gt


# --------------------------------------------------------------------------- readouts

def _graph_arrays(net):
    return (np.ascontiguousarray(net.threshold, np.float32), np.ascontiguousarray(net.src, np.uint32),
            np.ascontiguousarray(net.tgt, np.uint32), np.ascontiguousarray(net.weight, np.float32),
            np.ascontiguousarray(net.delay, np.uint32))


def readout_feedforward(net, ptr, neuron, weight, readout):
    """oracle/readout_oracle.c's feedforward readout: (scores float32 [S, R], spikes uint64 [S, R])."""
    thr, src, tgt, w, _ = _graph_arrays(net)
    ptr = np.ascontiguousarray(ptr, np.uint64)
    neuron = np.ascontiguousarray(neuron, np.uint32)
    weight = np.ascontiguousarray(weight, np.float32)
    readout = np.ascontiguousarray(readout, np.uint32)
    S, R = len(ptr) - 1, len(readout)
    scores = np.zeros((S, R), np.float32)
    spikes = np.zeros((S, R), np.uint64)
    rc = lib().roo_feedforward(len(thr), _p(thr), len(src), _p(src), _p(tgt), _p(w), S, _p(ptr), _p(neuron),
                               _p(weight), R, _p(readout), _p(scores), _p(spikes))
    if rc:
        raise MemoryError("roo_feedforward failed")
    return scores, spikes


def readout_recurrent(net, ptr, neuron, weight, readout, hidden, *, iterations=1, feedback_decay=0.5,
                      top_k_feedback_readouts=1, top_k_hidden=0, use_scores_not_spikes=True):
    """oracle/readout_oracle.c's recurrent readout: (scores, spikes, time_count)."""
    thr, src, tgt, w, dl = _graph_arrays(net)
    ptr = np.ascontiguousarray(ptr, np.uint64)
    neuron = np.ascontiguousarray(neuron, np.uint32)
    weight = np.ascontiguousarray(weight, np.float32)
    readout = np.ascontiguousarray(readout, np.uint32)
    hidden = np.ascontiguousarray(hidden, np.uint32)
    S, R = len(ptr) - 1, len(readout)
    scores = np.zeros((S, R), np.float32)
    spikes = np.zeros((S, R), np.uint64)
    T = np.zeros(1, np.uint64)
    rc = lib().roo_recurrent(len(thr), _p(thr), len(src), _p(src), _p(tgt), _p(w), _p(dl), S, _p(ptr), _p(neuron),
                             _p(weight), R, _p(readout), len(hidden), _p(hidden), iterations, feedback_decay,
                             top_k_feedback_readouts, top_k_hidden, int(use_scores_not_spikes), _p(scores),
                             _p(spikes), _p(T))
    if rc:
        raise MemoryError("roo_recurrent failed")
    return scores, spikes, int(T[0])
