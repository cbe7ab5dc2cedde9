"""snncuda_amd — MI355X (gfx950) backend for SNNCuda's per-tick spiking step.

Python access to the C ABI in ``include/snncuda_hip.h`` (``libsnncuda_hip.so`` in this package
directory, built by ``__graft_entry__.build()``). Mirrors the reference's CPU operator interface
(``NetworkPropagator``: inject / run_until / counters / state readers,
reference ``include/snncuda/runtime/NetworkPropagator.h:14-47``) and its ctypes module
(``python/snncuda/__init__.py``: ``version()``, ``cuda_available()``).

There is no CPU fallback: without the built library every entry point raises
``SNNCudaLibraryError``.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
from typing import Dict, Optional, Sequence

import numpy as np

PKG = pathlib.Path(__file__).resolve().parent
LIB_NAME = "libsnncuda_hip.so"

SNN_OK, SNN_EINVAL, SNN_EDEVICE, SNN_ECAPACITY, SNN_ENODEVICE = 0, -1, -2, -3, -5


class SNNCudaLibraryError(RuntimeError):
    """The native library could not be loaded."""


class SNNError(RuntimeError):
    """A C-ABI call failed; ``code`` is the negative status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class CapacityError(SNNError):
    """A device capacity was exceeded during a run (results of that run are invalid)."""


_lib: Optional[ctypes.CDLL] = None


class _Desc(ctypes.Structure):
    _fields_ = [("n_neurons", ctypes.c_uint64), ("threshold", ctypes.c_void_p), ("window", ctypes.c_void_p),
                ("sim_thr", ctypes.c_void_p), ("max_refs", ctypes.c_void_p), ("n_synapses", ctypes.c_uint64),
                ("source", ctypes.c_void_p), ("target", ctypes.c_void_p), ("weight", ctypes.c_void_p),
                ("max_weight", ctypes.c_void_p), ("delay", ctypes.c_void_p), ("compartment", ctypes.c_void_p),
                ("receptor", ctypes.c_void_p), ("plastic", ctypes.c_void_p), ("code_ptr", ctypes.c_void_p),
                ("codes", ctypes.c_void_p)]


class Options(ctypes.Structure):
    _fields_ = [("max_events_per_tick", ctypes.c_uint64), ("fire_log_capacity", ctypes.c_uint64),
                ("raster_capacity", ctypes.c_uint64), ("track_pre_count", ctypes.c_int32),
                ("device", ctypes.c_int32), ("learned_pattern_slots", ctypes.c_uint64),
                ("pattern_device_pages", ctypes.c_uint64), ("pattern_runs", ctypes.c_uint64)]


class _EiDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("n_sources", ctypes.c_uint64), ("fanout", ctypes.c_void_p),
                ("source_fanout", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("n_exc", ctypes.c_uint64),
                ("w_exc", ctypes.c_float), ("w_inh", ctypes.c_float), ("delay_max", ctypes.c_uint32),
                ("n_comps", ctypes.c_uint32), ("comps", ctypes.c_uint8 * 4), ("plastic", ctypes.c_int32),
                ("threshold", ctypes.c_float), ("source_threshold", ctypes.c_float), ("window", ctypes.c_uint32),
                ("sim_thr", ctypes.c_float), ("max_refs", ctypes.c_uint32), ("n_codes", ctypes.c_uint32),
                ("code_max", ctypes.c_uint32)]


class Schedule(ctypes.Structure):
    _fields_ = [("partition_shift", ctypes.c_uint32), ("scatter_workgroups", ctypes.c_uint32),
                ("graph_ticks", ctypes.c_uint32), ("fold_threads", ctypes.c_uint32)]


class RecurrentOptions(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint32), ("feedback_decay", ctypes.c_float),
                ("top_k_feedback_readouts", ctypes.c_int32), ("top_k_hidden", ctypes.c_int32),
                ("use_scores_not_spikes", ctypes.c_int32)]


GRAPH_NONE = 0xFFFFFFFF  # a readout / hidden entry that is not a neuron of the connectome


def library_path() -> pathlib.Path:
    return pathlib.Path(os.environ.get("SNNCUDA_HIP_LIBRARY", PKG / LIB_NAME))


def _load() -> ctypes.CDLL:
    path = library_path()
    try:
        L = ctypes.CDLL(str(path))
    except OSError as err:
        raise SNNCudaLibraryError(
            f"cannot load {path}: {err}. Build it with `python -c 'import __graft_entry__ as g; g.build()'`."
        ) from err
    P, U64, I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    L.snncuda_version.restype = ctypes.c_char_p
    L.snncuda_cuda_available.restype = I
    L.snn_device_count.restype = I
    L.snn_last_error.restype = ctypes.c_char_p
    L.snn_net_create.argtypes = [ctypes.POINTER(_Desc), ctypes.POINTER(Options), ctypes.POINTER(P)]
    L.snn_net_create_ei.argtypes = [ctypes.POINTER(_EiDesc), ctypes.POINTER(Options), ctypes.POINTER(P)]
    L.snn_net_destroy.argtypes = [P]
    L.snn_net_inject.argtypes = [P, P, P, P, U64]
    L.snn_net_run_until.argtypes = [P, U64]
    L.snn_net_run_until_static.argtypes = [P, U64]
    L.snn_net_counts.argtypes = [P, P, P, P]
    L.snn_net_read_tick_counts.argtypes = [P, U64, U64, P, P]
    L.snn_net_read_raster.argtypes = [P, P, P, U64]
    L.snn_net_read_neurons.argtypes = [P] * 8
    L.snn_net_read_synapses.argtypes = [P] * 8
    L.snn_net_read_plasticity.argtypes = [P] * 3
    L.snn_net_get_schedule.argtypes = [P, ctypes.POINTER(Schedule)]
    L.snn_net_set_schedule.argtypes = [P, ctypes.POINTER(Schedule)]
    L.snn_net_autotune.argtypes = [P, ctypes.c_uint32]
    L.snn_net_tune_log.argtypes = [P, ctypes.c_char_p, U64]
    L.snn_batch_create.argtypes = [P, U64, ctypes.POINTER(Options), ctypes.POINTER(P)]
    L.snn_net_record_fire_events.argtypes = [P, I]
    L.snn_net_read_fire_events.argtypes = [P, P, P, U64]
    L.snn_net_read_inject_log.argtypes = [P, P, P, P, P, P, U64]
    L.snn_adjoint_create.argtypes = [P, ctypes.POINTER(_Desc), U64, ctypes.POINTER(P)]
    L.snn_adjoint_destroy.argtypes = [P]
    L.snn_adjoint_sizes.argtypes = [P, P, P, P, P]
    L.snn_adjoint_read_events.argtypes = [P] * 6
    L.snn_adjoint_forward.argtypes = [P, P, P, P]
    L.snn_adjoint_backward.argtypes = [P, P, P, P]
    L.snn_batch_create_ei.argtypes = [P, U64, ctypes.POINTER(Options), ctypes.POINTER(P)]
    L.snn_batch_destroy.argtypes = [P]
    L.snn_batch_network.argtypes = [P, ctypes.POINTER(P)]
    L.snn_batch_offsets.argtypes = [P, P, P, P]
    L.snn_batch_inject.argtypes = [P, P, P, P, P, U64]
    L.snn_batch_run_until.argtypes = [P, U64]
    L.snn_batch_net_counts.argtypes = [P, P, P]
    L.snn_net_set_profiling.argtypes = [P, I]
    L.snn_net_kernel_times.argtypes = [P, P, P, I, ctypes.c_char_p, U64]
    L.snn_net_pattern_paging.argtypes = [P, P, P, P]
    L.snn_net_pattern_paging.restype = I
    L.snn_graph_create.argtypes = [ctypes.POINTER(_Desc), I, ctypes.POINTER(P)]
    L.snn_graph_destroy.argtypes = [P]
    L.snn_graph_max_delay.argtypes = [P, P]
    L.snn_graph_feedforward.argtypes = [P, U64, P, P, P, U64, P, P, P]
    L.snn_graph_recurrent.argtypes = [P, U64, P, P, P, U64, P, U64, P, ctypes.POINTER(RecurrentOptions), P, P, P]
    for name in ("snn_graph_create", "snn_graph_destroy", "snn_graph_max_delay", "snn_graph_feedforward",
                 "snn_graph_recurrent"):
        getattr(L, name).restype = I
    for name in ("snn_net_create", "snn_net_create_ei", "snn_net_destroy", "snn_net_inject", "snn_net_run_until", "snn_net_run_until_static", "snn_net_counts",
                 "snn_net_read_tick_counts", "snn_net_read_raster", "snn_net_read_neurons",
                 "snn_net_read_synapses", "snn_net_read_plasticity", "snn_net_set_profiling",
                 "snn_net_kernel_times", "snn_net_get_schedule", "snn_net_set_schedule", "snn_net_autotune",
                 "snn_net_tune_log", "snn_batch_create", "snn_batch_create_ei", "snn_batch_destroy",
                 "snn_batch_network", "snn_net_record_fire_events", "snn_net_read_fire_events",
                 "snn_net_read_inject_log", "snn_adjoint_create", "snn_adjoint_destroy", "snn_adjoint_sizes",
                 "snn_adjoint_read_events", "snn_adjoint_forward", "snn_adjoint_backward",
                 "snn_batch_offsets", "snn_batch_inject", "snn_batch_run_until", "snn_batch_net_counts"):
        getattr(L, name).restype = I
    return L


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def version() -> str:
    """snncuda_version() (reference python/snncuda/__init__.py:58-59)."""
    return lib().snncuda_version().decode("utf-8")


def cuda_available() -> bool:
    """snncuda_cuda_available(): compile-time accelerator flag, as in the reference."""
    return bool(lib().snncuda_cuda_available())


def device_count() -> int:
    return int(lib().snn_device_count())


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(rc: int) -> None:
    if rc == SNN_OK:
        return
    msg = lib().snn_last_error().decode("utf-8", "replace")
    if rc == SNN_ECAPACITY:
        raise CapacityError(rc, msg)
    raise SNNError(rc, msg)


class Network:
    """A connectome resident on one MI355X, driven like ``NetworkPropagator``.

    ``net`` is any object with the connectome-order arrays of ``snncuda_amd.workloads.Net``.
    """

    def __init__(self, net=None, *, max_events_per_tick: int = 0, fire_log_capacity: int = 0,
                 raster_capacity: int = 0, track_pre_count: bool = True, device: int = 0, spec=None,
                 learned_pattern_slots: int = 0, pattern_device_pages: int = 0, pattern_runs: int = 0):
        L = lib()
        o = Options(max_events_per_tick, fire_log_capacity, raster_capacity, int(track_pre_count), device,
                    learned_pattern_slots, pattern_device_pages, pattern_runs)
        self.track_pre_count = track_pre_count
        if spec is not None:
            self._create_from_spec(spec, o)
            return
        d, keep = _desc(net)
        self.n, self.e = int(d.n_neurons), int(d.n_synapses)
        h = ctypes.c_void_p()
        _check(L.snn_net_create(ctypes.byref(d), ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._owned = True
        del keep  # the library copied everything

    @classmethod
    def _borrowed(cls, handle, n: int, e: int, track_pre_count: bool) -> "Network":
        """A view of a handle owned elsewhere (a batch's union network); close() leaves it alone."""
        self = cls.__new__(cls)
        self._h, self._owned = handle, False
        self.n, self.e, self.track_pre_count = n, e, track_pre_count
        return self

    def _create_from_spec(self, spec: dict, o: "Options") -> None:
        """Generate a workloads.ei_network on the device (snn_net_create_ei)."""
        d, fan = _ei_desc(spec)
        h = ctypes.c_void_p()
        _check(lib().snn_net_create_ei(ctypes.byref(d), ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._owned = True
        self.n = int(spec["n"] + spec["n_sources"])
        self.e = int(fan.astype(np.uint64).sum() + np.uint64(spec["n_sources"]) * np.uint64(spec["source_fanout"]))

    def close(self) -> None:
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                lib().snn_net_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # NetworkPropagator::inject / run_until
    def inject(self, neuron, tick, weight) -> None:
        nb = np.atleast_1d(np.asarray(neuron, dtype=np.uint32))
        tk = np.atleast_1d(np.asarray(tick, dtype=np.uint64))
        wt = np.atleast_1d(np.asarray(weight, dtype=np.float32))
        nb, tk, wt = np.broadcast_arrays(nb, tk, wt)
        nb, tk, wt = (np.ascontiguousarray(x) for x in (nb, tk, wt))
        _check(lib().snn_net_inject(self._h, _p(nb), _p(tk), _p(wt), nb.shape[0]))

    def run_until(self, inclusive_tick: int) -> None:
        _check(lib().snn_net_run_until(self._h, int(inclusive_tick)))

    def run_until_static(self, inclusive_tick: int) -> None:
        """run_until without device allocations: the segment's inject tables and counters are staged
        in the library's fixed per-device array (snn_net_run_until_static)."""
        _check(lib().snn_net_run_until_static(self._h, int(inclusive_tick)))

    def run_commands(self, commands: Sequence) -> None:
        """Replay (kind, neuron, tick, weight) commands: kind 0 inject, 1 run_until. Consecutive
        injects are batched into one call (same order)."""
        batch = []
        for k, n, t, w in commands:
            if k == 0:
                batch.append((n, t, w))
                continue
            if batch:
                arr = np.array(batch, dtype=np.float64)
                self.inject(arr[:, 0].astype(np.uint32), arr[:, 1].astype(np.uint64), arr[:, 2].astype(np.float32))
                batch = []
            self.run_until(t)
        if batch:
            arr = np.array(batch, dtype=np.float64)
            self.inject(arr[:, 0].astype(np.uint32), arr[:, 1].astype(np.uint64), arr[:, 2].astype(np.float32))

    def pattern_paging(self) -> Dict[str, int]:
        """The learned-pattern pool's pages, the device cache's pages (0: the whole pool on the device)
        and the page copies made so far."""
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().snn_net_pattern_paging(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return dict(pages=a.value, device_pages=b.value, moves=c.value)

    def counts(self):
        d, f, t = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().snn_net_counts(self._h, ctypes.byref(d), ctypes.byref(f), ctypes.byref(t)))
        return d.value, f.value, t.value

    @property
    def delivered(self) -> int:
        return self.counts()[0]

    @property
    def fired(self) -> int:
        return self.counts()[1]

    def tick_counts(self, first: int = 0, n: Optional[int] = None):
        if n is None:
            n = self.counts()[2] - first
        d = np.zeros(n, np.uint64)
        f = np.zeros(n, np.uint64)
        _check(lib().snn_net_read_tick_counts(self._h, first, n, _p(d), _p(f)))
        return d, f

    def raster(self) -> np.ndarray:
        """Fires in processing order (tick-major, reference order within a tick)."""
        n = ctypes.c_uint64()
        _check(lib().snn_net_read_raster(self._h, ctypes.byref(n), None, 0))
        out = np.zeros(n.value, np.uint32)
        _check(lib().snn_net_read_raster(self._h, ctypes.byref(n), _p(out), out.shape[0]))
        return out

    def neurons(self) -> Dict[str, np.ndarray]:
        n = self.n
        o = dict(membrane=np.zeros(n, np.float32), current=np.zeros((4, n), np.float32),
                 tau=np.zeros((4, n), np.float32), spike_count=np.zeros(n, np.uint64),
                 last_fired=np.zeros(n, np.int64), t_match=np.zeros(n, np.uint64), t_learned=np.zeros(n, np.uint64))
        _check(lib().snn_net_read_neurons(self._h, *[_p(o[k]) for k in (
            "membrane", "current", "tau", "spike_count", "last_fired", "t_match", "t_learned")]))
        return o

    def synapses(self) -> Dict[str, np.ndarray]:
        e = self.e
        o = dict(weight=np.zeros(e, np.float32), upd_count=np.zeros(e, np.uint64), last_pre=np.zeros(e, np.int64),
                 last_post=np.zeros(e, np.int64), c_match=np.zeros(e, np.uint64), c_learned=np.zeros(e, np.uint64))
        if self.track_pre_count:
            o["pre_count"] = np.zeros(e, np.uint64)
        _check(lib().snn_net_read_synapses(self._h, _p(o["weight"]), _p(o.get("pre_count")), _p(o["upd_count"]),
                                           _p(o["last_pre"]), _p(o["last_post"]), _p(o["c_match"]),
                                           _p(o["c_learned"])))
        return o

    def plasticity(self) -> Dict[str, np.ndarray]:
        """Changed STDP updates per synapse (connectome order) split into delta > 0 (ltp) and the
        rest (ltd); ltp + ltd == synapses()["upd_count"]."""
        o = dict(ltp=np.zeros(self.e, np.uint64), ltd=np.zeros(self.e, np.uint64))
        _check(lib().snn_net_read_plasticity(self._h, _p(o["ltp"]), _p(o["ltd"])))
        return o

    # recording for the fixed-spike adjoint
    def record_fire_events(self, on: bool = True) -> None:
        """Keep each fire's event index and every delivered inject (switch on before the first tick)."""
        _check(lib().snn_net_record_fire_events(self._h, int(on)))

    def fire_events(self) -> np.ndarray:
        n = ctypes.c_uint64()
        _check(lib().snn_net_read_fire_events(self._h, ctypes.byref(n), None, 0))
        out = np.zeros(n.value, np.uint32)
        _check(lib().snn_net_read_fire_events(self._h, ctypes.byref(n), _p(out), out.shape[0]))
        return out

    def inject_log(self) -> Dict[str, np.ndarray]:
        n = ctypes.c_uint64()
        _check(lib().snn_net_read_inject_log(self._h, ctypes.byref(n), None, None, None, None, 0))
        m = n.value
        o = dict(neuron=np.zeros(m, np.uint32), tick=np.zeros(m, np.uint64), stage=np.zeros(m, np.uint64),
                 weight=np.zeros(m, np.float32))
        _check(lib().snn_net_read_inject_log(self._h, ctypes.byref(n), *[_p(o[k]) for k in ("neuron", "tick", "stage",
                                                                                              "weight")], m))
        return o

    # launch schedule and autotuner (results are identical under every schedule)
    def schedule(self) -> Dict[str, int]:
        sc = Schedule()
        _check(lib().snn_net_get_schedule(self._h, ctypes.byref(sc)))
        return dict(partition_shift=sc.partition_shift, scatter_workgroups=sc.scatter_workgroups,
                    graph_ticks=sc.graph_ticks, fold_threads=sc.fold_threads)

    def set_schedule(self, **kw) -> None:
        cur = self.schedule()
        cur.update({k: v for k, v in kw.items() if v is not None})
        sc = Schedule(cur["partition_shift"], cur["scatter_workgroups"], cur["graph_ticks"], cur["fold_threads"])
        _check(lib().snn_net_set_schedule(self._h, ctypes.byref(sc)))

    def autotune(self, trial_ticks: int = 32) -> None:
        """Tune the schedule on the first ticks of the next long enough run_until call."""
        _check(lib().snn_net_autotune(self._h, trial_ticks))

    def tune_log(self) -> str:
        buf = ctypes.create_string_buffer(1 << 16)
        lib().snn_net_tune_log(self._h, buf, 1 << 16)
        return buf.value.decode()

    KERNELS = ("plan", "segments", "scan_segments", "count", "scan_partitions", "scatter", "fold", "scan_fires",
               "fire_write", "post", "advance")

    def set_profiling(self, on: bool, kernels=None, every: int = 1) -> None:
        """Time the step kernels with HIP events (all of them, or only those named in `kernels`, on
        every `every`-th tick); clears the totals that kernel_times() reports."""
        mode = int(bool(on))
        if on and (kernels is not None or every > 1):
            mode = 0x40000000 | (min(max(int(every), 1), 255) << 16)
            for k in (kernels if kernels is not None else self.KERNELS):
                mode |= 1 << self.KERNELS.index(k)
        _check(lib().snn_net_set_profiling(self._h, mode))

    def kernel_times(self) -> Dict[str, tuple]:
        """{kernel: (total ms, launches)} over runs since set_profiling(True)."""
        cap = 16
        ms = np.zeros(cap, np.float64)
        nl = np.zeros(cap, np.uint64)
        names = ctypes.create_string_buffer(512)
        m = lib().snn_net_kernel_times(self._h, _p(ms), _p(nl), cap, names, 512)
        keys = names.value.decode().split(",")
        return {keys[i]: (float(ms[i]), int(nl[i])) for i in range(m)}


__all__ = ["SNNCudaLibraryError", "SNNError", "CapacityError", "Network", "Options", "version", "cuda_available",
           "device_count", "lib", "library_path"]


def _ei_desc(spec: dict):
    """snn_ei_desc of a workloads.ei_network spec; returns (desc, fan-out array to keep alive)."""
    fan = np.ascontiguousarray(np.broadcast_to(np.asarray(spec["fanout"], np.uint32), (spec["n"],)))
    comps = (ctypes.c_uint8 * 4)(*(list(spec["compartments"]) + [0] * 4)[:4])
    d = _EiDesc(spec["n"], spec["n_sources"], _p(fan), spec["source_fanout"], spec["seed"], spec["n_exc"],
                spec["w_exc"], spec["w_inh"], spec["delay_max"], len(spec["compartments"]), comps,
                int(spec["plastic"]), spec["threshold"], 1.0, spec["window"], spec["sim_thr"], spec["max_refs"],
                int(spec.get("n_codes", 1)), int(spec.get("code_max", 0)))
    return d, fan


def _desc(net):
    """snn_net_desc over C-contiguous copies of a workloads.Net's arrays; returns (desc, arrays to keep
    alive while the desc is used)."""
    a = [np.ascontiguousarray(x, dtype=dt) for x, dt in (
        (net.threshold, np.float32), (net.window, np.uint64), (net.sim_thr, np.float32),
        (net.max_refs, np.uint64), (net.src, np.uint32), (net.tgt, np.uint32), (net.weight, np.float32),
        (net.max_weight, np.float32), (net.delay, np.uint32), (net.comp, np.uint8),
        (net.receptor, np.uint8), (net.plastic, np.uint8), (net.code_ptr, np.uint64), (net.codes, np.uint32))]
    d = _Desc(int(a[0].shape[0]), _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), int(a[4].shape[0]), *[_p(x) for x in a[4:]])
    return d, a


def merge_inputs(samples, n_neurons: int):
    """Samples of (neuron, weight) inputs merged as the session merges them: per sample one entry per
    distinct neuron in order of first appearance, its weights summed in input order (float32); neurons
    outside [0, n_neurons) are skipped. Returns (ptr uint64[S+1], neuron uint32, weight float32)."""
    ptr, xs, ws = [0], [], []
    for sample in samples:
        at = {}
        for x, w in sample:
            x = int(x)
            if x < 0 or x >= n_neurons:
                continue
            if x in at:
                ws[at[x]] = np.float32(ws[at[x]] + np.float32(w))
            else:
                at[x] = len(xs)
                xs.append(x)
                ws.append(np.float32(w))
        ptr.append(len(xs))
    return (np.asarray(ptr, np.uint64), np.asarray(xs, np.uint32).reshape(-1),
            np.asarray(ws, np.float32).reshape(-1))


class Graph:
    """A connectome's fixed-weight edges on the device for the non-spiking readouts of
    ``CudaInferenceSession`` (snn_graph_*; reference ``run_feedforward_batch`` /
    ``run_recurrent_feedforward_batch``, ``include/snncuda/backends/CudaBackend.h:113-119``)."""

    def __init__(self, net, device: int = 0):
        d, keep = _desc(net)
        self.n, self.e = int(d.n_neurons), int(d.n_synapses)
        h = ctypes.c_void_p()
        _check(lib().snn_graph_create(ctypes.byref(d), device, ctypes.byref(h)))
        self._h = h
        del keep

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().snn_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def max_delay(self) -> int:
        v = ctypes.c_uint32()
        _check(lib().snn_graph_max_delay(self._h, ctypes.byref(v)))
        return int(v.value)

    def feedforward(self, ptr, neuron, weight, readout):
        """Merged inputs (``merge_inputs``) and readout neuron indices (``GRAPH_NONE`` for an unknown id)
        -> (scores float32 [S, R], spikes uint64 [S, R])."""
        ptr = np.ascontiguousarray(ptr, np.uint64)
        neuron = np.ascontiguousarray(neuron, np.uint32)
        weight = np.ascontiguousarray(weight, np.float32)
        readout = np.ascontiguousarray(readout, np.uint32)
        S, R = len(ptr) - 1, len(readout)
        scores = np.zeros((S, R), np.float32)
        spikes = np.zeros((S, R), np.uint64)
        _check(lib().snn_graph_feedforward(self._h, S, _p(ptr), _p(neuron), _p(weight), R, _p(readout),
                                           _p(scores), _p(spikes)))
        return scores, spikes

    def recurrent(self, ptr, neuron, weight, readout, hidden, *, iterations: int = 1, feedback_decay: float = 0.5,
                  top_k_feedback_readouts: int = 1, top_k_hidden: int = 0, use_scores_not_spikes: bool = True):
        """-> (scores float32 [S, R], spikes uint64 [S, R], time_count)."""
        ptr = np.ascontiguousarray(ptr, np.uint64)
        neuron = np.ascontiguousarray(neuron, np.uint32)
        weight = np.ascontiguousarray(weight, np.float32)
        readout = np.ascontiguousarray(readout, np.uint32)
        hidden = np.ascontiguousarray(hidden, np.uint32)
        S, R = len(ptr) - 1, len(readout)
        scores = np.zeros((S, R), np.float32)
        spikes = np.zeros((S, R), np.uint64)
        ticks = ctypes.c_uint64()
        o = RecurrentOptions(iterations, feedback_decay, top_k_feedback_readouts, top_k_hidden,
                             int(use_scores_not_spikes))
        _check(lib().snn_graph_recurrent(self._h, S, _p(ptr), _p(neuron), _p(weight), R, _p(readout), len(hidden),
                                         _p(hidden), ctypes.byref(o), _p(scores), _p(spikes), ctypes.byref(ticks)))
        return scores, spikes, int(ticks.value)


class Batch:
    """Independent networks stepped together (snn_batch_*): one set of launches per tick for all of
    them. ``union`` is the union network (a borrowed ``Network``) for state readers, schedules and
    the autotuner; ``net_neurons`` / ``net_synapses`` / ``net_raster`` slice it per network."""

    def __init__(self, nets=None, *, max_events_per_tick: int = 0, fire_log_capacity: int = 0,
                 raster_capacity: int = 0, track_pre_count: bool = True, device: int = 0, specs=None,
                 learned_pattern_slots: int = 0, pattern_device_pages: int = 0, pattern_runs: int = 0):
        """``nets``: workloads.Net objects (host arrays, snn_batch_create); or ``specs``: ei_network
        specs generated on the device (snn_batch_create_ei)."""
        o = Options(max_events_per_tick, fire_log_capacity, raster_capacity, int(track_pre_count), device,
                    learned_pattern_slots, pattern_device_pages, pattern_runs)
        descs, keep = [], []
        h = ctypes.c_void_p()
        if specs is not None:
            for sp in specs:
                d, a = _ei_desc(sp)
                descs.append(d)
                keep.append(a)
            arr = (_EiDesc * len(descs))(*descs)
            _check(lib().snn_batch_create_ei(arr, len(descs), ctypes.byref(o), ctypes.byref(h)))
        else:
            for net in nets:
                d, a = _desc(net)
                descs.append(d)
                keep.append(a)
            arr = (_Desc * len(descs))(*descs)
            _check(lib().snn_batch_create(arr, len(descs), ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        del keep
        self.n_nets = len(descs)
        self.neuron_off = np.zeros(self.n_nets + 1, np.uint64)
        self.synapse_off = np.zeros(self.n_nets + 1, np.uint64)
        _check(lib().snn_batch_offsets(self._h, None, _p(self.neuron_off), _p(self.synapse_off)))
        u = ctypes.c_void_p()
        _check(lib().snn_batch_network(self._h, ctypes.byref(u)))
        self.union = Network._borrowed(u, int(self.neuron_off[-1]), int(self.synapse_off[-1]), track_pre_count)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.union._h = None
            lib().snn_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def inject(self, net, neuron, tick, weight) -> None:
        arrs = np.broadcast_arrays(np.atleast_1d(np.asarray(net, np.uint32)), np.atleast_1d(np.asarray(neuron, np.uint32)),
                                   np.atleast_1d(np.asarray(tick, np.uint64)), np.atleast_1d(np.asarray(weight, np.float32)))
        k, nb, tk, wt = (np.ascontiguousarray(x) for x in arrs)
        _check(lib().snn_batch_inject(self._h, _p(k), _p(nb), _p(tk), _p(wt), k.shape[0]))

    def run_until(self, inclusive_tick: int) -> None:
        _check(lib().snn_batch_run_until(self._h, int(inclusive_tick)))

    def run_commands(self, commands_per_net: Sequence[Sequence]) -> None:
        """Replay each network's (kind, neuron, tick, weight) command list on the shared clock so that
        every network sees its injects at the same scheduler poll tick as in a run of its own
        (workloads.merge_commands)."""
        from .workloads import merge_commands
        cmds, owner = merge_commands(commands_per_net, self.neuron_off)
        batch = []
        i = 0
        for k, n, t, w in cmds:
            if k == 0:
                batch.append((owner[i], n - int(self.neuron_off[owner[i]]), t, w))
                i += 1
                continue
            if batch:
                a = np.array(batch, dtype=np.float64)
                self.inject(a[:, 0], a[:, 1], a[:, 2].astype(np.uint64), a[:, 3].astype(np.float32))
                batch = []
            self.run_until(t)
        if batch:
            a = np.array(batch, dtype=np.float64)
            self.inject(a[:, 0], a[:, 1], a[:, 2].astype(np.uint64), a[:, 3].astype(np.float32))

    def net_counts(self):
        """(delivered[n_nets], fired[n_nets]); delivered counts external injects, as the reference does."""
        d = np.zeros(self.n_nets, np.uint64) if self.union.track_pre_count else None
        f = np.zeros(self.n_nets, np.uint64)
        _check(lib().snn_batch_net_counts(self._h, _p(d), _p(f)))
        return d, f

    def _slices(self, k):
        return (slice(int(self.neuron_off[k]), int(self.neuron_off[k + 1])),
                slice(int(self.synapse_off[k]), int(self.synapse_off[k + 1])))

    def net_neurons(self, k: int, neurons=None) -> Dict[str, np.ndarray]:
        a = neurons if neurons is not None else self.union.neurons()
        ns, _ = self._slices(k)
        return {key: (v[:, ns] if v.ndim == 2 else v[ns]) for key, v in a.items()}

    def net_synapses(self, k: int, synapses=None) -> Dict[str, np.ndarray]:
        a = synapses if synapses is not None else self.union.synapses()
        _, es = self._slices(k)
        return {key: v[es] for key, v in a.items()}

    def net_raster(self, k: int, raster=None) -> np.ndarray:
        """Network k's fires in processing order, as network-local neuron indices."""
        r = raster if raster is not None else self.union.raster()
        lo, hi = int(self.neuron_off[k]), int(self.neuron_off[k + 1])
        return (r[(r >= lo) & (r < hi)] - lo).astype(np.uint32)
