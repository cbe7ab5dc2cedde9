"""ctypes access to the CPU model of the device step (tests/emu/emu_step.cpp). Test-only."""
import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "libsnn_emu.so"
_lib = None


def build() -> None:
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(LIB),
                    str(HERE / "emu_step.cpp")], check=True)


def lib():
    global _lib
    if _lib is None:
        src = HERE / "emu_step.cpp"
        hdrs = list((HERE.parent.parent / "snncuda_amd" / "csrc").glob("*.h"))
        if not LIB.exists() or any(p.stat().st_mtime > LIB.stat().st_mtime for p in [src] + hdrs):
            build()
        L = ctypes.CDLL(str(LIB))
        P = ctypes.c_void_p
        L.emu_create.restype = P
        L.emu_create.argtypes = [ctypes.c_uint64] + [P] * 4 + [ctypes.c_uint64] + [P] * 10 + \
            [ctypes.c_uint64] * 5 + [ctypes.c_char_p, ctypes.c_uint64]
        L.emu_destroy.argtypes = [P]
        L.emu_inject.argtypes = [P, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float]
        L.emu_run_until.argtypes = [P, ctypes.c_uint64]
        L.emu_run_until.restype = ctypes.c_uint32
        L.emu_ticks.argtypes = [P]
        L.emu_ticks.restype = ctypes.c_uint64
        L.emu_tick_counts.argtypes = [P, P, P]
        L.emu_raster.argtypes = [P, P, ctypes.c_uint64]
        L.emu_raster.restype = ctypes.c_uint64
        L.emu_neurons.argtypes = [P] * 6
        L.emu_synapses.argtypes = [P] * 5
        L.emu_sort_by_g.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.emu_temporal_check.argtypes = [P, P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.c_uint32]
        L.emu_temporal_check.restype = ctypes.c_uint32
        L.emu_pack_check.argtypes = [ctypes.c_uint64] + [P] * 4 + [ctypes.c_uint64] + [P] * 10 + \
            [ctypes.c_uint32, ctypes.c_char_p, ctypes.c_uint64]
        L.emu_pack_check.restype = ctypes.c_int
        L.emu_partition_shift.argtypes = [ctypes.c_uint64, ctypes.c_char_p, ctypes.c_uint64]
        L.emu_partition_shift.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Emu:
    """The device step's logic on the CPU, with NetworkPropagator-like calls."""

    def __init__(self, net, ev_cap=1 << 20, fire_cap=1 << 16, seed=1, ref_cap=0, pattern_runs=0):
        L = lib()
        self._keep = [np.ascontiguousarray(a) for a in (
            net.threshold, net.window, net.sim_thr, net.max_refs, net.src, net.tgt, net.weight,
            net.max_weight, net.delay, net.comp, net.receptor, net.plastic, net.code_ptr, net.codes)]
        k = self._keep
        err = ctypes.create_string_buffer(512)
        self.n, self.e = net.n_neurons, net.n_synapses
        self.h = L.emu_create(self.n, *[_p(a) for a in k[:4]], self.e, *[_p(a) for a in k[4:]],
                              ev_cap, fire_cap, seed, ref_cap, pattern_runs, err, 512)
        if not self.h:
            raise ValueError(err.value.decode())
        self.flags = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().emu_destroy(self.h)
            self.h = None

    def run_commands(self, commands):
        for k, n, t, w in commands:
            if k == 0:
                lib().emu_inject(self.h, n, t, w)
            else:
                self.flags |= lib().emu_run_until(self.h, t)

    def tick_counts(self):
        nt = lib().emu_ticks(self.h)
        d = np.zeros(nt, np.uint64); f = np.zeros(nt, np.uint64)
        lib().emu_tick_counts(self.h, _p(d), _p(f))
        return d, f

    def raster(self):
        n = lib().emu_raster(self.h, None, 0)
        out = np.zeros(n, np.uint32)
        lib().emu_raster(self.h, _p(out), n)
        return out

    def neurons(self):
        n = self.n
        o = dict(membrane=np.zeros(n, np.float32), current=np.zeros((4, n), np.float32),
                 spike_count=np.zeros(n, np.uint64), t_match=np.zeros(n, np.uint64), t_learned=np.zeros(n, np.uint64))
        lib().emu_neurons(self.h, *[_p(o[k]) for k in ("membrane", "current", "spike_count", "t_match", "t_learned")])
        return o

    def synapses(self):
        e = self.e
        o = dict(weight=np.zeros(e, np.float32), pre_count=np.zeros(e, np.uint64),
                 c_match=np.zeros(e, np.uint64), c_learned=np.zeros(e, np.uint64))
        lib().emu_synapses(self.h, *[_p(o[k]) for k in ("weight", "pre_count", "c_match", "c_learned")])
        return o


def sort_by_g(words, narrow_g, small=False):
    """The device's per-target sort, run on the CPU, of entries given as rows of 2, 3 or 4 uint32
    words (8-, 12- or 16-byte entries, gk first); returns the sorted rows.
    narrow_g: every g is below 2^27 (the key-and-place networks may be used).
    small: the chunked fold's variant (networks of at most 8 entries)."""
    a = np.ascontiguousarray(words, np.uint32).copy()
    lib().emu_sort_by_g(_p(a), a.shape[0], int(narrow_g) | (2 if small else 0), a.shape[1])
    return a


def temporal_check(ticks, times, window, sim, max_refs, cap, runs=8):
    """Index of the first tick where the per-tick temporal pattern update differs from the same
    observations applied one at a time, with a ring of `cap` learned-pattern slots (len(ticks) when
    they agree)."""
    ticks = np.ascontiguousarray(ticks, np.uint32)
    times = np.ascontiguousarray(times, np.uint32)
    return lib().emu_temporal_check(_p(ticks), _p(times), len(ticks), window, sim, max_refs, cap, runs)


def pack_check(net, threads):
    """The host pack on `threads` threads against the sequential reference: (code, message); code 0
    when every output field is equal (or both raise the same message)."""
    keep = [np.ascontiguousarray(a) for a in (
        net.threshold, net.window, net.sim_thr, net.max_refs, net.src, net.tgt, net.weight,
        net.max_weight, net.delay, net.comp, net.receptor, net.plastic, net.code_ptr, net.codes)]
    msg = ctypes.create_string_buffer(512)
    rc = lib().emu_pack_check(net.n_neurons, *[_p(a) for a in keep[:4]], net.n_synapses,
                              *[_p(a) for a in keep[4:]], threads, msg, 512)
    return rc, msg.value.decode()


def partition_shift(n):
    """(shift, error message) of the partition size snn_net_create picks for n neurons."""
    buf = ctypes.create_string_buffer(512)
    r = lib().emu_partition_shift(int(n), buf, 512)
    return r, buf.value.decode()
