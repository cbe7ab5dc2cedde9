"""Fixed-spike adjoint of the per-tick step's subthreshold dynamics on the MI355X (snn_adjoint_*).

The dynamics are the step's receptor / dendrite / soma integration with every event time and every
reset held at those of a recorded run; the per-tick somatic trace is v(t) = sum_k alpha_k c_k(t).
Parameters: synapse weights (connectome order), alpha[4] (soma, basal, apical, inhibitory coupling)
and tau[4] (AMPA, NMDA, GABA_A, GABA_B time constants). ``SomaticTrace`` is the
torch.autograd.Function over it: forward and backward run the hand-written HIP kernels on the
tensors' device memory (no host copies), with checkpoints every 64 ticks recomputed in backward.

    run = Network(net); run.record_fire_events(); run.run_commands(net.commands)
    adj = Adjoint(run, net, ticks)
    trace = somatic_trace(adj, weight, params)   # weight: cuda f32 [E], params: cuda f32 [8]
    loss = ((trace - target) ** 2).mean(); loss.backward()
"""
from __future__ import annotations

import ctypes
from typing import Dict, Sequence

import numpy as np

from . import Network, _check, _desc, _p, lib

DEFAULT_PARAMS = (1.0, 1.0, 1.0, 1.0, 5.0, 100.0, 10.0, 150.0)


class Adjoint:
    """Event lists of a recorded run plus the forward / backward kernels."""

    def __init__(self, run: Network, net, ticks: int):
        d, keep = _desc(net)
        h = ctypes.c_void_p()
        _check(lib().snn_adjoint_create(run._h, ctypes.byref(d), int(ticks), ctypes.byref(h)))
        del keep
        self._h = h
        n, e, t, m = (ctypes.c_uint64() for _ in range(4))
        _check(lib().snn_adjoint_sizes(self._h, ctypes.byref(n), ctypes.byref(e), ctypes.byref(t), ctypes.byref(m)))
        self.n, self.e, self.T, self.n_events = n.value, e.value, t.value, m.value
        self.token = 0  # forwards so far: backward needs the checkpoints of its own forward

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().snn_adjoint_destroy(self._h)
            self._h = None
        self._weight = None

    def __del__(self):
        self.close()

    def events(self) -> Dict[str, np.ndarray]:
        m = self.n_events
        o = dict(ptr=np.zeros(self.n + 1, np.uint64), tick=np.zeros(m, np.uint32), synapse=np.zeros(m, np.uint32),
                 weight=np.zeros(m, np.float32), reset=np.zeros(m, np.uint8))
        _check(lib().snn_adjoint_read_events(self._h, *[_p(o[k]) for k in ("ptr", "tick", "synapse", "weight", "reset")]))
        return o

    def forward(self, weight, params: Sequence[float], trace) -> None:
        """weight: cuda float32 [E] tensor; trace: cuda float32 [T, N] tensor (written)."""
        import torch
        pr = np.ascontiguousarray(np.asarray(params, np.float32).reshape(8))
        assert weight.is_cuda and weight.dtype == torch.float32 and weight.is_contiguous() and weight.numel() == self.e
        assert trace.is_cuda and trace.dtype == torch.float32 and trace.is_contiguous() and trace.numel() == self.T * self.n
        torch.cuda.synchronize(weight.device)
        _check(lib().snn_adjoint_forward(self._h, ctypes.c_void_p(weight.data_ptr()), _p(pr),
                                         ctypes.c_void_p(trace.data_ptr())))
        # backward reads the weights through this tensor's memory: hold it until the next forward
        self._weight = weight
        self.token += 1

    def backward(self, grad_trace, grad_weight) -> np.ndarray:
        """grad_trace: cuda float32 [T, N]; grad_weight: cuda float32 [E] (written). Returns dL/dparams."""
        import torch
        assert grad_trace.is_cuda and grad_trace.is_contiguous() and grad_trace.numel() == self.T * self.n
        assert grad_weight.is_cuda and grad_weight.is_contiguous() and grad_weight.numel() == self.e
        torch.cuda.synchronize(grad_trace.device)
        gp = np.zeros(8, np.float64)
        _check(lib().snn_adjoint_backward(self._h, ctypes.c_void_p(grad_trace.data_ptr()),
                                          ctypes.c_void_p(grad_weight.data_ptr()), _p(gp)))
        return gp


def _torch_function():
    import torch

    class SomaticTrace(torch.autograd.Function):
        """trace[T, N] = v(t) of every neuron; differentiable in weight [E] and params [8]."""

        @staticmethod
        def forward(ctx, adj: Adjoint, weight, params):
            w = weight.detach().contiguous()
            pr = params.detach().to("cpu", torch.float32).numpy()
            trace = torch.empty((adj.T, adj.n), device=w.device, dtype=torch.float32)
            adj.forward(w, pr, trace)
            ctx.adj, ctx.token, ctx.pr = adj, adj.token, pr
            ctx.save_for_backward(w, params)
            return trace

        @staticmethod
        def backward(ctx, grad_trace):
            adj = ctx.adj
            w, params = ctx.saved_tensors
            if adj.token != ctx.token:  # another forward replaced the checkpoints: recompute ours
                scratch = torch.empty((adj.T, adj.n), device=w.device, dtype=torch.float32)
                adj.forward(w, ctx.pr, scratch)
            gw = torch.empty_like(w)
            gp = adj.backward(grad_trace.contiguous().to(torch.float32), gw)
            return None, gw, torch.as_tensor(gp, dtype=params.dtype, device=params.device)

    return SomaticTrace


_fn = None


def somatic_trace(adj: Adjoint, weight, params):
    """Differentiable somatic traces [T, N] (torch.autograd.Function over the HIP adjoint)."""
    global _fn
    if _fn is None:
        _fn = _torch_function()
    return _fn.apply(adj, weight, params)
