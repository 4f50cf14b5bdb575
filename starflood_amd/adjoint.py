"""Differentiable KDK integration of PP gravity for PyTorch.

`kdk(x0, v0, m, nsteps, ckpt_every)` runs the library's fast-policy KDK step
(FP32 kicks, FP64 drift) `nsteps` times on the GPU and returns (x_N, v_N);
its backward pass is the library's discrete adjoint (one force VJP per step,
segments of `ckpt_every` steps recomputed from checkpoints).  Tensors are
float32, contiguous, shape (N, 3) / (N,), on the current HIP device; epsilon,
G, dt and acc_mode are the library parameters at the time of the forward
call (`starflood_amd._lib.set_params`).

The arithmetic is all in the C ABI (sf_adjoint_forward / sf_adjoint_backward,
sf_force_dev, sf_force_vjp_dev); torch only supplies device memory and the
stream.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L


def _check_state(*ts: torch.Tensor) -> None:
    for t in ts:
        if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("expected contiguous float32 tensors on the GPU")


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


class _Tape:
    """Owns an sf_tape*; freed with the autograd graph."""

    def __init__(self, lib, ptr: int):
        self.lib, self.ptr = lib, ptr

    def __del__(self):
        if self.ptr:
            self.lib.sf_tape_free(self.ptr)
            self.ptr = 0


class KDK(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, nsteps: int, ckpt_every: int):
        _check_state(x0, v0, m)
        N = m.shape[0]
        if x0.shape != (N, 3) or v0.shape != (N, 3):
            raise ValueError("x0 and v0 must have shape (N, 3)")
        lib = L.load()
        xN = torch.empty_like(x0)
        vN = torch.empty_like(v0)
        tape = C.c_void_p()
        L.check(lib.sf_adjoint_forward(N, x0.data_ptr(), v0.data_ptr(), m.data_ptr(), nsteps, ckpt_every,
                                       xN.data_ptr(), vN.data_ptr(), _stream(), C.byref(tape)),
                "sf_adjoint_forward")
        ctx.tape = _Tape(lib, tape.value)
        ctx.N = N
        return xN, vN

    @staticmethod
    def backward(ctx, gxN, gvN):
        lib = ctx.tape.lib
        gx = gxN.contiguous() if gxN is not None else None
        gv = gvN.contiguous() if gvN is not None else None
        ref = gx if gx is not None else gv
        gx0 = torch.empty((ctx.N, 3), dtype=torch.float32, device=ref.device)
        gv0 = torch.empty_like(gx0)
        L.check(lib.sf_adjoint_backward(ctx.tape.ptr, gx.data_ptr() if gx is not None else None,
                                        gv.data_ptr() if gv is not None else None, gx0.data_ptr(), gv0.data_ptr(),
                                        _stream()), "sf_adjoint_backward")
        return gx0, gv0, None, None, None


def kdk(x0: torch.Tensor, v0: torch.Tensor, m: torch.Tensor, nsteps: int, ckpt_every: int = 20):
    """(x_N, v_N) after nsteps KDK steps, differentiable in x0 and v0."""
    return KDK.apply(x0, v0, m, nsteps, ckpt_every)


def force(x: torch.Tensor, m: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """(acc, pot) of the fast PP kernel on device tensors."""
    _check_state(x, m)
    lib = L.load()
    acc = torch.empty_like(x)
    pot = torch.empty_like(m)
    L.check(lib.sf_force_dev(m.shape[0], x.data_ptr(), m.data_ptr(), acc.data_ptr(), pot.data_ptr(), _stream()),
            "sf_force_dev")
    return acc, pot


def force_vjp(x: torch.Tensor, m: torch.Tensor, mu: torch.Tensor) -> torch.Tensor:
    """(d acc / d x)^T mu."""
    _check_state(x, m, mu)
    lib = L.load()
    g = torch.empty_like(x)
    L.check(lib.sf_force_vjp_dev(m.shape[0], x.data_ptr(), m.data_ptr(), mu.data_ptr(), g.data_ptr(), _stream()),
            "sf_force_vjp_dev")
    return g
