"""Differentiable K17 time stepping: the hand-written adjoint kernel behind torch.autograd.

``k17_steps(f0, lattice, steps, checkpoint_every)`` runs `steps` collide+stream steps of one
Lattice starting from the population buffer f0 and returns the final buffer. Its backward runs
the reverse sweep with vf_k17_adjoint_box / vf_k17_adjoint (hand-written reverse mode of the
FP32 collision): the forward pass keeps a copy of the buffer every `checkpoint_every` steps; the
backward replays one segment at a time from its checkpoint, keeps the segment's states, and
applies the adjoint steps in reverse order with the parity each forward step used. Memory:
(steps / K + K) buffers of the storage type. The adjoint itself is FP32.

Objectives are ordinary torch functions of the returned buffer; ``pre_collision_index`` gives
the gather that turns a buffer into the (27, N) pre-collision populations of every node, and
``kinetic_energy`` is the FP64 sum of 1/2 |j|^2 / rho over the fluid nodes
(src/gpu/core/PostProcessor/KineticEnergyAnalyzer.cu:99-153).

The differentiable step covers the collision, streaming and solid nodes (no-slip delay
bounce-back); lattices with BC kernels (velocity, pressure, explicit no-slip) are rejected.
No counterpart in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import capi
from .grid import CX, CY, CZ, INV
from .lattice import Lattice, stream_handle


def _param_copy(L: Lattice, state: torch.Tensor, even: bool) -> capi.LBMParameter:
    p = capi.LBMParameter.from_buffer_copy(L.param)
    p.distributions = ctypes.c_void_p(state.data_ptr())
    p.isEvenTimestep = 1 if even else 0
    return p


def _check_differentiable(L: Lattice):
    for name, bc in L.bcs.items():
        if bc.n > 0:
            raise NotImplementedError(f"the differentiable step has no adjoint of the {name} BC kernel")


def _forward_step(L: Lattice, buf: torch.Tensor, even: bool, stream):
    p = _param_copy(L, buf, even)
    if L.box:
        capi.call("vf_k17_run_box", ctypes.byref(p), capi.COLLISION_DEFAULT, stream_handle(stream))
    else:
        capi.call("vf_k17_run_on_indices", ctypes.byref(p), ctypes.c_void_p(L.fluid.data_ptr()),
                  int(L.grid.fluid.size), capi.COLLISION_DEFAULT, stream_handle(stream))


def _adjoint_step(L: Lattice, state: torch.Tensor, lam: torch.Tensor, even: bool, shift_adj, stream):
    p = _param_copy(L, state, even)
    sa = None if shift_adj is None else ctypes.c_void_p(shift_adj.data_ptr())
    fn = "vf_k17_adjoint_box" if L.box else "vf_k17_adjoint"
    capi.call(fn, ctypes.byref(p), ctypes.c_void_p(lam.data_ptr()), sa, stream_handle(stream))


class _K17Steps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, L, steps, checkpoint_every, even0):
        if f0.dtype != L.f.dtype or f0.numel() != L.f.numel() or not f0.is_cuda:
            raise ValueError("f0 must be a device buffer of the lattice's storage type and size")
        stream = torch.cuda.current_stream()
        buf = f0.detach().clone()
        checkpoints = []
        even = even0
        for i in range(steps):
            if i % checkpoint_every == 0:
                checkpoints.append((i, even, buf.clone()))
            _forward_step(L, buf, even, stream)
            even = not even
        ctx.L, ctx.steps, ctx.K, ctx.even0 = L, steps, checkpoint_every, even0
        ctx.checkpoints = checkpoints
        return buf

    @staticmethod
    def backward(ctx, grad_out):
        L, steps, K = ctx.L, ctx.steps, ctx.K
        stream = torch.cuda.current_stream()
        lam = grad_out.detach().to(torch.float32).contiguous().clone()
        for start, even_start, ckpt in reversed(ctx.checkpoints):
            n = min(K, steps - start)
            states = [ckpt]
            work = ckpt.clone()
            even = even_start
            for j in range(n - 1):
                _forward_step(L, work, even, stream)
                even = not even
                states.append(work.clone())
            for j in reversed(range(n)):
                _adjoint_step(L, states[j], lam, even_start ^ (j % 2 == 1), None, stream)
            del states, work
        ctx.checkpoints = None
        return lam.to(grad_out.dtype), None, None, None, None


def k17_steps(f0: torch.Tensor, lattice: Lattice, steps: int, checkpoint_every: int = 10,
              even: bool | None = None) -> torch.Tensor:
    """`steps` collide+stream steps of `lattice` from buffer f0 (differentiable in f0).

    The lattice's own buffer and parity are not touched; `even` is the parity of the first
    step (default: the lattice's current parity). The result is a new buffer whose parity is
    flipped `steps` times.
    """
    _check_differentiable(lattice)
    if steps < 0 or checkpoint_every < 1:
        raise ValueError("steps must be >= 0 and checkpoint_every >= 1")
    even0 = lattice.even if even is None else bool(even)
    return _K17Steps.apply(f0, lattice, steps, checkpoint_every, even0)


def pre_collision_index(grid, even: bool) -> np.ndarray:
    """(27, N) flat indices into a 27*N buffer of every node's pre-collision populations."""
    N = grid.N
    out = np.empty((27, N), dtype=np.int64)
    k = np.arange(N)
    for d in range(27):
        loc = k
        if CX[d] < 0:
            loc = grid.nbx[loc].astype(np.int64)
        if CY[d] < 0:
            loc = grid.nby[loc].astype(np.int64)
        if CZ[d] < 0:
            loc = grid.nbz[loc].astype(np.int64)
        slot = d if even else INV[d]
        out[d] = slot * N + loc
    return out


_CXYZ = None


def node_moments(buf: torch.Tensor, index: torch.Tensor):
    """drho (N) and momentum j (3, N) in FP64 from a buffer and its pre-collision index."""
    global _CXYZ
    f = buf.reshape(-1)[index].to(torch.float64)  # (27, N)
    if _CXYZ is None or _CXYZ.device != f.device:
        _CXYZ = torch.tensor(np.stack([CX, CY, CZ]), dtype=torch.float64, device=f.device)
    drho = f.sum(0)
    j = _CXYZ @ f
    return drho, j


def kinetic_energy(buf: torch.Tensor, index: torch.Tensor, fluid: torch.Tensor | None = None) -> torch.Tensor:
    """FP64 sum over nodes (or the fluid subset) of 1/2 |j|^2 / (1 + drho)."""
    drho, j = node_moments(buf, index)
    ke = 0.5 * (j * j).sum(0) / (1.0 + drho)
    if fluid is not None:
        ke = ke[fluid]
    return ke.sum()
