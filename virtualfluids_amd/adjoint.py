"""Differentiable K17 time stepping: the hand-written adjoint kernels behind torch.autograd.

``k17_steps(f0, lattice, steps, checkpoint_every, bc_factory, velocity, rho_bc)`` runs
`steps` time steps of one Lattice from the population buffer f0 with the order of
UpdateGrid27 (collision, velocity BC, no-slip BC, geometry BC, parity swap, pressure BC;
src/gpu/core/Calculation/UpdateGrid27.cpp:54-96) and returns the final buffer. Its backward
sweeps the steps in reverse with the adjoint kernels (vf_k17_adjoint[_box] and the BC
adjoints): pressure BC, then geometry, no-slip and velocity, then the collision. The forward keeps a
copy of the buffer every `checkpoint_every` steps and steps in place, replaying one HIP graph
per parity; the backward replays one segment at a time from its checkpoint (out of place on
the dense box, so the replay writes each intermediate state once and copies nothing) and walks
the segment backwards. Memory: steps/K + K + 1
buffers of the storage type. Gradients flow to f0 and, when given, to the wall velocities of
the velocity BC (3, n), the outlet densities of the pressure BC (n,) and the Brinkman porosity
of every node (N,).

Objectives are ordinary torch functions of the returned buffer; ``pre_collision_index`` gives
the gather that turns a buffer into the (27, N) pre-collision populations of every node, and
``kinetic_energy`` is the FP64 sum of 1/2 |j|^2 / rho (src/gpu/core/PostProcessor/KineticEnergyAnalyzer.cu:99-153).
No counterpart in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import capi
from .grid import CX, CY, CZ, INV
from .graphs import StepGraphs as _StepGraphs
from .lattice import Lattice, stream_handle
from .simulation import BoundaryConditionFactory


def _param(L: Lattice, state: torch.Tensor, even: bool) -> capi.LBMParameter:
    p = capi.LBMParameter.from_buffer_copy(L.param)
    p.distributions = ctypes.c_void_p(state.data_ptr())
    p.isEvenTimestep = 1 if even else 0
    return p


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _bc_snapshot(L: Lattice):
    """Detached copies of the parameter arrays a step reads from the lattice: the velocity BC's
    wall velocities, the pressure BC's outlet densities and the porosity field (None where the
    lattice has none)."""
    vel, pr = L.bcs.get("velocity"), L.bcs.get("pressure")
    por = getattr(L, "porosity", None)
    return (None if vel is None or vel.vel is None else [v.detach().clone() for v in vel.vel],
            None if pr is None or pr.rho_bc is None else pr.rho_bc.detach().clone(),
            None if por is None else por.detach().clone())


def _bc_restore(L: Lattice, snap):
    """Write a _bc_snapshot back into the lattice's parameter arrays."""
    vel, rho, por = snap
    if vel is not None:
        for dst, src in zip(L.bcs["velocity"].vel, vel):
            dst.copy_(src)
    if rho is not None:
        L.bcs["pressure"].rho_bc.copy_(rho)
    if por is not None or getattr(L, "porosity", None) is not None:
        L.set_porosity(por)


def _bc_struct(bc, post):
    """The BC list's struct, with postCollision pointing at `post` (27 * n FP32) when given."""
    if post is None:
        return bc.struct
    s = capi.QBoundaryConditions.from_buffer_copy(bc.struct)
    s.postCollision = ctypes.c_void_p(post.data_ptr())
    return s


def _storage_input(t: torch.Tensor, L: Lattice, what: str) -> torch.Tensor:
    """A working copy of an initial buffer in the lattice's storage type. The buffer may hold the
    storage type itself or FP32 (for 16-bit lattices: then the outputs and the gradient stay FP32)."""
    if t.numel() != L.f.numel() or not t.is_cuda or t.dtype not in (L.f.dtype, torch.float32):
        raise ValueError(f"{what} must be a device buffer of the lattice's size, in its storage type or FP32")
    return t.detach().to(L.f.dtype, copy=True)


class Stepper:
    """One time step of a Lattice on any buffer, and its adjoint, with the BC functions a
    BoundaryConditionFactory selects (the same selection and checks as Simulation)."""

    def __init__(self, L: Lattice, bc_factory: BoundaryConditionFactory | None = None):
        f = bc_factory or BoundaryConditionFactory()
        self.L = L
        self._launch_groups = None
        self.velocity = f.get_velocity_boundary_condition_post()
        self.noslip = f.get_no_slip_boundary_condition_post()
        self.pressure = f.get_pressure_boundary_condition_pre()
        self.geometry = f.get_geometry_boundary_condition_post()
        for name, fn, label in (("velocity", self.velocity, "velocityBoundaryConditionPost"),
                                ("pressure", self.pressure, "directionalPressureBoundaryConditionPre")):
            bc = L.bcs.get(name)
            if bc is not None and bc.n > 0 and fn is None:
                raise RuntimeError(f"The boundary condition {label} was not set!")

    def _bc(self, name):
        bc = self.L.bcs.get(name)
        return bc if bc is not None and bc.n > 0 else None

    def _groups(self):
        """The post-collision launches of a step, (name, launcher, DeviceBC) in launch order
        (simulation.post_launch_groups); built on first use."""
        if self._launch_groups is None:
            from .simulation import post_launch_groups

            self._launch_groups = post_launch_groups(self.L, [("velocity", self.velocity), ("noslip", self.noslip),
                                                              ("geometry", self.geometry)])
        return self._launch_groups

    def post_lists(self):
        """(name, node count) of the post-collision launches a step makes, in launch order (a
        launch over two joined lists is named "a+b")."""
        return [(name, bc.n) for name, _, bc in self._groups()]

    def forward(self, buf: torch.Tensor, even: bool, out: torch.Tensor | None, stream, save=None) -> torch.Tensor:
        """One step from `buf` (parity `even`); into `out` when given, else in place. Returns
        the buffer holding the result. save: {BC list name: FP32 tensor of 27 * n} to receive
        the post-collision values of that list's nodes (the adjoint's linearisation point)."""
        L, s = self.L, stream_handle(stream)
        if out is not None:
            if L.box:
                capi.call("vf_k17_run_box_into", ctypes.byref(_param(L, buf, even)), _vp(out), s)
            else:
                out.copy_(buf)
                capi.call("vf_k17_run_on_indices", ctypes.byref(_param(L, out, even)), _vp(L.fluid),
                          int(L.grid.fluid.size), capi.COLLISION_DEFAULT, s)
            target = out
        else:
            target = buf
            p = _param(L, buf, even)
            if L.box:
                capi.call("vf_k17_run_box", ctypes.byref(p), capi.COLLISION_DEFAULT, s)
            else:
                capi.call("vf_k17_run_on_indices", ctypes.byref(p), _vp(L.fluid), int(L.grid.fluid.size),
                          capi.COLLISION_DEFAULT, s)
        p = _param(L, target, even)
        for name, fn, bc in self._groups():
            post = save.get(name) if save else None
            capi.call(fn, ctypes.byref(p), ctypes.byref(_bc_struct(bc, post)), s)
        bc = self._bc("pressure")
        if self.pressure and bc is not None:
            capi.call(self.pressure, ctypes.byref(_param(L, target, not even)), ctypes.byref(bc.struct), s)
        return target

    def adjoint(self, before: torch.Tensor, after: torch.Tensor, lam: torch.Tensor, even: bool, ubar, rbar, padj,
                stream, saved=None):
        """Reverse of one step: `before`/`after` are the states around it, `even` its parity;
        lam (FP32) is updated in place; ubar (3*n) / rbar (n) / padj (N) accumulate the adjoints
        of the wall velocities, outlet densities and porosity. saved: the post-collision values
        forward(save=...) stored for this step (the post BC adjoints then recompute nothing)."""
        L, s = self.L, stream_handle(stream)
        bc = self._bc("pressure")
        if self.pressure and bc is not None:
            capi.call(self.pressure + "Adjoint", ctypes.byref(_param(L, after, not even)), ctypes.byref(bc.struct),
                      _vp(lam), _vp(rbar), int(bc.kN_distinct), s)
        p = _param(L, before, even)
        for name, fn, bc in reversed(self._groups()):
            post = saved.get(name) if saved else None
            grad = ubar if name == "velocity" else None
            capi.call(fn + "Adjoint", ctypes.byref(p), ctypes.byref(_bc_struct(bc, post)), _vp(lam), _vp(grad), s)
        fn = "vf_k17_adjoint_box" if L.box else "vf_k17_adjoint"
        capi.call(fn, ctypes.byref(p), _vp(lam), None, _vp(padj), s)


class _K17Steps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, velocity, rho_bc, porosity, stepper, steps, checkpoint_every, even0, graphs, observe,
                observe_from, keep_bytes):
        L = stepper.L
        # in place on one buffer: the same arithmetic as the out-of-place replay of the backward
        buf = _storage_input(f0, L, "f0")
        if velocity is not None:
            bc = L.bcs["velocity"]
            for i in range(3):
                bc.vel[i].copy_(velocity[i].detach().to(torch.float32))
        if rho_bc is not None:
            L.bcs["pressure"].rho_bc.copy_(rho_bc.detach().to(torch.float32))
        if porosity is not None:
            L.set_porosity(porosity.detach())
        # the values this forward used: the backward replays with them, whatever the lattice
        # holds by then (another segment or objective may have set its own)
        ctx.params = _bc_snapshot(L)
        run = _StepGraphs(lambda ev, st: stepper.forward(buf, ev, None, st), graphs)
        checkpoints = []
        # the trailing segments whose states fit keep_bytes are stepped out of place and kept, so
        # the backward walks them without a replay
        lists = stepper.post_lists()
        step_bytes = buf.numel() * buf.element_size() + sum(27 * nb * 4 for _, nb in lists)
        kept_from = steps
        while kept_from > 0:
            start = (kept_from - 1) // checkpoint_every * checkpoint_every
            if (steps - start) * step_bytes > keep_bytes:
                break
            kept_from = start
        kept = {}
        stream = torch.cuda.current_stream()
        even = even0
        J = torch.zeros((), dtype=torch.float64, device=f0.device)
        i = 0
        while i < steps:
            if i % checkpoint_every == 0:
                checkpoints.append((i, even, buf.clone()))
            if i >= kept_from:
                # one segment out of place, every state and post-collision BC value kept
                n = min(checkpoint_every, steps - i)
                states, posts = [checkpoints[-1][2]], []
                for j in range(n):
                    posts.append({name: torch.empty(27 * nb, dtype=torch.float32, device=buf.device)
                                  for name, nb in lists})
                    states.append(stepper.forward(states[-1], even, torch.empty_like(buf), stream, save=posts[j]))
                    even = not even
                    if observe is not None and i + j >= observe_from:
                        J += observe(states[-1], even).detach().to(torch.float64)
                kept[i] = (states, posts)
                buf = states[-1]
                i += n
                continue
            if observe is not None and i >= observe_from:
                run(even)  # one step: the observation follows every step from here on
                even = not even
                J += observe(buf, even).detach().to(torch.float64)
                i += 1
                continue
            # up to the next checkpoint or the first observed step
            n = min(checkpoint_every - i % checkpoint_every, steps - i,
                    (observe_from - i) if observe is not None else steps)
            even = run.run(n, even)
            i += n
        run.close()
        ctx.stepper, ctx.steps, ctx.K = stepper, steps, checkpoint_every
        ctx.checkpoints = checkpoints
        ctx.kept = kept
        ctx.observe, ctx.observe_from = observe, observe_from
        ctx.has_velocity, ctx.has_rho = velocity is not None, rho_bc is not None
        ctx.has_porosity = porosity is not None
        ctx.in_dtype = f0.dtype
        # a kept final state stays the backward's: the caller gets a copy
        return buf.to(f0.dtype, copy=bool(kept)), J

    @staticmethod
    def backward(ctx, grad_out, gJ):
        L = ctx.stepper.L
        current = _bc_snapshot(L)
        _bc_restore(L, ctx.params)
        try:
            return _K17Steps._backward(ctx, grad_out, gJ)
        finally:
            _bc_restore(L, current)

    @staticmethod
    def _backward(ctx, grad_out, gJ):
        stepper, steps, K = ctx.stepper, ctx.steps, ctx.K
        L = stepper.L
        stream = torch.cuda.current_stream()
        if grad_out is None:
            grad_out = torch.zeros(ctx.checkpoints[0][2].shape, dtype=ctx.in_dtype, device=L.device)
        lam = grad_out.detach().to(torch.float32).contiguous().clone()
        observe = ctx.observe if gJ is not None else None
        vel_bc, pr_bc = L.bcs.get("velocity"), L.bcs.get("pressure")
        ubar = torch.zeros(3 * vel_bc.n, dtype=torch.float32, device=lam.device) if ctx.has_velocity else None
        rbar = torch.zeros(pr_bc.n, dtype=torch.float32, device=lam.device) if ctx.has_rho else None
        padj = torch.zeros(L.N, dtype=torch.float32, device=lam.device) if ctx.has_porosity else None
        lists = stepper.post_lists()
        for start, even_start, ckpt in reversed(ctx.checkpoints):
            n = min(K, steps - start)
            if start in ctx.kept:
                states, posts = ctx.kept.pop(start)
            else:
                states = [ckpt]
                # the post-collision values of the BC nodes, step by step, for the BC adjoints
                posts = [{name: torch.empty(27 * nb, dtype=torch.float32, device=lam.device) for name, nb in lists}
                         for _ in range(n)]
                even = even_start
                for j in range(n):
                    states.append(stepper.forward(states[-1], even, torch.empty_like(ckpt), stream, save=posts[j]))
                    even = not even
            for j in reversed(range(n)):
                even_j = even_start ^ (j % 2 == 1)
                if observe is not None and start + j >= ctx.observe_from:
                    # the observed objective of the state after this step
                    _observe_grad(observe, states[j + 1], not even_j, lam, gJ)
                stepper.adjoint(states[j], states[j + 1], lam, even_j, ubar, rbar, padj, stream, saved=posts[j])
            del states, posts
        ctx.checkpoints = None
        ctx.kept = None
        gv = ubar.view(3, -1) if ubar is not None else None
        return lam.to(ctx.in_dtype), gv, rbar, padj, None, None, None, None, None, None, None, None


def k17_steps(f0: torch.Tensor, lattice: Lattice, steps: int, checkpoint_every: int = 10, even: bool | None = None,
              bc_factory: BoundaryConditionFactory | None = None, velocity: torch.Tensor | None = None,
              rho_bc: torch.Tensor | None = None, porosity: torch.Tensor | None = None,
              graphs: bool = True, observe=None, observe_from: int = 0, keep_bytes: int = 0):
    """`steps` time steps of `lattice` from buffer f0, differentiable in f0 and, when given, in
    the velocity BC's wall velocities `velocity` (3, n), the pressure BC's densities `rho_bc`
    and the Brinkman porosity per node `porosity` (N).

    f0 holds the lattice's storage type or FP32: a 16-bit lattice then stores f0 rounded, steps
    in its storage type and returns the final buffer and the gradient with respect to f0 in
    FP32. The backward replays with the wall velocities, outlet densities and porosity this
    call used, so segments with different values can be chained and differentiated together.

    The lattice's own buffer and parity are not touched (its BC parameter arrays take the given
    values); `even` is the parity of the first step (default: the lattice's current parity).
    graphs: the forward sweep replays HIP graphs of the steps (False: every launch from the
    host; the results are the same bit for bit).
    observe(buf, even) -> scalar tensor, a differentiable function of the buffer and its parity
    (KineticEnergyObservation, or any torch function such as a momentum-exchange drag), is
    summed over the states after steps observe_from .. steps-1; then the call returns (buffer,
    observed sum in FP64), and the sum's gradient enters the reverse sweep step by step, so a
    time average needs no stored states beyond the checkpoints. Without observe it returns the
    buffer.

    keep_bytes: device memory the call may hold, beyond the checkpoints, for the states of its
    trailing checkpoint segments (0: none, every segment is replayed in the backward). The
    segments that fit, counted from the last, are stepped out of place in the forward and keep
    every state and the BC nodes' post-collision values, so the backward walks them without a
    replay; a segment costs checkpoint_every x (buffer + 27 x 4 B per BC node) bytes. The
    gradient is the same bit for bit.
    """
    if steps < 0 or checkpoint_every < 1 or keep_bytes < 0:
        raise ValueError("steps must be >= 0, checkpoint_every >= 1 and keep_bytes >= 0")
    stepper = Stepper(lattice, bc_factory)
    if velocity is not None and lattice.bcs.get("velocity") is None:
        raise ValueError("the lattice has no velocity BC nodes")
    if rho_bc is not None and lattice.bcs.get("pressure") is None:
        raise ValueError("the lattice has no pressure BC nodes")
    even0 = lattice.even if even is None else bool(even)
    if not (torch.is_grad_enabled() and any(t is not None and t.requires_grad for t in (f0, velocity, rho_bc, porosity))):
        keep_bytes = 0  # no backward will run: nothing to keep
    out, J = _K17Steps.apply(f0, velocity, rho_bc, porosity, stepper, steps, checkpoint_every, even0, graphs,
                             observe, observe_from, int(keep_bytes))
    return out if observe is None else (out, J)


class _TwoLevelSteps(torch.autograd.Function):
    """Coarse steps of a coarse/fine pair (UpdateGrid27::updateGrid for two levels: the fine level
    steps twice, the coarse level steps, then fine -> coarse and coarse -> fine), differentiable
    in both initial buffers and in the coarse velocity BC values, the coarse outlet densities and
    the porosity of either level. Backward per coarse step, in reverse: the coarse-to-fine and
    fine-to-coarse interface adjoints, the coarse step's adjoint, the two fine steps' adjoints."""

    @staticmethod
    def forward(ctx, fc0, ff0, velocity, rho_bc, por_c, por_f, sc, sf, cells, steps, K, even0, observe, observe_from,
                graphs, in_place, keep_bytes):
        Lc, Lf = sc.L, sf.L
        # in place on one buffer per level: the same arithmetic as the out-of-place replay
        C, F = _storage_input(fc0, Lc, "fc0"), _storage_input(ff0, Lf, "ff0")
        if velocity is not None:
            for i in range(3):
                Lc.bcs["velocity"].vel[i].copy_(velocity[i].detach().to(torch.float32))
        if rho_bc is not None:
            Lc.bcs["pressure"].rho_bc.copy_(rho_bc.detach().to(torch.float32))
        if por_c is not None:
            Lc.set_porosity(por_c.detach())
        if por_f is not None:
            Lf.set_porosity(por_f.detach())
        ctx.params = (_bc_snapshot(Lc), _bc_snapshot(Lf))
        run = _StepGraphs(lambda ev, st: _two_level_step_in_place(sc, sf, cells, C, F, ev, st), graphs)
        checkpoints = []
        # the trailing segments whose states fit keep_bytes are stepped out of place and kept, so
        # the backward walks them without a replay (the states of a coarse step: two fine
        # buffers, the coarse buffer, the interface copies where the writes are not in place, and
        # the BC lists' post-collision values)
        lists_c, lists_f = sc.post_lists(), sf.post_lists()
        step_bytes = (2 * F.numel() * F.element_size() + C.numel() * C.element_size() +
                      (0 if in_place else C.numel() * C.element_size() + F.numel() * F.element_size()) +
                      4 * 27 * (sum(nb for _, nb in lists_c) + 2 * sum(nb for _, nb in lists_f)))
        kept_from = steps
        while kept_from > 0:
            start = (kept_from - 1) // K * K
            if (steps - start) * step_bytes > keep_bytes:
                break
            kept_from = start
        ctx.kept = {}
        stream = torch.cuda.current_stream()
        even = even0
        J = torch.zeros((), dtype=torch.float64, device=fc0.device)
        i = 0
        while i < steps:
            if i % K == 0:
                # a kept segment needs no copy of its first state: it keeps all of them
                checkpoints.append((i, even, C.clone(), F.clone()) if i < kept_from else (i, even, None, None))
            if i >= kept_from:
                # a kept segment: every state of it, as the backward's replay would record it
                n = min(K, steps - i)
                states, posts = [], []
                for j in range(n):
                    pj = ({name: torch.empty(27 * nb, dtype=torch.float32, device=C.device) for name, nb in lists_c},
                          {name: torch.empty(27 * nb, dtype=torch.float32, device=C.device) for name, nb in lists_f},
                          {name: torch.empty(27 * nb, dtype=torch.float32, device=C.device) for name, nb in lists_f})
                    st = _two_level_step(sc, sf, cells, C, F, even, stream, pj, in_place)
                    states.append((even, C, F) + st)
                    posts.append(pj)
                    C, F = st[-2], st[-1]
                    even = not even
                    if observe is not None and i + j >= observe_from:
                        J += observe(C, even).detach().to(torch.float64)
                ctx.kept[i] = (states, posts)
                i += n
                continue
            if observe is not None and i >= observe_from:
                run(even)  # one step: the observation follows every step from here on
                even = not even
                J += observe(C, even).detach().to(torch.float64)
                i += 1
                continue
            # up to the next checkpoint or the first observed step
            n = min(K - i % K, steps - i, (observe_from - i) if observe is not None else steps)
            even = run.run(n, even)
            i += n
        run.close()
        ctx.sc, ctx.sf, ctx.cells, ctx.steps, ctx.K = sc, sf, cells, steps, K
        ctx.in_place = in_place
        ctx.observe, ctx.observe_from = observe, observe_from
        ctx.checkpoints = checkpoints
        ctx.flags = (velocity is not None, rho_bc is not None, por_c is not None, por_f is not None)
        ctx.in_dtypes = (fc0.dtype, ff0.dtype)
        # the outputs are copies where they are kept states
        return C.to(fc0.dtype, copy=bool(ctx.kept)), F.to(ff0.dtype, copy=bool(ctx.kept)), J

    @staticmethod
    def backward(ctx, gC, gF, gJ):
        Lc, Lf = ctx.sc.L, ctx.sf.L
        current = (_bc_snapshot(Lc), _bc_snapshot(Lf))
        _bc_restore(Lc, ctx.params[0])
        _bc_restore(Lf, ctx.params[1])
        try:
            return _TwoLevelSteps._backward(ctx, gC, gF, gJ)
        finally:
            _bc_restore(Lc, current[0])
            _bc_restore(Lf, current[1])

    @staticmethod
    def _backward(ctx, gC, gF, gJ):
        sc, sf, cells, steps, K = ctx.sc, ctx.sf, ctx.cells, ctx.steps, ctx.K
        Lc, Lf = sc.L, sf.L
        stream = torch.cuda.current_stream()
        dev = Lc.device
        if gC is None:
            gC = torch.zeros(Lc.f.shape, dtype=torch.float32, device=dev)
        if gF is None:
            gF = torch.zeros(Lf.f.shape, dtype=torch.float32, device=dev)
        lc = gC.detach().to(torch.float32).contiguous().clone()
        lf = gF.detach().to(torch.float32).contiguous().clone()
        has_u, has_r, has_pc, has_pf = ctx.flags
        ubar = torch.zeros(3 * Lc.bcs["velocity"].n, device=dev) if has_u else None
        rbar = torch.zeros(Lc.bcs["pressure"].n, device=dev) if has_r else None
        pcb = torch.zeros(Lc.N, device=dev) if has_pc else None
        pfb = torch.zeros(Lf.N, device=dev) if has_pf else None
        s = stream_handle(stream)
        cf, fc = cells
        cf_sources, nwork = cf.sources(Lc.grid)
        cf_work = torch.empty(max(nwork, 1), dtype=torch.float32, device=dev)
        lists_c, lists_f = sc.post_lists(), sf.post_lists()

        def post_buffers(lists):
            return {name: torch.empty(27 * nb, dtype=torch.float32, device=dev) for name, nb in lists}

        for start, even_start, C0, F0 in reversed(ctx.checkpoints):
            n = min(K, steps - start)
            if start in ctx.kept:
                states, posts = ctx.kept.pop(start)
            else:
                states = []
                # the post-collision values of every level's BC nodes, step by step, for the BC
                # adjoints
                posts = [(post_buffers(lists_c), post_buffers(lists_f), post_buffers(lists_f)) for _ in range(n)]
                C, F, even = C0, F0, even_start
                for j in range(n):
                    st = _two_level_step(sc, sf, cells, C, F, even, stream, posts[j], ctx.in_place)
                    states.append((even, C, F) + st)
                    C, F = st[-2], st[-1]
                    even = not even
            for j, (even, C0s, F0s, F1, F2, C1, C2, _, _) in reversed(list(enumerate(states))):
                post_c, post_f1, post_f2 = posts[j]
                if ctx.observe is not None and start + j >= ctx.observe_from:
                    # the observed objective of the state after this step
                    _observe_grad(ctx.observe, C2, not even, lc, gJ)
                pc = _param(Lc, C2, not even)
                pf = _param(Lf, F2, True)
                if cf.n:
                    # the transposed interpolation gathered per coarse source node (no atomics)
                    capi.call("vf_ScaleCoarseToFineCompressibleAdjointGather", ctypes.byref(pc), ctypes.byref(pf),
                              ctypes.byref(cf.struct), _neigh(cf), ctypes.byref(cf_sources), _vp(cf_work), _vp(lc),
                              _vp(lf), s)
                if fc.n:
                    capi.call("vf_ScaleFineToCoarseCompressibleAdjoint", ctypes.byref(pc), ctypes.byref(pf),
                              ctypes.byref(fc.struct), _neigh(fc), _vp(lc), _vp(lf), s)
                sc.adjoint(C0s, C1, lc, even, ubar, rbar, pcb, stream, saved=post_c)
                sf.adjoint(F1, F2, lf, False, None, None, pfb, stream, saved=post_f2)
                sf.adjoint(F0s, F1, lf, True, None, None, pfb, stream, saved=post_f1)
            del states, posts
        ctx.checkpoints = None
        ctx.kept = None
        gv = ubar.view(3, -1) if ubar is not None else None
        return (lc.to(ctx.in_dtypes[0]), lf.to(ctx.in_dtypes[1]), gv, rbar, pcb, pfb, None, None, None, None, None,
                None, None, None, None, None, None)


def _observe_grad(observe, buf, even, lam, gJ):
    """lam += gJ * d observe(buf, even) / d buf: on the observation's own cells when it names them
    (cells / value), else through a dense FP32 copy of the buffer."""
    with torch.enable_grad():
        if hasattr(observe, "cells"):
            cells = observe.cells(even)
            vals = buf.reshape(-1)[cells].detach().to(torch.float32).requires_grad_(True)
            (g,) = torch.autograd.grad(observe.value(vals, even), vals)
            # gJ stays on the device (no host synchronisation per observed step)
            lam.index_add_(0, cells.reshape(-1), g.reshape(-1).to(torch.float32) * gJ.to(torch.float32))
        else:
            x = buf.detach().to(torch.float32).requires_grad_(True)
            (g,) = torch.autograd.grad(observe(x, even), x)
            lam.add_(g.to(torch.float32) * gJ.to(torch.float32))


def _neigh(cells):
    return ctypes.byref(cells.neigh) if cells.offsets is not None else None


def _cell_nodes(grid, base) -> np.ndarray:
    """The 8 nodes of the cells at `base` (+x, +y, +z neighbours of the grid), flattened."""
    b = np.asarray(base, dtype=np.int64)
    n1, n2, n4 = grid.nbx[b], grid.nby[b], grid.nbz[b]
    n3 = grid.nby[n1]
    return np.concatenate([b, n1, n2, n3, n4, grid.nbz[n1], grid.nbz[n2], grid.nbz[n3]]).astype(np.int64)


def interface_writes_disjoint(coarse_grid, fine_grid, cf, fc) -> bool:
    """Whether a coarse step's interface writes leave every value that the reverse sweep reads
    from the states before them unchanged, so that the replay may apply them in place: the
    coarse nodes fine -> coarse writes are not the source nodes kN of a coarse pressure BC, and the
    fine nodes coarse -> fine writes are neither the cell nodes fine -> coarse reads nor the kN of a
    fine pressure BC. Each node's cell is a disjoint set of addresses, so disjoint node sets mean
    disjoint memory. cf, fc: InterfaceCells (host arrays)."""
    def kN(grid):
        bc = grid.bcs.get("pressure")
        return np.asarray(bc.kN, dtype=np.int64) if bc is not None and bc.n and bc.kN is not None else np.zeros(0, np.int64)

    if np.intersect1d(np.asarray(fc.coarse, dtype=np.int64), kN(coarse_grid)).size:
        return False
    cf_written = _cell_nodes(fine_grid, cf.fine) if cf.n else np.zeros(0, np.int64)
    fc_read = _cell_nodes(fine_grid, fc.fine) if fc.n else np.zeros(0, np.int64)
    return not (np.intersect1d(cf_written, fc_read).size or np.intersect1d(cf_written, kN(fine_grid)).size)


def _two_level_step(sc, sf, cells, C, F, even, stream, posts=None, in_place=False):
    """One coarse step out of place: returns (F1, F2, C1, C2, C_next, F_next) where F1, F2 follow
    the two fine steps, C1 the coarse step, C2 = C_next the fine -> coarse write and F_next the
    coarse -> fine write. posts: (coarse, first fine, second fine) save dicts of Stepper.forward.
    in_place: the interface writes go into C1 and F2 themselves (C2 is C1, F_next is F2), which
    interface_writes_disjoint allows; else into copies."""
    Lc, Lf = sc.L, sf.L
    cf, fc = cells
    s = stream_handle(stream)
    pc_, pf1, pf2 = posts if posts is not None else (None, None, None)
    F1 = sf.forward(F, True, torch.empty_like(F), stream, save=pf1)
    F2 = sf.forward(F1, False, torch.empty_like(F), stream, save=pf2)
    C1 = sc.forward(C, even, torch.empty_like(C), stream, save=pc_)
    C2 = C1 if in_place else C1.clone()
    F3 = F2 if in_place else F2.clone()
    pc = _param(Lc, C2, not even)
    pf = _param(Lf, F2, True)
    if fc.n:
        capi.call("vf_ScaleFineToCoarseCompressible", ctypes.byref(pc), ctypes.byref(pf), ctypes.byref(fc.struct),
                  _neigh(fc), s)
    pf3 = _param(Lf, F3, True)
    if cf.n:
        capi.call("vf_ScaleCoarseToFineCompressible", ctypes.byref(pc), ctypes.byref(pf3), ctypes.byref(cf.struct),
                  _neigh(cf), s)
    return F1, F2, C1, C2, C2, F3


def _two_level_step_in_place(sc, sf, cells, C, F, even, stream):
    """One coarse step in place on C and F, in the order of _two_level_step."""
    Lc, Lf = sc.L, sf.L
    cf, fc = cells
    sf.forward(F, True, None, stream)
    sf.forward(F, False, None, stream)
    sc.forward(C, even, None, stream)
    s = stream_handle(stream)
    pc = _param(Lc, C, not even)
    pf = _param(Lf, F, True)
    if fc.n:
        capi.call("vf_ScaleFineToCoarseCompressible", ctypes.byref(pc), ctypes.byref(pf), ctypes.byref(fc.struct),
                  _neigh(fc), s)
    if cf.n:
        capi.call("vf_ScaleCoarseToFineCompressible", ctypes.byref(pc), ctypes.byref(pf), ctypes.byref(cf.struct),
                  _neigh(cf), s)


def k17_two_level_steps(fc0: torch.Tensor, ff0: torch.Tensor, coarse: Lattice, fine: Lattice, interface,
                        steps: int, checkpoint_every: int = 10, even: bool | None = None,
                        bc_factory: BoundaryConditionFactory | None = None,
                        fine_bc_factory: BoundaryConditionFactory | None = None, velocity: torch.Tensor | None = None,
                        rho_bc: torch.Tensor | None = None, porosity_coarse: torch.Tensor | None = None,
                        porosity_fine: torch.Tensor | None = None, observe=None, observe_from: int = 0,
                        graphs: bool = True, keep_bytes: int = 0):
    """`steps` coarse steps of a two-level pair from (fc0, ff0), differentiable in both buffers and,
    when given, in the coarse velocity BC values (3, n), the coarse outlet densities and the
    porosity of either level. interface: (cf, fc) InterfaceCells of the pair (NestedBox.cf,
    NestedBox.fc). The fine level takes its steps with the even parity first, as after every
    coarse step; `even` is the coarse parity of the first step (default: the lattice's).

    fc0 / ff0 hold each level's storage type or FP32 (16-bit levels then return their final
    buffers and gradients in FP32). The backward replays with the BC values and porosities this
    call used, whatever the lattices hold by then.

    observe(buf, even) -> scalar tensor, a differentiable function of a coarse buffer and its
    parity, is summed over the coarse states after steps observe_from .. steps-1 and returned as
    the third output (FP64); its gradient enters the reverse sweep step by step, so a time
    average over thousands of steps needs no stored states beyond the checkpoints.
    graphs: the forward sweep replays one HIP graph per coarse parity (False: every launch from
    the host; the results are the same bit for bit).
    keep_bytes: device memory the forward may keep for the backward. The trailing checkpoint
    segments whose coarse steps' states (two fine buffers, the coarse buffer and the BC lists'
    post-collision values per coarse step) fit in it are stepped out of place and keep every
    state, so the backward walks them without a replay; the outputs, the observed sum and the
    gradients are the same bit for bit. Ignored when nothing requires a gradient.
    Returns (coarse buffer, fine buffer, observed sum)."""
    from .refinement import DeviceInterfaceCells

    if steps < 0 or checkpoint_every < 1 or keep_bytes < 0:
        raise ValueError("steps must be >= 0, checkpoint_every >= 1 and keep_bytes >= 0")
    if fine.storage != coarse.storage:
        raise ValueError("both levels must use the same storage type")
    # a coarse step cycles through both levels' populations
    for L in (coarse, fine):
        L.set_working_set(coarse.state_bytes + fine.state_bytes)
    sc = Stepper(coarse, bc_factory)
    sf = Stepper(fine, fine_bc_factory)
    cf, fc = interface
    cells = (DeviceInterfaceCells(cf, coarse.device), DeviceInterfaceCells(fc, coarse.device))
    even0 = coarse.even if even is None else bool(even)
    # the replay of the backward applies the interface in place where that changes nothing it reads
    in_place = interface_writes_disjoint(coarse.grid, fine.grid, cf, fc)
    inputs = (fc0, ff0, velocity, rho_bc, porosity_coarse, porosity_fine)
    if not torch.is_grad_enabled() or not any(t is not None and t.requires_grad for t in inputs):
        keep_bytes = 0
    return _TwoLevelSteps.apply(fc0, ff0, velocity, rho_bc, porosity_coarse, porosity_fine, sc, sf, cells, steps,
                                checkpoint_every, even0, observe, observe_from, graphs, in_place, int(keep_bytes))


# ---- any number of levels ---------------------------------------------------------------------
def _refresh_interface(steppers, cells, level, C, F, F_out, even, stream):
    """The interface refresh after a step of `level` (parity `even`): fine -> coarse into C from
    the finer level's state F, then coarse -> fine from C into F_out (F itself, or a copy)."""
    Lc, Lf = steppers[level].L, steppers[level + 1].L
    cf, fc = cells[level]
    s = stream_handle(stream)
    pc = _param(Lc, C, not even)
    if fc.n:
        capi.call("vf_ScaleFineToCoarseCompressible", ctypes.byref(pc), ctypes.byref(_param(Lf, F, True)),
                  ctypes.byref(fc.struct), _neigh(fc), s)
    if cf.n:
        capi.call("vf_ScaleCoarseToFineCompressible", ctypes.byref(pc), ctypes.byref(_param(Lf, F_out, True)),
                  ctypes.byref(cf.struct), _neigh(cf), s)


def _level_step_in_place(steppers, cells, bufs, level, even, stream):
    """One step of `level` in place on bufs, as UpdateGrid27::updateGrid recurses
    (src/gpu/core/Calculation/UpdateGrid27.cpp:54-96): the next finer level takes two steps (even
    parity first), this level steps, then the interface to the finer level is refreshed."""
    finer = level + 1 < len(steppers)
    if finer:
        _level_step_in_place(steppers, cells, bufs, level + 1, True, stream)
        _level_step_in_place(steppers, cells, bufs, level + 1, False, stream)
    steppers[level].forward(bufs[level], even, None, stream)
    if finer:
        _refresh_interface(steppers, cells, level, bufs[level], bufs[level + 1], bufs[level + 1], even, stream)


def _level_step_record(steppers, cells, bufs, level, even, stream, post_lists):
    """One step of `level` out of place from the states bufs (levels level, level + 1, ...):
    the same arithmetic as _level_step_in_place, keeping what the reverse sweep reads. Returns
    (record, states after the step) with record = (even, state before, state after the step,
    state after the interface writes, finer state the fine -> coarse write read, post-collision
    BC values, records of the finer level's two steps)."""
    finer = level + 1 < len(steppers)
    subs = []
    fine_states = bufs[1:]
    if finer:
        for ev in (True, False):
            rec, fine_states = _level_step_record(steppers, cells, fine_states, level + 1, ev, stream, post_lists)
            subs.append(rec)
    S = bufs[0]
    posts = {name: torch.empty(27 * nb, dtype=torch.float32, device=S.device) for name, nb in post_lists[level]}
    C1 = steppers[level].forward(S, even, torch.empty_like(S), stream, save=posts)
    if not finer:
        return (even, S, C1, C1, None, posts, subs), [C1]
    F2 = fine_states[0]
    C2, F3 = C1.clone(), F2.clone()
    _refresh_interface(steppers, cells, level, C2, F2, F3, even, stream)
    return (even, S, C1, C2, F2, posts, subs), [C2, F3] + fine_states[1:]


def _level_adjoint(steppers, cells, sources, rec, lams, grads, level, stream):
    """Reverse of one recorded step of `level`: the interface adjoints (coarse -> fine gathered per
    coarse source node, then fine -> coarse), this level's step adjoint, then the finer level's
    two steps in reverse. lams: the FP32 adjoint of every level, updated in place; grads[level]:
    (wall velocity, outlet density, porosity) adjoint accumulators or None."""
    even, S, C1, C2, F2, posts, subs = rec
    if level + 1 < len(steppers):
        Lc, Lf = steppers[level].L, steppers[level + 1].L
        cf, fc = cells[level]
        s = stream_handle(stream)
        pc, pf = _param(Lc, C2, not even), _param(Lf, F2, True)
        if cf.n:
            src, work = sources[level]
            capi.call("vf_ScaleCoarseToFineCompressibleAdjointGather", ctypes.byref(pc), ctypes.byref(pf),
                      ctypes.byref(cf.struct), _neigh(cf), ctypes.byref(src), _vp(work), _vp(lams[level]),
                      _vp(lams[level + 1]), s)
        if fc.n:
            capi.call("vf_ScaleFineToCoarseCompressibleAdjoint", ctypes.byref(pc), ctypes.byref(pf),
                      ctypes.byref(fc.struct), _neigh(fc), _vp(lams[level]), _vp(lams[level + 1]), s)
    ubar, rbar, padj = grads[level]
    steppers[level].adjoint(S, C1, lams[level], even, ubar, rbar, padj, stream, saved=posts)
    for sub in reversed(subs):
        _level_adjoint(steppers, cells, sources, sub, lams, grads, level + 1, stream)


class _MultiLevelSteps(torch.autograd.Function):
    """Coarse steps of a level hierarchy (any number of levels), differentiable in every level's
    initial buffer, the coarsest level's wall velocities and outlet densities and every level's
    porosity. Inputs after `meta`: the n initial buffers, velocity, rho_bc, the n porosities."""

    @staticmethod
    def forward(ctx, meta, *args):
        steppers, cells, steps, K, even0, observe, observe_from, graphs, keep_bytes = meta
        n = len(steppers)
        f0s, velocity, rho_bc, pors = args[:n], args[n], args[n + 1], args[n + 2:]
        lats = [s.L for s in steppers]
        bufs = [_storage_input(f, L, f"level {i} buffer") for i, (f, L) in enumerate(zip(f0s, lats))]
        if velocity is not None:
            for i in range(3):
                lats[0].bcs["velocity"].vel[i].copy_(velocity[i].detach().to(torch.float32))
        if rho_bc is not None:
            lats[0].bcs["pressure"].rho_bc.copy_(rho_bc.detach().to(torch.float32))
        for L, p in zip(lats, pors):
            if p is not None:
                L.set_porosity(p.detach())
        # the values this forward used: the backward replays with them
        ctx.params = [_bc_snapshot(L) for L in lats]
        run = _StepGraphs(lambda ev, st: _level_step_in_place(steppers, cells, bufs, 0, ev, st), graphs)
        checkpoints = []
        # the trailing segments whose records fit keep_bytes are stepped out of place and kept, so
        # the backward walks them without a replay. A step of level l (2^l per coarse step) keeps
        # the level's state after it, the BC lists' post-collision values and, above the finest
        # level, the level's state after the interface writes and a copy of the finer level's.
        post_lists = [s.post_lists() for s in steppers]
        nbytes = [b.numel() * b.element_size() for b in bufs]
        step_bytes = sum(2 ** l * (nbytes[l] + 4 * 27 * sum(nb for _, nb in post_lists[l]) +
                                   (nbytes[l] + nbytes[l + 1] if l + 1 < n else 0)) for l in range(n))
        kept_from = steps
        while kept_from > 0:
            start = (kept_from - 1) // K * K
            if (steps - start) * step_bytes > keep_bytes:
                break
            kept_from = start
        ctx.kept = {}
        stream = torch.cuda.current_stream()
        even = even0
        J = torch.zeros((), dtype=torch.float64, device=bufs[0].device)
        i = 0
        while i < steps:
            if i % K == 0:
                # a kept segment needs no copy of its first states: it keeps all of them
                checkpoints.append((i, even, [b.clone() for b in bufs] if i < kept_from else None))
            if i >= kept_from:
                records = []
                for j in range(min(K, steps - i)):
                    rec, bufs = _level_step_record(steppers, cells, bufs, 0, even, stream, post_lists)
                    records.append(rec)
                    even = not even
                    if observe is not None and i + j >= observe_from:
                        J += observe(bufs[0], even).detach().to(torch.float64)
                ctx.kept[i] = records
                i += len(records)
                continue
            if observe is not None and i >= observe_from:
                run(even)  # one step: the observation follows every step from here on
                even = not even
                J += observe(bufs[0], even).detach().to(torch.float64)
                i += 1
                continue
            n_run = min(K - i % K, steps - i, (observe_from - i) if observe is not None else steps)
            even = run.run(n_run, even)
            i += n_run
        run.close()
        ctx.meta = meta
        ctx.checkpoints = checkpoints
        ctx.flags = (velocity is not None, rho_bc is not None, [p is not None for p in pors])
        ctx.in_dtypes = [f.dtype for f in f0s]
        # the outputs are copies where they are kept states
        return tuple(b.to(f.dtype, copy=bool(ctx.kept)) for b, f in zip(bufs, f0s)) + (J,)

    @staticmethod
    def backward(ctx, *grads):
        steppers = ctx.meta[0]
        current = [_bc_snapshot(s.L) for s in steppers]
        for s, p in zip(steppers, ctx.params):
            _bc_restore(s.L, p)
        try:
            return _MultiLevelSteps._backward(ctx, grads[:-1], grads[-1])
        finally:
            for s, p in zip(steppers, current):
                _bc_restore(s.L, p)

    @staticmethod
    def _backward(ctx, gbufs, gJ):
        steppers, cells, steps, K, even0, observe, observe_from, graphs, _ = ctx.meta
        n = len(steppers)
        lats = [s.L for s in steppers]
        dev = lats[0].device
        stream = torch.cuda.current_stream()
        lams = [(torch.zeros(L.f.shape, dtype=torch.float32, device=dev) if g is None
                 else g.detach().to(torch.float32).contiguous().clone()) for g, L in zip(gbufs, lats)]
        has_u, has_r, has_p = ctx.flags
        grads = [(torch.zeros(3 * L.bcs["velocity"].n, device=dev) if (i == 0 and has_u) else None,
                  torch.zeros(L.bcs["pressure"].n, device=dev) if (i == 0 and has_r) else None,
                  torch.zeros(L.N, device=dev) if has_p[i] else None) for i, L in enumerate(lats)]
        sources = []
        for (cf, _), L in zip(cells, lats):
            src, nwork = cf.sources(L.grid)
            sources.append((src, torch.empty(max(nwork, 1), dtype=torch.float32, device=dev)))
        post_lists = [s.post_lists() for s in steppers]
        if gJ is not None and observe is None:
            gJ = None
        for start, even_start, bufs0 in reversed(ctx.checkpoints):
            if start in ctx.kept:
                records, states = ctx.kept.pop(start), None
            else:
                records, states, even = [], bufs0, even_start
                for _ in range(min(K, steps - start)):
                    rec, states = _level_step_record(steppers, cells, states, 0, even, stream, post_lists)
                    records.append(rec)
                    even = not even
            for j in reversed(range(len(records))):
                rec = records[j]
                if gJ is not None and start + j >= observe_from:
                    # the observed objective of the coarse state after this step
                    _observe_grad(observe, rec[3], not rec[0], lams[0], gJ)
                _level_adjoint(steppers, cells, sources, rec, lams, grads, 0, stream)
            del records, states
        ctx.checkpoints = None
        ctx.kept = None
        ubar, rbar = grads[0][0], grads[0][1]
        return ((None,) + tuple(lam.to(dt) for lam, dt in zip(lams, ctx.in_dtypes))
                + (ubar.view(3, -1) if ubar is not None else None, rbar) + tuple(g[2] for g in grads))


def k17_multilevel_steps(f0s, lattices, interfaces, steps: int, checkpoint_every: int = 10, even: bool | None = None,
                         bc_factories=None, velocity: torch.Tensor | None = None, rho_bc: torch.Tensor | None = None,
                         porosity=None, observe=None, observe_from: int = 0, graphs: bool = True,
                         keep_bytes: int = 0):
    """`steps` coarse steps of a hierarchy of any number of levels (level 0 coarsest) from the
    buffers f0s, differentiable in every buffer and, when given, in the coarsest level's velocity BC
    values (3, n) and outlet densities and in each level's porosity (porosity: one tensor or None
    per level). interfaces: one (cf, fc) InterfaceCells pair per pair of neighbouring levels.

    A coarse step follows UpdateGrid27::updateGrid (src/gpu/core/Calculation/UpdateGrid27.cpp:54-96):
    every finer level takes two steps per step of the level above it (even parity first, so every
    finer level starts with the even parity), then the level steps and refreshes the interface to
    its finer level (fine -> coarse, coarse -> fine). The forward works in place and replays HIP
    graphs of the coarse step (graphs=False: launches from the host, the same bit for bit); the
    backward replays each checkpoint segment out of place, keeping every level's intermediate
    states, and walks it in reverse: per step of a level, the interface adjoints, the level's step
    adjoint, then its finer level's two steps. For two levels it computes what k17_two_level_steps
    computes. f0s hold each level's storage type or FP32; observe / observe_from as in
    k17_two_level_steps, on the coarsest level; keep_bytes as in k17_two_level_steps (a kept coarse
    step holds every level's records: per step of level l its state after the step, its BC lists'
    post-collision values and, above the finest level, its state after the interface writes and a
    copy of the finer level's). Returns (list of final buffers, observed sum)."""
    from .refinement import DeviceInterfaceCells

    n = len(lattices)
    if n < 2 or len(interfaces) != n - 1 or len(f0s) != n:
        raise ValueError("a hierarchy needs two or more levels, one buffer per level and one interface per level pair")
    if steps < 0 or checkpoint_every < 1 or keep_bytes < 0:
        raise ValueError("steps must be >= 0, checkpoint_every >= 1 and keep_bytes >= 0")
    if any(L.storage != lattices[0].storage for L in lattices):
        raise ValueError("every level must use the same storage type")
    pors = list(porosity) if porosity is not None else [None] * n
    if len(pors) != n:
        raise ValueError("porosity: one tensor or None per level")
    if velocity is not None and lattices[0].bcs.get("velocity") is None:
        raise ValueError("the coarsest level has no velocity BC nodes")
    if rho_bc is not None and lattices[0].bcs.get("pressure") is None:
        raise ValueError("the coarsest level has no pressure BC nodes")
    # a coarse step cycles through every level's populations
    total = sum(L.state_bytes for L in lattices)
    for L in lattices:
        L.set_working_set(total)
    facs = list(bc_factories) if bc_factories is not None else [None] * n
    steppers = [Stepper(L, f) for L, f in zip(lattices, facs)]
    cells = [(DeviceInterfaceCells(cf, L.device), DeviceInterfaceCells(fc, L.device))
             for (cf, fc), L in zip(interfaces, lattices)]
    even0 = lattices[0].even if even is None else bool(even)
    inputs = list(f0s) + [velocity, rho_bc] + pors
    if not torch.is_grad_enabled() or not any(t is not None and t.requires_grad for t in inputs):
        keep_bytes = 0  # no backward will run: nothing to keep
    meta = (steppers, cells, steps, checkpoint_every, even0, observe, observe_from, graphs, int(keep_bytes))
    out = _MultiLevelSteps.apply(meta, *f0s, velocity, rho_bc, *pors)
    return list(out[:n]), out[n]


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


_CXYZ = {}


def node_moments(buf: torch.Tensor, index: torch.Tensor):
    """drho (N) and momentum j (3, N) in FP64 from a buffer and its pre-collision index."""
    f = buf.reshape(-1)[index].to(torch.float64)  # (27, N)
    c = _CXYZ.get(f.device)
    if c is None:
        c = _CXYZ[f.device] = torch.tensor(np.stack([CX, CY, CZ]), dtype=torch.float64, device=f.device)
    return f.sum(0), c @ f


def kinetic_energy(buf: torch.Tensor, index: torch.Tensor, nodes: torch.Tensor | None = None) -> torch.Tensor:
    """FP64 sum over all nodes (or the given node subset) of 1/2 |j|^2 / (1 + drho)."""
    if nodes is not None:
        index = index[:, nodes]
    drho, j = node_moments(buf, index)
    return (0.5 * (j * j).sum(0) / (1.0 + drho)).sum()


class KineticEnergyObservation:
    """scale * kinetic energy (FP64 sum of 1/2 |j|^2 / (1 + drho)) of a node subset of a state: an
    observation for k17_two_level_steps that reads only the 27 cells of each of those nodes, so
    its gradient is scattered into the adjoint without touching the rest of the buffer."""

    def __init__(self, grid, nodes, scale: float = 1.0, device="cuda"):
        nodes = np.asarray(nodes, dtype=np.int64)
        self.index = {e: torch.from_numpy(pre_collision_index(grid, e)[:, nodes]).to(device) for e in (True, False)}
        self.scale = float(scale)

    def cells(self, even: bool) -> torch.Tensor:
        """(27, M) flat indices of the cells the observation reads in a state of parity `even`."""
        return self.index[bool(even)]

    def value(self, f: torch.Tensor, even: bool) -> torch.Tensor:
        """The observation from the values of its cells, (27, M)."""
        f = f.to(torch.float64)
        c = _CXYZ.get(f.device)
        if c is None:
            c = _CXYZ[f.device] = torch.tensor(np.stack([CX, CY, CZ]), dtype=torch.float64, device=f.device)
        drho, j = f.sum(0), c @ f
        return self.scale * (0.5 * (j * j).sum(0) / (1.0 + drho)).sum()

    def __call__(self, buf: torch.Tensor, even: bool) -> torch.Tensor:
        return self.value(buf.reshape(-1)[self.index[bool(even)]], even)


class MomentumExchangeObservation:
    """scale * component `axis` of the momentum-exchange force on a BC list (momentum_exchange):
    an observation for k17_steps / k17_two_level_steps that reads only the link cells, so its
    gradient is scattered into the adjoint without touching the rest of the buffer."""

    def __init__(self, grid, bc_name: str, axis: int = 0, scale: float = 1.0, device="cuda"):
        self.index, self.weight = {}, {}
        for e in (True, False):
            out_idx, in_idx, c = momentum_exchange_index(grid, bc_name, e)
            self.index[e] = torch.from_numpy(np.concatenate([out_idx, in_idx])).to(device)
            self.weight[e] = torch.from_numpy(np.concatenate([c[axis], c[axis]])).to(device)
        self.scale = float(scale)

    def cells(self, even: bool) -> torch.Tensor:
        """Flat indices of the cells the observation reads in a state of parity `even`."""
        return self.index[bool(even)]

    def value(self, f: torch.Tensor, even: bool) -> torch.Tensor:
        """The observation from the values of its cells."""
        return self.scale * (self.weight[bool(even)] * f.to(torch.float64)).sum()

    def __call__(self, buf: torch.Tensor, even: bool) -> torch.Tensor:
        return self.value(buf.reshape(-1)[self.index[bool(even)]], even)


def _cell_index(grid, k: np.ndarray, d: int, even: bool) -> np.ndarray:
    """Flat buffer index of the cell holding pre-collision direction d of nodes k."""
    loc = k.astype(np.int64)
    if CX[d] < 0:
        loc = grid.nbx[loc].astype(np.int64)
    if CY[d] < 0:
        loc = grid.nby[loc].astype(np.int64)
    if CZ[d] < 0:
        loc = grid.nbz[loc].astype(np.int64)
    slot = d if even else INV[d]
    return slot * grid.N + loc


def momentum_exchange_index(grid, bc_name: str, even: bool):
    """Cells of every active link of a BC list in a state of parity `even` taken right after a
    step: for link d of node n, the population that left n towards the wall (now in the
    pre-collision d cell of the node across the link) and the one the BC sent back (n's
    pre-collision inv(d) cell). Returns (out_idx, in_idx, c) with c the link directions (3, M)."""
    bc = grid.bcs[bc_name]
    q = bc.q27.reshape(27, bc.n)
    x, y, z = grid.coords()
    outs, ins, cs = [], [], []
    for d in range(1, 27):
        act = (q[d] >= 0) & (q[d] <= 1)
        if not np.any(act):
            continue
        n = bc.k[act].astype(np.int64)
        m = grid.index((x[n] + CX[d]) % grid.nx, (y[n] + CY[d]) % grid.ny, (z[n] + CZ[d]) % grid.nz)
        outs.append(_cell_index(grid, m, d, even))
        ins.append(_cell_index(grid, n, INV[d], even))
        cs.append(np.repeat(np.array([[CX[d]], [CY[d]], [CZ[d]]]), n.size, axis=1))
    return np.concatenate(outs), np.concatenate(ins), np.concatenate(cs, axis=1).astype(np.float64)


def momentum_exchange(buf: torch.Tensor, out_idx: torch.Tensor, in_idx: torch.Tensor, c: torch.Tensor) -> torch.Tensor:
    """Force on the boundary (3,) in FP64 from the momentum its links exchanged in the last step:
    sum over links of c_d (f_d out + f_inv(d) back), the momentum-exchange form of the drag and
    lift sums of src/gpu/core/PostProcessor/DragLift.cu:50-300."""
    flat = buf.reshape(-1)
    return (c * (flat[out_idx].to(torch.float64) + flat[in_idx].to(torch.float64))).sum(1)
