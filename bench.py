"""Bandwidth benchmark of the K17 time step on one MI355X.

Workload: Taylor-Green vortex, D3Q27 cumulant K17, 256^3 fully periodic, FP32 shifted
populations and FP32 arithmetic, plain forward step (collide + esoteric-twist stream in
place), u0 = 0.05, Re = 1600. A step is one launch of the dense-box K17 kernel over all
16,777,216 nodes; inputs are resident in HBM before timing starts.

Prints ONE JSON line: MLUPS, the kernel's HBM roofline (algorithmic bytes 2*27*4 = 216 B
per node update, divided by the event-timed average launch duration), the FP16/BF16
storage variants measured the same way, the adjoint step (kernel alone, and a whole
checkpointed forward + backward through k17_steps) against the forward step, a measured
copy bandwidth, and the CPU oracle timed on the host cores on a bounded sample; beside it the
full-size sphere-channel and two-level gradients, the reference's published cavity benchmark,
and the other configurations' GPU rates with the CPU oracle on the same cases
(`configs_gpu_cpu`: the 64^3 cavity, Taylor-Green 512^3 in FP32/FP16/BF16).

    python bench.py [--gpus 1] [--steps 2000] [--warmup 20]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, AMD public spec
METRIC = "MLUPS + %HBM roofline, D3Q27 K17 256^3; FP16-store and adjoint vs plain FP32"


def time_steps(L, steps, warmup):
    """Event-timed K17 launches on torch's current stream (the stream the kernel runs on)."""
    import torch

    for _ in range(warmup):
        L.collide()
        L.swap()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        L.collide()
        L.swap()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / 1000.0


def time_adjoint(L, steps, warmup):
    """Event-timed vf_k17_adjoint_box launches (FP32 adjoint buffer, state fixed)."""
    import ctypes

    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import stream_handle

    lam = torch.full_like(L.f, 1e-3, dtype=torch.float32)
    lp = ctypes.c_void_p(lam.data_ptr())

    def step():
        capi.call("vf_k17_adjoint_box", L.pref(), lp, None, None, stream_handle())
        L.swap()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        step()
    stop.record()
    torch.cuda.synchronize()
    del lam
    return start.elapsed_time(stop) / 1000.0


def time_gradient(L, steps, checkpoint_every):
    """Wall time of k17_steps forward + backward (checkpoints, replay, adjoint) over `steps`."""
    import torch

    from virtualfluids_amd.adjoint import k17_steps

    x = L.f.detach().clone().requires_grad_(True)
    out = k17_steps(x, L, 2, checkpoint_every=1)
    out.backward(torch.ones_like(out))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = k17_steps(x, L, steps, checkpoint_every=checkpoint_every)
    out.backward(torch.ones_like(out))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    del out, x
    torch.cuda.empty_cache()
    return dt


def sphere_channel_gradient(steps=2000, checkpoint_every=100, average_last=200, omega=1.8, keep_fraction=0.6):
    """Channel flow past a sphere at full size (192 x 96 x 96 fluid nodes, sphere of radius 12
    at (48, 48, 48), velocity inflow 0.04, pressure outlet, no-slip walls, interpolated
    bounce-back on the sphere) with a seeded Gaussian random porosity field in [0, 0.1]
    (correlation length 8). Objective: the x drag by momentum exchange averaged over the last
    `average_last` of `steps` steps; its gradient with respect to the porosity field through
    k17_steps, checkpointing every `checkpoint_every` steps. Times a forward-only run of the
    same objective and two forward + backward runs: every segment replayed in the backward, and
    with keep_bytes = keep_fraction of the free device memory, where the trailing segments that
    fit keep their states from the forward (the same gradient bit for bit)."""
    import numpy as np
    import torch

    from virtualfluids_amd.adjoint import MomentumExchangeObservation, k17_steps
    from virtualfluids_amd.grid import gaussian_random_field, sphere_channel
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory

    g = sphere_channel((192, 96, 96), (48, 48, 48), 12.0, 0.04)
    phi0 = gaussian_random_field((g.nx, g.ny, g.nz), 8.0, seed=20261202, low=0.0, high=0.1)
    F = BoundaryConditionFactory
    fac = F()
    fac.set_velocity_boundary_condition(F.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_geometry_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(F.PressureBC.PressureNonEquilibriumCompressible)
    L = Lattice(g, omega)
    zeros = np.zeros(g.N)
    L.init_equilibrium(zeros, zeros, zeros, zeros)
    spin_up = steps - average_last
    # x drag by momentum exchange on the sphere's links, averaged over the window
    drag = MomentumExchangeObservation(g, "geometry", axis=0, scale=1.0 / average_last)

    def objective(phi, n=steps, observe_from=spin_up, keep=0):
        return k17_steps(L.f.detach().clone(), L, n, checkpoint_every=checkpoint_every, even=True,
                         bc_factory=fac, porosity=phi, observe=drag, observe_from=observe_from, keep_bytes=keep)[1]

    phi = torch.from_numpy(phi0.astype(np.float32)).cuda()
    # untimed warm-up over two checkpoint segments, forward and backward: first-use costs (code
    # loading, graph capture, allocator growth) stay out of the timed runs
    w = phi.clone().requires_grad_(True)
    objective(w, 2 * checkpoint_every, checkpoint_every).backward()
    del w
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        J0 = float(objective(phi))
    torch.cuda.synchronize()
    t_forward = time.perf_counter() - t0

    def gradient(keep):
        p = phi.detach().clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        objective(p, keep=keep).backward()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, p.grad

    t_recompute, grad_recompute = gradient(0)
    keep = int(keep_fraction * torch.cuda.mem_get_info()[0])
    t_gradient, grad = gradient(keep)
    seg_bytes = checkpoint_every * 4 * 27 * (g.N + sum(bc.n for name, bc in L.bcs.items() if name != "pressure"))
    kept = min(-(-steps // checkpoint_every), keep // seg_bytes)
    out = {"case": "sphere channel 192x96x96 fluid, r 12, inflow 0.04, Gaussian porosity [0, 0.1] (corr. 8), "
                   f"omega {omega}, {steps} steps, checkpoint every {checkpoint_every}, "
                   f"drag averaged over the last {average_last}",
           "nodes": int(g.N), "drag": round(J0, 6), "grad_l2": float(grad.double().norm()),
           "forward_s": round(t_forward, 3), "gradient_s": round(t_gradient, 3),
           "gradient_vs_forward": round(t_gradient / t_forward, 3),
           "reverse_vs_forward": round((t_gradient - t_forward) / t_forward, 3),
           "keep_GB": round(keep / 1e9, 1), "kept_segments": int(kept),
           "segments": -(-steps // checkpoint_every),
           "recompute": {"gradient_s": round(t_recompute, 3),
                         "reverse_vs_forward": round((t_recompute - t_forward) / t_forward, 3)},
           "kept_equals_recompute": bool(torch.equal(grad, grad_recompute))}
    del L, phi, grad, grad_recompute
    torch.cuda.empty_cache()
    return out


def two_level_gradient(steps=4000, average_last=1000, checkpoint_every=50, storage="fp16", modes=(8, 4),
                       keep_fraction=0.6):
    """Two-level 2:1 nested refinement: coarse channel of 256 x 128 x 128 fluid nodes (velocity
    inflow, pressure outlet, no-slip walls), Re 500 on the coarse level (inflow 0.03, channel
    height 128), a fine box of 128^3 nodes over the coarse cells [64, 128) x [32, 96)^2 around a
    porous obstacle (a seeded Gaussian random porosity in [0, 0.1], correlation length 8 fine
    nodes, inside a ball of radius 24 fine nodes), FP16 shifted storage on both levels. The
    inflow profile is u_x(y, z) = sum_mn a_mn sin(m pi y^) sin(n pi z^) over 8 x 4 modes, started
    from the sine series of the uniform 0.03 inflow. Objective: the kinetic energy of the last
    coarse fluid plane before the outlet, averaged over the last `average_last` coarse steps; its
    gradient with respect to the 32 coefficients through k17_two_level_steps. Times a
    forward-only run and two forward + backward runs: every segment replayed in the backward,
    and (keep_fraction > 0) with keep_bytes = keep_fraction of the free device memory, where the
    trailing segments that fit keep their states from the forward (the same gradient bit for
    bit)."""
    import numpy as np
    import torch

    from virtualfluids_amd.adjoint import KineticEnergyObservation, k17_two_level_steps
    from virtualfluids_amd.grid import channel, gaussian_random_field
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.refinement import nested_box
    from virtualfluids_amd.simulation import BoundaryConditionFactory

    u0, re = 0.03, 500.0
    nu = u0 * 128 / re
    omega_c, omega_f = 1.0 / (3 * nu + 0.5), 1.0 / (6 * nu + 0.5)
    nb = nested_box(None, (65, 33, 33), (129, 97, 97), coarse=channel((256, 128, 128), inflow=u0))
    xf, yf, zf = nb.fine.coords()
    r2 = (xf - 63.5) ** 2 + (yf - 63.5) ** 2 + (zf - 63.5) ** 2
    grf = gaussian_random_field((nb.fine.nx, nb.fine.ny, nb.fine.nz), 8.0, seed=20261202, low=0.0, high=0.1)
    phi = np.where(r2 <= 24.0 ** 2, grf, 0.0).astype(np.float32)
    F = BoundaryConditionFactory
    fac = F()
    fac.set_velocity_boundary_condition(F.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(F.PressureBC.PressureNonEquilibriumCompressible)
    C = Lattice(nb.coarse, omega_c, storage=storage, level=0)
    Fl = Lattice(nb.fine, omega_f, storage=storage, level=1)
    for L in (C, Fl):
        z = np.zeros(L.N)
        L.init_equilibrium(z, np.where(L.grid.geo == 19, u0, 0.0), z, z)
    # inflow profile basis on the inflow nodes
    vel = nb.coarse.bcs["velocity"]
    _, yk, zk = nb.coarse.coords()
    yh = (yk[vel.k] - 0.5) / 128.0
    zh = (zk[vel.k] - 0.5) / 128.0
    basis = np.stack([np.sin(m * np.pi * yh) * np.sin(n * np.pi * zh)
                      for m in range(1, modes[0] + 1) for n in range(1, modes[1] + 1)])  # (32, nbc)
    a0 = np.array([u0 * 16 / (np.pi ** 2 * m * n) if (m % 2 and n % 2) else 0.0
                   for m in range(1, modes[0] + 1) for n in range(1, modes[1] + 1)])
    B = torch.from_numpy(basis.astype(np.float32)).cuda()
    x, _, _ = nb.coarse.coords()
    plane = np.nonzero((x == nb.coarse.nx - 3) & (nb.coarse.geo == 19))[0]
    outlet_ke = KineticEnergyObservation(nb.coarse, plane, scale=1.0 / average_last)
    phi_t = torch.from_numpy(phi).cuda()
    zero = torch.zeros(vel.n, device="cuda")

    def objective(a, n=steps, observe_from=steps - average_last, keep=0):
        ux = a.to(torch.float32) @ B
        U = torch.stack([ux, zero, zero])
        return k17_two_level_steps(C.f.detach().clone(), Fl.f.detach().clone(), C, Fl, (nb.cf, nb.fc), n,
                                   checkpoint_every=checkpoint_every, even=True, bc_factory=fac, velocity=U,
                                   porosity_fine=phi_t,
                                   observe=outlet_ke, observe_from=observe_from, keep_bytes=keep)[2]

    a = torch.from_numpy(a0).cuda()
    # untimed warm-up over two checkpoint segments, forward and backward (first-use costs)
    w = a.clone().requires_grad_(True)
    objective(w, 2 * checkpoint_every, checkpoint_every).backward()
    del w
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        J0 = float(objective(a))
    torch.cuda.synchronize()
    t_forward = time.perf_counter() - t0
    def gradient(keep):
        x = a.detach().clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        objective(x, keep=keep).backward()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, x.grad.double().cpu().numpy()

    t_recompute, g_recompute = gradient(0)
    keep = int(keep_fraction * torch.cuda.mem_get_info()[0])
    t_gradient, g = gradient(keep) if keep_fraction > 0 else (t_recompute, g_recompute)
    # the states a kept coarse step holds (two fine buffers, the coarse buffer, the BC values)
    lists = 27 * 4 * sum(bc.n for name, bc in C.bcs.items() if name != "pressure")
    step_bytes = 2 * Fl.f.numel() * Fl.f.element_size() + C.f.numel() * C.f.element_size() + lists
    segments = -(-steps // checkpoint_every)
    kept = min(segments, keep // (checkpoint_every * step_bytes)) if keep_fraction > 0 else 0
    nodes = int(nb.coarse.fluid.size + nb.fine.N)
    out = {"case": f"two-level channel 256x128x128 coarse + 128^3 fine around a porous obstacle, Re 500, {storage} "
                   f"storage, {steps} coarse steps, checkpoint every {checkpoint_every}, outlet kinetic energy "
                   f"averaged over the last {average_last}; gradient w.r.t. {len(a0)} inflow Fourier coefficients",
           "fluid_nodes": nodes, "objective": J0, "grad_l2": float(np.linalg.norm(g)),
           "grad_first4": [float(v) for v in g[:4]], "_grad": g,
           "forward_s": round(t_forward, 3), "gradient_s": round(t_gradient, 3),
           "gradient_vs_forward": round(t_gradient / t_forward, 3),
           "reverse_vs_forward": round((t_gradient - t_forward) / t_forward, 3),
           "forward_MLUPS": round((nb.coarse.fluid.size + 2 * nb.fine.N) * steps / t_forward / 1e6, 1)}
    if keep_fraction > 0:
        out.update({"keep_GB": round(keep / 1e9, 1), "kept_segments": int(kept), "segments": segments,
                    "recompute": {"gradient_s": round(t_recompute, 3),
                                  "reverse_vs_forward": round((t_recompute - t_forward) / t_forward, 3)},
                    "kept_equals_recompute": bool(np.array_equal(g, g_recompute))})
    del C, Fl
    torch.cuda.empty_cache()
    return out


# Time-to-solution and NUPS the reference publishes for its GPU DrivenCavity app (K17, 10,000
# steps; VirtualFluids.pdf p.12, Table 4), in seconds and millions of node updates per second
PUBLISHED_CAVITY = {"NVIDIA A100-PCIE-40GB": (1.73, 1658.6), "Tesla V100-SXM2-32GB": (1.47, 1950.1),
                    "Tesla V100-PCIE-16GB": (1.44, 2001.5), "NVIDIA RTX 5090": (0.27, 10852.3)}


def published_cavity(steps=10000):
    """The reference's published GPU benchmark, its DrivenCavity app
    (apps/gpu/DrivenCavity/DrivenCavity.cpp:68-162): K17, 64 nodes per side (66^3 nodes with the
    wall layer, the node count its NUPS figures imply), lid on z-max moving diagonally at 0.05
    (vx = vy = 0.05 / sqrt 2), Re 1000 (nu = 64 * 0.05 / 1000), NoSlipBounceBack and
    VelocityBounceBack, FP32, 10,000 steps through Simulation.run (HIP-graph replay), no
    output files. Timed from the first step to the last, initialisation excluded."""
    import numpy as np
    import torch

    from virtualfluids_amd.grid import d0M0, d0P0, d00M, d00P, dM00, dP00, walled_box
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory, Simulation

    n, u_lid, re = 64, 0.05, 1000.0
    walls = {s: ("noslip",) for s in (dP00, dM00, d0P0, d0M0, d00M)}
    walls[d00P] = ("velocity", (u_lid / np.sqrt(2.0), u_lid / np.sqrt(2.0), 0.0))
    g = walled_box((n, n, n), walls)
    nu = n * u_lid / re
    F = BoundaryConditionFactory
    fac = F()
    fac.set_velocity_boundary_condition(F.VelocityBC.VelocityBounceBack)
    fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipBounceBack)
    L = Lattice(g, 1.0 / (3.0 * nu + 0.5))
    z = np.zeros(g.N)
    L.init_equilibrium(z, z, z, z)
    sim = Simulation(L, fac)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(steps)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    L.macroscopic()
    umax = float(torch.sqrt(L.vx ** 2 + L.vy ** 2 + L.vz ** 2).max())
    out = {"case": "K17 lid-driven cavity, 64 nodes per side + wall layer (66^3), diagonal lid 0.05, Re 1000, "
                   "NoSlipBounceBack + VelocityBounceBack, FP32, 10,000 steps",
           "nodes": int(g.N), "steps": steps, "time_to_solution_s": round(t, 4),
           "MNUPS": round(g.N * steps / t / 1e6, 1), "max_speed": round(umax, 5),
           "published": {k: {"time_to_solution_s": v[0], "MNUPS": v[1]} for k, v in PUBLISHED_CAVITY.items()}}
    del sim, L
    torch.cuda.empty_cache()
    return out


def configs_gpu_cpu(cavity_steps=1000, n3=512, gpu_steps=100, cpu_seconds=5.0):
    """The other configurations' throughput on the GPU with the CPU oracle (FP64 build, OpenMP
    on the host cores) on the same case beside it, each CPU leg bounded to about cpu_seconds of
    steps. Config 1: lid-driven cavity, 64^3 fluid nodes and the wall layer, lid 0.05 in x on the
    z-max face, Re 100, NoSlipBounceBack + VelocityBounceBack, from rest, `cavity_steps` steps
    (FP32 on the GPU through Simulation.run). Config 3: Taylor-Green n3^3 periodic, FP32, FP16 and
    BF16 storage on the GPU (`gpu_steps` event-timed steps each); the CPU on a few steps."""
    import numpy as np
    import torch

    import oracle
    from virtualfluids_amd.grid import (d0M0, d0P0, d00M, d00P, dM00, dP00, periodic_box, taylor_green_fields,
                                        taylor_green_omega, walled_box)
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory, Simulation

    out = {}
    # config 1
    n, u_lid = 64, 0.05
    walls = {side: ("noslip",) for side in (dP00, dM00, d0P0, d0M0, d00M)}
    walls[d00P] = ("velocity", (u_lid, 0.0, 0.0))
    g = walled_box((n, n, n), walls)
    omega = 1.0 / (3.0 * (u_lid * n / 100.0) + 0.5)
    F = BoundaryConditionFactory
    fac = F()
    fac.set_velocity_boundary_condition(F.VelocityBC.VelocityBounceBack)
    fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipBounceBack)
    L = Lattice(g, omega)
    z = np.zeros(g.N)
    L.init_equilibrium(z, z, z, z)
    sim = Simulation(L, fac)
    sim.run(20)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(cavity_steps)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    del sim, L
    C = oracle.Lattice(g.nbx, g.nby, g.nbz)
    C.init_eq(z, z, z, z)
    vel, ns = g.bcs["velocity"], g.bcs["noslip"]
    steps, t0 = 0, time.perf_counter()
    while steps < 2 or (time.perf_counter() - t0 < cpu_seconds and steps < cavity_steps):
        C.collide(g.fluid, omega)
        C.bc_post(1, vel.k, vel.q27, omega, vel.velocity)  # VelocityBounceBack
        C.bc_post(0, ns.k, ns.q27, omega)                  # NoSlipBounceBack
        C.swap()
        steps += 1
    t_cpu = time.perf_counter() - t0
    del C
    out["config1_cavity64"] = {
        "case": f"lid-driven cavity 64^3 + wall layer ({g.N} nodes), lid 0.05, Re 100, bounce-back walls",
        "gpu_MLUPS": round(g.N * cavity_steps / t_gpu / 1e6, 1), "gpu_steps": cavity_steps, "gpu_dtype": "f32",
        "cpu_MLUPS": round(g.N * steps / t_cpu / 1e6, 2), "cpu_steps": steps, "cpu_dtype": "f64 (oracle)"}
    # config 3
    grid = periodic_box(n3, n3, n3)
    fields = taylor_green_fields(grid, 0.05)
    omega = taylor_green_omega(n3, 0.05, 1600.0)
    rec = {"case": f"Taylor-Green {n3}^3 periodic, u0 0.05, Re 1600"}
    for storage in ("fp32", "fp16", "bf16"):
        L = Lattice(grid, omega, storage=storage, limiters=(1e6, 1e6, 1e6))
        L.init_equilibrium(*fields)
        dt = time_steps(L, gpu_steps, 10)
        rec[f"gpu_{storage}_MLUPS"] = round(grid.N * gpu_steps / dt / 1e6, 1)
        del L
        torch.cuda.empty_cache()
    rec["gpu_steps"] = gpu_steps
    C = oracle.Lattice(grid.nbx, grid.nby, grid.nbz)
    C.init_eq(*fields)
    even, steps, t0 = True, 0, time.perf_counter()
    while steps < 1 or (time.perf_counter() - t0 < cpu_seconds and steps < 20):
        oracle.collide_box(C.f, n3, n3, n3, omega, even, limiters=(1e6, 1e6, 1e6))
        even = not even
        steps += 1
    t_cpu = time.perf_counter() - t0
    del C
    rec["cpu_MLUPS"] = round(grid.N * steps / t_cpu / 1e6, 2)
    rec["cpu_steps"] = steps
    rec["cpu_dtype"] = "f64 (oracle)"
    out["config3_tgv"] = rec
    return out


def copy_bandwidth():
    """Device-to-device copy of 2 GiB, (read + write bytes) / time, by two kernels: the
    library's aligned non-temporal 16-byte copy (vf_stream_copy) and the runtime's blit copy
    (torch copy_). Returns (stream copy GB/s, blit copy GB/s)."""
    import ctypes

    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import stream_handle

    n = 1 << 29
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)

    def stream_copy():
        capi.call("vf_stream_copy", ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                  ctypes.c_uint64(4 * n), stream_handle())

    def blit_copy():
        b.copy_(a)

    out = []
    for fn in (stream_copy, blit_copy):
        b.zero_()
        fn()
        if not torch.equal(a, b):
            raise RuntimeError(f"copy bandwidth probe: {fn.__name__} left the destination different")
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        torch.cuda.synchronize()
        out.append(2 * 4 * n / (s.elapsed_time(e) / 1000.0 / reps) / 1e9)
    del a, b
    torch.cuda.empty_cache()
    return tuple(out)


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def inplace27_bandwidth(n=256, reps=20, caps=(0, 2, 3, 4)):
    """The in-place 27-stream read-modify-write yardstick (vf_stream_inplace27) on a 27 * n^3 FP32
    buffer, the population buffer of the headline workload, at several caps of resident
    workgroups per CU (0: none): {cap: 216 B per node / time} and the best cap."""
    import ctypes

    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import stream_handle

    N = n ** 3
    buf = torch.zeros(27 * N, dtype=torch.float32, device="cuda")
    out = {}
    for cap in caps:
        def run():
            capi.call("vf_stream_inplace27", ctypes.c_void_p(buf.data_ptr()), ctypes.c_uint64(4 * N), cap,
                      stream_handle())

        for _ in range(3):
            run()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            run()
        e.record()
        torch.cuda.synchronize()
        out[cap] = round(216.0 * N / (s.elapsed_time(e) / 1000.0 / reps) / 1e9, 1)
    del buf
    torch.cuda.empty_cache()
    best = max(out, key=out.get)
    return {"GBps_by_workgroups_per_CU": {str(k): v for k, v in out.items()}, "best_workgroups_per_CU": best,
            "best_GBps": out[best]}


def row_pattern_bandwidth(L, reps=20):
    """The FP32 row kernel's own memory pattern without the collision, as a yardstick: the same
    launch on the same lattice with a fluid mask that marks every node solid, so every node takes
    the kernel's solid-node path and writes back the 27 values it read (the same aligned loads
    from the four rows a cell touches, the same lane and LDS exchange of the nine x-shifted
    directions, the same stores; plus one mask byte per node). 216 B per node / time, GB/s."""
    import ctypes

    import torch

    solid = torch.zeros(L.N, dtype=torch.uint8, device=L.device)
    saved = L.param.fluidMask
    L.param.fluidMask = ctypes.c_void_p(solid.data_ptr())
    try:
        t = time_steps(L, reps, 3)
    finally:
        L.param.fluidMask = saved
    return round(216.0 * L.N * reps / t / 1e9, 1)


def cpu_baseline(n=256, target_s=8.0):
    """The CPU oracle (the C++ restatement of the reference's K17 step, OpenMP over z-slabs) on the
    headline workload, Taylor-Green 256^3 periodic, in its FP64 build (the reference's CPU
    precision) and its FP32 build, each timed for about `target_s` seconds of steps after one
    untimed step."""
    import numpy as np

    import oracle
    from virtualfluids_amd.grid import periodic_box, taylor_green_fields, taylor_green_omega

    grid = periodic_box(n, n, n)
    omega = taylor_green_omega(n, 0.05, 1600.0)
    fields = taylor_green_fields(grid, 0.05)
    out = {}
    sample = None
    for dtype in (np.float64, np.float32):
        L = oracle.Lattice(grid.nbx, grid.nby, grid.nbz, dtype=dtype)
        L.init_eq(*fields)
        if dtype == np.float64:
            # node-major populations of the first 2^18 nodes, for the reference's own collision
            sample = np.ascontiguousarray(L.f.reshape(27, -1)[:, :1 << 18].T)
        even = True
        oracle.collide_box(L.f, n, n, n, omega, even, limiters=(1e6, 1e6, 1e6))
        even = not even
        steps, t0 = 0, time.perf_counter()
        while steps < 2 or (time.perf_counter() - t0 < target_s and steps < 400):
            oracle.collide_box(L.f, n, n, n, omega, even, limiters=(1e6, 1e6, 1e6))
            even = not even
            steps += 1
        dt = time.perf_counter() - t0
        out[np.dtype(dtype).name] = (grid.N * steps / dt / 1e6, steps, dt)
        del L
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = min(threads, os.cpu_count() or threads)
    f64, f32 = out["float64"], out["float32"]
    rec = {"value": round(f64[0], 3), "unit": "MLUPS", "cores": threads, "kind": "port",
           "sample": f"CPU oracle (C++ restatement of the reference's K17 step, OpenMP), Taylor-Green {n}^3 periodic: "
                     f"FP64 build {f64[1]} steps in {f64[2]:.1f} s; FP32 build {f32[1]} steps in {f32[2]:.1f} s",
           "fp32_build_MLUPS": round(f32[0], 3), "nproc": os.cpu_count(), "cpu_model": _cpu_model()}
    if oracle.ref("f64") is not None and sample is not None:
        # the reference's own collision (runK17CompressibleNavierStokes compiled from its src/lbm
        # headers into oracle/_ref, FP64, one thread) on the same workload's nodes, collision only
        calls, t0 = 0, time.perf_counter()
        while calls < 2 or time.perf_counter() - t0 < target_s / 2:
            oracle.ref_k17_nodes(sample, omega, limiters=(1e6, 1e6, 1e6))
            calls += 1
        dt = time.perf_counter() - t0
        rec["reference_collision"] = {
            "value": round(sample.shape[0] * calls / dt / 1e6, 3), "unit": "MLUPS", "cores": 1, "kind": "reference",
            "sample": f"the reference's K17 collision function (FP64 build of its own headers), {sample.shape[0]} "
                      f"nodes of the {n}^3 Taylor-Green field x {calls} calls in {dt:.1f} s, one thread, no streaming"}
    return rec


def traffic_from_profile():
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", "k17_row_fp32_256_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def velocity_field(L):
    """(3, N) FP32 velocity of the current state on the device (macroscopic output kernel)."""
    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import stream_handle

    capi.call("vf_CalculateMacroscopicQuantitiesCompressible", L.pref(), stream_handle())
    return torch.stack([L.vx, L.vy, L.vz]).clone()


def rel_l2(a, b) -> float:
    return float(((a.double() - b.double()) ** 2).sum().sqrt() / (b.double() ** 2).sum().sqrt())


def storage_precision(n=64, steps=200):
    """Velocity error of each storage type against the FP64 oracle on the same lattice: Taylor-Green
    n^3, Re 1600, u0 0.05, limiter off, `steps` steps; relative L2 of the velocity field."""
    import numpy as np
    import torch

    import oracle
    from virtualfluids_amd.grid import periodic_box, taylor_green_fields, taylor_green_omega
    from virtualfluids_amd.lattice import Lattice

    grid = periodic_box(n, n, n)
    omega = taylor_green_omega(n, 0.05, 1600.0)
    fields = taylor_green_fields(grid, 0.05)
    ref = oracle.Lattice(grid.nbx, grid.nby, grid.nbz)
    ref.init_eq(*fields)
    for _ in range(steps):
        oracle.collide_box(ref.f, n, n, n, omega, ref.even, limiters=(1e6, 1e6, 1e6))
        ref.swap()
    _, vx, vy, vz, _ = ref.macroscopic()
    u64 = torch.from_numpy(np.stack([vx, vy, vz]))
    out = {"case": f"Taylor-Green {n}^3, Re 1600, u0 0.05, {steps} steps, velocity relative L2 against the FP64 oracle"}
    for storage in ("fp32", "fp16", "bf16"):
        L = Lattice(grid, omega, storage=storage, limiters=(1e6, 1e6, 1e6))
        L.init_equilibrium(*fields)
        for _ in range(steps):
            L.collide()
            L.swap()
        out[storage] = rel_l2(velocity_field(L).cpu(), u64)
        del L
    return out


def interface_kernels(reps=200):
    """The refinement interface of the two-level case alone (258 x 130 x 130 coarse and 128^3
    fine FP16 lattices, the nested box of two_level_gradient): each interface kernel launched
    `reps` times back to back on its full cell list and on prefixes of it (1/2, 1/4, 1/8, 1/16),
    with the time per launch; and, counted on the host from the cell lists, the distinct 64-byte
    segments the fine-to-coarse kernel's 27 gathers per lane touch (each 64-lane wave holds 8
    cells; summed over the waves) and the distinct 128-byte lines of the launch."""
    import numpy as np
    import torch

    from virtualfluids_amd.grid import CX, CY, CZ, channel
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.refinement import (DeviceInterfaceCells, GridScalingFactory, GridScalingKernelManager,
                                              InterfaceCells, nested_box)

    nb = nested_box(None, (65, 33, 33), (129, 97, 97), coarse=channel((256, 128, 128), inflow=0.03))
    C = Lattice(nb.coarse, 1.8, storage="fp16", level=0)
    F = Lattice(nb.fine, 1.7, storage="fp16", level=1)
    for L in (C, F):
        z = np.zeros(L.N)
        L.init_equilibrium(z, np.full(L.N, 0.03), z, z)
    fac = GridScalingFactory()
    fac.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
    man = GridScalingKernelManager(2, fac, nb.fc.n, nb.cf.n)
    out = {"case": "refinement interface of the two-level channel, FP16, cells and us per launch",
           "fine_to_coarse": {}, "coarse_to_fine": {}}
    for frac in (1.0, 0.5, 0.25, 0.125, 0.0625):
        for key, cells, run in (("fine_to_coarse", nb.fc, man.run_fine_to_coarse),
                                ("coarse_to_fine", nb.cf, man.run_coarse_to_fine)):
            n = max(1, int(cells.n * frac))
            dc = DeviceInterfaceCells(InterfaceCells(cells.coarse[:n].copy(), cells.fine[:n].copy(), None), "cuda")
            for _ in range(5):
                run(C, F, dc)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                run(C, F, dc)
            e1.record()
            torch.cuda.synchronize()
            out[key][str(n)] = round(e0.elapsed_time(e1) / reps * 1e3, 2)
    # the fine-to-coarse kernel's gathers: lane 8 c + q reads the 27 populations of corner q of
    # cell c's fine source cell, direction d from the cell node moved by +1 where c_d < 0
    g = nb.fine

    def moved(k, bits):
        if bits & 1:
            k = g.nbx[k]
        if bits & 2:
            k = g.nby[k]
        if bits & 4:
            k = g.nbz[k]
        return k.astype(np.int64)

    base = nb.fc.fine.astype(np.int64)
    corners = np.stack([moved(base, q) for q in range(8)], 1).reshape(-1)
    waves = (corners.size + 63) // 64
    pad = waves * 64 - corners.size
    segments, lines = 0, 0
    for d in range(27):
        sel = (1 if CX[d] < 0 else 0) | (2 if CY[d] < 0 else 0) | (4 if CZ[d] < 0 else 0)
        addr = (d * g.N + moved(corners, sel)) * 2
        lines += np.unique(addr // 128).size
        a = np.concatenate([addr, np.repeat(addr[-1:], pad)]).reshape(waves, 64) // 64
        a.sort(axis=1)
        segments += int(waves + (np.diff(a, axis=1) != 0).sum())
    out["fine_to_coarse_gathers"] = {"cells": int(nb.fc.n), "waves": int(waves), "segments_64B": segments,
                                     "lines_128B": int(lines), "line_MB": round(lines * 128 / 1e6, 1),
                                     "population_MB": round(nb.fc.n * 8 * 27 * 2 / 1e6, 2)}
    return out


def run_case(case, steps=None):
    """One full-size case alone (profiling runs): its record, with `steps` steps when given (the
    sphere channel averages its drag over the last tenth, the two-level channel its outlet energy
    over the last quarter; checkpoint intervals as in the full runs)."""
    if case == "sphere":
        return sphere_channel_gradient() if steps is None else \
            sphere_channel_gradient(steps=steps, average_last=max(1, steps // 10))
    if case == "two-level":
        out = two_level_gradient() if steps is None else \
            two_level_gradient(steps=steps, average_last=max(1, steps // 4))
        out.pop("_grad")
        return out
    if case == "interface":
        return interface_kernels()
    if case == "configs":
        return configs_gpu_cpu()
    return published_cavity() if steps is None else published_cavity(steps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steady-steps", type=int, default=500,
                    help="a longer steady window per storage type, reported beside the main one")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing")
    ap.add_argument("--adjoint-steps", type=int, default=200, help="timed adjoint kernel steps")
    ap.add_argument("--no-sphere", action="store_true", help="skip the full-size sphere-channel gradient")
    ap.add_argument("--no-two-level", action="store_true", help="skip the full-size two-level gradient")
    ap.add_argument("--no-cavity", action="store_true", help="skip the reference's published cavity benchmark")
    ap.add_argument("--only", choices=["fp32", "fp16", "bf16"], default=None,
                    help="time one storage type only (for profiling runs)")
    ap.add_argument("--only-case", choices=["sphere", "two-level", "cavity", "interface", "configs"], default=None,
                    help="run one full-size case alone and print its record (for profiling runs)")
    ap.add_argument("--case-steps", type=int, default=None,
                    help="steps of --only-case (default: the case's full length)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (one GPU owns the domain); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if args.only_case is not None:
        print(json.dumps(run_case(args.only_case, args.case_steps)), flush=True)
        return

    import torch

    from virtualfluids_amd.grid import periodic_box, taylor_green_fields, taylor_green_omega
    from virtualfluids_amd.lattice import Lattice

    n = args.n
    grid = periodic_box(n, n, n)
    N = grid.N
    u0 = 0.05
    omega = taylor_green_omega(n, u0, 1600.0)
    fields = taylor_green_fields(grid, u0)
    lim = (1e6, 1e6, 1e6)  # the Taylor-Green app switches the limiter off

    results = {}
    velocity = {}
    adjoint = None
    copy_gbps = blit_gbps = rmw = None
    for storage in (["fp32", "fp16", "bf16"] if args.only is None else [args.only]):
        L = Lattice(grid, omega, storage=storage, limiters=lim)
        L.init_equilibrium(*fields)
        if copy_gbps is None:
            # the copy and in-place yardsticks right before the first timed steps, with the
            # lattice already set up: the GPU then enters the timed window busy, not from idle
            copy_gbps, blit_gbps = copy_bandwidth()
            rmw = inplace27_bandwidth()
        # the longer steady window first: after idle the GPU's power management takes some 30
        # launches of this kernel to settle (per-launch times rise by up to 30 % and come back),
        # so the main window below starts in the settled state
        steady = None
        if args.steady_steps > 0:
            ts = time_steps(L, args.steady_steps, 0)
            steady = {"steps": args.steady_steps, "MLUPS": round(N * args.steady_steps / ts / 1e6, 1),
                      "ms_per_step": round(1e3 * ts / args.steady_steps, 4)}
        # the same window for every storage type: --warmup untimed steps, then --steps timed
        t = time_steps(L, args.steps, args.warmup)
        velocity[storage] = velocity_field(L)
        bytes_per_node = 2 * 27 * L.f.element_size()
        r = {"t": t, "steps": args.steps, "mlups": N * args.steps / t / 1e6, "ms": 1e3 * t / args.steps,
             "gbps": bytes_per_node * N * args.steps / t / 1e9, "bytes_per_node": bytes_per_node}
        if steady is not None:
            r["steady"] = steady
        r["mass_drift"] = float(L.moment_sums()[0])
        if storage == "fp32":
            # after the timed window: the same kernel without the collision on the same state
            r["row_pattern_GBps"] = row_pattern_bandwidth(L)
        results[storage] = r
        if storage == "fp32" and not args.no_adjoint:
            ta = time_adjoint(L, args.adjoint_steps, 5)
            tg = time_gradient(L, 40, 10)
            t_fwd = t / args.steps
            adjoint = {"kernel_ms_per_step": round(1e3 * ta / args.adjoint_steps, 4),
                       "kernel_vs_forward": round((ta / args.adjoint_steps) / t_fwd, 3),
                       "kernel_GBps": round(3 * 27 * 4 * N * args.adjoint_steps / ta / 1e9, 1),
                       "gradient_ms_per_step": round(1e3 * tg / 40, 4),
                       "gradient_vs_forward": round((tg / 40) / t_fwd, 3),
                       "gradient_run": "k17_steps 40 steps, checkpoint every 10: forward, replay, adjoint"}
        del L
        torch.cuda.empty_cache()

    main_key = "fp32" if "fp32" in results else args.only
    r = results[main_key]
    traffic = traffic_from_profile() if main_key == "fp32" else None
    variants = {}
    for k, v in results.items():
        variants[k] = {"MLUPS": round(v["mlups"], 1), "GBps": round(v["gbps"], 1), "ms_per_step": round(v["ms"], 4),
                       "steps": v["steps"], "warmup": args.warmup,
                       "vs_fp32": round(v["mlups"] / results["fp32"]["mlups"], 3) if "fp32" in results else None}
        if "steady" in v:
            variants[k]["steady"] = v["steady"]
            if "fp32" in results and "steady" in results["fp32"]:
                variants[k]["steady"]["vs_fp32"] = round(v["steady"]["MLUPS"] / results["fp32"]["steady"]["MLUPS"], 3)
        if "fp32" in velocity and k != "fp32":
            # the storage type's velocity error after the same steps from the same initial field
            variants[k]["u_rel_l2_vs_fp32_storage"] = rel_l2(velocity[k], velocity["fp32"])
    del velocity
    precision = storage_precision() if args.only is None else None
    two_level = None
    if not (args.no_two_level or args.no_adjoint):
        two_level = two_level_gradient()
        ref32 = two_level_gradient(storage="fp32", keep_fraction=0.0)
        g16, g32 = two_level.pop("_grad"), ref32.pop("_grad")
        import numpy as np

        two_level["fp32_storage"] = {k: ref32[k] for k in ("objective", "grad_l2", "forward_s", "gradient_s",
                                                           "gradient_vs_forward", "reverse_vs_forward",
                                                           "forward_MLUPS")}
        two_level["fp16_vs_fp32_gradient_rel_l2"] = float(np.linalg.norm(g16 - g32) / np.linalg.norm(g32))
        two_level["fp16_vs_fp32_objective_rel"] = abs(two_level["objective"] / ref32["objective"] - 1.0)
    line = {
        "metric": METRIC,
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Taylor-Green vortex D3Q27 K17 {n}^3 periodic, {main_key} storage, "
                               f"plain forward (dense-box kernel)", "nodes": N, "u0": u0, "Re": 1600,
                   "omega": round(omega, 6), "storage": main_key},
        "roofline": {"bound": "hbm", "achieved": round(r["gbps"], 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(r["gbps"] / HBM_PEAK_GBPS, 4), "traffic": traffic},
        "storage_variants": variants,
        "storage_precision_vs_fp64": precision,
        "adjoint": adjoint,
        "sphere_channel_gradient": None if (args.no_sphere or args.no_adjoint) else sphere_channel_gradient(),
        "two_level_gradient": two_level,
        "published_cavity": None if args.no_cavity else published_cavity(),
        "configs_gpu_cpu": None if (args.no_cpu or args.only is not None) else configs_gpu_cpu(),
        "hbm_copy_GBps": round(copy_gbps, 1),
        "blit_copy_GBps": round(blit_gbps, 1),
        "inplace27": rmw,
        "vs_inplace27": round(r["gbps"] / rmw["best_GBps"], 4) if main_key == "fp32" else None,
        # the kernel against its own access pattern without the collision (row_pattern_bandwidth)
        "row_pattern_GBps": r.get("row_pattern_GBps"),
        "vs_row_pattern": round(r["gbps"] / r["row_pattern_GBps"], 4) if "row_pattern_GBps" in r else None,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
