"""GPU adjoint of the K17 collide+stream step and the differentiable time step.

* vf_k17_adjoint_box / vf_k17_adjoint (hand-written FP32 reverse mode) against the oracle's
  FP64 reference adjoint (forward-mode dual numbers through its own collision) on the same
  stored state: every storage type, both parities, masked and periodic lattices, and the
  velocity-shift adjoint. Tolerance: 3e-5 of the largest adjoint entry (FP32 reverse mode; the
  limiter's derivative amplifies rounding).
* k17_steps + an FP64 kinetic-energy objective: the gradient matches FP64 central finite
  differences of the oracle's forward run (relative error < 1e-4, the stated bound) and the
  oracle's FP64 multi-step adjoint, and does not depend on the checkpoint interval.
"""
import ctypes

import numpy as np
import pytest

from helpers import near_equilibrium, pack_pre
from virtualfluids_amd.grid import periodic_box, walled_box

pytestmark = pytest.mark.gpu


def _grid(kind):
    if kind == "periodic":
        return periodic_box(24, 10, 8)
    if kind == "walled":
        return walled_box((22, 8, 6), {})
    # long_walled: rows of four waves, the last one partial (the row kernel's LDS hand-over);
    # long_periodic and split_walled: 130 = 128 + 2 and 258 = 256 + 2, whole waves in the row
    # kernel and two tail columns in the scalar kernel
    if kind == "long_walled":
        return walled_box((228, 4, 3), {})
    if kind == "split_walled":
        return walled_box((256, 3, 2), {})
    return periodic_box(130, 4, 3)


# box: the automatic choice (the one-node-per-thread kernel for these widths); box_row: the row
# kernel (buffer addressing) over the whole waves of each row plus the scalar kernel on split-off
# columns; box_global: the same with 64-bit addresses; box_scalar: the one-node-per-thread kernel;
# indexed: fluid-node list and neighbour arrays (pinned through vf_lbm_parameter.kernelChoice)
ADJ_CHOICE = {"box": {}, "box_global": {"rowAddressing": 2, "adjointKernel": 3}, "box_scalar": {"adjointKernel": 2},
              "box_row": {"adjointKernel": 3}, "indexed": {}}


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("path", ["box", "box_global", "box_scalar", "box_row", "indexed"])
@pytest.mark.parametrize("even", [True, False])
@pytest.mark.parametrize("kind", ["periodic", "walled", "long_periodic", "long_walled", "split_walled"])
def test_adjoint_kernel_matches_reference_adjoint(gpu, orc, storage, path, even, kind):
    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import Lattice

    if kind.startswith(("long", "split")) and path in ("box_scalar", "indexed") and storage != "fp32":
        pytest.skip("long rows exercise the row kernel; the other paths are covered on the short boxes")
    choice = ADJ_CHOICE[path]
    path = "indexed" if path == "indexed" else "box"
    grid = _grid(kind)
    rng = np.random.default_rng(500 + 2 * even + (path == "box") + 4 * len(storage))
    omega = 1.8
    L = Lattice(grid, omega, storage=storage, box=(path == "box"))
    for k, v in choice.items():
        setattr(L.param.kernelChoice, k, v)
    L.even = even
    L.set_populations(pack_pre(near_equilibrium(rng, grid.N, u_amp=0.05, noise=1e-2), grid, even))
    lam0 = rng.normal(0, 1, 27 * grid.N)
    lam = torch.from_numpy(lam0.astype(np.float32)).cuda()
    shift = torch.zeros(3 * grid.N, dtype=torch.float32, device="cuda")
    fn = "vf_k17_adjoint_box" if path == "box" else "vf_k17_adjoint"
    print(f"{fn} {storage} even={even} {kind}", flush=True)
    capi.call(fn, L.pref(), ctypes.c_void_p(lam.data_ptr()), ctypes.c_void_p(shift.data_ptr()), None, None)
    got = lam.cpu().numpy().astype(np.float64)
    got_shift = shift.cpu().numpy().reshape(3, grid.N)

    A = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    A.f[:] = L.populations()  # the stored (rounded) state
    A.even = even
    ref = lam0.astype(np.float32).astype(np.float64)
    ref_shift = np.zeros((3, grid.N))
    orc.adjoint_step(A, ref, grid.fluid, omega, shift_adjoint=ref_shift)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    err_s = np.abs(got_shift - ref_shift).max() / np.abs(ref_shift).max()
    print(f"  rel err adjoint {err:.2e} shift {err_s:.2e}")
    assert err < 3e-5 and err_s < 3e-4
    if "walled" in kind:  # cells only solid nodes own keep their adjoint
        assert np.any(got == lam0.astype(np.float32))


def _objective_setup(orc, steps):
    import torch

    from virtualfluids_amd.adjoint import pre_collision_index
    from virtualfluids_amd.lattice import Lattice

    grid = periodic_box(16, 12, 10)
    rng = np.random.default_rng(600)
    f0 = pack_pre(near_equilibrium(rng, grid.N, u_amp=0.05, noise=2e-3), grid, True)
    omega = 1.9
    L = Lattice(grid, omega)
    even_final = steps % 2 == 0
    idx = torch.from_numpy(pre_collision_index(grid, even_final)).cuda()
    return grid, f0, omega, L, idx, even_final


def _gpu_gradient(L, f0, idx, steps, K):
    import torch

    from virtualfluids_amd.adjoint import k17_steps, kinetic_energy

    x = torch.from_numpy(f0.astype(np.float32)).cuda().requires_grad_(True)
    J = kinetic_energy(k17_steps(x, L, steps, checkpoint_every=K, even=True), idx)
    J.backward()
    return float(J.detach()), x.grad.cpu().numpy().astype(np.float64)


def _oracle_J(orc, grid, buf, omega, steps, even_final):
    from virtualfluids_amd.adjoint import pre_collision_index

    L = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    L.f[:] = buf
    states = []
    for _ in range(steps):
        states.append((L.f.copy(), L.even))
        L.collide(grid.fluid, omega)
        L.swap()
    f = L.f[pre_collision_index(grid, even_final)]
    drho = f.sum(0)
    from virtualfluids_amd.grid import CX, CY, CZ

    j = np.stack([CX, CY, CZ]).astype(np.float64) @ f
    return 0.5 * ((j * j).sum(0) / (1 + drho)).sum(), states, f, drho, j


def test_gradient_matches_fp64_finite_differences(gpu, orc):
    steps = 24
    grid, f0, omega, L, idx, even_final = _objective_setup(orc, steps)
    J, grad = _gpu_gradient(L, f0, idx, steps, K=5)
    rng = np.random.default_rng(601)
    v = rng.normal(0, 1, f0.shape)
    h = 1e-7  # resolves the limiter's curvature near x = 0; FP64 round-off stays ~1e-8 relative
    Jp = _oracle_J(orc, grid, f0 + h * v, omega, steps, even_final)[0]
    Jm = _oracle_J(orc, grid, f0 - h * v, omega, steps, even_final)[0]
    fd = (Jp - Jm) / (2 * h)
    ad = float(grad @ v)
    print(f"J={J:.6e} <grad, v>: adjoint {ad:.10e} FD {fd:.10e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(ad - fd) < 1e-4 * abs(fd)


def test_gradient_matches_fp64_reference_adjoint(gpu, orc):
    from virtualfluids_amd.grid import CX, CY, CZ

    steps = 12
    grid, f0, omega, L, idx, even_final = _objective_setup(orc, steps)
    _, grad = _gpu_gradient(L, f0, idx, steps, K=4)
    J, states, f, drho, j = _oracle_J(orc, grid, f0, omega, steps, even_final)
    # dJ/df_d of 1/2 |j|^2 / (1 + drho) at every node, scattered to the cells
    u = j / (1 + drho)
    dJ = np.stack([CX, CY, CZ]).T.astype(np.float64) @ u - 0.5 * (u * u).sum(0)
    from virtualfluids_amd.adjoint import pre_collision_index

    lam = np.zeros(27 * grid.N)
    lam[pre_collision_index(grid, even_final)] = dJ
    A = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    for state, even in reversed(states):
        A.f[:] = state
        A.even = even
        orc.adjoint_step(A, lam, grid.fluid, omega)
    err = np.abs(grad - lam).max() / np.abs(lam).max()
    print(f"gradient vs FP64 reference adjoint: rel err {err:.2e}")
    assert err < 1e-4


def test_checkpoint_interval_does_not_change_gradient(gpu, orc):
    steps = 9
    grid, f0, omega, L, idx, _ = _objective_setup(orc, steps)
    grads = [_gpu_gradient(L, f0, idx, steps, K)[1] for K in (1, 4, 9, 20)]
    for g in grads[1:]:
        assert np.array_equal(g, grads[0])


# ---- boundary conditions -------------------------------------------------------------------------
POST_KINDS = {  # C entry point -> oracle kind
    "vf_NoSlipBounceBack": 0, "vf_VelocityBounceBack": 1, "vf_NoSlipInterpolatedCompressible": 2,
    "vf_VelocityInterpolatedCompressible": 3, "vf_VelocityWithPressureInterpolatedCompressible": 4,
}


def _duct(q_random, seed=47):
    from virtualfluids_amd.grid import channel

    g = channel((14, 7, 6), inflow=0.03)
    if q_random:
        rng = np.random.default_rng(seed)
        for bc in g.bcs.values():
            bc.q27 = np.where(bc.q27 >= 0, rng.uniform(0.15, 0.85, bc.q27.shape), bc.q27)
    return g


@pytest.mark.parametrize("fn", sorted(POST_KINDS))
@pytest.mark.parametrize("even", [True, False])
@pytest.mark.parametrize("box", [True, False])  # (x, y, z) addressing or neighbour arrays
@pytest.mark.parametrize("saved", [False, True])  # recomputed or stored post-collision values
@pytest.mark.parametrize("rows", [0, 1])  # layout hint: a lane per direction, or a thread per node
def test_post_bc_adjoint_kernel_matches_reference(gpu, orc, fn, even, box, saved, rows):
    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.adjoint import _bc_struct
    from virtualfluids_amd.lattice import Lattice

    kind = POST_KINDS[fn]
    g = _duct(q_random=kind >= 2)
    bc_name = "velocity" if kind in (1, 3, 4) else "noslip"
    rng = np.random.default_rng(700 + kind + 10 * even)
    omega = 1.8
    L = Lattice(g, omega, box=box)
    L.even = even
    L.set_populations(pack_pre(near_equilibrium(rng, g.N, u_amp=0.04, noise=5e-3), g, even))
    lam0 = rng.normal(0, 1, 27 * g.N).astype(np.float32)
    lam = torch.from_numpy(lam0).cuda()
    bc = L.bcs[bc_name]
    bc.struct.nodesAlongRows = rows
    ubar = torch.zeros(3 * bc.n, dtype=torch.float32, device="cuda")
    post = None
    if saved:
        # the forward BC kernel stores the post-collision values it reads, on a copy of the state
        before = L.f.clone()
        L.collide()
        post = torch.full((27 * bc.n,), float("nan"), dtype=torch.float32, device="cuda")
        capi.call(fn, L.pref(), ctypes.byref(_bc_struct(bc, post)), None)
        L.f.copy_(before)
    print(f"{fn}Adjoint even={even} nodes={bc.n} saved={saved} rows={rows}", flush=True)
    capi.call(fn + "Adjoint", L.pref(), ctypes.byref(_bc_struct(bc, post)), ctypes.c_void_p(lam.data_ptr()),
              ctypes.c_void_p(ubar.data_ptr()), None)
    got = lam.cpu().numpy().astype(np.float64)

    P = orc.Lattice(g.nbx, g.nby, g.nbz)
    P.f[:] = L.populations()
    P.even = even
    P.collide(g.fluid, omega)  # the post-collision state the forward BC kernel reads
    ref = lam0.astype(np.float64)
    ref_u = np.zeros((bc.n, 3))
    nodes = g.bcs[bc_name]
    orc.bc_post_adjoint(kind, P, ref, nodes.k, nodes.q27, omega, velocity=nodes.velocity,
                        ubar=ref_u if kind in (1, 3, 4) else None)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"  rel err {err:.2e}")
    assert err < 2e-5
    if kind in (1, 3, 4):
        gu = ubar.cpu().numpy().reshape(3, bc.n).T
        assert np.abs(gu - ref_u).max() < 1e-4 * np.abs(ref_u).max()


@pytest.mark.parametrize("even", [True, False])
@pytest.mark.parametrize("box", [True, False])
@pytest.mark.parametrize("distinct", [1, 0])  # plain read-modify-writes, or atomic adds
@pytest.mark.parametrize("rows", [0, 1])  # layout hint: a lane per direction, or a thread per node
def test_pressure_bc_adjoint_kernel_matches_reference(gpu, orc, even, box, distinct, rows):
    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import Lattice

    g = _duct(False)
    rng = np.random.default_rng(720 + even)
    L = Lattice(g, 1.8, box=box)
    L.even = even
    L.set_populations(pack_pre(near_equilibrium(rng, g.N, u_amp=0.04, noise=5e-3), g, even))
    lam0 = rng.normal(0, 1, 27 * g.N).astype(np.float32)
    lam = torch.from_numpy(lam0).cuda()
    bc = L.bcs["pressure"]
    assert bc.kN_distinct
    bc.struct.nodesAlongRows = rows
    rbar = torch.zeros(bc.n, dtype=torch.float32, device="cuda")
    capi.call("vf_PressureNonEquilibriumCompressibleAdjoint", L.pref(), ctypes.byref(bc.struct),
              ctypes.c_void_p(lam.data_ptr()), ctypes.c_void_p(rbar.data_ptr()), distinct, None)
    A = orc.Lattice(g.nbx, g.nby, g.nbz)
    A.f[:] = L.populations()
    A.even = even
    ref = lam0.astype(np.float64)
    ref_r = np.zeros(bc.n)
    pr = g.bcs["pressure"]
    orc.bc_pressure_adjoint(A, ref, pr.k, pr.kN, pr.rho_bc, pr.direction, rhobar=ref_r)
    err = np.abs(lam.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"pressure adjoint even={even} rows={rows}: rel err {err:.2e}")
    assert err < 2e-5
    assert np.abs(rbar.cpu().numpy() - ref_r).max() < 1e-4 * np.abs(ref_r).max()


@pytest.mark.parametrize("velocity_bc,noslip_bc", [
    ("VelocityBounceBack", "NoSlipBounceBack"),
    ("VelocityInterpolatedCompressible", "NoSlipInterpolatedCompressible"),
])
def test_duct_gradients_match_fp64_finite_differences(gpu, orc, velocity_bc, noslip_bc):
    """20 steps of a duct with velocity inflow, pressure outlet and walls: the gradient of the
    FP64 kinetic energy with respect to the initial populations, the inflow wall velocities and
    the outlet densities, against FP64 central differences of the oracle's forward run."""
    import torch

    from helpers import ORACLE_POST_BC
    from virtualfluids_amd.adjoint import k17_steps, kinetic_energy, pre_collision_index
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory

    g = _duct(q_random=velocity_bc.startswith("VelocityInterp"))
    vel, ns, pr = g.bcs["velocity"], g.bcs["noslip"], g.bcs["pressure"]
    omega, steps = 1.8, 20
    rng = np.random.default_rng(730)
    f0 = pack_pre(near_equilibrium(rng, g.N, u_amp=0.03, noise=2e-3), g, True)
    U0 = np.array(vel.velocity, dtype=np.float32).astype(np.float64)
    R0 = np.full(pr.n, 0.002, dtype=np.float32).astype(np.float64)
    fluid = g.geo == 19

    fac = BoundaryConditionFactory()
    fac.set_velocity_boundary_condition(BoundaryConditionFactory.VelocityBC[velocity_bc])
    fac.set_no_slip_boundary_condition(BoundaryConditionFactory.NoSlipBC[noslip_bc])
    fac.set_pressure_boundary_condition(BoundaryConditionFactory.PressureBC.PressureNonEquilibriumCompressible)
    L = Lattice(g, omega)
    idx = torch.from_numpy(pre_collision_index(g, steps % 2 == 0)).cuda()
    nodes = torch.from_numpy(np.nonzero(fluid)[0]).cuda()
    x = torch.from_numpy(f0.astype(np.float32)).cuda().requires_grad_(True)
    U = torch.from_numpy(U0.astype(np.float32)).cuda().requires_grad_(True)
    R = torch.from_numpy(R0.astype(np.float32)).cuda().requires_grad_(True)
    J = kinetic_energy(k17_steps(x, L, steps, checkpoint_every=6, even=True, bc_factory=fac, velocity=U, rho_bc=R),
                       idx, nodes)
    J.backward()

    def oracle_J(buf, Uv, Rv):
        A = orc.Lattice(g.nbx, g.nby, g.nbz)
        A.f[:] = buf
        for _ in range(steps):
            A.collide(g.fluid, omega)
            A.bc_post(ORACLE_POST_BC[velocity_bc], vel.k, vel.q27, omega, tuple(Uv))
            A.bc_post(ORACLE_POST_BC[noslip_bc], ns.k, ns.q27, omega)
            A.swap()
            A.bc_pressure(pr.k, pr.kN, Rv, pr.direction)
        f = A.f[pre_collision_index(g, A.even)]
        from virtualfluids_amd.grid import CX, CY, CZ

        j = np.stack([CX, CY, CZ]).astype(np.float64) @ f
        ke = 0.5 * (j * j).sum(0) / (1 + f.sum(0))
        return ke[fluid].sum()

    h = 1e-7
    for name, grad, base in (("f0", x.grad, f0), ("U", U.grad, U0), ("rhoBC", R.grad, R0)):
        v = rng.normal(0, 1, base.shape)
        if name == "f0":
            Jp, Jm = oracle_J(f0 + h * v, U0, R0), oracle_J(f0 - h * v, U0, R0)
        elif name == "U":
            Jp, Jm = oracle_J(f0, U0 + h * v, R0), oracle_J(f0, U0 - h * v, R0)
        else:
            Jp, Jm = oracle_J(f0, U0, R0 + h * v), oracle_J(f0, U0, R0 - h * v)
        fd = (Jp - Jm) / (2 * h)
        ad = float((grad.cpu().numpy().astype(np.float64) * v).sum())
        print(f"{velocity_bc}: d J / d {name}: adjoint {ad:.8e} FD {fd:.8e} rel {abs(ad - fd) / abs(fd):.2e}")
        assert abs(ad - fd) < 1e-4 * abs(fd)


def _duct_oracle_J(orc, g, buf, omega, segments, fluid, velocity_bc="VelocityBounceBack"):
    """FP64 kinetic energy after consecutive segments [(steps, U (3, n))] of the duct run
    (velocity BC, bounce-back walls, pressure outlet at drho = 0)."""
    from helpers import ORACLE_POST_BC
    from virtualfluids_amd.adjoint import pre_collision_index
    from virtualfluids_amd.grid import CX, CY, CZ

    vel, ns, pr = g.bcs["velocity"], g.bcs["noslip"], g.bcs["pressure"]
    A = orc.Lattice(g.nbx, g.nby, g.nbz)
    A.f[:] = buf
    for n, U in segments:
        for _ in range(n):
            A.collide(g.fluid, omega)
            A.bc_post(ORACLE_POST_BC[velocity_bc], vel.k, vel.q27, omega, tuple(U))
            A.bc_post(ORACLE_POST_BC["NoSlipBounceBack"], ns.k, ns.q27, omega)
            A.swap()
            A.bc_pressure(pr.k, pr.kN, np.asarray(pr.rho_bc, dtype=np.float64), pr.direction)
    f = A.f[pre_collision_index(g, A.even)]
    j = np.stack([CX, CY, CZ]).astype(np.float64) @ f
    return (0.5 * (j * j).sum(0) / (1 + f.sum(0)))[fluid].sum()


def test_chained_segments_with_different_velocities(gpu, orc):
    """Two k17_steps segments with different inflow velocities, chained and differentiated
    together: each segment's backward must replay with its own wall velocities (the lattice holds
    the second segment's when both run backward). Gradients with respect to the initial state and
    to both velocity sets against FP64 central differences of the oracle's run."""
    import torch

    from virtualfluids_amd.adjoint import k17_steps, kinetic_energy, pre_collision_index
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory

    g = _duct(False)
    vel = g.bcs["velocity"]
    omega, n1, n2 = 1.8, 9, 8
    rng = np.random.default_rng(740)
    f0 = pack_pre(near_equilibrium(rng, g.N, u_amp=0.03, noise=2e-3), g, True)
    base = np.array(vel.velocity, dtype=np.float64)
    U1 = (base * 1.5).astype(np.float32).astype(np.float64)
    U2 = (base * 0.4 + 0.01 * rng.normal(0, 1, base.shape)).astype(np.float32).astype(np.float64)
    fluid = g.geo == 19
    F = BoundaryConditionFactory
    fac = F()
    fac.set_velocity_boundary_condition(F.VelocityBC.VelocityBounceBack)
    fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipBounceBack)
    fac.set_pressure_boundary_condition(F.PressureBC.PressureNonEquilibriumCompressible)
    L = Lattice(g, omega)
    x = torch.from_numpy(f0.astype(np.float32)).cuda().requires_grad_(True)
    u1 = torch.from_numpy(U1.astype(np.float32)).cuda().requires_grad_(True)
    u2 = torch.from_numpy(U2.astype(np.float32)).cuda().requires_grad_(True)
    mid = k17_steps(x, L, n1, checkpoint_every=4, even=True, bc_factory=fac, velocity=u1)
    out = k17_steps(mid, L, n2, checkpoint_every=3, even=(n1 % 2 == 0), bc_factory=fac, velocity=u2)
    idx = torch.from_numpy(pre_collision_index(g, (n1 + n2) % 2 == 0)).cuda()
    nodes = torch.from_numpy(np.nonzero(fluid)[0]).cuda()
    J = kinetic_energy(out, idx, nodes)
    J.backward()
    # the lattice's velocity arrays hold the second segment's values again after the backward
    assert np.array_equal(torch.stack(L.bcs["velocity"].vel).cpu().numpy(), U2.astype(np.float32))

    h = 1e-7
    for name, grad in (("f0", x.grad), ("U1", u1.grad), ("U2", u2.grad)):
        v = rng.normal(0, 1, f0.shape if name == "f0" else base.shape)
        if name == "f0":
            args_p = (f0 + h * v, [(n1, U1), (n2, U2)])
            args_m = (f0 - h * v, [(n1, U1), (n2, U2)])
        elif name == "U1":
            args_p = (f0, [(n1, U1 + h * v), (n2, U2)])
            args_m = (f0, [(n1, U1 - h * v), (n2, U2)])
        else:
            args_p = (f0, [(n1, U1), (n2, U2 + h * v)])
            args_m = (f0, [(n1, U1), (n2, U2 - h * v)])
        fd = (_duct_oracle_J(orc, g, args_p[0], omega, args_p[1], fluid)
              - _duct_oracle_J(orc, g, args_m[0], omega, args_m[1], fluid)) / (2 * h)
        ad = float((grad.cpu().numpy().astype(np.float64) * v).sum())
        print(f"chained segments: d J / d {name}: adjoint {ad:.8e} FD {fd:.8e} rel {abs(ad - fd) / abs(fd):.2e}")
        assert abs(ad - fd) < 1e-4 * abs(fd)


@pytest.mark.parametrize("storage,tol", [("fp16", 2e-2), ("bf16", 1.5e-1)])
def test_sixteen_bit_lattice_takes_and_returns_fp32(gpu, orc, storage, tol):
    """A 16-bit lattice stepped from an FP32 f0: the final buffer and the gradient with respect to
    f0 are FP32 (no 16-bit rounding of the gradient). The gradient is the adjoint of the stepping
    at the 16-bit states, so against FP64 central differences of the oracle (no storage rounding)
    it agrees to the storage precision: tolerance 2e-2 (FP16), 0.15 (BF16) relative."""
    import torch

    from virtualfluids_amd.adjoint import k17_steps, kinetic_energy, pre_collision_index
    from virtualfluids_amd.lattice import Lattice

    steps = 12
    grid = periodic_box(16, 12, 10)
    rng = np.random.default_rng(750)
    f0 = pack_pre(near_equilibrium(rng, grid.N, u_amp=0.05, noise=2e-3), grid, True)
    omega = 1.9
    L = Lattice(grid, omega, storage=storage)
    idx = torch.from_numpy(pre_collision_index(grid, steps % 2 == 0)).cuda()
    x = torch.from_numpy(f0.astype(np.float32)).cuda().requires_grad_(True)
    out = k17_steps(x, L, steps, checkpoint_every=5, even=True)
    assert out.dtype == torch.float32
    J = kinetic_energy(out, idx)
    J.backward()
    assert x.grad.dtype == torch.float32
    grad = x.grad.cpu().numpy().astype(np.float64)
    v = rng.normal(0, 1, f0.shape)
    h = 1e-7
    Jp = _oracle_J(orc, grid, f0 + h * v, omega, steps, steps % 2 == 0)[0]
    Jm = _oracle_J(orc, grid, f0 - h * v, omega, steps, steps % 2 == 0)[0]
    fd = (Jp - Jm) / (2 * h)
    ad = float(grad @ v)
    rel = abs(ad - fd) / abs(fd)
    print(f"{storage} lattice, FP32 f0: <grad, v> adjoint {ad:.8e} FD {fd:.8e} rel {rel:.2e}")
    assert rel < tol


@pytest.mark.parametrize("keep_segments", [1, 2, 10])  # trailing segments kept (10: all of them)
def test_kept_segments_give_the_same_gradient(gpu, orc, keep_segments):
    """keep_bytes: the trailing segments that fit are stepped out of place in the forward and walked
    by the backward without a replay. The duct with velocity inflow, pressure outlet, walls, a
    porosity field and an observed kinetic energy: outputs, observed sum and every gradient equal
    to those of the pure-recompute run bit for bit."""
    import torch

    from virtualfluids_amd.adjoint import KineticEnergyObservation, k17_steps
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory

    g = _duct(True)
    omega, steps, K = 1.7, 11, 3
    rng = np.random.default_rng(760)
    f0 = pack_pre(near_equilibrium(rng, g.N, u_amp=0.03, noise=2e-3), g, True).astype(np.float32)
    F = BoundaryConditionFactory
    fac = F()
    fac.set_velocity_boundary_condition(F.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(F.PressureBC.PressureNonEquilibriumCompressible)
    L = Lattice(g, omega)
    U = np.array(g.bcs["velocity"].velocity, dtype=np.float32)
    rho = rng.normal(0, 1e-3, g.bcs["pressure"].n).astype(np.float32)
    phi = rng.uniform(0, 0.05, g.N).astype(np.float32)
    lists = sum(27 * bc.n * 4 for name, bc in L.bcs.items() if name in ("velocity", "noslip"))
    seg = K * (27 * g.N * 4 + lists)

    def run(keep):
        x = torch.from_numpy(f0).cuda().requires_grad_(True)
        u = torch.from_numpy(U).cuda().requires_grad_(True)
        r = torch.from_numpy(rho).cuda().requires_grad_(True)
        p = torch.from_numpy(phi).cuda().requires_grad_(True)
        obs = KineticEnergyObservation(g, np.nonzero(g.geo == 19)[0])
        out, J = k17_steps(x, L, steps, checkpoint_every=K, even=True, bc_factory=fac, velocity=u, rho_bc=r,
                           porosity=p, observe=obs, observe_from=4, keep_bytes=keep)
        (J + (out.double() ** 2).sum() * 1e-3).backward()
        return [t.detach().cpu().numpy() for t in (out, J, x.grad, u.grad, r.grad, p.grad)]

    print(f"kept segments: {keep_segments} of {-(-steps // K)}", flush=True)
    ref = run(0)
    got = run(keep_segments * seg)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
