"""Gradients through two refinement levels on the GPU (k17_two_level_steps).

A channel (velocity inflow, pressure outlet, interpolated no-slip walls) with a nested fine box
and a random porosity field on the fine level, started from a uniform flow. Objective: the
kinetic energy of a coarse plane downstream of the fine box, averaged over the last coarse
steps. Its gradients with respect
to the inflow velocities and the fine porosity match FP64 central differences of the oracle's
two-level run (the stated 1e-4 bound); they do not depend on the checkpoint interval.
"""
import numpy as np
import pytest

from helpers import ORACLE_POST_BC
from virtualfluids_amd.grid import channel, gaussian_random_field
from virtualfluids_amd.refinement import nested_box

pytestmark = pytest.mark.gpu

NU = 0.01
OMEGA_C, OMEGA_F = 1.0 / (3 * NU + 0.5), 1.0 / (6 * NU + 0.5)
SPIN, AVG = 14, 4


def _case():
    coarse = channel((20, 12, 12), inflow=0.03)
    nb = nested_box(None, (6, 3, 3), (11, 9, 9), coarse=coarse)
    x, y, z = nb.coarse.coords()
    plane = np.nonzero((x == 15) & (nb.coarse.geo == 19))[0]  # downstream of the fine box
    phi = gaussian_random_field((nb.fine.nx, nb.fine.ny, nb.fine.nz), 2.0, seed=1500, low=0.0, high=0.05)
    return nb, plane, phi.astype(np.float32).astype(np.float64)


def _oracle_objective(orc, nb, plane, f0c, f0f, vx, phi):
    from virtualfluids_amd.adjoint import pre_collision_index

    C = orc.Lattice(nb.coarse.nbx, nb.coarse.nby, nb.coarse.nbz)
    F = orc.Lattice(nb.fine.nbx, nb.fine.nby, nb.fine.nbz)
    C.f[:] = f0c
    F.f[:] = f0f
    g = nb.coarse
    vel, ns, pr = g.bcs["velocity"], g.bcs["noslip"], g.bcs["pressure"]
    velocity = (vx, vel.velocity[1], vel.velocity[2])
    zc, zf = [np.zeros(nb.cf.n)] * 3, [np.zeros(nb.fc.n)] * 3
    idx = {e: pre_collision_index(g, e) for e in (True, False)}
    J = 0.0
    for step in range(SPIN + AVG):
        for _ in range(2):
            F.collide(nb.fine.fluid, OMEGA_F, porosity=phi)
            F.swap()
        C.collide(g.fluid, OMEGA_C)
        C.bc_post(ORACLE_POST_BC["VelocityInterpolatedCompressible"], vel.k, vel.q27, OMEGA_C, velocity)
        C.bc_post(ORACLE_POST_BC["NoSlipInterpolatedCompressible"], ns.k, ns.q27, OMEGA_C)
        C.swap()
        C.bc_pressure(pr.k, pr.kN, pr.rho_bc, pr.direction)
        orc.scale(True, C, F, C.even, nb.fc.coarse, nb.fc.fine, zf, OMEGA_C, OMEGA_F)
        orc.scale(False, C, F, C.even, nb.cf.coarse, nb.cf.fine, zc, OMEGA_C, OMEGA_F)
        if step >= SPIN:
            f = C.f[idx[C.even][:, plane]]
            drho = f.sum(0)
            j = np.stack([(f * c[:, None]).sum(0) for c in _C])
            J += (0.5 * (j * j).sum(0) / (1.0 + drho)).sum() / AVG
    return J


from virtualfluids_amd.grid import CX, CY, CZ  # noqa: E402

_C = [np.asarray(CX, float), np.asarray(CY, float), np.asarray(CZ, float)]


@pytest.mark.parametrize("K,observe", [(3, False), (100, False), (5, "callable"), (4, "cells")])
def test_two_level_gradients_match_fp64_differences(gpu, orc, K, observe):
    import torch

    from virtualfluids_amd.adjoint import (KineticEnergyObservation, k17_two_level_steps, kinetic_energy,
                                           pre_collision_index)
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory

    nb, plane, phi0 = _case()
    Fac = BoundaryConditionFactory
    fac = Fac()
    fac.set_velocity_boundary_condition(Fac.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(Fac.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(Fac.PressureBC.PressureNonEquilibriumCompressible)
    C = Lattice(nb.coarse, OMEGA_C, level=0)
    F = Lattice(nb.fine, OMEGA_F, level=1)
    for L in (C, F):
        z = np.zeros(L.N)
        L.init_equilibrium(z, np.full(L.N, 0.03), z, z)
    f0c, f0f = C.populations().astype(np.float64), F.populations().astype(np.float64)
    vel = nb.coarse.bcs["velocity"]
    u0 = np.stack([np.asarray(v, dtype=np.float64) for v in vel.velocity])
    U = torch.from_numpy(u0.astype(np.float32)).cuda().requires_grad_(True)
    phi = torch.from_numpy(phi0.astype(np.float32)).cuda().requires_grad_(True)
    idx = {e: torch.from_numpy(pre_collision_index(nb.coarse, e)).cuda() for e in (True, False)}
    nodes = torch.from_numpy(plane).cuda()

    if observe:
        # the time average through the observation hook, in one call: a plain callable (dense
        # gradient) or an observation that names its cells (sparse gradient)
        fn = (KineticEnergyObservation(nb.coarse, plane, scale=1.0 / AVG) if observe == "cells"
              else (lambda b, e: kinetic_energy(b, idx[e], nodes) / AVG))
        _, _, J = k17_two_level_steps(C.f.detach().clone(), F.f.detach().clone(), C, F, (nb.cf, nb.fc), SPIN + AVG,
                                      checkpoint_every=K, even=True, bc_factory=fac, velocity=U, porosity_fine=phi,
                                      observe=fn, observe_from=SPIN)
    else:
        c, f, _ = k17_two_level_steps(C.f.detach().clone(), F.f.detach().clone(), C, F, (nb.cf, nb.fc), SPIN,
                                      checkpoint_every=K, even=True, bc_factory=fac, velocity=U, porosity_fine=phi)
        even = SPIN % 2 == 0
        J = 0.0
        for _ in range(AVG):
            c, f, _ = k17_two_level_steps(c, f, C, F, (nb.cf, nb.fc), 1, checkpoint_every=K, even=even,
                                          bc_factory=fac, velocity=U, porosity_fine=phi)
            even = not even
            J = J + kinetic_energy(c, idx[even], nodes) / AVG
    J.backward()
    gU = U.grad.cpu().numpy().astype(np.float64)
    gphi = phi.grad.cpu().numpy().astype(np.float64)
    J_or = _oracle_objective(orc, nb, plane, f0c, f0f, u0[0], phi0)
    print(f"K={K} observe={observe}: J gpu {float(J.detach()):.10e} oracle {J_or:.10e}")
    assert abs(float(J.detach()) - J_or) < 1e-5 * abs(J_or)

    rng = np.random.default_rng(1501)
    h = 1e-7
    v = rng.normal(size=u0.shape[1])
    fd = (_oracle_objective(orc, nb, plane, f0c, f0f, u0[0] + h * v, phi0) -
          _oracle_objective(orc, nb, plane, f0c, f0f, u0[0] - h * v, phi0)) / (2 * h)
    ad = float(gU[0] @ v)
    print(f"  dJ/dU_x . v: adjoint {ad:.8e} FD {fd:.8e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(ad - fd) < 1e-4 * abs(fd)
    w = rng.normal(size=phi0.size)
    fd = (_oracle_objective(orc, nb, plane, f0c, f0f, u0[0], phi0 + h * w) -
          _oracle_objective(orc, nb, plane, f0c, f0f, u0[0], phi0 - h * w)) / (2 * h)
    ad = float(gphi @ w)
    print(f"  dJ/dphi_fine . w: adjoint {ad:.8e} FD {fd:.8e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(ad - fd) < 1e-4 * abs(fd)


@pytest.mark.parametrize("storage,tol", [("fp16", 2e-2)])
def test_two_level_sixteen_bit_gradients_match_fp64_differences(gpu, orc, storage, tol):
    """The two-level gradients with FP16 shifted storage on both levels (the storage of the
    two-level channel benchmark), from FP32 initial buffers: the adjoint linearises the
    FP16-stored trajectory, whose states differ from the FP64 run by the storage rounding
    (2^-11 relative per store), so against FP64 central differences of the oracle the gradients
    agree to the stated tolerance (2e-2 relative), not to the FP32 bound of 1e-4."""
    import torch

    from virtualfluids_amd.adjoint import KineticEnergyObservation, k17_two_level_steps
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory

    nb, plane, phi0 = _case()
    Fac = BoundaryConditionFactory
    fac = Fac()
    fac.set_velocity_boundary_condition(Fac.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(Fac.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(Fac.PressureBC.PressureNonEquilibriumCompressible)
    # FP32 initial buffers (the uniform flow), stepped on 16-bit lattices
    init = []
    for g, om, lev in ((nb.coarse, OMEGA_C, 0), (nb.fine, OMEGA_F, 1)):
        L32 = Lattice(g, om, level=lev)
        z = np.zeros(L32.N)
        L32.init_equilibrium(z, np.full(L32.N, 0.03), z, z)
        init.append(L32.f.detach().clone())
    C = Lattice(nb.coarse, OMEGA_C, storage=storage, level=0)
    F = Lattice(nb.fine, OMEGA_F, storage=storage, level=1)
    f0c, f0f = (t.cpu().numpy().astype(np.float64) for t in init)
    vel = nb.coarse.bcs["velocity"]
    u0 = np.stack([np.asarray(v, dtype=np.float64) for v in vel.velocity])
    U = torch.from_numpy(u0.astype(np.float32)).cuda().requires_grad_(True)
    phi = torch.from_numpy(phi0.astype(np.float32)).cuda().requires_grad_(True)
    x0c = init[0].clone().requires_grad_(True)
    obs = KineticEnergyObservation(nb.coarse, plane, scale=1.0 / AVG)
    c, f, J = k17_two_level_steps(x0c, init[1].clone(), C, F, (nb.cf, nb.fc), SPIN + AVG, checkpoint_every=5,
                                  even=True, bc_factory=fac, velocity=U, porosity_fine=phi, observe=obs,
                                  observe_from=SPIN)
    assert c.dtype == f.dtype == torch.float32
    J.backward()
    assert x0c.grad.dtype == torch.float32
    J_or = _oracle_objective(orc, nb, plane, f0c, f0f, u0[0], phi0)
    print(f"{storage}: J gpu {float(J.detach()):.10e} oracle {J_or:.10e}")
    assert abs(float(J.detach()) - J_or) < tol * abs(J_or)
    rng = np.random.default_rng(1502)
    h = 1e-7
    v = rng.normal(size=u0.shape[1])
    fd = (_oracle_objective(orc, nb, plane, f0c, f0f, u0[0] + h * v, phi0) -
          _oracle_objective(orc, nb, plane, f0c, f0f, u0[0] - h * v, phi0)) / (2 * h)
    ad = float(U.grad.cpu().numpy().astype(np.float64)[0] @ v)
    print(f"  {storage} dJ/dU_x . v: adjoint {ad:.8e} FD {fd:.8e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(ad - fd) < tol * abs(fd)
    w = rng.normal(size=phi0.size)
    fd = (_oracle_objective(orc, nb, plane, f0c, f0f, u0[0], phi0 + h * w) -
          _oracle_objective(orc, nb, plane, f0c, f0f, u0[0], phi0 - h * w)) / (2 * h)
    ad = float(phi.grad.cpu().numpy().astype(np.float64) @ w)
    print(f"  {storage} dJ/dphi_fine . w: adjoint {ad:.8e} FD {fd:.8e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(ad - fd) < tol * abs(fd)


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("keep_segments", [1, 2, 10])
def test_two_level_kept_segments_give_the_same_gradient(gpu, storage, keep_segments):
    """keep_bytes on two levels: the trailing segments that fit are stepped out of place in the
    forward and walked by the backward without a replay. The channel with a nested fine box,
    porosity on both levels and the observed plane energy: outputs, observed sum and every
    gradient (both initial buffers, the inflow velocities, both porosity fields) equal to those of
    the pure-recompute run bit for bit (10 segments: all of them kept)."""
    import torch

    from virtualfluids_amd.adjoint import KineticEnergyObservation, k17_two_level_steps
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory as Fac

    nb, plane, phi = _case()
    steps, K = 9, 2
    fac = Fac()
    fac.set_velocity_boundary_condition(Fac.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(Fac.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(Fac.PressureBC.PressureNonEquilibriumCompressible)
    Lc = Lattice(nb.coarse, OMEGA_C, storage=storage, level=0)
    Lf = Lattice(nb.fine, OMEGA_F, storage=storage, level=1)
    for L in (Lc, Lf):
        z = np.zeros(L.N)
        L.init_equilibrium(z, np.full(L.N, 0.03), z, z)
    rng = np.random.default_rng(1640)
    pc0 = rng.uniform(0, 0.03, nb.coarse.N).astype(np.float32)
    pf0 = phi.astype(np.float32)
    u0 = np.stack([np.asarray(v, dtype=np.float32) for v in nb.coarse.bcs["velocity"].velocity])
    lists = 27 * 4 * sum(bc.n for name, bc in Lc.bcs.items() if name != "pressure")
    step_bytes = 2 * Lf.f.numel() * Lf.f.element_size() + Lc.f.numel() * Lc.f.element_size() + lists
    seg = K * (step_bytes + Lc.f.numel() * Lc.f.element_size() + Lf.f.numel() * Lf.f.element_size())

    def run(keep):
        xc = Lc.f.detach().float().clone().requires_grad_(True)
        xf = Lf.f.detach().float().clone().requires_grad_(True)
        U = torch.from_numpy(u0).cuda().requires_grad_(True)
        pc = torch.from_numpy(pc0).cuda().requires_grad_(True)
        pf = torch.from_numpy(pf0).cuda().requires_grad_(True)
        obs = KineticEnergyObservation(nb.coarse, plane, scale=0.25)
        c, f, J = k17_two_level_steps(xc, xf, Lc, Lf, (nb.cf, nb.fc), steps, checkpoint_every=K, even=True,
                                      bc_factory=fac, velocity=U, porosity_coarse=pc, porosity_fine=pf,
                                      observe=obs, observe_from=5, keep_bytes=keep)
        (J + 1e-3 * (c.double() ** 2).sum() + 1e-3 * (f.double() ** 2).sum()).backward()
        return [t.detach().cpu().numpy() for t in (c, f, J, xc.grad, xf.grad, U.grad, pc.grad, pf.grad)]

    print(f"{storage}: kept segments {keep_segments} of {-(-steps // K)}", flush=True)
    ref = run(0)
    got = run(keep_segments * seg)
    for name, a, b in zip(("c", "f", "J", "gc", "gf", "gU", "gpc", "gpf"), ref, got):
        assert np.array_equal(a, b), name
