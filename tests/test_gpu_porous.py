"""Brinkman porosity on the GPU and the sphere-channel gradient (a reduced config 4).

* one porous collide+stream step (box and indexed kernels) against the oracle's porous
  collision (FP32 bar 2e-7);
* the adjoint kernels' porosity adjoint against the oracle's dual-number adjoint (3e-5 of the
  largest entry);
* a duct with a sphere (interpolated no-slip with exact ray-sphere q), velocity inflow,
  pressure outlet and a seeded Gaussian porosity field: the gradient of the drag averaged over
  the last steps with respect to the porosity field matches FP64 central differences of the
  oracle's run (1e-4, the stated bound).
"""
import ctypes

import numpy as np
import pytest

from helpers import near_equilibrium, pack_pre
from virtualfluids_amd.grid import gaussian_random_field, periodic_box, walled_box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("box", [True, False])
@pytest.mark.parametrize("kind", ["periodic", "walled"])
def test_porous_step_matches_oracle(gpu, orc, box, kind):
    from virtualfluids_amd.lattice import Lattice

    g = periodic_box(24, 10, 8) if kind == "periodic" else walled_box((22, 8, 6), {})
    rng = np.random.default_rng(800 + box)
    phi = gaussian_random_field((g.nx, g.ny, g.nz), 2.0, seed=801)
    buf = pack_pre(near_equilibrium(rng, g.N, u_amp=0.05, noise=5e-3), g, True)
    L = Lattice(g, 1.8, box=box)
    L.set_porosity(phi)
    L.set_populations(buf)
    L.collide()
    ref = orc.Lattice(g.nbx, g.nby, g.nbz)
    ref.f[:] = buf.astype(np.float32)
    ref.collide(g.fluid, 1.8, porosity=phi.astype(np.float32).astype(np.float64))
    err = np.abs(L.populations() - ref.f).max()
    print(f"porous step box={box} {kind}: max err {err:.2e}")
    assert err < 2e-7


@pytest.mark.parametrize("box", [True, False])
def test_porosity_adjoint_matches_reference(gpu, orc, box):
    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import Lattice

    g = walled_box((22, 8, 6), {})
    rng = np.random.default_rng(810 + box)
    phi = gaussian_random_field((g.nx, g.ny, g.nz), 2.0, seed=811)
    L = Lattice(g, 1.8, box=box)
    L.set_porosity(phi)
    L.set_populations(pack_pre(near_equilibrium(rng, g.N, u_amp=0.05, noise=5e-3), g, True))
    lam0 = rng.normal(0, 1, 27 * g.N).astype(np.float32)
    lam = torch.from_numpy(lam0).cuda()
    padj = torch.zeros(g.N, dtype=torch.float32, device="cuda")
    fn = "vf_k17_adjoint_box" if box else "vf_k17_adjoint"
    capi.call(fn, L.pref(), ctypes.c_void_p(lam.data_ptr()), None, ctypes.c_void_p(padj.data_ptr()), None)
    A = orc.Lattice(g.nbx, g.nby, g.nbz)
    A.f[:] = L.populations()
    ref = lam0.astype(np.float64)
    ref_p = np.zeros(g.N)
    orc.adjoint_step(A, ref, g.fluid, 1.8, porosity=phi.astype(np.float32).astype(np.float64), porosity_adjoint=ref_p)
    err = np.abs(lam.cpu().numpy() - ref).max() / np.abs(ref).max()
    err_p = np.abs(padj.cpu().numpy() - ref_p).max() / np.abs(ref_p).max()
    print(f"porosity adjoint box={box}: rel err {err:.2e}, porosity {err_p:.2e}")
    assert err < 3e-5 and err_p < 1e-4


# per_step: the averaging window as single-step k17_steps calls; observe: one k17_steps call
# whose observation hook sums the drag over the window; observe_cells: the same with the
# MomentumExchangeObservation, whose gradient is scattered onto the link cells only
@pytest.mark.parametrize("form", ["per_step", "observe", "observe_cells"])
def test_sphere_channel_drag_gradient_wrt_porosity(gpu, orc, form):
    import torch

    from helpers import ORACLE_POST_BC
    from virtualfluids_amd.adjoint import k17_steps, momentum_exchange, momentum_exchange_index
    from virtualfluids_amd.grid import sphere_channel
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory

    g = sphere_channel((32, 14, 14), (11, 6.5, 6.5), 3.5, 0.04)
    phi0 = gaussian_random_field((g.nx, g.ny, g.nz), 3.0, seed=20261202).astype(np.float32).astype(np.float64)
    omega, spin_up, avg = 1.8, 30, 10
    F = BoundaryConditionFactory
    fac = F()
    fac.set_velocity_boundary_condition(F.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_geometry_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(F.PressureBC.PressureNonEquilibriumCompressible)

    L = Lattice(g, omega)
    zeros = np.zeros(g.N)
    L.init_equilibrium(zeros, zeros, zeros, zeros)
    f0 = L.populations().astype(np.float64)
    idx = [torch.from_numpy(a).cuda() for a in momentum_exchange_index(g, "geometry", True)]
    idx_odd = [torch.from_numpy(a).cuda() for a in momentum_exchange_index(g, "geometry", False)]
    phi = torch.from_numpy(phi0.astype(np.float32)).cuda().requires_grad_(True)
    x = L.f.detach().clone()
    if form == "observe":
        def drag(buf, even):
            return momentum_exchange(buf, *(idx if even else idx_odd))[0] / avg

        _, J = k17_steps(x, L, spin_up + avg, checkpoint_every=10, even=True, bc_factory=fac, porosity=phi,
                         observe=drag, observe_from=spin_up)
    elif form == "observe_cells":
        from virtualfluids_amd.adjoint import MomentumExchangeObservation

        obs = MomentumExchangeObservation(g, "geometry", axis=0, scale=1.0 / avg)
        _, J = k17_steps(x, L, spin_up + avg, checkpoint_every=10, even=True, bc_factory=fac, porosity=phi,
                         observe=obs, observe_from=spin_up)
    else:
        s = k17_steps(x, L, spin_up, checkpoint_every=10, even=True, bc_factory=fac, porosity=phi)
        even = spin_up % 2 == 0
        J = 0.0
        for _ in range(avg):
            s = k17_steps(s, L, 1, checkpoint_every=1, even=even, bc_factory=fac, porosity=phi)
            even = not even
            i = idx if even else idx_odd
            J = J + momentum_exchange(s, *i)[0] / avg
    J.backward()
    grad = phi.grad.cpu().numpy().astype(np.float64)

    vel, ns, pr, geo = g.bcs["velocity"], g.bcs["noslip"], g.bcs["pressure"], g.bcs["geometry"]
    mx = [momentum_exchange_index(g, "geometry", e) for e in (False, True)]

    def oracle_J(ph):
        A = orc.Lattice(g.nbx, g.nby, g.nbz)
        A.f[:] = f0
        total = 0.0
        for step in range(spin_up + avg):
            A.collide(g.fluid, omega, porosity=ph)
            A.bc_post(ORACLE_POST_BC["VelocityInterpolatedCompressible"], vel.k, vel.q27, omega, vel.velocity)
            A.bc_post(ORACLE_POST_BC["NoSlipInterpolatedCompressible"], ns.k, ns.q27, omega)
            A.bc_post(ORACLE_POST_BC["NoSlipInterpolatedCompressible"], geo.k, geo.q27, omega)
            A.swap()
            A.bc_pressure(pr.k, pr.kN, pr.rho_bc, pr.direction)
            if step >= spin_up:
                o, i_, c = mx[int(A.even)]
                total += (c[0] * (A.f[o] + A.f[i_])).sum() / avg
        return total

    rng = np.random.default_rng(820)
    h = 1e-6
    for trial in range(2):
        v = rng.normal(0, 1, g.N)
        fd = (oracle_J(phi0 + h * v) - oracle_J(phi0 - h * v)) / (2 * h)
        ad = float(grad @ v)
        print(f"drag {float(J):.6e}: <dJ/dphi, v> adjoint {ad:.8e} FD {fd:.8e} rel {abs(ad - fd) / abs(fd):.2e}")
        assert abs(ad - fd) < 1e-4 * abs(fd)
