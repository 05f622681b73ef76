"""Any number of refinement levels on the GPU: MultiLevelSimulation's steps and the gradients of
k17_multilevel_steps.

Three levels forward: every level's populations after several coarse steps of MultiLevelSimulation
(HIP graphs and host launches) against the FP64 oracle's run of the same hierarchy.

Two levels: the same outputs, observed sum and gradients as k17_two_level_steps, bit for bit.
Three levels (multilevel_reference's case: a channel with a fine box nested in it and a finer
box nested in the fine box): the GPU objective matches the oracle's FP64 run, and every gradient
matches the FP64 reference adjoint, itself pinned by FP64 differences on the CPU.
"""
import numpy as np
import pytest

from multilevel_reference import omega as _omega
from virtualfluids_amd.grid import channel
from virtualfluids_amd.refinement import nested_box

pytestmark = pytest.mark.gpu

SPIN, AVG = 16, 4


def _factory():
    from virtualfluids_amd.simulation import BoundaryConditionFactory as Fac

    fac = Fac()
    fac.set_velocity_boundary_condition(Fac.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(Fac.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(Fac.PressureBC.PressureNonEquilibriumCompressible)
    return fac


def _lattices(grids, limiters=(0.01, 0.01, 0.01)):
    from virtualfluids_amd.lattice import Lattice

    lats = []
    for lev, g in enumerate(grids):
        L = Lattice(g, _omega(lev), limiters=limiters, level=lev)
        z = np.zeros(L.N)
        L.init_equilibrium(z, np.full(L.N, 0.03), z, z)
        lats.append(L)
    return lats


def test_two_levels_equal_the_two_level_stepper(gpu, orc):
    """k17_multilevel_steps on two levels against k17_two_level_steps: outputs, observed sum and
    every gradient (initial buffers, inflow velocities, both porosities) equal bit for bit."""
    import torch

    from virtualfluids_amd.adjoint import KineticEnergyObservation, k17_multilevel_steps, k17_two_level_steps

    coarse = channel((20, 12, 12), inflow=0.03)
    nb = nested_box(None, (6, 3, 3), (11, 9, 9), coarse=coarse)
    lats = _lattices([nb.coarse, nb.fine])
    x, _, _ = nb.coarse.coords()
    plane = np.nonzero((x == 15) & (nb.coarse.geo == 19))[0]
    rng = np.random.default_rng(1610)
    pc0 = rng.uniform(0, 0.03, nb.coarse.N).astype(np.float32)
    pf0 = rng.uniform(0, 0.05, nb.fine.N).astype(np.float32)
    u0 = np.stack([np.asarray(v, dtype=np.float32) for v in nb.coarse.bcs["velocity"].velocity])
    fac = _factory()

    def run(multi):
        xc = lats[0].f.detach().clone().requires_grad_(True)
        xf = lats[1].f.detach().clone().requires_grad_(True)
        U = torch.from_numpy(u0).cuda().requires_grad_(True)
        pc = torch.from_numpy(pc0).cuda().requires_grad_(True)
        pf = torch.from_numpy(pf0).cuda().requires_grad_(True)
        obs = KineticEnergyObservation(nb.coarse, plane, scale=0.25)
        if multi:
            (c, f), J = k17_multilevel_steps([xc, xf], lats, [(nb.cf, nb.fc)], 9, checkpoint_every=4, even=True,
                                             bc_factories=[fac, None], velocity=U, porosity=[pc, pf], observe=obs,
                                             observe_from=5)
        else:
            c, f, J = k17_two_level_steps(xc, xf, lats[0], lats[1], (nb.cf, nb.fc), 9, checkpoint_every=4, even=True,
                                          bc_factory=fac, velocity=U, porosity_coarse=pc, porosity_fine=pf,
                                          observe=obs, observe_from=5)
        (J + 1e-3 * (c.double() ** 2).sum() + 1e-3 * (f.double() ** 2).sum()).backward()
        return [t.detach().cpu().numpy() for t in (c, f, J, xc.grad, xf.grad, U.grad, pc.grad, pf.grad)]

    print("two levels: multilevel against two-level stepper", flush=True)
    ref = run(False)
    got = run(True)
    for name, a, b in zip(("c", "f", "J", "gc", "gf", "gU", "gpc", "gpf"), ref, got):
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("K,limiter", [(3, 0.01), (100, 0.01), (3, 0.1)])
def test_three_level_gradients_match_the_fp64_reference_adjoint(gpu, orc, K, limiter):
    """Three levels from perturbed uniform flows, porosity on the two finer levels: the GPU's
    observed sum against the oracle's FP64 run, and every gradient (inflow velocities, both
    porosity fields, the three initial buffers) against the FP64 reference adjoint of
    multilevel_reference, which FP64 differences pin to 4e-7 (test_multilevel_adjoint_cpu).

    Bounds (relative L2 error of the whole gradient): 1e-5 for the inflow velocities and the
    coarse level's buffer; for the finer levels' porosities and buffers 1e-2 with the K17 default
    quadric limiter L = 0.01 and 2e-3 with L = 0.1. The finer levels' gradients are stiff in the
    state: the limiter x N / (|x| + L) has a jump in its second derivative wherever a cumulant
    crosses zero, so the derivative along a finest-level cell moves by 10 % over 1e-7 (measured
    with FP64 differences), and the FP32 trajectory's rounding moves the gradient by about
    5e-3 / 7e-4 at L = 0.01 / 0.1 (the error falls with the limiter's curvature; the same
    reference at the coarse level agrees to 1.5e-6). With L = 1 the unlimited relaxation blows up."""
    import torch

    import multilevel_reference as ref
    from virtualfluids_amd.adjoint import KineticEnergyObservation, k17_multilevel_steps

    grids, interfaces, plane, phi2 = ref.three_levels()
    rng = np.random.default_rng(1620)
    phi1 = rng.uniform(0, 0.05, grids[1].N).astype(np.float32).astype(np.float64)
    f0s = ref.initial_states(grids, 1621)
    lim = (limiter,) * 3
    lats = _lattices(grids, lim)
    vel = grids[0].bcs["velocity"]
    u0 = np.stack([np.asarray(v, dtype=np.float32) for v in vel.velocity])
    xs = [torch.from_numpy(f.astype(np.float32)).cuda().requires_grad_(True) for f in f0s]
    U = torch.from_numpy(u0).cuda().requires_grad_(True)
    p1 = torch.from_numpy(phi1.astype(np.float32)).cuda().requires_grad_(True)
    p2 = torch.from_numpy(phi2.astype(np.float32)).cuda().requires_grad_(True)
    obs = KineticEnergyObservation(grids[0], plane, scale=1.0 / AVG)
    print(f"three levels K={K} limiter={limiter}: nodes {[g.N for g in grids]}", flush=True)
    _, J = k17_multilevel_steps(xs, lats, interfaces, SPIN + AVG, checkpoint_every=K, even=True,
                                bc_factories=[_factory(), None, None], velocity=U, porosity=[None, p1, p2],
                                observe=obs, observe_from=SPIN)
    J.backward()
    J_ref, lams, ubar, padj = ref.gradient(orc, grids, interfaces, plane, f0s, u0[0].astype(np.float64),
                                           [None, phi1, phi2], SPIN + AVG, AVG, limiters=lim)
    print(f"  J gpu {float(J.detach()):.10e} reference {J_ref:.10e}")
    assert abs(float(J.detach()) - J_ref) < 1e-5 * abs(J_ref)
    for name, got, want in (
            ("inflow u_x", U.grad[0], ubar[:, 0]),
            ("porosity level 1", p1.grad, padj[1]),
            ("porosity level 2", p2.grad, padj[2]),
            ("initial state level 0", xs[0].grad, lams[0]),
            ("initial state level 1", xs[1].grad, lams[1]),
            ("initial state level 2", xs[2].grad, lams[2])):
        g = got.detach().cpu().numpy().astype(np.float64)
        err = np.linalg.norm(g - want) / np.linalg.norm(want)
        tol = 1e-5 if name in ("inflow u_x", "initial state level 0") else (1e-2 if limiter < 0.05 else 2e-3)
        print(f"  {name}: relative L2 error against the FP64 reference adjoint {err:.2e} (bound {tol:g})")
        assert err < tol, name


def test_three_level_steps_match_oracle(gpu, orc):
    """Six coarse steps (24 on the finest level) of MultiLevelSimulation on three levels with the
    channel's interpolated BCs and pressure outlet, against multilevel_reference's FP64 run from the
    same FP32 initial buffers: every cell of every level within 2e-6; HIP-graph replays and host
    launches equal bit for bit."""
    import torch

    import multilevel_reference as ref
    from virtualfluids_amd.refinement import GridScalingFactory
    from virtualfluids_amd.simulation import MultiLevelSimulation

    grids, interfaces, plane, _ = ref.three_levels()
    f0s = ref.initial_states(grids, 1630)
    fac = GridScalingFactory()
    fac.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
    out = []
    for graphs in (True, False):
        lats = _lattices(grids)
        for L, f in zip(lats, f0s):
            L.f.copy_(torch.from_numpy(f.astype(np.float32)))
        MultiLevelSimulation(lats, interfaces, [_factory(), None, None], scaling_factory=fac).run(6, graphs=graphs)
        out.append([L.populations() for L in lats])
    vx = np.asarray(grids[0].bcs["velocity"].velocity[0], dtype=np.float64)
    _, want = ref.run(orc, grids, interfaces, plane, f0s, vx, [None, None, None], 6, 1, final_states=True)
    for lev in range(3):
        assert np.array_equal(out[0][lev], out[1][lev]), lev
        err = np.abs(out[0][lev] - want[lev]).max()
        print(f"three levels, 6 coarse steps: level {lev} max |GPU - FP64 oracle| {err:.2e}")
        assert err < 2e-6, lev


@pytest.mark.parametrize("keep_segments", [1, 3])
def test_three_level_kept_segments_give_the_same_gradient(gpu, keep_segments):
    """keep_bytes on three levels: the trailing segments that fit are stepped out of place in the
    forward and walked by the backward without a replay. Outputs, observed sum and every gradient
    (the three initial buffers, the inflow velocities, both porosity fields) equal to those of the
    pure-recompute run bit for bit (3 segments: all of them kept)."""
    import torch

    import multilevel_reference as ref
    from virtualfluids_amd.adjoint import KineticEnergyObservation, k17_multilevel_steps

    grids, interfaces, plane, phi2 = ref.three_levels()
    rng = np.random.default_rng(1650)
    phi1 = rng.uniform(0, 0.05, grids[1].N).astype(np.float32)
    f0s = ref.initial_states(grids, 1651)
    lats = _lattices(grids)
    u0 = np.stack([np.asarray(v, dtype=np.float32) for v in grids[0].bcs["velocity"].velocity])
    steps, K = 7, 3
    lists = [27 * 4 * sum(bc.n for name, bc in L.bcs.items() if name != "pressure") for L in lats]
    nbytes = [27 * L.N * 4 for L in lats]
    step_bytes = sum(2 ** l * (nbytes[l] + lists[l] + (nbytes[l] + nbytes[l + 1] if l < 2 else 0)) for l in range(3))

    def run(keep):
        xs = [torch.from_numpy(f.astype(np.float32)).cuda().requires_grad_(True) for f in f0s]
        U = torch.from_numpy(u0).cuda().requires_grad_(True)
        p1 = torch.from_numpy(phi1).cuda().requires_grad_(True)
        p2 = torch.from_numpy(phi2.astype(np.float32)).cuda().requires_grad_(True)
        obs = KineticEnergyObservation(grids[0], plane, scale=0.25)
        outs, J = k17_multilevel_steps(xs, lats, interfaces, steps, checkpoint_every=K, even=True,
                                       bc_factories=[_factory(), None, None], velocity=U, porosity=[None, p1, p2],
                                       observe=obs, observe_from=3, keep_bytes=keep)
        (J + sum(1e-3 * (o.double() ** 2).sum() for o in outs)).backward()
        return ([o.detach().cpu().numpy() for o in outs] + [J.detach().cpu().numpy()] +
                [t.grad.cpu().numpy() for t in xs + [U, p1, p2]])

    print(f"three levels: kept segments {keep_segments} of {-(-steps // K)}", flush=True)
    a = run(0)
    b = run(keep_segments * K * step_bytes)
    for i, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), i
