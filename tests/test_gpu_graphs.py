"""HIP-graph replay of the forward sweep against eager launches, bit for bit.

* One level (a channel with velocity inflow, pressure outlet, interpolated no-slip walls and a
  porosity field; dense-box and indexed addressing) and a two-level pair: the step replayed from
  its graph and the same step launched eagerly on a second copy leave identical states after
  every step, with eager torch work and a synchronisation between the replays.
* k17_steps / k17_two_level_steps with graphs=True and graphs=False: identical outputs,
  observed sums and gradients.
"""
import numpy as np
import pytest

from virtualfluids_amd.grid import channel, gaussian_random_field
from virtualfluids_amd.refinement import nested_box

pytestmark = pytest.mark.gpu


def _factory():
    from virtualfluids_amd.simulation import BoundaryConditionFactory as F

    fac = F()
    fac.set_velocity_boundary_condition(F.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(F.PressureBC.PressureNonEquilibriumCompressible)
    return fac


def _channel_lattice(box, storage="fp32"):
    import torch

    from virtualfluids_amd.lattice import Lattice

    g = channel((24, 10, 8), inflow=0.03)
    L = Lattice(g, 1.7, storage=storage, box=box)
    z = np.zeros(g.N)
    L.init_equilibrium(z, np.where(g.geo == 19, 0.03, 0.0), z, z)
    phi = gaussian_random_field((g.nx, g.ny, g.nz), 2.0, seed=1700, low=0.0, high=0.05)
    L.set_porosity(torch.from_numpy(phi.astype(np.float32)).cuda())
    return L


@pytest.mark.parametrize("box", [True, False])
def test_single_level_replay_matches_eager_steps(gpu, box):
    import torch

    from virtualfluids_amd.adjoint import Stepper
    from virtualfluids_amd.graphs import StepGraphs as _StepGraphs

    L = _channel_lattice(box)
    st = Stepper(L, _factory())
    A, B = L.f.clone(), L.f.clone()
    run = _StepGraphs(lambda ev, s: st.forward(A, ev, None, s))
    even = True
    for i in range(9):
        run(even)
        st.forward(B, even, None, torch.cuda.current_stream())
        torch.cuda.synchronize()
        assert torch.equal(A, B), f"step {i} (graphs: {sorted(run.graphs)})"
        _ = (B.float() * 2).sum().item()  # eager work between the replays
        even = not even
    assert len(run.graphs) == 2


def _pair(storage="fp16"):
    import torch

    from virtualfluids_amd.lattice import Lattice

    nb = nested_box(None, (6, 3, 3), (11, 9, 9), coarse=channel((20, 12, 12), inflow=0.03))
    C = Lattice(nb.coarse, 1.6, storage=storage, level=0)
    F = Lattice(nb.fine, 1.45, storage=storage, level=1)
    for L in (C, F):
        z = np.zeros(L.N)
        L.init_equilibrium(z, np.where(L.grid.geo == 19, 0.03, 0.0), z, z)
    phi = gaussian_random_field((nb.fine.nx, nb.fine.ny, nb.fine.nz), 2.0, seed=1710, low=0.0, high=0.05)
    F.set_porosity(torch.from_numpy(phi.astype(np.float32)).cuda())
    return nb, C, F


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_two_level_replay_matches_eager_steps(gpu, storage):
    import torch

    from virtualfluids_amd.adjoint import Stepper, _two_level_step_in_place
    from virtualfluids_amd.graphs import StepGraphs as _StepGraphs
    from virtualfluids_amd.refinement import DeviceInterfaceCells

    nb, Lc, Lf = _pair(storage)
    sc, sf = Stepper(Lc, _factory()), Stepper(Lf)
    cells = (DeviceInterfaceCells(nb.cf, Lc.device), DeviceInterfaceCells(nb.fc, Lc.device))
    Ca, Fa, Cb, Fb = Lc.f.clone(), Lf.f.clone(), Lc.f.clone(), Lf.f.clone()
    run = _StepGraphs(lambda ev, s: _two_level_step_in_place(sc, sf, cells, Ca, Fa, ev, s))
    even = True
    for i in range(7):
        run(even)
        _two_level_step_in_place(sc, sf, cells, Cb, Fb, even, torch.cuda.current_stream())
        torch.cuda.synchronize()
        assert torch.equal(Ca, Cb) and torch.equal(Fa, Fb), f"coarse step {i}"
        _ = Fb.float().abs().max().item()
        even = not even


# 13 steps every 4: single-step graphs; 45 steps every 20: graphs of 16 steps inside segments
@pytest.mark.parametrize("steps,K", [(13, 4), (45, 20)])
def test_k17_steps_same_with_and_without_graphs(gpu, steps, K):
    import torch

    from virtualfluids_amd.adjoint import k17_steps

    L = _channel_lattice(True)
    U0 = torch.stack(L.bcs["velocity"].vel).clone()
    out = []
    for graphs in (True, False):
        x = L.f.detach().clone().requires_grad_(True)
        U = U0.clone().requires_grad_(True)
        y = k17_steps(x, L, steps, checkpoint_every=K, even=True, bc_factory=_factory(), velocity=U, graphs=graphs)
        w = torch.linspace(-1, 1, y.numel(), device="cuda").reshape(y.shape)
        (y.float() * w).sum().backward()
        out.append((y.detach(), x.grad, U.grad))
    for a, b in zip(*out):
        assert torch.equal(a, b)


@pytest.mark.parametrize("steps,K,observe_from", [(9, 4, 5), (41, 20, 37)])
def test_two_level_steps_same_with_and_without_graphs(gpu, steps, K, observe_from):
    import torch

    from virtualfluids_amd.adjoint import KineticEnergyObservation, k17_two_level_steps

    nb, Lc, Lf = _pair("fp32")
    x, _, _ = nb.coarse.coords()
    # a plane the inflow's perturbation reaches within the run
    obs = KineticEnergyObservation(nb.coarse, np.nonzero((x == 3) & (nb.coarse.geo == 19))[0], scale=0.25)
    U0 = torch.stack(Lc.bcs["velocity"].vel).clone()
    out = []
    for graphs in (True, False):
        U = U0.clone().requires_grad_(True)
        C, F, J = k17_two_level_steps(Lc.f.clone(), Lf.f.clone(), Lc, Lf, (nb.cf, nb.fc), steps,
                                      checkpoint_every=K, even=True, bc_factory=_factory(), velocity=U, observe=obs,
                                      observe_from=observe_from, graphs=graphs)
        J.backward()
        out.append((C.detach(), F.detach(), J.detach(), U.grad))
    for name, a, b in zip(("coarse", "fine", "objective", "inflow gradient"), *out):
        print(name, float((a.double() - b.double()).abs().max()), float(a.double().abs().max()))
    # the forward sweep is the same bit for bit; the reverse sweep's interface adjoints add into
    # shared source cells with atomics, whose order (and last bit) varies from run to run
    for name, a, b in zip(("coarse", "fine", "objective"), out[0][:3], out[1][:3]):
        assert torch.equal(a, b), name
    ga, gb = out[0][3].double(), out[1][3].double()
    assert float((ga - gb).abs().max()) <= 1e-6 * float(ga.abs().max())
    assert float(out[0][2]) > 0 and float(out[0][3].abs().sum()) > 0


@pytest.mark.parametrize("box", [True, False])
def test_simulation_run_same_with_and_without_graphs(gpu, box):
    """Simulation.run replays one graph per parity: chunks of steps (eager work and a
    synchronisation between them) leave the same state as host-launched steps."""
    import torch

    from virtualfluids_amd.simulation import Simulation

    A, B = _channel_lattice(box), _channel_lattice(box)
    sa, sb = Simulation(A, _factory()), Simulation(B, _factory())
    # 7 steps: per-parity graphs only; 37 and 40: graphs of 16 steps, then single steps
    for chunk, n in enumerate((7, 37, 7, 40)):
        sa.run(n, graphs=True)
        sb.run(n, graphs=False)
        torch.cuda.synchronize()
        assert A.even == B.even and sa.t == sb.t
        assert torch.equal(A.f, B.f), f"chunk {chunk}"
        _ = A.f.float().abs().sum().item()


def test_multilevel_run_same_with_and_without_graphs(gpu):
    """MultiLevelSimulation.run replays pairs of coarse steps from one graph: runs of 7 and 10
    coarse steps (eager work and a synchronisation between them) leave the same fields on both
    levels as host-launched steps."""
    import torch

    from virtualfluids_amd.refinement import GridScalingFactory
    from virtualfluids_amd.simulation import MultiLevelSimulation

    runs = []
    for _ in range(2):
        nb, C, F = _pair("fp16")
        sf = GridScalingFactory()
        sf.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
        runs.append((nb, C, F, MultiLevelSimulation([C, F], [(nb.cf, nb.fc)], [_factory(), None],
                                                    scaling_factory=sf)))
    (_, Ca, Fa, ma), (_, Cb, Fb, mb) = runs
    # a coarse step cycles through both levels: the collisions pick their kernel from the sum
    assert Ca.param.workingSetBytes == Fa.param.workingSetBytes == Ca.state_bytes + Fa.state_bytes
    for n in (7, 10, 3):
        ma.run(n, graphs=True)
        mb.run(n, graphs=False)
        torch.cuda.synchronize()
        assert ma.t == mb.t and [lv.t for lv in ma.levels] == [lv.t for lv in mb.levels]
        assert Ca.even == Cb.even and Fa.even == Fb.even
        assert torch.equal(Ca.f, Cb.f) and torch.equal(Fa.f, Fb.f), f"after {n} steps"
        _ = Fa.f.float().abs().max().item()
