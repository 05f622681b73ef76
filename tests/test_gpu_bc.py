"""GPU parity of the boundary-condition kernels and the full step ordering (C ABI).

The GPU runs the reference's step order through the Python BoundaryConditionFactory /
Simulation; the FP64 oracle runs the same order (tests/helpers.oracle_run). Errors are
stated per field: L2 and max of drho, and of u relative to the driving velocity.
"""
import numpy as np
import pytest

from helpers import oracle_run
from virtualfluids_amd.grid import cavity, channel

pytestmark = pytest.mark.gpu


def _gpu_run(grid, omega, steps, velocity_bc, noslip_bc, pressure_bc=None, box=True, storage="fp32",
             limiters=(0.01, 0.01, 0.01), rows=None):
    """rows: the layout hint of every BC list (None: what the host derives from the lists; 0: a
    lane per direction for the interpolated and pressure BCs; 1: a thread per node)."""
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory as F, Simulation

    L = Lattice(grid, omega, storage=storage, box=box, limiters=limiters)
    if rows is not None:
        for bc in L.bcs.values():
            bc.struct.nodesAlongRows = rows
    N = grid.N
    L.init_equilibrium(np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N))
    f = F()
    if velocity_bc:
        f.set_velocity_boundary_condition(getattr(F.VelocityBC, velocity_bc))
    f.set_no_slip_boundary_condition(getattr(F.NoSlipBC, noslip_bc or "NoSlipDelayBounceBack"))
    if pressure_bc:
        f.set_pressure_boundary_condition(getattr(F.PressureBC, pressure_bc))
    sim = Simulation(L, f)
    sim.run(steps)
    return L


def _errors(grid, gpu_fields, ref_fields, u_ref):
    fl = grid.geo == 19
    g_rho, g_ux, g_uy, g_uz = gpu_fields[:4]
    r_rho, r_ux, r_uy, r_uz = ref_fields[:4]
    du = np.sqrt((g_ux - r_ux) ** 2 + (g_uy - r_uy) ** 2 + (g_uz - r_uz) ** 2)[fl]
    return {
        "u_L2": float(np.sqrt((du ** 2).mean()) / u_ref),
        "u_max": float(du.max() / u_ref),
        "rho_L2": float(np.sqrt(((g_rho - r_rho)[fl] ** 2).mean())),
        "rho_max": float(np.abs(g_rho - r_rho)[fl].max()),
    }


@pytest.mark.parametrize("velocity_bc,noslip_bc", [
    ("VelocityBounceBack", "NoSlipBounceBack"),
    ("VelocityBounceBack", None),  # NoSlipDelayBounceBack: solid nodes reflect
    ("VelocityInterpolatedCompressible", "NoSlipInterpolatedCompressible"),
    ("VelocityWithPressureInterpolatedCompressible", "NoSlipInterpolatedCompressible"),
])
@pytest.mark.parametrize("box", [True, False])
@pytest.mark.parametrize("rows", [None, 0, 1])
def test_cavity_bcs_match_oracle(gpu, orc, velocity_bc, noslip_bc, box, rows):
    grid = cavity(24, lid_velocity=0.05)
    omega, steps = 1.8, 200
    ref = oracle_run(orc, grid, omega, steps, velocity_bc=velocity_bc, noslip_bc=noslip_bc)
    L = _gpu_run(grid, omega, steps, velocity_bc, noslip_bc, box=box, rows=rows)
    e = _errors(grid, L.macroscopic(), ref.macroscopic(grid.geo), 0.05)
    print(velocity_bc, noslip_bc, box, rows, e)
    assert e["u_max"] < 1e-4 and e["rho_max"] < 1e-6


@pytest.mark.parametrize("velocity_bc,noslip_bc", [
    ("VelocityBounceBack", "NoSlipBounceBack"),
    ("VelocityInterpolatedCompressible", "NoSlipInterpolatedCompressible"),
])
@pytest.mark.parametrize("rows", [None, 0, 1])
def test_channel_pressure_outlet_matches_oracle(gpu, orc, velocity_bc, noslip_bc, rows):
    grid = channel((32, 8, 8), inflow=0.02)
    omega, steps = 1.8, 300
    ref = oracle_run(orc, grid, omega, steps, velocity_bc=velocity_bc, noslip_bc=noslip_bc, pressure=True)
    L = _gpu_run(grid, omega, steps, velocity_bc, noslip_bc, pressure_bc="PressureNonEquilibriumCompressible",
                 rows=rows)
    e = _errors(grid, L.macroscopic(), ref.macroscopic(grid.geo), 0.02)
    print("channel", velocity_bc, rows, e)
    assert e["u_max"] < 1e-4 and e["rho_max"] < 1e-6


def test_lid_driven_cavity_64_config(gpu, orc):
    """The 64^3 lid-driven cavity: lid 0.05 in x on z-max, bounce-back no-slip on the other five
    faces, Re = 100 (nu = 0.05*64/100 = 0.032, omega = 1/(3 nu + 1/2) = 1.678), limiter 0.01,
    rest initial state, 1000 steps; GPU FP32 against the FP64 oracle.

    At this omega the reference's third-order rate OxyyMxzz is negative (pole at 14/9, zero at
    7/4): the limiter holds the growing modes, and the run amplifies round-off. The bar is
    therefore the reference arithmetic's own FP32 deviation: the oracle's FP32 build on the
    same input, against the same FP64 run."""
    n, steps = 64, 1000
    grid = cavity(n, lid_velocity=0.05)
    omega = 1.0 / (3 * 0.05 * n / 100.0 + 0.5)
    kw = dict(velocity_bc="VelocityBounceBack", noslip_bc="NoSlipBounceBack")
    ref64 = oracle_run(orc, grid, omega, steps, **kw).macroscopic(grid.geo)
    ref32 = oracle_run(orc, grid, omega, steps, dtype=np.float32, **kw).macroscopic(grid.geo)
    L = _gpu_run(grid, omega, steps, "VelocityBounceBack", "NoSlipBounceBack")
    e_gpu = _errors(grid, L.macroscopic(), ref64, 0.05)
    e_cpu32 = _errors(grid, [a.astype(np.float64) for a in ref32], ref64, 0.05)
    print("cavity64 gpu-fp32 vs fp64", e_gpu)
    print("cavity64 cpu-fp32 vs fp64", e_cpu32)
    for k in ("u_L2", "u_max", "rho_L2", "rho_max"):
        assert e_gpu[k] < 3.0 * e_cpu32[k] + 1e-6, (k, e_gpu[k], e_cpu32[k])
    assert e_gpu["u_L2"] < 1e-2


def test_lid_driven_cavity_64_well_posed_rates(gpu, orc):
    """The same cavity at omega = 1.8, where every K17 rate lies in (0, 2): GPU FP32 tracks
    the FP64 oracle to FP32 round-off over 1000 steps."""
    n, steps = 64, 1000
    grid = cavity(n, lid_velocity=0.05)
    kw = dict(velocity_bc="VelocityBounceBack", noslip_bc="NoSlipBounceBack")
    ref64 = oracle_run(orc, grid, 1.8, steps, **kw).macroscopic(grid.geo)
    L = _gpu_run(grid, 1.8, steps, "VelocityBounceBack", "NoSlipBounceBack")
    e = _errors(grid, L.macroscopic(), ref64, 0.05)
    print("cavity64 omega=1.8", e)
    assert e["u_L2"] < 2e-5 and e["u_max"] < 2e-4 and e["rho_max"] < 2e-6


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("row_kernel", [False, True])
def test_fused_bounce_back_matches_separate_launches(gpu, storage, graphs, row_kernel):
    """Simulation on a walled box with VelocityBounceBack and NoSlipBounceBack: the collision that
    applies the bounce-back lists in its own launch (vf_k17_run_box_bounce_back) leaves every
    population bit-identical to the collision followed by the two BC launches; pinned to the row
    kernel, the entry point takes the separate launches itself."""
    from virtualfluids_amd import capi
    from virtualfluids_amd.grid import walled_box, d0M0, d0P0, d00M, d00P, dM00, dP00
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory as F, Simulation

    walls = {s: ("noslip",) for s in (dP00, dM00, d0P0, d0M0, d00M)}
    walls[d00P] = ("velocity", (0.035, 0.02, 0.0))
    grid = walled_box((30, 20, 12), walls)
    rng = np.random.default_rng(17)
    N = grid.N
    init = (rng.normal(0, 0.01, N), rng.normal(0, 0.02, N), rng.normal(0, 0.02, N), rng.normal(0, 0.02, N))
    f = F()
    f.set_velocity_boundary_condition(F.VelocityBC.VelocityBounceBack)
    f.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipBounceBack)
    out = []
    for fuse in (True, False):
        L = Lattice(grid, 1.7, storage=storage)
        if row_kernel:
            L.param.kernelChoice.boxKernel = capi.BOX_KERNEL_ROW
        L.init_equilibrium(*init)
        sim = Simulation(L, f, fuse_bounce_back=fuse)
        assert (sim.fused is not None) == fuse
        print(f"{storage} graphs={graphs} row_kernel={row_kernel} fused={fuse}", flush=True)
        sim.run(37, graphs=graphs)
        out.append(L.populations())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_joined_bc_launches_match_separate_launches(gpu, storage):
    """A sphere channel whose wall and sphere lists share the no-slip launcher: the step that
    launches them as one list (post_launch_groups) against the two launches, through Simulation
    with HIP graphs, the Stepper's out-of-place step and its adjoint step: every population, the
    adjoint buffer and the inflow-velocity adjoint bit-identical."""
    import torch

    from virtualfluids_amd.adjoint import Stepper
    from virtualfluids_amd.grid import sphere_channel
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory as F, Simulation

    grid = sphere_channel((30, 16, 16), (11, 8, 8), 3.6, 0.04)
    fac = F()
    fac.set_velocity_boundary_condition(F.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_geometry_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(F.PressureBC.PressureNonEquilibriumCompressible)
    rng = np.random.default_rng(29)
    N = grid.N
    init = (rng.normal(0, 0.01, N), 0.04 + rng.normal(0, 0.01, N), rng.normal(0, 0.01, N), rng.normal(0, 0.01, N))
    lam0 = rng.normal(0, 1, 27 * N).astype(np.float32)
    out = []
    for joined in (True, False):
        L = Lattice(grid, 1.8, storage=storage)
        L.init_equilibrium(*init)
        sim = Simulation(L, fac)
        st = Stepper(L, fac)
        if not joined:
            separate = [(name, fn, L.bcs[name]) for name, fn in
                        (("velocity", sim.velocity_fn), ("noslip", sim.noslip_fn), ("geometry", sim.geometry_fn))]
            sim.groups = separate
            st._launch_groups = [(name, fn, bc) for name, fn, bc in separate]
        sim.run(41, graphs=True)
        assert [name for name, _, _ in sim.groups] == (["velocity", "noslip+geometry"] if joined else
                                                      ["velocity", "noslip", "geometry"])
        posts = {name: torch.zeros(27 * n, device="cuda") for name, n in st.post_lists()}
        before = L.f.clone()
        after = st.forward(before, True, torch.empty_like(before), None, save=posts)
        lam = torch.from_numpy(lam0).cuda()
        ubar = torch.zeros(3 * grid.bcs["velocity"].n, device="cuda")
        rbar = torch.zeros(grid.bcs["pressure"].n, device="cuda")
        st.adjoint(before, after, lam, True, ubar, rbar, None, None, saved=posts)
        torch.cuda.synchronize()
        print(f"{storage} joined={joined}", flush=True)
        out.append([L.populations(), after.cpu().numpy(), lam.cpu().numpy(), ubar.cpu().numpy()])
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
