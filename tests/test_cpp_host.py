"""The C++ host layer (include/virtualfluids_amd.hpp) through its test program.

CPU: factory and manager semantics, mirroring the reference's own unit tests of
BoundaryConditionFactory, BoundaryConditionKernelManager and GridScalingKernelManager.
GPU: cases run through UpdateGrid27 / KernelFactory / the BC and scaling factories in C++ give
bit-identical fields to the Python driver (same kernels, same order), which the other GPU tests
hold to the oracle.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from cpp_case import HOST_TEST, read_output, run_case, write_case


def _binary():
    if not os.path.exists(HOST_TEST):
        pytest.fail(f"{HOST_TEST} is missing: build with virtualfluids_amd/build.py")
    return HOST_TEST


def test_ctypes_struct_layout_matches_c():
    """Every field offset and struct size the ctypes binding assumes equals the C compiler's."""
    from virtualfluids_amd import capi

    r = subprocess.run([_binary(), "layout"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    structs = {"vf_lbm_parameter": capi.LBMParameter, "vf_q_boundary_conditions": capi.QBoundaryConditions,
               "vf_interpolation_cells": capi.InterpolationCells,
               "vf_interpolation_cell_neighbor": capi.InterpolationCellNeighbor}
    for line in r.stdout.split("\n"):
        if not line.strip():
            continue
        key, value = line.split()
        head, field = key.split(".")
        if head == "sizeof":
            assert ctypes.sizeof(structs[field]) == int(value), key
        else:
            assert getattr(structs[head], field).offset == int(value), key


def test_factories_and_managers():
    r = subprocess.run([_binary(), "factories"], capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "0 failure(s)" in r.stdout


@pytest.mark.gpu
def test_cavity_through_cpp_update_grid(gpu, tmp_path):
    from virtualfluids_amd.grid import cavity
    from virtualfluids_amd.lattice import Lattice, omega_from_viscosity
    from virtualfluids_amd.simulation import BoundaryConditionFactory, Simulation

    nu, steps = 0.02, 150
    g = cavity(20, lid_velocity=0.05)
    zeros = np.zeros(g.N)
    init = (zeros, zeros, zeros, zeros)
    write_case(str(tmp_path), [(g, init)], nu, velocity_bc="VelocityInterpolatedCompressible",
               noslip_bc="NoSlipInterpolatedCompressible")
    print(run_case(str(tmp_path), steps))
    got = read_output(str(tmp_path), 0)

    L = Lattice(g, omega_from_viscosity(nu))
    L.init_equilibrium(*init)
    f = BoundaryConditionFactory()
    f.set_velocity_boundary_condition(BoundaryConditionFactory.VelocityBC.VelocityInterpolatedCompressible)
    f.set_no_slip_boundary_condition(BoundaryConditionFactory.NoSlipBC.NoSlipInterpolatedCompressible)
    Simulation(L, f).run(steps)
    want = L.macroscopic()[:4]
    fluid = g.geo == 19
    for a, b in zip(got, want):
        assert np.array_equal(a[fluid], b[fluid])
    assert np.abs(want[1][fluid]).max() > 1e-3  # the lid has driven a flow


@pytest.mark.gpu
def test_channel_pressure_outlet_through_cpp(gpu, tmp_path):
    from virtualfluids_amd.grid import channel
    from virtualfluids_amd.lattice import Lattice, omega_from_viscosity
    from virtualfluids_amd.simulation import BoundaryConditionFactory, Simulation

    nu, steps = 0.03, 120
    g = channel((24, 10, 10), inflow=0.03)
    zeros = np.zeros(g.N)
    init = (zeros, zeros, zeros, zeros)
    write_case(str(tmp_path), [(g, init)], nu, velocity_bc="VelocityBounceBack", noslip_bc="NoSlipBounceBack",
               pressure_bc=True)
    run_case(str(tmp_path), steps)
    got = read_output(str(tmp_path), 0)
    L = Lattice(g, omega_from_viscosity(nu))
    L.init_equilibrium(*init)
    f = BoundaryConditionFactory()
    f.set_velocity_boundary_condition(BoundaryConditionFactory.VelocityBC.VelocityBounceBack)
    f.set_no_slip_boundary_condition(BoundaryConditionFactory.NoSlipBC.NoSlipBounceBack)
    f.set_pressure_boundary_condition(BoundaryConditionFactory.PressureBC.PressureNonEquilibriumCompressible)
    Simulation(L, f).run(steps)
    want = L.macroscopic()[:4]
    fluid = g.geo == 19
    for a, b in zip(got, want):
        assert np.array_equal(a[fluid], b[fluid])


@pytest.mark.gpu
def test_two_levels_through_cpp_update_grid(gpu, tmp_path):
    from helpers import shear_wave_init
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.refinement import GridScalingFactory, nested_box
    from virtualfluids_amd.simulation import MultiLevelSimulation

    nu, steps = 0.01, 40
    nb = nested_box((12, 12, 24), (2, 2, 6), (9, 9, 18))
    k, init_c, init_f = shear_wave_init(nb, 0.03)
    write_case(str(tmp_path), [(nb.coarse, init_c), (nb.fine, init_f)], nu, interfaces=[(nb.cf, nb.fc)])
    run_case(str(tmp_path), steps)
    got_c, got_f = read_output(str(tmp_path), 0), read_output(str(tmp_path), 1)

    C = Lattice(nb.coarse, 1.0 / (3 * nu + 0.5), level=0)
    F = Lattice(nb.fine, 1.0 / (6 * nu + 0.5), level=1)
    C.init_equilibrium(*init_c)
    F.init_equilibrium(*init_f)
    fac = GridScalingFactory()
    fac.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
    MultiLevelSimulation([C, F], [(nb.cf, nb.fc)], scaling_factory=fac).run(steps)
    valid = nb.coarse_valid()
    for a, b in zip(got_c, C.macroscopic()[:4]):
        assert np.array_equal(a[valid], b[valid])
    for a, b in zip(got_f, F.macroscopic()[:4]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_three_levels_through_cpp_update_grid(gpu, tmp_path):
    """Three levels (a channel with interpolated BCs and a pressure outlet, a box nested in it and
    one nested in that): the C++ UpdateGrid27's recursion gives the Python MultiLevelSimulation's
    fields bit for bit."""
    import multilevel_reference as ref
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.refinement import GridScalingFactory
    from virtualfluids_amd.simulation import BoundaryConditionFactory as F, MultiLevelSimulation

    nu, steps = ref.NU, 8
    grids, interfaces, _, _ = ref.three_levels()
    inits = [(np.zeros(g.N), np.full(g.N, 0.03), np.zeros(g.N), np.zeros(g.N)) for g in grids]
    write_case(str(tmp_path), list(zip(grids, inits)), nu, velocity_bc="VelocityInterpolatedCompressible",
               noslip_bc="NoSlipInterpolatedCompressible", pressure_bc=True, interfaces=interfaces)
    run_case(str(tmp_path), steps)
    lats = []
    for lev, (g, init) in enumerate(zip(grids, inits)):
        L = Lattice(g, ref.omega(lev), level=lev)
        L.init_equilibrium(*init)
        lats.append(L)
    fac = F()
    fac.set_velocity_boundary_condition(F.VelocityBC.VelocityInterpolatedCompressible)
    fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
    fac.set_pressure_boundary_condition(F.PressureBC.PressureNonEquilibriumCompressible)
    scaling = GridScalingFactory()
    scaling.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
    MultiLevelSimulation(lats, interfaces, [fac, None, None], scaling_factory=scaling).run(steps)
    for lev, (g, L) in enumerate(zip(grids, lats)):
        fluid = g.geo == 19
        for a, b in zip(read_output(str(tmp_path), lev), L.macroscopic()[:4]):
            assert np.array_equal(a[fluid], b[fluid]), lev


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_checkpoint_and_restart_through_cpp(gpu, tmp_path, storage):
    """Two levels: 20 coarse steps with a binary checkpoint after 10, against a second run that
    starts from that checkpoint and makes the last 10 steps: identical fields. The Python
    restart reader, loading the C++ file into the Python multi-level driver, continues to the
    same fields as well."""
    import shutil

    from helpers import shear_wave_init
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.refinement import GridScalingFactory, nested_box
    from virtualfluids_amd.restart import BinaryRestartObject
    from virtualfluids_amd.simulation import MultiLevelSimulation

    nu = 0.01
    nb = nested_box((12, 12, 24), (2, 2, 6), (9, 9, 18))
    k, init_c, init_f = shear_wave_init(nb, 0.03)
    d = str(tmp_path / "case")
    write_case(d, [(nb.coarse, init_c), (nb.fine, init_f)], nu, storage=storage, interfaces=[(nb.cf, nb.fc)])
    ckpt = str(tmp_path / "shear_Restart_0_10")
    run_case(d, 20, restart=("checkpoint", 10, ckpt, "bin"))
    full = [read_output(d, 0), read_output(d, 1)]
    assert os.path.exists(ckpt + ".bin")
    shutil.rmtree(d)
    write_case(d, [(nb.coarse, init_c), (nb.fine, init_f)], nu, storage=storage, interfaces=[(nb.cf, nb.fc)])
    run_case(d, 10, restart=("restart", ckpt, "bin"))
    resumed = [read_output(d, 0), read_output(d, 1)]
    valid = nb.coarse_valid()
    for a, b in zip(full[0], resumed[0]):
        assert np.array_equal(a[valid], b[valid])
    for a, b in zip(full[1], resumed[1]):
        assert np.array_equal(a, b)

    C = Lattice(nb.coarse, 1.0 / (3 * nu + 0.5), storage=storage, level=0)
    F = Lattice(nb.fine, 1.0 / (6 * nu + 0.5), storage=storage, level=1)
    BinaryRestartObject().deserialize(ckpt, [C, F])
    fac = GridScalingFactory()
    fac.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
    MultiLevelSimulation([C, F], [(nb.cf, nb.fc)], scaling_factory=fac).run(10)
    for a, b in zip(full[0], C.macroscopic()[:4]):
        assert np.array_equal(a[valid], b[valid])
    for a, b in zip(full[1], F.macroscopic()[:4]):
        assert np.array_equal(a, b)
