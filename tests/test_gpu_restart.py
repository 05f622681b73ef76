"""Checkpoint and restart of a GPU run through the Python restart objects.

A cavity run of 20 steps against a run restarted from the checkpoint its first 10 steps wrote:
the binary file restarts bit for bit in every storage type (FP16/BF16 values are exact in
float), the ASCII file (6 significant digits, as the reference) to its rounding.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cavity(storage):
    from virtualfluids_amd.grid import cavity
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory, Simulation

    g = cavity(16, lid_velocity=0.05)
    L = Lattice(g, 1.8, storage=storage)
    zeros = np.zeros(g.N)
    L.init_equilibrium(zeros, zeros, zeros, zeros)
    f = BoundaryConditionFactory()
    f.set_velocity_boundary_condition(BoundaryConditionFactory.VelocityBC.VelocityInterpolatedCompressible)
    f.set_no_slip_boundary_condition(BoundaryConditionFactory.NoSlipBC.NoSlipInterpolatedCompressible)
    return L, Simulation(L, f)


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("fmt", ["bin", "txt"])
def test_restart_resumes_the_run(gpu, tmp_path, storage, fmt):
    from virtualfluids_amd.restart import ASCIIRestartObject, BinaryRestartObject, restart_file_name

    cls = BinaryRestartObject if fmt == "bin" else ASCIIRestartObject
    name = restart_file_name(str(tmp_path / "cavity"), 10)
    L, sim = _cavity(storage)
    sim.run(10)
    cls().serialize(name, [L])
    sim.run(10)
    want = L.populations()

    L2, sim2 = _cavity(storage)  # freshly initialised: even parity, as after 10 steps
    cls().deserialize(name, [L2])
    sim2.run(10)
    got = L2.populations()
    assert L2.even == L.even
    if fmt == "bin":
        np.testing.assert_array_equal(got, want)
    else:
        # 6 significant digits of |f - w| <= 0.1 leave < 1e-7 per value; ten steps keep that scale
        print(f"{storage} ascii restart: max diff {np.abs(got - want).max():.2e}")
        assert np.abs(got - want).max() < {"fp32": 1e-6, "fp16": 1e-4, "bf16": 1e-3}[storage]
    cls().delete_restart_file(name)
