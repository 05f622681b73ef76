"""GPU initialisation kernels against the oracle: the non-equilibrium init on the dense-box
addressing and on neighbour arrays (with neighborInverse), every storage type, a walled box
(solid markers) and a smooth field. Tolerance: FP32 arithmetic on O(0.05) populations (2e-7),
plus one storage rounding for FP16/BF16."""
import numpy as np
import pytest

from virtualfluids_amd.grid import periodic_box, walled_box

pytestmark = pytest.mark.gpu

TOL = {"fp32": 2e-7, "fp16": 0.06 * 2.0 ** -11 + 2e-7, "bf16": 0.06 * 2.0 ** -8 + 2e-7}


def _field(g):
    x, y, z = g.coords()
    k = 2 * np.pi / np.array([g.nx, g.ny, g.nz])
    ux = 0.04 * np.sin(k[0] * x) * np.cos(k[1] * y)
    uy = -0.03 * np.cos(k[0] * x) * np.sin(k[2] * z)
    uz = 0.02 * np.sin(k[1] * y + k[2] * z)
    drho = 0.01 * np.cos(k[0] * x + k[1] * y)
    return drho, ux, uy, uz


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("addressing", ["arrays", "box"])
@pytest.mark.parametrize("kind", ["periodic", "walled"])
def test_init_neq_matches_oracle(gpu, orc, storage, addressing, kind):
    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import Lattice, _ptr, _u32, stream_handle

    g = periodic_box(20, 12, 10) if kind == "periodic" else walled_box((18, 10, 8), {})
    omega = 1.75
    fields = _field(g)
    # box lattices address by (x, y, z); indexed lattices pass the neighbour arrays
    L = Lattice(g, omega, storage=storage, box=(addressing == "box"))
    for dst, src in zip((L.rho, L.vx, L.vy, L.vz), fields):
        dst.copy_(__import__("torch").from_numpy(src.astype(np.float32)))
    nbinv = _u32(g.neighbor_inverse(), L.device)
    assert (L.param.neighborX is None) == (addressing == "box")
    print(f"init neq {storage} {addressing} {kind}", flush=True)
    capi.call("vf_InitNavierStokesCompressibleNonEquilibrium", L.pref(), _ptr(nbinv), stream_handle())
    got = L.populations()
    ref = orc.Lattice(g.nbx, g.nby, g.nbz)
    ref.init_neq(*(f.astype(np.float32).astype(np.float64) for f in fields), g.neighbor_inverse(), omega, geo=g.geo)
    err = np.abs(got - ref.f).max()
    print(f"  max err {err:.3e}")
    assert err < TOL[storage]


def test_init_neq_through_lattice_method(gpu, orc):
    from virtualfluids_amd.lattice import Lattice

    g = periodic_box(16, 8, 8)
    fields = _field(g)
    L = Lattice(g, 1.6)
    L.init_non_equilibrium(*fields)
    ref = orc.Lattice(g.nbx, g.nby, g.nbz)
    ref.init_neq(*(f.astype(np.float32).astype(np.float64) for f in fields), g.neighbor_inverse(), 1.6)
    assert np.abs(L.populations() - ref.f).max() < 2e-7
