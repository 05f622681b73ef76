"""FP64 post-processing reductions on the GPU against the oracle, through the C ABI.

* vf_DragLiftPostD27 / vf_DragLiftPreD27, node by node, against the oracle's restatement of
  DragLift.cu evaluated on the GPU's own state (same stored values; sums of FP32 inputs in
  FP64, so they agree to double rounding);
* vf_DragLiftSum (one pass on the state after the step) against the node sum of the two halves;
* vf_KineticEnergy against the oracle's KineticEnergyAnalyzer restatement on the GPU's own
  macroscopic fields, on neighbour arrays and on the dense-box addressing.
"""
import ctypes

import numpy as np
import pytest

from helpers import near_equilibrium, pack_pre
from virtualfluids_amd.grid import sphere_channel, walled_box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("box", [True, False])
def test_drag_lift_halves_and_sum(gpu, orc, storage, box):
    from virtualfluids_amd.lattice import Lattice

    g = sphere_channel((20, 12, 12), (8, 5.5, 5.5), 3.0, 0.03)
    geo = g.bcs["geometry"]
    rng = np.random.default_rng(950 + box)
    omega = 1.8
    L = Lattice(g, omega, storage=storage, box=box)
    L.set_populations(pack_pre(near_equilibrium(rng, g.N, u_amp=0.03, noise=2e-3), g, True))
    L.collide()
    post_gpu = L.drag_lift_half("geometry", post=True)
    ref = orc.Lattice(g.nbx, g.nby, g.nbz)
    ref.even = L.even
    ref.f[:] = L.populations()
    post_orc = orc.drag_lift(ref, geo.k, geo.q27, post=True)
    L.apply_bc("vf_NoSlipInterpolatedCompressible", "geometry")
    L.swap()
    pre_gpu = L.drag_lift_half("geometry", post=False)
    ref.even = L.even
    ref.f[:] = L.populations()
    pre_orc = orc.drag_lift(ref, geo.k, geo.q27, post=False)
    scale = np.abs(post_orc).max()
    print(f"{storage} box={box}: post {np.abs(post_gpu - post_orc).max():.2e} pre {np.abs(pre_gpu - pre_orc).max():.2e}"
          f" of {scale:.2e}")
    np.testing.assert_allclose(post_gpu, post_orc, rtol=0, atol=1e-15 * scale)
    np.testing.assert_allclose(pre_gpu, pre_orc, rtol=0, atol=1e-15 * scale)
    F = L.drag_force("geometry")
    F_halves = (pre_gpu - post_gpu).sum(1)
    print(f"force {F} vs halves {F_halves}")
    np.testing.assert_allclose(F, F_halves, rtol=1e-13, atol=1e-15 * scale)
    # fixed summation order: the same value on every call
    np.testing.assert_array_equal(L.drag_force("geometry"), F)


def test_drag_sum_with_no_nodes_is_zero(gpu):
    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import Lattice, stream_handle

    g = walled_box((6, 5, 4), {})
    L = Lattice(g, 1.7)
    bc = capi.QBoundaryConditions()
    bc.numberOfBCnodes = 0
    out = torch.full((3,), 7.0, dtype=torch.float64, device="cuda")
    capi.call("vf_DragLiftSum", L.pref(), ctypes.byref(bc), ctypes.c_void_p(out.data_ptr()), stream_handle())
    torch.cuda.synchronize()
    assert out.cpu().numpy().tolist() == [0.0, 0.0, 0.0]


@pytest.mark.parametrize("kind", ["walled", "periodic_box_addressing"])
def test_kinetic_energy_matches_oracle(gpu, orc, kind):
    import torch

    from virtualfluids_amd.grid import periodic_box
    from virtualfluids_amd.lattice import Lattice

    g = walled_box((10, 7, 6), {}) if kind == "walled" else periodic_box(16, 6, 5)
    rng = np.random.default_rng(960)
    L = Lattice(g, 1.7, box=(kind != "walled"))  # walled: the neighbour-array addressing
    L.set_populations(pack_pre(near_equilibrium(rng, g.N, u_amp=0.04, noise=1e-3), g, True))
    assert (L.param.neighborX is None) == (kind != "walled")
    if kind != "walled":
        # dense-box addressing: no neighbour arrays reach the kernel
        g.geo[rng.random(g.N) < 0.1] = 2  # a few solid nodes exclude themselves and their neighbours
        L.geo.copy_(torch.from_numpy(g.geo.astype(np.uint32).view(np.int32)).cuda())
    e, count = L.kinetic_energy()
    drho, vx, vy, vz, _ = (t.cpu().numpy() for t in (L.rho, L.vx, L.vy, L.vz, L.press))
    ref = orc.kinetic_energy(g.nbx, g.nby, g.nbz, g.neighbor_inverse(), g.geo, vx, vy, vz, drho)
    print(f"{kind}: E {e:.15e} vs {ref[0]:.15e}, count {count} vs {ref[1]}")
    assert count == int(ref[1])
    assert abs(e - ref[0]) <= 1e-13 * abs(ref[0])
    if kind == "walled":
        assert count == (10 - 2) * (7 - 2) * (6 - 2)


@pytest.mark.parametrize("nbytes", [16, 4096 + 48, 64 << 20])
def test_stream_copy_copies_exactly(gpu, nbytes):
    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import stream_handle

    a = torch.randint(0, 255, (nbytes + 32,), dtype=torch.uint8, device="cuda")
    b = torch.zeros_like(a)
    capi.call("vf_stream_copy", ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
              ctypes.c_uint64(nbytes), stream_handle())
    torch.cuda.synchronize()
    assert torch.equal(a[:nbytes], b[:nbytes])
    assert int(b[nbytes:].abs().sum()) == 0  # nothing beyond the requested bytes


@pytest.mark.parametrize("stream_bytes", [1024, 3 * 4096, 4 << 20])
@pytest.mark.parametrize("wgs_per_cu", [0, 2])
def test_stream_inplace27_swaps_opposite_streams(gpu, stream_bytes, wgs_per_cu):
    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.grid import INV
    from virtualfluids_amd.lattice import stream_handle

    n = stream_bytes // 4
    a = torch.randn(27 * n + 64, dtype=torch.float32, device="cuda")
    ref = a.clone()
    capi.call("vf_stream_inplace27", ctypes.c_void_p(a.data_ptr()), ctypes.c_uint64(stream_bytes), wgs_per_cu,
              stream_handle())
    torch.cuda.synchronize()
    want = ref[:27 * n].view(27, n)[torch.as_tensor(INV, device="cuda")].reshape(-1)
    assert torch.equal(a[:27 * n], want)
    assert torch.equal(a[27 * n:], ref[27 * n:])  # nothing beyond the 27 streams
    capi.call("vf_stream_inplace27", ctypes.c_void_p(a.data_ptr()), ctypes.c_uint64(stream_bytes), wgs_per_cu,
              stream_handle())
    torch.cuda.synchronize()
    assert torch.equal(a, ref)
