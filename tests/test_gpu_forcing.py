"""GPU parity of the forced collision templates against the FP64 oracle (through the C ABI).

* a global force (vf_lbm_parameter.forces, scaled by 1/2 forceFactor) on the Default template,
  dense-box and indexed kernels, levels 0-2;
* the per-node body force of ApplyBodyForce, AllFeatures and SubDomainBorder
  (K17CompressibleNavierStokes.cu:127-137, RunCollision.cuh:87-99): the shift is
  (forces + bodyForce[k]) * 1/2 * forceFactor, the body force is reset to zero, and the
  AllFeatures / SubDomainBorder templates also write the macroscopic values.
Bar: 2e-7 absolute on the populations (FP32 arithmetic against FP64 on |f| <~ 0.06), the
macroscopic values 1e-6, the body-force arrays exactly zero afterwards.
"""
import numpy as np
import pytest

from helpers import near_equilibrium, pack_pre
from virtualfluids_amd import capi
from virtualfluids_amd.grid import periodic_box

pytestmark = pytest.mark.gpu

FORCES = (3e-5, -2e-5, 1.5e-5)


@pytest.mark.parametrize("box", [True, False])
@pytest.mark.parametrize("level", [0, 2])
@pytest.mark.parametrize("even", [True, False])
def test_global_force_matches_oracle(gpu, orc, box, level, even):
    from virtualfluids_amd.lattice import Lattice

    grid = periodic_box(40, 6, 5)
    rng = np.random.default_rng(300 + level + 2 * even)
    buf = pack_pre(near_equilibrium(rng, grid.N, u_amp=0.04, noise=5e-3), grid, even)
    omega = 1.7
    ff = 1.0 / 2 ** level
    ref = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    ref.even = even
    ref.f[:] = buf
    ref.collide(grid.fluid, omega, force=tuple(np.float32(F) * np.float32(0.5) * np.float32(ff) for F in FORCES))
    L = Lattice(grid, omega, box=box, forces=FORCES, level=level)
    L.even = even
    L.set_populations(buf)
    L.collide()
    print(f"global force box={box} level={level} even={even}", flush=True)
    err = np.abs(L.populations() - ref.f).max()
    assert err < 2e-7, err


@pytest.mark.parametrize("template", ["ApplyBodyForce", "AllFeatures", "SubDomainBorder"])
@pytest.mark.parametrize("level", [0, 1])
@pytest.mark.parametrize("even", [True, False])
def test_body_force_templates_match_oracle(gpu, orc, template, level, even):
    import torch

    from virtualfluids_amd.lattice import Lattice

    code = {"ApplyBodyForce": capi.COLLISION_APPLY_BODY_FORCE, "AllFeatures": capi.COLLISION_ALL_FEATURES,
            "SubDomainBorder": capi.COLLISION_SUBDOMAIN_BORDER}[template]
    macro = template != "ApplyBodyForce"
    grid = periodic_box(24, 7, 6)
    rng = np.random.default_rng(400 + level + 2 * even + 7 * code)
    buf = pack_pre(near_equilibrium(rng, grid.N, u_amp=0.04, noise=5e-3), grid, even)
    omega = 1.6
    ff = 1.0 / 2 ** level
    bf = [rng.normal(0.0, 5e-5, grid.N).astype(np.float32) for _ in range(3)]
    # the oracle in FP64 on the FP32 inputs the GPU reads
    bf64 = [b.astype(np.float64) for b in bf]
    frc64 = np.array(FORCES, dtype=np.float32).astype(np.float64)
    ref = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    ref.even = even
    ref.f[:] = buf
    mref = [np.zeros(grid.N) for _ in range(4)] if macro else None
    ref.collide_body_force(grid.fluid, omega, frc64, ff, bf64, macro=mref)

    L = Lattice(grid, omega, box=False, forces=FORCES, level=level)
    L.even = even
    L.set_populations(buf)
    L.set_body_force(*bf)
    if macro:
        for t in (L.rho, L.vx, L.vy, L.vz):
            t.fill_(123.0)
    L.collide(template=code)
    torch.cuda.synchronize()
    print(f"{template} level={level} even={even}", flush=True)
    err = np.abs(L.populations() - ref.f).max()
    assert err < 2e-7, err
    for b in L.body_force:
        assert not bool(b.any()), "bodyForce must be reset to zero after the collision"
    if macro:
        for got, want in zip((L.rho, L.vx, L.vy, L.vz), mref):
            assert np.abs(got.cpu().numpy() - want).max() < 1e-6
    else:
        assert float(L.rho.abs().max()) == 0.0  # ApplyBodyForce writes no macroscopic values


def test_body_force_needs_arrays(gpu):
    from virtualfluids_amd.lattice import Lattice

    grid = periodic_box(8, 4, 4)
    L = Lattice(grid, 1.5, box=False)
    with pytest.raises(RuntimeError, match="bodyForce"):
        L.collide(template=capi.COLLISION_APPLY_BODY_FORCE)
    with pytest.raises(RuntimeError, match="Invalid CollisionTemplate"):
        L.collide(template=7)
