"""The fused BC adjoint (vf_k17_adjoint_box_bc: the post-collision BC adjoints applied inside the
collision adjoint's launch) against the separate launches (vf_<BC>Adjoint per list, then
vf_k17_adjoint_box), both from the post-collision values the replay stored: the adjoint of the
populations, of the wall velocities and of the porosity agree to FP32 rounding order. Sphere
channels (interpolated velocity inflow, no-slip walls and sphere, pressure outlet) in the
one-node-per-thread and the row adjoint kernel (rows of one wave), and a cavity with bounce-back
walls."""
import numpy as np
import pytest

from helpers import near_equilibrium, pack_pre

pytestmark = pytest.mark.gpu

CASES = ["cavity_bb", "sphere_node", "sphere_row"]


def _setup(case, storage):
    from virtualfluids_amd.grid import cavity, gaussian_random_field, sphere_channel
    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.simulation import BoundaryConditionFactory

    F = BoundaryConditionFactory
    fac = F()
    if case == "cavity_bb":
        g = cavity(10, 0.05)
        fac.set_velocity_boundary_condition(F.VelocityBC.VelocityBounceBack)
        fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipBounceBack)
    else:
        # 62 fluid nodes + 2 walls = 64: one-wave rows, which the adjoint row kernel takes
        fx = 62 if case == "sphere_row" else 30
        g = sphere_channel((fx, 14, 12), (fx // 3, 7, 6), 3.5, 0.04)
        fac.set_velocity_boundary_condition(F.VelocityBC.VelocityInterpolatedCompressible)
        fac.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
        fac.set_geometry_boundary_condition(F.NoSlipBC.NoSlipInterpolatedCompressible)
        fac.set_pressure_boundary_condition(F.PressureBC.PressureNonEquilibriumCompressible)
    L = Lattice(g, 1.8, storage=storage)
    if case != "cavity_bb":
        L.set_porosity(gaussian_random_field((g.nx, g.ny, g.nz), 3.0, seed=3, low=0.0, high=0.05))
    rng = np.random.default_rng(11)
    L.set_populations(pack_pre(near_equilibrium(rng, g.N, u_amp=0.03, noise=3e-3), g, True))
    return L, fac


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("even", [True, False])
def test_fused_bc_adjoint_matches_separate_launches(gpu, case, storage, even):
    import torch

    from virtualfluids_amd.adjoint import Stepper

    L, fac = _setup(case, storage)
    fused, plain = Stepper(L, fac), Stepper(L, fac, fuse_adjoint=False)
    assert fused._fused_adjoint() is not None, "the fused BC adjoint must apply to this case"
    before = L.f.clone()
    posts = {name: torch.zeros(27 * n, dtype=torch.float32, device="cuda") for name, n in fused.post_lists()}
    after = fused.forward(before, even, torch.empty_like(before), None, save=posts)
    rng = np.random.default_rng(23 + even)
    lam0 = torch.from_numpy(rng.normal(0, 1, 27 * L.N).astype(np.float32)).cuda()
    out = {}
    for label, st in (("fused", fused), ("plain", plain)):
        lam = lam0.clone()
        vel = L.bcs.get("velocity")
        ubar = torch.zeros(3 * vel.n, dtype=torch.float32, device="cuda")
        padj = torch.zeros(L.N, dtype=torch.float32, device="cuda") if getattr(L, "porosity", None) is not None else None
        rbar = torch.zeros(L.bcs["pressure"].n, device="cuda") if L.bcs.get("pressure") is not None and \
            L.bcs["pressure"].n else None
        print(f"{case} {storage} even={even} {label}", flush=True)
        st.adjoint(before, after, lam, even, ubar, rbar, padj, None, saved=posts)
        torch.cuda.synchronize()
        out[label] = (lam.double(), ubar.double(), None if padj is None else padj.double())
    lf, lp = out["fused"][0], out["plain"][0]
    err = float((lf - lp).abs().max() / lp.abs().max())
    print(f"  adjoint rel err {err:.2e}")
    assert err < 2e-6
    uf, up = out["fused"][1], out["plain"][1]
    assert float((uf - up).abs().max()) <= 2e-6 * float(up.abs().max()) + 1e-12
    if out["plain"][2] is not None:
        pf, pp = out["fused"][2], out["plain"][2]
        assert float((pf - pp).abs().max()) <= 2e-6 * float(pp.abs().max()) + 1e-12
