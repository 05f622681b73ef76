"""GPU parity of the grid-refinement interface kernels and the two-level time step.

The interface kernels (vf_ScaleCoarseToFineCompressible, vf_ScaleFineToCoarseCompressible)
are called through the C ABI on a nested box and compared with the FP64 oracle applied to
the same stored input. Tolerances: FP32 arithmetic on O(0.05) shifted populations rebuilt
from 8-node sums (the oracle's own FP32 build stays within 1e-8 of FP64 here; 5e-7 leaves room
for a different summation order), plus one storage rounding of the output for FP16
(0.06 * 2^-11) and BF16 (0.06 * 2^-8).
"""
import numpy as np
import pytest

from helpers import near_equilibrium, oracle_multilevel, pack_pre, product_equilibrium, shear_wave_init
from virtualfluids_amd.refinement import nested_box

pytestmark = pytest.mark.gpu

NU = 0.01
OMEGA_C, OMEGA_F = 1.0 / (3 * NU + 0.5), 1.0 / (6 * NU + 0.5)
TOL = {"fp32": 5e-7, "fp16": 0.06 * 2.0 ** -11 + 5e-7, "bf16": 0.06 * 2.0 ** -8 + 5e-7}


def _levels(nb, storage="fp32", box=True):
    from virtualfluids_amd.lattice import Lattice

    # box lattices address by (x, y, z); indexed lattices pass the neighbour arrays
    return (Lattice(nb.coarse, OMEGA_C, storage=storage, level=0, box=box),
            Lattice(nb.fine, OMEGA_F, storage=storage, level=1, box=box))


def _oracle_levels(orc, nb):
    return (orc.Lattice(nb.coarse.nbx, nb.coarse.nby, nb.coarse.nbz),
            orc.Lattice(nb.fine.nbx, nb.fine.nby, nb.fine.nbz))


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("fine_to_coarse", [False, True])
@pytest.mark.parametrize("even_coarse", [True, False])
@pytest.mark.parametrize("variant", ["arrays", "offsets", "box_addressing"])
def test_scale_kernel_matches_oracle(gpu, orc, storage, fine_to_coarse, even_coarse, variant):
    import ctypes

    from virtualfluids_amd import capi
    from virtualfluids_amd.refinement import DeviceInterfaceCells, InterfaceCells

    nb = nested_box((12, 10, 14), (2, 1, 3), (9, 8, 10))
    rng = np.random.default_rng(300 + 4 * fine_to_coarse + 2 * even_coarse + len(variant))
    fc0 = near_equilibrium(rng, nb.coarse.N, u_amp=0.04, noise=5e-3)
    ff0 = near_equilibrium(rng, nb.fine.N, u_amp=0.04, noise=5e-3)
    cells = nb.fc if fine_to_coarse else nb.cf
    offs = [np.zeros(cells.n)] * 3
    if variant == "offsets":
        offs = [rng.uniform(-0.5, 0.5, cells.n) for _ in range(3)]
        cells = InterfaceCells(cells.coarse, cells.fine, tuple(offs))

    C, F = _levels(nb, storage, box=(variant == "box_addressing"))
    C.even = even_coarse
    C.set_populations(pack_pre(fc0, nb.coarse, even_coarse))
    F.set_populations(pack_pre(ff0, nb.fine, True))
    oc, of = _oracle_levels(orc, nb)
    oc.even = even_coarse
    oc.f[:] = C.populations()  # oracle input = the stored (rounded) values
    of.f[:] = F.populations()

    dc = DeviceInterfaceCells(cells, C.device)
    neigh = ctypes.byref(dc.neigh) if dc.offsets is not None else None
    for L in (C, F):  # box_addressing: no neighbour arrays reach the kernel
        assert (L.param.neighborX is None) == (variant == "box_addressing")
    fn = "vf_ScaleFineToCoarseCompressible" if fine_to_coarse else "vf_ScaleCoarseToFineCompressible"
    print(f"{fn} {storage} even_coarse={even_coarse} {variant} cells={cells.n}", flush=True)
    capi.call(fn, C.pref(), F.pref(), ctypes.byref(dc.struct), neigh, None)
    orc.scale(fine_to_coarse, oc, of, even_coarse, cells.coarse, cells.fine, offs, OMEGA_C, OMEGA_F)

    got_c, got_f = C.populations(), F.populations()
    err_c = np.abs(got_c - oc.f).max()
    err_f = np.abs(got_f - of.f).max()
    print(f"  max err coarse {err_c:.3e} fine {err_f:.3e}")
    assert err_c < TOL[storage] and err_f < TOL[storage]
    # the written level really changed (the kernel ran) and the other one is untouched
    changed = np.abs((got_c if fine_to_coarse else got_f) - pack_pre(fc0 if fine_to_coarse else ff0,
                                                                     nb.coarse if fine_to_coarse else nb.fine,
                                                                     even_coarse if fine_to_coarse else True))
    assert changed.max() > 1e-3


def test_scale_kernel_argument_errors(gpu):
    import ctypes

    from virtualfluids_amd import capi
    from virtualfluids_amd.refinement import DeviceInterfaceCells

    nb = nested_box((12, 12, 12), (2, 2, 2), (9, 9, 9))
    C, F = _levels(nb, "fp32")
    F16 = _levels(nb, "fp16")[1]
    dc = DeviceInterfaceCells(nb.cf, C.device)
    with pytest.raises(capi.VFError, match="same population storage"):
        capi.call("vf_ScaleCoarseToFineCompressible", C.pref(), F16.pref(), ctypes.byref(dc.struct), None, None)
    with pytest.raises(capi.VFError, match="interface cells"):
        capi.call("vf_ScaleFineToCoarseCompressible", C.pref(), F.pref(), None, None, None)


def test_two_level_steps_match_oracle(gpu, orc):
    """Five coarse steps (ten fine steps) of the whole two-level update against the oracle."""
    from virtualfluids_amd.refinement import GridScalingFactory
    from virtualfluids_amd.simulation import MultiLevelSimulation

    nb = nested_box((12, 12, 16), (2, 2, 3), (9, 9, 12))
    k, init_c, init_f = shear_wave_init(nb, 0.03)
    C, F = _levels(nb)
    C.init_equilibrium(*init_c)
    F.init_equilibrium(*init_f)
    fac = GridScalingFactory()
    fac.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
    sim = MultiLevelSimulation([C, F], [(nb.cf, nb.fc)], scaling_factory=fac)
    sim.run(5)
    oc, of = oracle_multilevel(orc, nb, OMEGA_C, OMEGA_F, 5, init_c, init_f)
    assert C.even == oc.even and F.even == of.even
    # solid coarse nodes (under the fine box) are left alone by both, so every slot compares
    err_c = np.abs(C.populations() - oc.f).max()
    err_f = np.abs(F.populations() - of.f).max()
    print(f"two-level 5 steps: err coarse {err_c:.3e} fine {err_f:.3e}")
    assert err_c < 5e-7 and err_f < 5e-7


def _oracle_multilevel_rounded(orc, nb, steps, buf_c, buf_f, rounding):
    """The two-level update of helpers.oracle_multilevel in FP64 arithmetic, with every level's
    buffer rounded to the 16-bit storage type after each kernel (`rounding`: numpy dtype), as
    the GPU stores it."""
    C, F = _oracle_levels(orc, nb)
    C.f[:] = buf_c
    F.f[:] = buf_f

    def rnd(L):
        L.f[:] = L.f.astype(rounding).astype(np.float64)

    zc = [np.zeros(nb.cf.n)] * 3
    zf = [np.zeros(nb.fc.n)] * 3
    for _ in range(steps):
        for _ in range(2):
            F.collide(nb.fine.fluid, OMEGA_F)
            F.swap()
            rnd(F)
        C.collide(nb.coarse.fluid, OMEGA_C)
        C.swap()
        rnd(C)
        orc.scale(True, C, F, C.even, nb.fc.coarse, nb.fc.fine, zf, OMEGA_C, OMEGA_F)
        rnd(C)
        orc.scale(False, C, F, C.even, nb.cf.coarse, nb.cf.fine, zc, OMEGA_C, OMEGA_F)
        rnd(F)
    return C, F


def test_two_level_steps_match_oracle_fp16(gpu, orc):
    """Five coarse steps of the two-level update with FP16 shifted storage on both levels (the
    storage of the two-level channel benchmark) against the oracle from the same stored initial
    state: (a) the oracle rounding each level to FP16 after every kernel, as the GPU stores it,
    agrees to 4 FP16 units in the last place of the largest population (0.06 * 2^-10: FP32
    against FP64 arithmetic flips a rounding now and then, and a flipped unit spreads through
    the later steps); (b) the oracle without storage rounding agrees to 3e-4 absolute (ten fine
    and five coarse roundings of at most half a unit, 0.06 * 2^-12, with margin) and to 2e-3 in
    the relative L2 norm of the velocity field."""
    from virtualfluids_amd.refinement import GridScalingFactory
    from virtualfluids_amd.simulation import MultiLevelSimulation

    nb = nested_box((12, 12, 16), (2, 2, 3), (9, 9, 12))
    k, init_c, init_f = shear_wave_init(nb, 0.03)
    C, F = _levels(nb, "fp16")
    C.init_equilibrium(*init_c)
    F.init_equilibrium(*init_f)
    buf_c, buf_f = C.populations().astype(np.float64), F.populations().astype(np.float64)
    fac = GridScalingFactory()
    fac.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
    MultiLevelSimulation([C, F], [(nb.cf, nb.fc)], scaling_factory=fac).run(5)
    got_c, got_f = C.populations(), F.populations()
    rc, rf = _oracle_multilevel_rounded(orc, nb, 5, buf_c, buf_f, np.float16)
    err_c, err_f = np.abs(got_c - rc.f).max(), np.abs(got_f - rf.f).max()
    print(f"two-level FP16, rounding oracle: err coarse {err_c:.3e} fine {err_f:.3e}")
    assert err_c < 0.06 * 2.0 ** -10 * 4 and err_f < 0.06 * 2.0 ** -10 * 4
    oc, of = _oracle_multilevel_rounded(orc, nb, 5, buf_c, buf_f, np.float64)
    err_c, err_f = np.abs(got_c - oc.f).max(), np.abs(got_f - of.f).max()
    print(f"two-level FP16, FP64 oracle: err coarse {err_c:.3e} fine {err_f:.3e}")
    assert err_c < 3e-4 and err_f < 3e-4
    valid = nb.coarse_valid()
    for L, R, sel in ((C, oc, valid), (F, of, slice(None))):
        g = L.macroscopic()
        r = R.macroscopic()
        du = np.sqrt(sum(((g[i] - r[i])[sel] ** 2).sum() for i in (1, 2, 3)))
        un = np.sqrt(sum((r[i][sel] ** 2).sum() for i in (1, 2, 3)))
        print(f"  rel L2(u) {du / un:.2e}")
        assert du / un < 2e-3


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
def test_uniform_flow_stays_uniform(gpu, storage):
    from virtualfluids_amd.refinement import GridScalingFactory
    from virtualfluids_amd.simulation import MultiLevelSimulation

    nb = nested_box((16, 16, 16), (3, 4, 5), (12, 12, 13))
    u = (0.04, -0.02, 0.03)
    feq = product_equilibrium(0.01, u)
    C, F = _levels(nb, storage)
    C.set_populations(pack_pre(np.repeat(feq, nb.coarse.N, axis=0), nb.coarse, True))
    F.set_populations(pack_pre(np.repeat(feq, nb.fine.N, axis=0), nb.fine, True))
    fac = GridScalingFactory()
    fac.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
    MultiLevelSimulation([C, F], [(nb.cf, nb.fc)], scaling_factory=fac).run(20)
    tol = {"fp32": 2e-6, "fp16": 2e-3, "bf16": 1e-2}[storage]
    valid = nb.coarse_valid()
    for L, sel in ((C, valid), (F, slice(None))):
        r, vx, vy, vz, _ = L.macroscopic()
        print(storage, "drho", np.abs(r[sel] - 0.01).max(), "u", np.abs(vx[sel] - u[0]).max())
        for got, want in ((r, 0.01), (vx, u[0]), (vy, u[1]), (vz, u[2])):
            assert np.abs(got[sel] - want).max() < tol


def test_shear_wave_through_fine_box(gpu):
    """u_x = A sin(k z) decays as A exp(-nu k^2 t) on both levels (FP32, 600 coarse steps)."""
    from virtualfluids_amd.refinement import GridScalingFactory
    from virtualfluids_amd.simulation import MultiLevelSimulation

    nb = nested_box((16, 16, 64), (3, 3, 20), (13, 13, 44))
    amp, steps = 0.02, 600
    k, init_c, init_f = shear_wave_init(nb, amp)
    C, F = _levels(nb)
    C.init_equilibrium(*init_c)
    F.init_equilibrium(*init_f)
    fac = GridScalingFactory()
    fac.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
    MultiLevelSimulation([C, F], [(nb.cf, nb.fc)], scaling_factory=fac).run(steps)
    decay = amp * np.exp(-NU * k * k * steps)
    _, _, zc = nb.coarse.coords()
    _, _, zf = nb.fine_coords()
    valid = nb.coarse_valid()
    _, vxc, _, _, _ = C.macroscopic()
    _, vxf, _, _, _ = F.macroscopic()
    err_c = np.abs(vxc[valid] - decay * np.sin(k * zc[valid])).max() / decay
    err_f = np.abs(vxf - decay * np.sin(k * zf)).max() / decay
    print(f"shear wave 600 steps: rel err coarse {err_c:.2e} fine {err_f:.2e}")
    # the unrefined 64-node wave is within 1e-3 of the analytic decay after 600 steps
    assert err_c < 5e-3 and err_f < 5e-3
