"""GPU parity of the K17 collide+stream kernels against the FP64 CPU oracle.

All calls go through the C ABI (libvfamd.so). Tolerances (stated per test) are set by
FP32 arithmetic against the FP64 oracle and, for FP16/BF16, by the storage rounding of
the shifted populations (unit round-off 2^-11 and 2^-8 relative to |f_i|).
"""
import numpy as np
import pytest

from helpers import near_equilibrium, pack_pre, unpack_pre
from virtualfluids_amd.grid import CX, CY, CZ, periodic_box

pytestmark = pytest.mark.gpu


def _lattice(grid, omega, storage="fp32", box=True, limiters=(0.01, 0.01, 0.01)):
    from virtualfluids_amd.lattice import Lattice

    return Lattice(grid, omega, storage=storage, box=box, limiters=limiters)


def _choose(L, **kw):
    """Pin a kernel form through vf_lbm_parameter.kernelChoice (capi.KernelChoice fields)."""
    for k, v in kw.items():
        setattr(L.param.kernelChoice, k, v)
    return L


def _oracle_lattice(orc, grid):
    return orc.Lattice(grid.nbx, grid.nby, grid.nbz)


@pytest.mark.parametrize("dims", [(24, 20, 16), (80, 6, 5), (64, 8, 4), (256, 3, 2)])
@pytest.mark.parametrize("box", [True, False])
@pytest.mark.parametrize("even", [True, False])
def test_one_step_matches_oracle(gpu, orc, dims, box, even):
    grid = periodic_box(*dims)
    rng = np.random.default_rng(100 + sum(dims))
    f0 = near_equilibrium(rng, grid.N, u_amp=0.04, noise=5e-3)
    omega = 1.67
    buf0 = pack_pre(f0, grid, even)

    ref = _oracle_lattice(orc, grid)
    ref.even = even
    ref.f[:] = buf0
    ref.collide(grid.fluid, omega)

    L = _lattice(grid, omega, box=box)
    L.even = even
    L.set_populations(buf0)
    L.collide()
    got = L.populations()
    print(f"dims={dims} box={box} even={even}")
    err = np.abs(got - ref.f).max()
    # FP32 arithmetic on O(0.05) shifted populations: a few ulp of the inputs
    assert err < 2e-7, err


# row kernel widths: 0 = scalar kernel; 1/2/4 nodes per lane (rows of 1-4 waves, the last one
# possibly partial: 40, 130 and 260 nodes); the row
# kernel addresses the direction arrays through buffer descriptors up to 4 GiB per 14 slots
# and through 64-bit global addresses beyond (pinned here through kernelChoice.rowAddressing)
@pytest.mark.parametrize("addressing", ["buffer", "global"])
@pytest.mark.parametrize("width", [0, 1, 2, 4])
@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("dims", [(64, 5, 3), (128, 4, 3), (256, 3, 2), (512, 2, 2), (40, 5, 3), (130, 4, 3),
                                  (260, 3, 2)])
def test_row_kernel_widths_match_oracle(gpu, orc, width, storage, dims, addressing):
    """Every row-kernel width, storage type, addressing form and row length (1 to 4 waves per
    row, with the periodic wrap) against the FP64 oracle applied to the same stored (rounded)
    input."""
    from virtualfluids_amd import capi

    if width == 0 and addressing == "global":
        pytest.skip("the scalar kernel has one addressing form")
    choice = dict(rowAddressing=capi.ROW_ADDRESSING_BUFFER if addressing == "buffer" else capi.ROW_ADDRESSING_GLOBAL)
    if width == 0:
        choice["boxKernel"] = capi.BOX_KERNEL_NODE
    else:
        choice.update(boxKernel=capi.BOX_KERNEL_ROW, rowWidth=width)
    grid = periodic_box(*dims)
    rng = np.random.default_rng(200 + width + sum(dims))
    f0 = near_equilibrium(rng, grid.N, u_amp=0.04, noise=5e-3)
    for even in (True, False):
        L = _choose(_lattice(grid, 1.9, storage=storage), **choice)
        L.even = even
        L.set_populations(pack_pre(f0, grid, even))
        stored = L.populations().astype(np.float64)  # input after storage rounding
        ref = _oracle_lattice(orc, grid)
        ref.even = even
        ref.f[:] = stored
        ref.collide(grid.fluid, 1.9)
        L.collide()
        got = L.populations()
        # FP32: a few ulp; FP16/BF16: one storage rounding of the output (|f| < 0.06)
        tol = {"fp32": 2e-7, "fp16": 0.06 * 2.0 ** -11 + 2e-7, "bf16": 0.06 * 2.0 ** -8 + 2e-7}[storage]
        err = np.abs(got - ref.f).max()
        print(f"width={width} {storage} {addressing} dims={dims} even={even} err={err:.3e}")
        assert err < tol, err


def test_box_and_indexed_kernels_identical(gpu):
    grid = periodic_box(40, 12, 10)
    rng = np.random.default_rng(5)
    buf = pack_pre(near_equilibrium(rng, grid.N, noise=1e-2), grid, True)
    A = _lattice(grid, 1.8, box=True)
    B = _lattice(grid, 1.8, box=False)
    A.set_populations(buf)
    B.set_populations(buf)
    for _ in range(5):
        A.collide()
        A.swap()
        B.collide()
        B.swap()
    np.testing.assert_array_equal(A.populations(), B.populations())


def test_write_macro_vars(gpu, orc):
    grid = periodic_box(16, 8, 8)
    rng = np.random.default_rng(6)
    f0 = near_equilibrium(rng, grid.N, u_amp=0.05, noise=5e-3)
    L = _lattice(grid, 1.5)
    L.set_populations(pack_pre(f0, grid, True))
    L.collide(write_macro=True)
    import torch

    torch.cuda.synchronize()
    _, macro = orc.k17_nodes(f0, 1.5)
    for col, t in enumerate((L.rho, L.vx, L.vy, L.vz)):
        np.testing.assert_allclose(t.cpu().numpy(), macro[:, col], rtol=0, atol=1e-7)


def _tgv(grid, u0):
    x, y, z = grid.coords()
    k = 2 * np.pi / grid.nx
    ux = u0 * np.sin(k * x) * np.cos(k * y) * np.cos(k * z)
    uy = -u0 * np.cos(k * x) * np.sin(k * y) * np.cos(k * z)
    uz = np.zeros_like(ux)
    drho = 3.0 * (u0 ** 2 / 16.0) * (np.cos(2 * k * x) + np.cos(2 * k * y)) * (np.cos(2 * k * z) + 2.0)
    return drho, ux, uy, uz


@pytest.mark.parametrize("storage,tol_u", [("fp32", 2e-5), ("fp16", 5e-3), ("bf16", 4e-2)])
def test_taylor_green_100_steps(gpu, orc, storage, tol_u):
    """32^3 Taylor-Green vortex, 100 steps: relative L2 error of u against the FP64 oracle."""
    n, u0, steps = 32, 0.05, 100
    grid = periodic_box(n, n, n)
    nu = u0 * (n / (2 * np.pi)) / 1600.0
    omega = 1.0 / (3 * nu + 0.5)
    drho, ux, uy, uz = _tgv(grid, u0)

    ref = _oracle_lattice(orc, grid)
    ref.init_eq(drho, ux, uy, uz)
    for _ in range(steps):
        orc.collide_box(ref.f, n, n, n, omega, ref.even, limiters=(1e6, 1e6, 1e6))
        ref.swap()
    r_rho, r_ux, r_uy, r_uz, _ = ref.macroscopic()

    L = _lattice(grid, omega, storage=storage, limiters=(1e6, 1e6, 1e6))
    L.init_equilibrium(drho, ux, uy, uz)
    for _ in range(steps):
        L.collide()
        L.swap()
    g_rho, g_ux, g_uy, g_uz, _ = L.macroscopic()
    du = np.sqrt(((g_ux - r_ux) ** 2 + (g_uy - r_uy) ** 2 + (g_uz - r_uz) ** 2).mean())
    unorm = np.sqrt((r_ux ** 2 + r_uy ** 2 + r_uz ** 2).mean())
    rel_l2 = du / unorm
    max_rho = np.abs(g_rho - r_rho).max()
    print(f"{storage}: rel L2(u) = {rel_l2:.3e}, max|drho err| = {max_rho:.3e}")
    assert rel_l2 < tol_u
    assert max_rho < 50 * tol_u * u0 ** 2


def test_mass_and_momentum_sums_fp64(gpu, orc):
    grid = periodic_box(32, 16, 8)
    rng = np.random.default_rng(8)
    f0 = near_equilibrium(rng, grid.N, u_amp=0.03, noise=1e-3)
    L = _lattice(grid, 1.6)
    L.set_populations(pack_pre(f0, grid, True))
    f32 = unpack_pre(L.populations().astype(np.float64), grid, True)
    s = L.moment_sums()
    drho = f32.sum(1)
    jx, jy, jz = f32 @ CX, f32 @ CY, f32 @ CZ
    np.testing.assert_allclose(s[:4], [drho.sum(), jx.sum(), jy.sum(), jz.sum()], rtol=1e-12, atol=1e-12)
    ke = 0.5 * (jx ** 2 + jy ** 2 + jz ** 2) / (1 + drho)
    assert s[4] == pytest.approx(ke.sum(), rel=1e-12)
    # conservation over steps: FP32 arithmetic keeps the total mass to round-off
    m0 = s[0]
    for _ in range(20):
        L.collide()
        L.swap()
    assert abs(L.moment_sums()[0] - m0) < 1e-6 * grid.N * 1e-2


@pytest.mark.parametrize("addressing", ["buffer", "global"])
@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("dims,solid", [((128, 4, 3), False), ((128, 4, 3), True), ((40, 6, 5), True)])
def test_out_of_place_step_matches_in_place(gpu, storage, dims, solid, addressing):
    """vf_k17_run_box_into (row kernel for rows of whole waves, box kernel otherwise; with and
    without solid nodes) writes exactly the state the in-place step leaves, and leaves its
    input untouched."""
    import ctypes

    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import stream_handle

    # a lattice this small would take the scalar kernel (cache-resident working set): pin the row kernel
    choice = dict(boxKernel=capi.BOX_KERNEL_ROW,
                  rowAddressing=capi.ROW_ADDRESSING_BUFFER if addressing == "buffer" else capi.ROW_ADDRESSING_GLOBAL)
    grid = periodic_box(*dims)
    rng = np.random.default_rng(300 + sum(dims) + solid)
    if solid:
        grid.geo[rng.random(grid.N) < 0.15] = 2
        grid.fluid = np.nonzero(grid.geo == 19)[0].astype(np.uint32)
    f0 = near_equilibrium(rng, grid.N, u_amp=0.04, noise=5e-3)
    for even in (True, False):
        L = _choose(_lattice(grid, 1.8, storage=storage), **choice)
        assert (L.mask is not None) == solid
        L.even = even
        L.set_populations(pack_pre(f0, grid, even))
        before = L.f.clone()
        out = torch.full_like(L.f, float("nan"))
        capi.call("vf_k17_run_box_into", L.pref(), ctypes.c_void_p(out.data_ptr()), stream_handle())
        torch.cuda.synchronize()
        assert torch.equal(L.f, before)
        L.collide()
        torch.cuda.synchronize()
        assert torch.equal(out, L.f), f"{storage} {dims} solid={solid} even={even}"


# rows whose last wave carries a few lanes (130 at one node per lane; 258 and 260 at two; 194),
# with solid walls, a porosity field, in place and out of place
@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("case", ["periodic130", "periodic260", "walled258", "walled194_porous"])
def test_partial_rows_match_oracle(gpu, orc, storage, case):
    import ctypes

    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.adjoint import _param
    from virtualfluids_amd.grid import walled_box
    from virtualfluids_amd.lattice import stream_handle

    grid = {"periodic130": lambda: periodic_box(130, 4, 3), "periodic260": lambda: periodic_box(260, 3, 2),
            "walled258": lambda: walled_box((256, 3, 2), {}),
            "walled194_porous": lambda: walled_box((192, 3, 3), {})}[case]()
    rng = np.random.default_rng(230 + len(case))
    f0 = near_equilibrium(rng, grid.N, u_amp=0.04, noise=5e-3)
    phi = rng.uniform(0, 0.1, grid.N) if "porous" in case else None
    for even in (True, False):
        # a lattice this small would take the scalar kernel (cache-resident working set): pin the row kernel
        L = _choose(_lattice(grid, 1.9, storage=storage), boxKernel=capi.BOX_KERNEL_ROW)
        if phi is not None:
            L.set_porosity(torch.from_numpy(phi.astype(np.float32)).cuda())
        L.even = even
        L.set_populations(pack_pre(f0, grid, even))
        stored = L.populations().astype(np.float64)
        ref = _oracle_lattice(orc, grid)
        ref.even = even
        ref.f[:] = stored
        ref.collide(grid.fluid, 1.9, porosity=None if phi is None else phi.astype(np.float32).astype(np.float64))
        out = torch.empty_like(L.f)
        capi.call("vf_k17_run_box_into", ctypes.byref(_param(L, L.f, even)), ctypes.c_void_p(out.data_ptr()),
                  stream_handle())
        L.collide()
        torch.cuda.synchronize()
        got = L.populations()
        tol = {"fp32": 2e-7, "fp16": 0.06 * 2.0 ** -11 + 2e-7, "bf16": 0.06 * 2.0 ** -8 + 2e-7}[storage]
        err = np.abs(got - ref.f).max()
        print(f"{case} {storage} even={even} nx={grid.nx} err={err:.3e}")
        assert err < tol, err
        assert torch.equal(out, L.f), "out-of-place step differs from the in-place step"
