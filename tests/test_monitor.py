"""The conservation check flags NaN injection and a stale parity (CPU, on the oracle's FP32 build).

The oracle steps a periodic box with a net flow in FP32 arithmetic; the FP64 sums of the state
(the quantities vf_MomentSums reduces on the GPU) pass the check over many steps, and fail it
after a NaN lands in one population or after one step taken with the stale parity.
"""
import numpy as np
import pytest

from helpers import conserved_sums, equilibrium, near_equilibrium, pack_pre
from virtualfluids_amd.grid import periodic_box, walled_box
from virtualfluids_amd.monitor import ConservationCheck, ConservationError


def _box(orc, seed=3):
    """A periodic box moving with (0.03, -0.02, 0.01) plus a zero-mean disturbance, FP32 oracle."""
    grid = periodic_box(12, 10, 8)
    rng = np.random.default_rng(seed)
    noise = near_equilibrium(rng, grid.N, u_amp=0.01, noise=2e-3)
    f0 = equilibrium(np.zeros(grid.N), np.tile([0.03, -0.02, 0.01], (grid.N, 1))) + (noise - noise.mean(0))
    L = orc.Lattice(grid.nbx, grid.nby, grid.nbz, dtype=np.float32)
    L.f[:] = pack_pre(f0, grid, True)
    return grid, L


def test_closed_lattice_passes_over_many_steps(orc):
    grid, L = _box(orc)
    chk = ConservationCheck(conserved_sums(L.f, grid, L.even), grid.fluid.size)
    for _ in range(200):
        L.collide(grid.fluid, 1.7)
        L.swap()
    p = chk.problems(conserved_sums(L.f, grid, L.even), 200)
    assert p == [], p


def test_nan_injection_is_flagged(orc):
    grid, L = _box(orc)
    chk = ConservationCheck(conserved_sums(L.f, grid, L.even), grid.fluid.size)
    for _ in range(5):
        L.collide(grid.fluid, 1.7)
        L.swap()
    L.f[13 * grid.N + 77] = np.nan
    for _ in range(3):  # the NaN spreads through the collision; the sums stay non-finite
        L.collide(grid.fluid, 1.7)
        L.swap()
    with pytest.raises(ConservationError, match="not finite"):
        chk.check(conserved_sums(L.f, grid, L.even), 8)


def test_stale_parity_is_flagged(orc):
    grid, L = _box(orc)
    chk = ConservationCheck(conserved_sums(L.f, grid, L.even), grid.fluid.size)
    for _ in range(4):
        L.collide(grid.fluid, 1.7)
        L.swap()
    L.collide(grid.fluid, 1.7)
    L.swap()
    L.swap()  # the parity flag misses one swap: the next step reads with the stale parity
    L.collide(grid.fluid, 1.7)
    L.swap()
    with pytest.raises(ConservationError, match="momentum"):
        chk.check(conserved_sums(L.f, grid, L.even), 6)


def test_bounce_back_box_passes(orc):
    """Mass and momentum of a resting closed box with bounce-back walls stay within the bounds."""
    from helpers import oracle_run
    from virtualfluids_amd.grid import d0M0, d0P0, d00M, d00P, dM00, dP00

    walls = {s: ("noslip",) for s in (dP00, dM00, d0P0, d0M0, d00P, d00M)}
    grid = walled_box((6, 5, 4), walls)
    L0 = oracle_run(orc, grid, 1.8, 0, dtype=np.float32)
    chk = ConservationCheck(conserved_sums(L0.f, grid, L0.even), grid.fluid.size)
    L = oracle_run(orc, grid, 1.8, 40, noslip_bc="NoSlipBounceBack", dtype=np.float32)
    chk.check(conserved_sums(L.f, grid, L.even), 40)
