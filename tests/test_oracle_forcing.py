"""Pins of the oracle's body-force templates (CPU only).

runCollision<..., applyBodyForce=true> (src/gpu/core/Utilities/RunCollision.cuh:87-99) shifts
the collision velocity by half a step of (forces + bodyForce[k]) * forceFactor and then clears
bodyForce[k]. The K17 collision (src/lbm/collision/K17CompressibleNavierStokes.h:164-166,
512-514) turns that shift into a momentum change of exactly rho_k (forces + bodyForce[k]) *
forceFactor per node and step, with the node's mass unchanged. A dropped 1/2, a force applied
without forceFactor or a body force taken from another node changes the momentum by O(F).
"""
import numpy as np
import pytest

from helpers import near_equilibrium, pack_pre, unpack_pre
from virtualfluids_amd.grid import CX, CY, CZ, periodic_box


def _unstream(pre_next, nx, ny, nz):
    """Post-collision populations per node from the next step's pre-collision ones (pull
    streaming pre'[x, d] = post[x - c_d, d], inverted)."""
    f = pre_next.reshape(nz, ny, nx, 27)
    out = np.empty_like(f)
    for d in range(27):
        out[..., d] = np.roll(f[..., d], shift=(-CZ[d], -CY[d], -CX[d]), axis=(0, 1, 2))
    return out.reshape(-1, 27)


@pytest.mark.parametrize("even", [True, False])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_body_force_momentum_change_is_exact(orc, even, level):
    grid = periodic_box(9, 7, 5)
    rng = np.random.default_rng(40 + level + 3 * even)
    f0 = near_equilibrium(rng, grid.N, u_amp=0.03, noise=4e-3)
    forces = np.array([2e-5, -3e-5, 1e-5])
    ff = 1.0 / 2 ** level
    bf = [rng.normal(0.0, 4e-5, grid.N) for _ in range(3)]
    bf_in = [b.copy() for b in bf]
    L = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    L.even = even
    L.f[:] = pack_pre(f0, grid, even)
    L.collide_body_force(grid.fluid, 1.6, forces, ff, bf)
    post = _unstream(unpack_pre(L.f, grid, not even), grid.nx, grid.ny, grid.nz)
    rho0 = 1.0 + f0.sum(1)
    c = np.stack([CX, CY, CZ], 1).astype(np.float64)
    dj = post @ c - f0 @ c
    expect = rho0[:, None] * (forces[None, :] + np.stack(bf_in, 1)) * ff
    assert np.abs(dj - expect).max() < 1e-15
    assert np.abs(post.sum(1) - f0.sum(1)).max() < 1e-15  # mass per node
    for b in bf:
        assert not b.any()  # reset after use


def test_body_force_equals_pinned_collision_per_node(orc):
    """Each node's collision equals the reference-pinned k17_nodes with the force
    (forces + bodyForce[k]) * 1/2 * forceFactor, bit for bit."""
    grid = periodic_box(6, 5, 4)
    rng = np.random.default_rng(5)
    f0 = near_equilibrium(rng, grid.N, u_amp=0.03, noise=4e-3)
    forces, ff = np.array([1e-4, 2e-5, -7e-5]), 0.5
    bf = [rng.normal(0.0, 1e-4, grid.N) for _ in range(3)]
    half = [(forces[i] + bf[i]) * 0.5 * ff for i in range(3)]
    L = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    L.f[:] = pack_pre(f0, grid, True)
    L.collide_body_force(grid.fluid, 1.7, forces, ff, [b.copy() for b in bf])
    post = _unstream(unpack_pre(L.f, grid, False), grid.nx, grid.ny, grid.nz)
    for k in range(0, grid.N, 7):
        ref, _ = orc.k17_nodes(f0[k:k + 1], 1.7, force=(half[0][k], half[1][k], half[2][k]))
        assert np.array_equal(post[k], ref[0]), k


def test_zero_body_force_is_the_global_force_path(orc):
    """With bodyForce = 0 the template equals the Default template with forces * 1/2 * forceFactor."""
    grid = periodic_box(5, 4, 6)
    rng = np.random.default_rng(6)
    buf = pack_pre(near_equilibrium(rng, grid.N, u_amp=0.03, noise=4e-3), grid, True)
    forces, ff = np.array([3e-5, -1e-5, 2e-5]), 0.25
    A = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    A.f[:] = buf
    A.collide_body_force(grid.fluid, 1.5, forces, ff, [np.zeros(grid.N) for _ in range(3)])
    B = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    B.f[:] = buf
    B.collide(grid.fluid, 1.5, force=tuple(forces * 0.5 * ff))
    assert np.array_equal(A.f, B.f)
