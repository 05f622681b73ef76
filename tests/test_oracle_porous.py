"""Brinkman porosity in the oracle's collision, pinned by closed forms (CPU).

The porosity phi enters the collision as the velocity scale (1 - phi/2) of the half-step
velocity shift, i.e. an acceleration -phi * j / rho applied with the second-order forcing of
the K17 kernel (the same route as the global force). For a uniform state, which streaming
leaves unchanged, this fixes the momentum after every step exactly:

    j_{t+1} = (1 - phi) j_t + rho g        (g: the global acceleration)

so a flow at rest driven by g approaches the Darcy velocity j / rho = g / phi along
j_t = rho g (1 - (1 - phi)^t) / phi, while the density is conserved. A dropped factor 1/2 in
the velocity scale, a sign error or a scale applied to the force instead of the momentum
each break these by O(phi) per step.
"""
import numpy as np
import pytest

from helpers import pack_pre, product_equilibrium, unpack_pre
from virtualfluids_amd.grid import CX, CY, CZ, W, periodic_box


def _momentum(L, grid):
    full = unpack_pre(L.f, grid, L.even) + W
    rho = full.sum(axis=1)
    j = np.stack([(full * c).sum(axis=1) for c in (CX, CY, CZ)], axis=1)
    return rho, j


@pytest.mark.parametrize("phi", [0.0, 0.03, 0.1])
@pytest.mark.parametrize("omega", [1.2, 1.9])
def test_uniform_flow_momentum_decays_by_one_minus_porosity(orc, phi, omega):
    grid = periodic_box(6, 5, 4)
    u0 = np.array([0.03, -0.02, 0.01])
    f0 = product_equilibrium(np.full(grid.N, 0.01), np.tile(u0, (grid.N, 1)))
    L = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    L.f[:] = pack_pre(f0, grid, True)
    rho0, j0 = _momentum(L, grid)
    por = np.full(grid.N, phi)
    for t in range(1, 6):
        L.collide(grid.fluid, omega, limiters=(1e6, 1e6, 1e6), porosity=por)
        L.swap()
        rho, j = _momentum(L, grid)
        np.testing.assert_allclose(rho, rho0, rtol=0, atol=1e-14)
        np.testing.assert_allclose(j, j0 * (1.0 - phi) ** t, rtol=0, atol=1e-15)


@pytest.mark.parametrize("phi", [0.05, 0.2])
def test_darcy_velocity_under_a_body_force(orc, phi):
    grid = periodic_box(4, 4, 4)
    g = np.array([2e-4, 0.0, -1e-4])
    L = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    L.f[:] = 0.0  # rest, rho = 1 (omega 1.9: outside the windows where K17's third-order rates
    # turn negative and rounding noise grows without the limiter)
    por = np.full(grid.N, phi)
    for t in range(1, 41):
        # the oracle takes the half-step velocity shift g/2 (forces * 1/2 * forceFactor)
        L.collide(grid.fluid, 1.9, limiters=(1e6, 1e6, 1e6), force=tuple(0.5 * g), porosity=por)
        L.swap()
        rho, j = _momentum(L, grid)
        expect = g * (1.0 - (1.0 - phi) ** t) / phi
        np.testing.assert_allclose(rho, 1.0, rtol=0, atol=1e-14)
        np.testing.assert_allclose(j, np.tile(expect, (grid.N, 1)), rtol=0, atol=2e-14)
    # close to the Darcy velocity g / phi after many relaxation times of the porous drag
    assert np.allclose(j[0], g / phi, rtol=(1.0 - phi) ** 40 * 1.01)
