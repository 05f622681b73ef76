"""Momentum-exchange force on a BC list (CPU, oracle states).

The index construction of virtualfluids_amd.adjoint.momentum_exchange_index is pinned against
a direct count on the oracle: the post-collision populations a step sends into the sphere
(read before the BC kernel runs) plus the ones the geometry BC writes back, summed with their
lattice directions. A symmetric flow past a centred sphere has no lift, and the drag points
downstream.
"""
import numpy as np

from helpers import ORACLE_POST_BC, near_equilibrium, pack_pre
from virtualfluids_amd.grid import CX, CY, CZ, INV, sphere_channel


def test_momentum_exchange_index_matches_direct_count(orc):
    from virtualfluids_amd.adjoint import _cell_index, momentum_exchange_index

    g = sphere_channel((20, 12, 12), (8, 5.5, 5.5), 3.0, 0.03)
    geo = g.bcs["geometry"]
    rng = np.random.default_rng(90)
    L = orc.Lattice(g.nbx, g.nby, g.nbz)
    L.f[:] = pack_pre(near_equilibrium(rng, g.N, u_amp=0.03, noise=2e-3), g, True)
    omega = 1.8
    L.collide(g.fluid, omega)
    post = L.f.copy()
    L.bc_post(ORACLE_POST_BC["NoSlipInterpolatedCompressible"], geo.k, geo.q27, omega)
    written = L.f.copy()
    L.swap()
    # direct: post-collision value of link d at node n sits in n's cell of pre(inv(d)) before
    # the swap; the BC's value for inv(d) goes to n's cell of pre(inv(d)) after the swap
    q = geo.q27.reshape(27, geo.n)
    F = np.zeros(3)
    for d in range(1, 27):
        act = (q[d] >= 0) & (q[d] <= 1)
        n = geo.k[act].astype(np.int64)
        fout = post[_cell_index(g, n, INV[d], True)]
        fin = written[_cell_index(g, n, INV[d], False)]
        F += np.array([CX[d], CY[d], CZ[d]]) * (fout + fin).sum()
    out_idx, in_idx, c = momentum_exchange_index(g, "geometry", L.even)
    F2 = (c * (L.f[out_idx] + L.f[in_idx])).sum(1)
    np.testing.assert_allclose(F2, F, rtol=1e-12, atol=1e-15)


def test_flow_past_sphere_drag_points_downstream(orc):
    """300 steps of inflow past a sphere centred in the duct's cross-section: the force points
    downstream and the mirror symmetry in y and z leaves no lift."""
    from helpers import oracle_run
    from virtualfluids_amd.adjoint import momentum_exchange_index

    g = sphere_channel((24, 13, 13), (9, 6, 6), 3.0, 0.04)
    L = oracle_run(orc, g, 1.8, 300, velocity_bc="VelocityInterpolatedCompressible",
                   noslip_bc="NoSlipInterpolatedCompressible", geometry_bc="NoSlipInterpolatedCompressible",
                   pressure=True)
    out_idx, in_idx, c = momentum_exchange_index(g, "geometry", L.even)
    F = (c * (L.f[out_idx] + L.f[in_idx])).sum(1)
    assert F[0] > 0
    assert abs(F[1]) < 1e-6 * F[0] and abs(F[2]) < 1e-6 * F[0]
