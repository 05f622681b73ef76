"""The oracle's non-equilibrium initialisation (InitNavierStokesCompressibleNonEquilibrium_Device).

The kernel is CUDA-only in the reference tree, so it is pinned by what its construction fixes
for any velocity field, checked at interior nodes of a linear field (exact central differences):
the non-equilibrium part carries no mass and no momentum, its off-diagonal second moments are
-2 S_ab / (3 omega) and its normal-stress differences -2 (S_aa - S_bb) / (3 omega) (first-order
Chapman-Enskog with the shear rate omega), and its trace is -(5/3) div u whatever omega (the
reference's bulk part with rate 1). With a uniform field it reduces to the plain equilibrium,
and solid nodes get the marker 96 in their rest slot.
"""
import numpy as np
import pytest

from helpers import equilibrium, unpack_pre
from virtualfluids_amd.grid import CX, CY, CZ, periodic_box, walled_box


def _linear_field(g, G, u0=(0.01, -0.02, 0.015)):
    x, y, z = g.coords()
    X = np.stack([x, y, z]).astype(np.float64)
    return [u0[i] + G[i] @ X for i in range(3)]


@pytest.mark.parametrize("omega", [1.2, 1.7, 1.95])
def test_neq_moments_follow_chapman_enskog(orc, omega):
    g = periodic_box(9, 9, 9)
    rng = np.random.default_rng(70)
    G = rng.normal(0, 0.004, (3, 3))
    ux, uy, uz = _linear_field(g, G)
    drho = np.full(g.N, 0.013)
    L = orc.Lattice(g.nbx, g.nby, g.nbz)
    L.init_neq(drho, ux, uy, uz, g.neighbor_inverse(), omega)
    f = unpack_pre(L.f, g, True)
    S = 0.5 * (G + G.T)
    C = [CX, CY, CZ]
    for k in g.index(np.array([3, 4, 5]), np.array([4, 4, 3]), np.array([5, 3, 4])):
        feq = equilibrium(drho[k:k + 1], np.array([[ux[k], uy[k], uz[k]]]))[0]
        neq = (f[k] - feq) / (1 + drho[k])
        assert abs(neq.sum()) < 1e-16
        for c in C:
            assert abs(c @ neq) < 1e-16
        for a, b in ((0, 1), (0, 2), (1, 2)):
            assert (C[a] * C[b]) @ neq == pytest.approx(-2 * S[a, b] / (3 * omega), rel=1e-12, abs=1e-17)
        P = [(C[a] * C[a]) @ neq for a in range(3)]
        assert P[0] - P[1] == pytest.approx(-2 * (S[0, 0] - S[1, 1]) / (3 * omega), rel=1e-12, abs=1e-17)
        assert P[0] - P[2] == pytest.approx(-2 * (S[0, 0] - S[2, 2]) / (3 * omega), rel=1e-12, abs=1e-17)
        assert sum(P) == pytest.approx(-5.0 / 3.0 * np.trace(G), rel=1e-12, abs=1e-17)


def test_neq_init_of_uniform_field_is_equilibrium_and_marks_solids(orc):
    g = walled_box((6, 5, 4), {})
    N = g.N
    u = (0.03, -0.01, 0.02)
    args = (np.full(N, 0.01), np.full(N, u[0]), np.full(N, u[1]), np.full(N, u[2]))
    A = orc.Lattice(g.nbx, g.nby, g.nbz)
    A.init_neq(*args, g.neighbor_inverse(), 1.6, geo=g.geo)
    B = orc.Lattice(g.nbx, g.nby, g.nbz)
    B.init_eq(*args, geo=g.geo)
    solid = g.geo == 15
    rest = np.arange(N)  # slot d000 = 0 holds node k's rest population at index k
    assert np.all(A.f[rest[solid]] == 96.0)
    B.f[rest[solid]] = 96.0
    np.testing.assert_array_equal(A.f, B.f)
