"""The oracle's two-level time step on a nested box (CPU only).

Pins the interface geometry (which coarse and fine cells are interpolated, the fine node
positions s + 1/4 + j/2 and the parity each level is read with) against what the physics
fixes: a uniform flow stays uniform, and a decaying shear wave u_x = A sin(k z) that crosses
the fine box follows A exp(-nu k^2 t) on both levels as closely as on the unrefined grid.
A fine box shifted by one fine node, or an FC cell that reads the outer fine layers, breaks
the second property by an order of magnitude.
"""
import numpy as np

from helpers import oracle_multilevel, oracle_run, product_equilibrium, shear_wave_init
from virtualfluids_amd.grid import periodic_box
from virtualfluids_amd.refinement import nested_box

NU = 0.01  # coarse-level viscosity; both relaxation rates (1.887, 1.786) lie above 7/4


def _omegas():
    return 1.0 / (3 * NU + 0.5), 1.0 / (6 * NU + 0.5)


def test_nested_box_interface_counts():
    nb = nested_box((16, 16, 16), (4, 4, 4), (12, 12, 12))
    assert (nb.fine.nx, nb.fine.ny, nb.fine.nz) == (16, 16, 16)
    # CF: surface of the 8^3 block of cells [4, 11]^3; FC: surface of the node block [6, 10]^3
    assert nb.cf.n == 8 ** 3 - 6 ** 3
    assert nb.fc.n == 5 ** 3 - 3 ** 3
    assert int((nb.coarse.geo == 15).sum()) == 3 ** 3
    fx, fy, fz = nb.fine_coords()
    cx, cy, cz = nb.coarse.coords()
    # a CF fine cell's base sits at +1/4 of its coarse base, an FC fine cell's at -1/4
    np.testing.assert_allclose(fx[nb.cf.fine], cx[nb.cf.coarse] + 0.25)
    np.testing.assert_allclose(fz[nb.fc.fine], cz[nb.fc.coarse] - 0.25)
    # FC cells read fine nodes j = 3 .. 12 of 0 .. 15: the two outer layers on each side are
    # CF nodes, and the layer next to them is spared as well (FC cells span two fine nodes)
    jx, jy, jz = nb.fine.coords()
    for j in (jx, jy, jz):
        assert j[nb.fc.fine].min() == 3 and j[nb.fc.fine].max() + 1 == 12


def test_uniform_flow_stays_uniform(orc):
    nb = nested_box((14, 14, 14), (3, 3, 3), (11, 11, 11))
    u = (0.04, -0.02, 0.03)
    # the collision's own equilibrium: interior nodes keep it and the interpolation rebuilds it
    feq = product_equilibrium(0.01, u)
    init_c = np.repeat(feq, nb.coarse.N, axis=0)
    init_f = np.repeat(feq, nb.fine.N, axis=0)
    om_c, om_f = _omegas()
    C, F = oracle_multilevel(orc, nb, om_c, om_f, 6, init_c, init_f)
    valid = nb.coarse_valid()
    r, vx, vy, vz, _ = C.macroscopic(nb.coarse.geo)
    for got, want in ((r, 0.01), (vx, u[0]), (vy, u[1]), (vz, u[2])):
        np.testing.assert_allclose(got[valid], want, rtol=0, atol=1e-13)
    r, vx, vy, vz, _ = F.macroscopic()
    for got, want in ((r, 0.01), (vx, u[0]), (vy, u[1]), (vz, u[2])):
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


def test_shear_wave_decays_through_the_fine_box(orc):
    nb = nested_box((8, 8, 32), (1, 1, 10), (6, 6, 22))
    amp, steps = 0.02, 300
    k, init_c, init_f = shear_wave_init(nb, amp)
    om_c, om_f = _omegas()
    C, F = oracle_multilevel(orc, nb, om_c, om_f, steps, init_c, init_f)
    decay = amp * np.exp(-NU * k * k * steps)
    _, _, zc = nb.coarse.coords()
    _, _, zf = nb.fine_coords()
    valid = nb.coarse_valid()
    _, vxc, _, _, _ = C.macroscopic(nb.coarse.geo)
    _, vxf, vyf, vzf, _ = F.macroscopic()
    err_c = np.abs(vxc[valid] - decay * np.sin(k * zc[valid])).max() / decay
    err_f = np.abs(vxf - decay * np.sin(k * zf)).max() / decay
    # the same wave on the unrefined coarse grid
    g = periodic_box(8, 8, 32)
    _, _, z0 = g.coords()
    L0 = oracle_run(orc, g, om_c, steps, init=(np.zeros(g.N), amp * np.sin(k * z0), np.zeros(g.N), np.zeros(g.N)))
    _, vx0, _, _, _ = L0.macroscopic()
    err_0 = np.abs(vx0 - decay * np.sin(k * z0)).max() / decay
    print(f"shear wave rel. errors: unrefined {err_0:.2e} coarse {err_c:.2e} fine {err_f:.2e}")
    assert err_0 < 5e-3
    assert err_c < 2 * err_0 + 2e-3
    assert err_f < 2 * err_0 + 2e-3
    assert np.abs(vyf).max() < 1e-3 * decay and np.abs(vzf).max() < 1e-3 * decay
