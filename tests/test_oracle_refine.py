"""Pins of the oracle's grid-refinement interpolation (CPU only).

* agreement with the reference's own InterpolationCoefficients / InterpolationCF /
  InterpolationFC headers compiled into oracle/_ref (1e-14 relative: the oracle sums the
  cell in a different order) and with golden vectors written by that build;
* a uniform equilibrium state is reproduced exactly (every gradient coefficient vanishes);
* a linear velocity field in equilibrium is reproduced at every fine position;
* the density and velocity of the rebuilt node are the polynomial values.
"""
import numpy as np
import pytest

from helpers import near_equilibrium, product_equilibrium

GOLD = np.load(__file__.replace("test_oracle_refine.py", "golden/k17_golden.npz"))
NAMES = None


def _pos(q):
    return (((q & 1) - 0.5) / 2, ((q >> 1 & 1) - 0.5) / 2, ((q >> 2 & 1) - 0.5) / 2)


def test_refinement_matches_reference_golden(orc):
    cells, offs = GOLD["refine_cells"], GOLD["refine_offsets"]
    for i in range(cells.shape[0]):
        for q in range(8):
            c, f = orc.interp_cell(cells[i], 1.7, offs[i], 0, 1.85, 0.5, _pos(q))
            np.testing.assert_allclose(c, GOLD["refine_coeffs"][i], rtol=1e-13, atol=1e-17)
            np.testing.assert_allclose(f, GOLD["refine_cf"][i, q], rtol=0, atol=1e-16)
        _, f = orc.interp_cell(cells[i], 1.85, offs[i], 1, 1.7, 2.0)
        np.testing.assert_allclose(f, GOLD["refine_fc"][i], rtol=0, atol=1e-16)


def test_refinement_matches_reference_build(orc):
    if orc.ref("f64") is None:
        pytest.skip("oracle/_ref not built here (needs the reference tree)")
    rng = np.random.default_rng(21)
    for trial in range(40):
        f8 = near_equilibrium(rng, 8, u_amp=0.06, noise=8e-3)
        off = tuple(rng.choice([-0.5, 0.0, 0.5], 3)) if trial % 2 else (0.0, 0.0, 0.0)
        pos = tuple(rng.uniform(-0.5, 0.5, 3))
        for mode, (om, eps) in ((0, (1.3, 0.5)), (1, (1.9, 2.0))):
            a, fa = orc.interp_cell(f8, 1.6, off, mode, om, eps, pos)
            b, fb = orc.ref_interp_cell(f8, 1.6, off, mode, om, eps, pos)
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-16)
            np.testing.assert_allclose(fa, fb, rtol=0, atol=1e-16)


def test_uniform_equilibrium_is_reproduced(orc):
    feq = product_equilibrium(0.013, [0.03, -0.02, 0.01])[0]
    f8 = np.tile(feq, (8, 1))
    c, _ = orc.interp_cell(f8, 1.6)
    names = orc.COEFF_NAMES
    for name, value in zip(names, c):
        if name in ("a000", "b000", "c000", "d000"):
            continue
        assert abs(value) < 1e-15, (name, value)
    assert c[names.index("a000")] == pytest.approx(0.03, rel=1e-14)
    assert c[names.index("d000")] == pytest.approx(0.013, rel=1e-14)
    for q in range(8):
        _, f = orc.interp_cell(f8, 1.6, mode=0, omega_dst=1.8, eps=0.5, pos=_pos(q))
        np.testing.assert_allclose(f, feq, rtol=0, atol=2e-16)
    _, f = orc.interp_cell(f8, 1.8, mode=1, omega_dst=1.6, eps=2.0)
    np.testing.assert_allclose(f, feq, rtol=0, atol=2e-16)


def test_linear_velocity_field_is_reproduced(orc):
    rng = np.random.default_rng(22)
    u0 = np.array([0.02, -0.01, 0.015])
    G = rng.normal(0, 0.01, (3, 3))
    corners = np.array([[(n & 1) - 0.5, (n >> 1 & 1) - 0.5, (n >> 2 & 1) - 0.5] for n in range(8)])
    u = u0 + corners @ G.T
    f8 = product_equilibrium(np.zeros(8), u)
    for q in range(8):
        p = np.array(_pos(q))
        _, f = orc.interp_cell(f8, 1.6, mode=0, omega_dst=1.8, eps=0.5, pos=tuple(p))
        m = orc.moments(f)
        np.testing.assert_allclose(m[5:8], u0 + G @ p, rtol=0, atol=1e-14)
