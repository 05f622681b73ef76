"""CPU pins of the adjoint (no reference counterpart; parity is pinned by mathematics).

* the oracle's reference adjoint (forward-mode dual numbers through its own collision) agrees
  with FP64 central finite differences of the collision;
* the GPU source of the hand-written reverse mode (vf_k17_adjoint.h), compiled for the host,
  agrees with that reference adjoint for limiter-active and limiter-free nodes, several relaxation
  rates and a body force, including the velocity-shift adjoint;
* the oracle's lattice adjoint step passes the dot-product test <w, J v> = <J^T w, v> over
  several steps of a masked lattice, with J v from FP64 central differences.
"""
import ctypes
import os

import numpy as np
import pytest

from helpers import near_equilibrium, pack_pre
from virtualfluids_amd.grid import walled_box

HOST_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libk17host.so")


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOST_LIB):
        pytest.fail(f"{HOST_LIB} is missing: build with virtualfluids_amd/build.py")
    return ctypes.CDLL(HOST_LIB)


CASES = [  # omega, limiter, noise, u_amp, body force
    (1.2, 0.01, 2e-2, 0.08, None),
    (1.99, 1e6, 5e-3, 0.05, None),
    (1.8, 1e-4, 2e-2, 0.05, (2e-4, -4e-4, 6e-4)),
    (0.8, 0.01, 1e-3, 0.1, (2e-3, 0.0, 0.0)),
]


def test_dual_adjoint_matches_central_differences(orc):
    rng = np.random.default_rng(40)
    fin = near_equilibrium(rng, 40, u_amp=0.06, noise=2e-2)
    g = rng.normal(0, 1, fin.shape)
    v = rng.normal(0, 1, fin.shape)
    for omega, lim in ((1.7, 0.01), (1.95, 1e6)):
        fb, _ = orc.k17_vjp(fin, g, omega, (lim,) * 3)
        h = 1e-6
        fp, _ = orc.k17_nodes(fin + h * v, omega, (lim,) * 3)
        fm, _ = orc.k17_nodes(fin - h * v, omega, (lim,) * 3)
        lhs = ((fp - fm) / (2 * h) * g).sum(1)
        rhs = (fb * v).sum(1)
        assert np.abs(lhs - rhs).max() < 2e-6 * np.abs(rhs).max()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_host_build_of_gpu_adjoint_matches_reference_adjoint(orc, host, case):
    omega, lim, noise, uamp, force = CASES[case]
    rng = np.random.default_rng(41 + case)
    n = 96
    fin = near_equilibrium(rng, n, u_amp=uamp, noise=noise).astype(np.float32)
    g = rng.normal(0, 1, (n, 27)).astype(np.float32)
    lim3 = np.full(3, lim)
    # the oracle takes the velocity shift, the host build the body force (shift = force / 2)
    shift = None if force is None else np.asarray(force)
    fb_ref, ub_ref = orc.k17_vjp(fin.astype(np.float64), g.astype(np.float64), omega, lim3, shift)
    fb = np.empty_like(fin)
    ub = np.empty((n, 3), dtype=np.float32)
    body = None if force is None else (2 * np.asarray(force)).astype(np.float32)
    host.host_k17_vjp(_p(fin), _p(g), ctypes.c_longlong(n), ctypes.c_float(omega), _p(lim3.astype(np.float32)),
                      _p(body), _p(fb), _p(ub))
    # FP32 reverse mode against FP64: relative to the largest entry (the limiter's derivative
    # N L / (|x| + L)^2 amplifies rounding when L is small)
    tol = 2e-5 if lim < 1e-3 else 2e-6
    assert np.abs(fb - fb_ref).max() < tol * np.abs(fb_ref).max()
    assert np.abs(ub - ub_ref).max() < 10 * tol * np.abs(ub_ref).max()


def test_host_build_of_gpu_collision_matches_oracle(orc, host):
    rng = np.random.default_rng(45)
    fin = near_equilibrium(rng, 64, u_amp=0.05, noise=5e-3).astype(np.float32)
    out = np.empty_like(fin)
    lim = np.full(3, 0.01, dtype=np.float32)
    host.host_k17_collide(_p(fin), _p(out), ctypes.c_longlong(64), ctypes.c_float(1.8), _p(lim), None)
    ref, _ = orc.k17_nodes(fin.astype(np.float64), 1.8, (0.01,) * 3)
    assert np.abs(out - ref).max() < 2e-7


def test_lattice_adjoint_dot_product(orc):
    """Four steps of a walled box (solid nodes never collide): <w, J v> (central differences,
    FP64) equals <J^T w, v> (oracle adjoint steps in reverse order)."""
    grid = walled_box((5, 4, 3), {})
    rng = np.random.default_rng(46)
    f0 = pack_pre(near_equilibrium(rng, grid.N, u_amp=0.05, noise=5e-3), grid, True)
    v = rng.normal(0, 1, f0.shape)
    w = rng.normal(0, 1, f0.shape)
    omega, steps = 1.85, 4

    def run(buf, keep=False):
        L = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
        L.f[:] = buf
        states = []
        for _ in range(steps):
            if keep:
                states.append((L.f.copy(), L.even))
            L.collide(grid.fluid, omega)
            L.swap()
        return L.f.copy(), states

    h = 1e-6
    jv = (run(f0 + h * v)[0] - run(f0 - h * v)[0]) / (2 * h)
    _, states = run(f0, keep=True)
    lam = w.copy()
    A = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    for state, even in reversed(states):
        A.f[:] = state
        A.even = even
        orc.adjoint_step(A, lam, grid.fluid, omega)
    lhs = float(w @ jv)
    rhs = float(lam @ v)
    assert abs(lhs - rhs) < 1e-7 * abs(rhs), (lhs, rhs)
    # solid nodes' cells that no fluid node reads keep their adjoint (identity)
    assert np.any(lam == w)
