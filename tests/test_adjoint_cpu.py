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

    h = 1e-7  # small enough to resolve the limiter's curvature near x = 0
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
    assert abs(lhs - rhs) < 1e-6 * abs(rhs), (lhs, rhs)
    # solid nodes' cells that no fluid node reads keep their adjoint (identity)
    assert np.any(lam == w)


def _channel(q_random):
    from virtualfluids_amd.grid import channel

    g = channel((7, 5, 4), inflow=0.03)
    if q_random:
        rng = np.random.default_rng(47)
        for bc in g.bcs.values():
            act = bc.q27 >= 0
            bc.q27 = np.where(act, rng.uniform(0.15, 0.85, bc.q27.shape), bc.q27)
    return g


@pytest.mark.parametrize("velocity_kind,noslip_kind,q_random", [
    (1, 0, False),   # VelocityBounceBack, NoSlipBounceBack
    (3, 2, True),    # VelocityInterpolatedCompressible, NoSlipInterpolatedCompressible
    (4, 2, True),    # VelocityWithPressureInterpolatedCompressible
])
def test_step_with_bcs_adjoint_dot_product(orc, velocity_kind, noslip_kind, q_random):
    """Three steps of a duct (velocity inflow, pressure outlet, walls): <w, J v> by FP64 central
    differences equals <J^T w, v> from the oracle's adjoint of the pressure BC, the post-collision
    BCs and the collision, and the inflow-velocity and outlet-density gradients match too."""
    g = _channel(q_random)
    vel, ns, pr = g.bcs["velocity"], g.bcs["noslip"], g.bcs["pressure"]
    rng = np.random.default_rng(48 + velocity_kind)
    f0 = pack_pre(near_equilibrium(rng, g.N, u_amp=0.03, noise=3e-3), g, True)
    omega, steps = 1.8, 3
    U0 = np.array(vel.velocity)
    R0 = np.asarray(pr.rho_bc, dtype=np.float64)

    def run(buf, U, R, keep=False):
        L = orc.Lattice(g.nbx, g.nby, g.nbz)
        L.f[:] = buf
        rec = []
        for _ in range(steps):
            s0 = (L.f.copy(), L.even)
            L.collide(g.fluid, omega)
            post = L.f.copy()
            L.bc_post(velocity_kind, vel.k, vel.q27, omega, tuple(U))
            L.bc_post(noslip_kind, ns.k, ns.q27, omega)
            L.swap()
            L.bc_pressure(pr.k, pr.kN, R, pr.direction)
            if keep:
                rec.append((s0, post, L.f.copy(), L.even))
        return L.f.copy(), rec

    v = rng.normal(0, 1, f0.shape)
    w = rng.normal(0, 1, f0.shape)
    vU = rng.normal(0, 1, U0.shape)
    vR = rng.normal(0, 1, R0.shape)
    # the limiter x N / (|x| + L) has a large curvature near x = 0: central differences need a
    # small step to resolve its derivative (FP64 round-off still far below the tolerance)
    h = 1e-7
    jv = (run(f0 + h * v, U0, R0)[0] - run(f0 - h * v, U0, R0)[0]) / (2 * h)
    jU = (run(f0, U0 + h * vU, R0)[0] - run(f0, U0 - h * vU, R0)[0]) / (2 * h)
    jR = (run(f0, U0, R0 + h * vR)[0] - run(f0, U0, R0 - h * vR)[0]) / (2 * h)
    _, rec = run(f0, U0, R0, keep=True)
    lam = w.copy()
    ubar = np.zeros((vel.n, 3))
    rbar = np.zeros(pr.n)
    A = orc.Lattice(g.nbx, g.nby, g.nbz)
    for (s0, even0), post, s1, even1 in reversed(rec):
        A.f[:] = s1
        A.even = even1
        orc.bc_pressure_adjoint(A, lam, pr.k, pr.kN, R0, pr.direction, rhobar=rbar)
        A.f[:] = post
        A.even = even0
        orc.bc_post_adjoint(noslip_kind, A, lam, ns.k, ns.q27, omega)
        orc.bc_post_adjoint(velocity_kind, A, lam, vel.k, vel.q27, omega, velocity=tuple(U0), ubar=ubar)
        A.f[:] = s0
        orc.adjoint_step(A, lam, g.fluid, omega)
    assert float(w @ jv) == pytest.approx(float(lam @ v), rel=1e-6)
    assert float(w @ jU) == pytest.approx(float((ubar.T * vU).sum()), rel=1e-6)
    assert float(w @ jR) == pytest.approx(float(rbar @ vR), rel=1e-6)


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4])
def test_host_build_of_gpu_bc_adjoint_matches_reference(orc, host, kind):
    """The GPU's per-node reverse mode of the five post-collision BC formulas against the
    oracle's dual-number reverse mode of the same formulas (random links and distances)."""
    rng = np.random.default_rng(60 + kind)
    for trial in range(20):
        post = near_equilibrium(rng, 1, u_amp=0.05, noise=5e-3)[0]
        q = np.where(rng.random(27) < 0.4, rng.uniform(0.05, 0.95, 27), -1.0)
        q[0] = -1.0
        U = rng.normal(0, 0.03, 3)
        G = rng.normal(0, 1, 27)
        omega = 1.85
        fb_ref, ub_ref = np.empty(27), np.empty(3)
        orc.lib().orc_bc_node_vjp_f64(ctypes.c_int(kind), _p(post), _p(q), _p(U), ctypes.c_double(omega), _p(G),
                                      _p(fb_ref), _p(ub_ref))
        fb, ub = np.empty(27, np.float32), np.empty(3, np.float32)
        host.host_bc_post_vjp(ctypes.c_int(kind), _p(post.astype(np.float32)), _p(q.astype(np.float32)),
                              ctypes.c_float(omega), _p(G.astype(np.float32)), _p(fb), _p(ub))
        scale = max(np.abs(fb_ref).max(), 1e-30)
        assert np.abs(fb - fb_ref).max() < 1e-5 * scale, (trial, np.abs(fb - fb_ref).max(), scale)
        if kind in (1, 3, 4):
            assert np.abs(ub - ub_ref).max() < 1e-5 * max(np.abs(ub_ref).max(), 1e-30)


@pytest.mark.parametrize("direction", [1, 2, 3, 4, 5, 6])
def test_host_build_of_gpu_pressure_adjoint_matches_reference(orc, host, direction):
    from virtualfluids_amd.grid import CX, CY, CZ

    rng = np.random.default_rng(70 + direction)
    ax, sgn = (direction - 1) // 2, (1 if direction % 2 == 1 else -1)
    c = [CX, CY, CZ][ax]
    for trial in range(10):
        fn = near_equilibrium(rng, 1, u_amp=0.05, noise=5e-3)[0]
        rbc = rng.normal(0, 0.01)
        G = np.where(c * sgn <= 0, rng.normal(0, 1, 27), 0.0)
        fnb_ref, rb_ref = np.empty(27), np.empty(1)
        orc.lib().orc_pressure_node_vjp_f64(_p(fn), ctypes.c_double(rbc), ctypes.c_int(direction), _p(G),
                                            _p(fnb_ref), _p(rb_ref))
        fnb, rb = np.empty(27, np.float32), np.empty(1, np.float32)
        host.host_pressure_vjp(_p(fn.astype(np.float32)), ctypes.c_float(rbc), _p(G.astype(np.float32)), _p(fnb),
                               _p(rb))
        assert np.abs(fnb - fnb_ref).max() < 1e-5 * np.abs(fnb_ref).max()
        assert abs(rb[0] - rb_ref[0]) < 1e-5 * abs(rb_ref[0])


def test_host_build_of_brinkman_adjoint_matches_reference(orc, host):
    """The velocity-scale (Brinkman) term of the GPU reverse mode against the dual numbers."""
    rng = np.random.default_rng(80)
    n = 64
    fin = near_equilibrium(rng, n, u_amp=0.05, noise=5e-3)
    g = rng.normal(0, 1, (n, 27))
    vs = 1.0 - 0.5 * rng.uniform(0.0, 0.1, n)
    lim = np.full(3, 0.01)
    fb_ref, sb_ref = np.empty_like(fin), np.empty(n)
    orc.lib().orc_k17_vjp_scaled_f64(_p(fin), _p(g), ctypes.c_longlong(n), ctypes.c_double(1.8), _p(lim), _p(vs),
                                     _p(fb_ref), _p(sb_ref))
    f32, g32 = fin.astype(np.float32), g.astype(np.float32)
    fb, sb = np.empty_like(f32), np.empty(n, dtype=np.float32)
    host.host_k17_vjp_scaled(_p(f32), _p(g32), ctypes.c_longlong(n), ctypes.c_float(1.8), _p(lim.astype(np.float32)),
                             _p(vs.astype(np.float32)), _p(fb), _p(sb))
    assert np.abs(fb - fb_ref).max() < 2e-6 * np.abs(fb_ref).max()
    assert np.abs(sb - sb_ref).max() < 2e-5 * np.abs(sb_ref).max()


def test_porosity_gradient_dot_product(orc):
    """Three steps of a periodic box with a Brinkman porosity field: the porosity adjoint
    accumulated by the oracle's adjoint steps equals FP64 central differences of <w, f(phi)>."""
    from virtualfluids_amd.grid import periodic_box

    g = periodic_box(6, 5, 4)
    rng = np.random.default_rng(81)
    f0 = pack_pre(near_equilibrium(rng, g.N, u_amp=0.05, noise=3e-3), g, True)
    phi = rng.uniform(0.0, 0.1, g.N)
    w = rng.normal(0, 1, f0.shape)
    v = rng.normal(0, 1, g.N)
    omega, steps = 1.8, 3

    def run(ph, keep=False):
        L = orc.Lattice(g.nbx, g.nby, g.nbz)
        L.f[:] = f0
        rec = []
        for _ in range(steps):
            if keep:
                rec.append((L.f.copy(), L.even))
            L.collide(g.fluid, omega, porosity=ph)
            L.swap()
        return L.f.copy(), rec

    h = 1e-7
    fd = float(w @ ((run(phi + h * v)[0] - run(phi - h * v)[0]) / (2 * h)))
    _, rec = run(phi, keep=True)
    lam = w.copy()
    padj = np.zeros(g.N)
    A = orc.Lattice(g.nbx, g.nby, g.nbz)
    for state, even in reversed(rec):
        A.f[:] = state
        A.even = even
        orc.adjoint_step(A, lam, g.fluid, omega, porosity=phi, porosity_adjoint=padj)
    assert float(padj @ v) == pytest.approx(fd, rel=1e-6)
