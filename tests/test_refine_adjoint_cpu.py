"""Reverse mode of the refinement interpolation (CPU).

The interpolation of one cell (vf_scaling.h: source moments -> coefficients -> rebuilt nodes)
and its hand-written vector-Jacobian product, compiled for the host, against the oracle: the
forward against the oracle's restatement (itself pinned to the reference build), the reverse
against forward-mode dual numbers through that restatement, and the dual numbers against FP64
central differences of the oracle's forward. Near-equilibrium random cells, with and without
wall offsets (the offsets switch the LaplaceRho term off), several relaxation rates.
"""
import ctypes
import os

import numpy as np
import pytest

from helpers import near_equilibrium

HOST_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libk17host.so")
F = ctypes.c_float


def _host():
    h = ctypes.CDLL(HOST_LIB)
    return h


def _cell(rng, scale=1.0):
    return near_equilibrium(rng, 8, u_amp=0.04 * scale, noise=3e-3)


def _host_forward(h, f8, ws, off, mode, wd):
    out = np.empty((8, 27) if mode == 0 else (27,), np.float32)
    src = np.ascontiguousarray(f8, np.float32)
    h.host_interp_cell(src.ctypes.data_as(ctypes.c_void_p), F(ws), F(off[0]), F(off[1]), F(off[2]), ctypes.c_int(mode),
                       F(wd), out.ctypes.data_as(ctypes.c_void_p))
    return out


def _host_vjp(h, f8, ws, off, mode, wd, g):
    out = np.empty((8, 27), np.float32)
    src = np.ascontiguousarray(f8, np.float32)
    gg = np.ascontiguousarray(g, np.float32)
    h.host_interp_cell_vjp(src.ctypes.data_as(ctypes.c_void_p), F(ws), F(off[0]), F(off[1]), F(off[2]),
                           ctypes.c_int(mode), F(wd), gg.ctypes.data_as(ctypes.c_void_p),
                           out.ctypes.data_as(ctypes.c_void_p))
    return out


def _oracle_forward(orc, f8, ws, off, mode, wd):
    if mode == 1:
        return orc.interp_cell(f8, ws, off, 1, wd, 2.0)[1]
    return np.stack([orc.interp_cell(f8, ws, off, 0, wd, 0.5, pos=tuple(0.25 if (q >> a) & 1 else -0.25
                                                                        for a in range(3)))[1] for q in range(8)])


CASES = [(0, 1.6, 1.8, (0.0, 0.0, 0.0)), (0, 1.3, 1.7, (0.1, -0.2, 0.05)), (1, 1.8, 1.6, (0.0, 0.0, 0.0)),
         (1, 1.7, 1.4, (-0.15, 0.0, 0.2))]


@pytest.mark.parametrize("mode,ws,wd,off", CASES)
def test_host_forward_matches_oracle(orc, mode, ws, wd, off):
    h = _host()
    rng = np.random.default_rng(1000 + mode)
    for _ in range(5):
        f8 = _cell(rng)
        f8 = f8.astype(np.float32).astype(np.float64)
        got = _host_forward(h, f8, ws, off, mode, wd)
        want = _oracle_forward(orc, f8, ws, off, mode, wd)
        assert np.abs(got - want).max() < 2e-7


@pytest.mark.parametrize("mode,ws,wd,off", CASES)
def test_dual_reference_matches_central_differences(orc, mode, ws, wd, off):
    rng = np.random.default_rng(1100 + mode)
    f8 = _cell(rng)
    g = rng.normal(size=(8, 27) if mode == 0 else (27,))
    vjp = orc.interp_cell_vjp(f8, ws, g, off, mode, wd)
    h = 1e-6
    for _ in range(3):
        v = rng.normal(size=(8, 27))
        fd = ((_oracle_forward(orc, f8 + h * v, ws, off, mode, wd) - _oracle_forward(orc, f8 - h * v, ws, off, mode, wd))
              * g).sum() / (2 * h)
        ad = (vjp * v).sum()
        assert abs(ad - fd) < 1e-7 * max(1.0, abs(fd)), (ad, fd)


@pytest.mark.parametrize("mode,ws,wd,off", CASES)
def test_host_reverse_matches_dual_numbers(orc, mode, ws, wd, off):
    h = _host()
    rng = np.random.default_rng(1200 + mode)
    worst = 0.0
    for _ in range(6):
        f8 = _cell(rng).astype(np.float32).astype(np.float64)
        g = rng.normal(size=(8, 27) if mode == 0 else (27,)).astype(np.float32).astype(np.float64)
        want = orc.interp_cell_vjp(f8, ws, g, off, mode, wd)
        got = _host_vjp(h, f8, ws, off, mode, wd, g)
        err = np.abs(got - want).max() / np.abs(want).max()
        worst = max(worst, err)
    print(f"mode {mode} omega {ws}->{wd} offset {off}: max rel err {worst:.2e}")
    assert worst < 1e-5


@pytest.mark.parametrize("fine_to_coarse", [False, True])
@pytest.mark.parametrize("even_coarse", [True, False])
@pytest.mark.parametrize("offsets", [False, True])
def test_lattice_scale_adjoint_dot_product(orc, fine_to_coarse, even_coarse, offsets):
    """The reference adjoint of a whole interface call (written cells cleared, transposed
    interpolation added to the source cells) against FP64 central differences of the oracle's
    forward call: <lam, T(x + h v) - T(x - h v)> / 2h = <T^T lam, v> over both levels."""
    from helpers import pack_pre
    from virtualfluids_amd.refinement import nested_box

    nb = nested_box((12, 12, 12), (2, 3, 2), (8, 9, 9))
    cells = nb.fc if fine_to_coarse else nb.cf
    rng = np.random.default_rng(1300 + 2 * fine_to_coarse + even_coarse + 4 * offsets)
    C0 = pack_pre(near_equilibrium(rng, nb.coarse.N, u_amp=0.04, noise=3e-3), nb.coarse, even_coarse)
    F0 = pack_pre(near_equilibrium(rng, nb.fine.N, u_amp=0.04, noise=3e-3), nb.fine, True)
    off = [rng.uniform(-0.3, 0.3, cells.n) for _ in range(3)] if offsets else [np.zeros(cells.n)] * 3
    wc, wf = 1.7, 1.5

    def forward(c, f):
        C = orc.Lattice(nb.coarse.nbx, nb.coarse.nby, nb.coarse.nbz)
        F = orc.Lattice(nb.fine.nbx, nb.fine.nby, nb.fine.nbz)
        C.f[:] = c
        F.f[:] = f
        C.even = even_coarse
        orc.scale(fine_to_coarse, C, F, even_coarse, cells.coarse, cells.fine, off, wc, wf)
        return C.f.copy(), F.f.copy()

    lc = rng.normal(size=C0.size)
    lf = rng.normal(size=F0.size)
    C = orc.Lattice(nb.coarse.nbx, nb.coarse.nby, nb.coarse.nbz)
    F = orc.Lattice(nb.fine.nbx, nb.fine.nby, nb.fine.nbz)
    C.f[:] = C0
    F.f[:] = F0
    ac, af = lc.copy(), lf.copy()
    orc.scale_adjoint(fine_to_coarse, C, F, even_coarse, cells.coarse, cells.fine, off if offsets else None, wc, wf,
                      ac, af)
    h = 1e-7  # central-difference truncation ~1e-9 absolute here (it scales as h^2)
    for _ in range(2):
        vc = rng.normal(size=C0.size)
        vf = rng.normal(size=F0.size)
        cp, fp = forward(C0 + h * vc, F0 + h * vf)
        cm, fm = forward(C0 - h * vc, F0 - h * vf)
        fd = ((cp - cm) @ lc + (fp - fm) @ lf) / (2 * h)
        ad = ac @ vc + af @ vf
        print(f"f2c={fine_to_coarse} even={even_coarse} off={offsets}: adjoint {ad:.10e} FD {fd:.10e}")
        assert abs(ad - fd) < 1e-7 * abs(fd) + 1e-8


def test_interface_writes_disjoint_detects_overlaps():
    """The replay of the two-level backward applies the interface writes in place only where the
    written nodes are disjoint from what the reverse sweep reads: nested boxes are, and an FC cell
    reading a CF-written fine node, or an FC write onto a pressure BC's source node, is caught."""
    from virtualfluids_amd.adjoint import interface_writes_disjoint
    from virtualfluids_amd.grid import channel
    from virtualfluids_amd.refinement import InterfaceCells, nested_box

    nb = nested_box((24, 20, 20), (6, 5, 5), (14, 13, 13))
    assert interface_writes_disjoint(nb.coarse, nb.fine, nb.cf, nb.fc)
    # an FC cell whose fine cell is a CF target
    fc = InterfaceCells(np.append(nb.fc.coarse, nb.fc.coarse[:1]), np.append(nb.fc.fine, nb.cf.fine[:1]))
    assert not interface_writes_disjoint(nb.coarse, nb.fine, nb.cf, fc)
    # a channel: the FC coarse targets stay clear of the outlet's kN, unless one is made to hit it
    nc = nested_box(None, (5, 4, 4), (11, 10, 10), coarse=channel((16, 12, 12), inflow=0.03))
    assert interface_writes_disjoint(nc.coarse, nc.fine, nc.cf, nc.fc)
    kn = int(nc.coarse.bcs["pressure"].kN[0])
    fc = InterfaceCells(np.append(nc.fc.coarse, kn).astype(np.uint32), np.append(nc.fc.fine, nc.fc.fine[:1]))
    assert not interface_writes_disjoint(nc.coarse, nc.fine, nc.cf, fc)
