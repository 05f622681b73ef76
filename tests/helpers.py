"""Shared input generators and eso-twist packing for the tests (seeded, numpy only)."""
import numpy as np

from virtualfluids_amd.grid import CX, CY, CZ, INV, W


def equilibrium(drho, u):
    """Shifted equilibrium populations, rows = nodes (shape (n, 27))."""
    drho = np.asarray(drho, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    cu = u[:, :1] * CX + u[:, 1:2] * CY + u[:, 2:3] * CZ
    usq = (u ** 2).sum(1)[:, None]
    return W * (drho[:, None] + (1 + drho[:, None]) * (3 * cu + 4.5 * cu ** 2 - 1.5 * usq))


def product_equilibrium(drho, u):
    """Shifted equilibrium of the cumulant collision: a product of 1D factors per axis, whose
    central moments are the ones K17 relaxes to (its collision leaves it unchanged)."""
    def phi(c, v):
        return np.where(c == 0, 2.0 / 3.0 - v * v, (1.0 / 3.0 + v * v + c * v) / 2.0)

    u = np.atleast_2d(u)
    rho = 1.0 + np.atleast_1d(drho)[:, None]
    return rho * phi(CX, u[:, :1]) * phi(CY, u[:, 1:2]) * phi(CZ, u[:, 2:3]) - W


def near_equilibrium(rng, n, u_amp=0.03, noise=1e-3, rho_amp=0.01):
    drho = rng.normal(0.0, rho_amp, n)
    u = rng.normal(0.0, u_amp, (n, 3))
    return equilibrium(drho, u) + W * rng.normal(0.0, noise, (n, 27))


def _read_nodes(grid, d):
    """Node at which direction d of every node is stored: +1 along each axis with c_d < 0."""
    loc = np.arange(grid.N)
    if CX[d] < 0:
        loc = grid.nbx[loc].astype(np.int64)
    if CY[d] < 0:
        loc = grid.nby[loc].astype(np.int64)
    if CZ[d] < 0:
        loc = grid.nbz[loc].astype(np.int64)
    return loc


def pack_pre(fnodes, grid, even=True, buf=None):
    """Place pre-collision populations (N, 27) into a 27*N eso-twist buffer."""
    N = grid.N
    if buf is None:
        buf = np.zeros(27 * N, dtype=np.float64)
    for d in range(27):
        s = d if even else INV[d]
        buf[s * N + _read_nodes(grid, d)] = fnodes[:, d]
    return buf


def unpack_pre(buf, grid, even=True):
    """Pre-collision populations (N, 27) a collision with this parity would read."""
    N = grid.N
    out = np.empty((N, 27), dtype=np.float64)
    for d in range(27):
        s = d if even else INV[d]
        out[:, d] = buf[s * N + _read_nodes(grid, d)]
    return out


def stream_periodic(post, nx, ny, nz):
    """Pull streaming on a periodic box: pre'[x, d] = post[x - c_d, d] (np.roll)."""
    f = post.reshape(nz, ny, nx, 27)
    out = np.empty_like(f)
    for d in range(27):
        out[..., d] = np.roll(f[..., d], shift=(CZ[d], CY[d], CX[d]), axis=(0, 1, 2))
    return out.reshape(-1, 27)


# ---- the oracle driven with UpdateGrid27's order (src/gpu/core/Calculation/UpdateGrid27.cpp:54-96) ----
ORACLE_POST_BC = {
    "VelocityBounceBack": 1, "VelocityInterpolatedCompressible": 3,
    "VelocityWithPressureInterpolatedCompressible": 4,
    "NoSlipBounceBack": 0, "NoSlipInterpolatedCompressible": 2,
}


def oracle_run(orc, grid, omega, steps, velocity_bc=None, noslip_bc=None, pressure=False, init=None,
               limiters=(0.01, 0.01, 0.01), dtype=np.float64, geometry_bc=None, porosity=None):
    """collision -> velocity BC -> no-slip BC -> geometry BC -> swap -> pressure BC, `steps` times."""
    L = orc.Lattice(grid.nbx, grid.nby, grid.nbz, dtype=dtype)
    N = grid.N
    if init is None:
        L.init_eq(np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N), geo=grid.geo)
    else:
        L.init_eq(*init, geo=grid.geo)
    vel = grid.bcs.get("velocity")
    ns = grid.bcs.get("noslip")
    pr = grid.bcs.get("pressure")
    geo = grid.bcs.get("geometry")
    for _ in range(steps):
        L.collide(grid.fluid, omega, limiters, porosity=porosity)
        if velocity_bc and vel is not None:
            L.bc_post(ORACLE_POST_BC[velocity_bc], vel.k, vel.q27, omega, vel.velocity)
        if noslip_bc and ns is not None:
            L.bc_post(ORACLE_POST_BC[noslip_bc], ns.k, ns.q27, omega)
        if geometry_bc and geo is not None:
            L.bc_post(ORACLE_POST_BC[geometry_bc], geo.k, geo.q27, omega)
        L.swap()
        if pressure and pr is not None:
            L.bc_pressure(pr.k, pr.kN, pr.rho_bc, pr.direction)
    return L


def oracle_multilevel(orc, nb, omega_c, omega_f, steps, init_c, init_f, limiters=(0.01, 0.01, 0.01),
                      dtype=np.float64):
    """Two levels of a NestedBox on the oracle, one coarse step = two fine steps, the coarse
    step, fine->coarse, then coarse->fine (UpdateGrid27::updateGrid)."""
    C = orc.Lattice(nb.coarse.nbx, nb.coarse.nby, nb.coarse.nbz, dtype=dtype)
    F = orc.Lattice(nb.fine.nbx, nb.fine.nby, nb.fine.nbz, dtype=dtype)
    for L, g, init in ((C, nb.coarse, init_c), (F, nb.fine, init_f)):
        if isinstance(init, np.ndarray):  # pre-collision populations (N, 27)
            L.f[:] = pack_pre(init, g, True)
        else:
            L.init_eq(*init, geo=g.geo)
    zc = [np.zeros(nb.cf.n)] * 3
    zf = [np.zeros(nb.fc.n)] * 3
    for _ in range(steps):
        for _ in range(2):
            F.collide(nb.fine.fluid, omega_f, limiters)
            F.swap()
        C.collide(nb.coarse.fluid, omega_c, limiters)
        C.swap()
        orc.scale(True, C, F, C.even, nb.fc.coarse, nb.fc.fine, zf, omega_c, omega_f)
        orc.scale(False, C, F, C.even, nb.cf.coarse, nb.cf.fine, zc, omega_c, omega_f)
    return C, F


def shear_wave_init(nb, amp):
    """u_x = amp sin(2 pi z / Lz) on both levels of a NestedBox (coarse lattice units)."""
    k = 2 * np.pi / nb.coarse.nz
    _, _, zc = nb.coarse.coords()
    _, _, zf = nb.fine_coords()
    Nc, Nf = nb.coarse.N, nb.fine.N
    init_c = (np.zeros(Nc), amp * np.sin(k * zc), np.zeros(Nc), np.zeros(Nc))
    init_f = (np.zeros(Nf), amp * np.sin(k * zf), np.zeros(Nf), np.zeros(Nf))
    return k, init_c, init_f


def conserved_sums(buf, grid, even):
    """FP64 sums of drho and momentum over the fluid nodes of a host buffer (the quantities
    vf_MomentSums reduces on the GPU), read with the given parity."""
    f = unpack_pre(np.asarray(buf, dtype=np.float64), grid, even)[grid.geo == 19]
    return np.array([f.sum(), (f @ CX).sum(), (f @ CY).sum(), (f @ CZ).sum()])
