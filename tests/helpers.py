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
               limiters=(0.01, 0.01, 0.01), dtype=np.float64):
    """collision -> velocity BC -> no-slip BC -> swap -> pressure BC, `steps` times."""
    L = orc.Lattice(grid.nbx, grid.nby, grid.nbz, dtype=dtype)
    N = grid.N
    if init is None:
        L.init_eq(np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N), geo=grid.geo)
    else:
        L.init_eq(*init, geo=grid.geo)
    vel = grid.bcs.get("velocity")
    ns = grid.bcs.get("noslip")
    pr = grid.bcs.get("pressure")
    for _ in range(steps):
        L.collide(grid.fluid, omega, limiters)
        if velocity_bc and vel is not None:
            L.bc_post(ORACLE_POST_BC[velocity_bc], vel.k, vel.q27, omega, vel.velocity)
        if noslip_bc and ns is not None:
            L.bc_post(ORACLE_POST_BC[noslip_bc], ns.k, ns.q27, omega)
        L.swap()
        if pressure and pr is not None:
            L.bc_pressure(pr.k, pr.kN, pr.rho_bc, pr.direction)
    return L
