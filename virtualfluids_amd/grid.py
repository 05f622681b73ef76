"""Structured grids in the reference's data layout.

Builds, on the host, the arrays the GridGenerator/GridProvider pair hands to the
time step (src/gpu/core/DataStructureInitializer/GridProvider.h:61-85):
+x/+y/+z neighbour arrays (src/gpu/GridGenerator/grid/GridImp.cpp:1012-1078),
node types (src/gpu/core/Calculation/Calculation.h:36-45), the fluid-node list, and
boundary-condition node lists with SoA sub-grid distances q (q = 0.5 on box sides,
src/gpu/GridGenerator/grid/BoundaryConditions/Side.cpp:149-167).

Node index of a box: k = x + nx*(y + ny*z). Boxes carry no dummy node 0; the
neighbour arrays wrap periodically, so a solid layer closes a walled box.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GEO_SOLID = 15
GEO_FLUID = 19

# D3Q27 lattice velocities in the order of src/lbm/constants/D3Q27.h:46-72
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, 1, -1, -1, -1, -1])
INV = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 26, 25, 24, 23, 22, 21, 20, 19])
W = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)

# the six sides by their D3Q27 direction (Side::getD3Q27Direction)
dP00, dM00, d0P0, d0M0, d00P, d00M = 1, 2, 3, 4, 5, 6


@dataclass
class BoundaryNodes:
    """A QforBoundaryConditions list: BC node indices and SoA q27[d*n + i] (-1 = no link)."""

    k: np.ndarray
    q27: np.ndarray
    velocity: tuple | None = None      # (Vx, Vy, Vz) per BC node
    kN: np.ndarray | None = None       # pressure BC neighbour nodes
    rho_bc: np.ndarray | None = None   # pressure BC drho per node
    direction: int = 0                 # pressure BC side

    @property
    def n(self) -> int:
        return int(self.k.size)


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    nbx: np.ndarray
    nby: np.ndarray
    nbz: np.ndarray
    geo: np.ndarray
    fluid: np.ndarray
    bcs: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return self.nx * self.ny * self.nz

    def coords(self):
        k = np.arange(self.N)
        return k % self.nx, (k // self.nx) % self.ny, k // (self.nx * self.ny)

    def index(self, x, y, z):
        return np.asarray(x) + self.nx * (np.asarray(y) + self.ny * np.asarray(z))

    def neighbor_inverse(self) -> np.ndarray:
        """The (-1,-1,-1) neighbour of every node (neighborInverse), periodic wrap."""
        x, y, z = self.coords()
        return self.index((x - 1) % self.nx, (y - 1) % self.ny, (z - 1) % self.nz).astype(np.uint32)

    @property
    def fluid_mask(self) -> np.ndarray:
        return (self.geo == GEO_FLUID).astype(np.uint8)


def _neighbours(nx, ny, nz):
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    x, y, z = x.transpose(2, 1, 0).ravel(), y.transpose(2, 1, 0).ravel(), z.transpose(2, 1, 0).ravel()
    idx = lambda a, b, c: (a % nx) + nx * ((b % ny) + ny * (c % nz))
    return (idx(x + 1, y, z).astype(np.uint32), idx(x, y + 1, z).astype(np.uint32),
            idx(x, y, z + 1).astype(np.uint32), x, y, z)


def periodic_box(nx: int, ny: int, nz: int) -> Grid:
    nbx, nby, nbz, *_ = _neighbours(nx, ny, nz)
    N = nx * ny * nz
    geo = np.full(N, GEO_FLUID, dtype=np.uint32)
    return Grid(nx, ny, nz, nbx, nby, nbz, geo, np.arange(N, dtype=np.uint32))


def walled_box(n: tuple, walls: dict, q: float = 0.5, periodic: tuple = (False, False, False)) -> Grid:
    """Fluid block of n = (fx, fy, fz) nodes with one solid layer on every non-periodic side.

    walls maps a side (dP00, dM00, d0P0, d0M0, d00P, d00M) to ("noslip",),
    ("velocity", (Ux, Uy, Uz)) or ("pressure", drho). A link belongs to the side whose
    plane it crosses (z sides first, then y, then x); every wall link gets sub-grid
    distance q. Returns the grid with bcs["noslip"], bcs["velocity"] and bcs["pressure"]
    lists (the pressure list holds the last fluid layer and its inward neighbours kN).
    """
    fx, fy, fz = n
    pad = [0 if p else 1 for p in periodic]
    nx, ny, nz = fx + 2 * pad[0], fy + 2 * pad[1], fz + 2 * pad[2]
    nbx, nby, nbz, x, y, z = _neighbours(nx, ny, nz)
    lo = np.array(pad)
    hi = np.array([nx - 1 - pad[0], ny - 1 - pad[1], nz - 1 - pad[2]])
    fluid_sel = (x >= lo[0]) & (x <= hi[0]) & (y >= lo[1]) & (y <= hi[1]) & (z >= lo[2]) & (z <= hi[2])
    geo = np.where(fluid_sel, GEO_FLUID, GEO_SOLID).astype(np.uint32)
    grid = Grid(nx, ny, nz, nbx, nby, nbz, geo, np.nonzero(fluid_sel)[0].astype(np.uint32))

    kf = np.nonzero(fluid_sel)[0]
    xf, yf, zf = x[kf], y[kf], z[kf]

    def side_of(d):
        tx, ty, tz = xf + CX[d], yf + CY[d], zf + CZ[d]
        s = np.zeros(kf.size, dtype=np.int64)
        if not periodic[0]:
            s = np.where(tx > hi[0], dP00, np.where(tx < lo[0], dM00, s))
        if not periodic[1]:
            s = np.where(ty > hi[1], d0P0, np.where(ty < lo[1], d0M0, s))
        if not periodic[2]:
            s = np.where(tz > hi[2], d00P, np.where(tz < lo[2], d00M, s))
        return s

    sides = np.stack([side_of(d) for d in range(27)])  # 27 x nfluid, 0 = no wall
    for side, spec in walls.items():
        if spec[0] != "pressure":
            continue
        # pressure side: the last fluid layer, extrapolated from the next layer inwards
        ax = (side - 1) // 2
        plus = side % 2 == 1
        coord = (xf, yf, zf)[ax]
        on = coord == (hi[ax] if plus else lo[ax])
        kb = kf[on]
        step = (1 if ax == 0 else (nx if ax == 1 else nx * ny))
        kn = kb - step if plus else kb + step
        rho = np.full(kb.size, spec[1] if len(spec) > 1 else 0.0)
        grid.bcs["pressure"] = BoundaryNodes(k=kb.astype(np.int32), q27=np.full(27 * kb.size, -1.0),
                                             kN=kn.astype(np.int32), rho_bc=rho, direction=side)
    for kind in ("noslip", "velocity"):
        active = np.zeros_like(sides, dtype=bool)
        ux = np.zeros(kf.size)
        uy = np.zeros(kf.size)
        uz = np.zeros(kf.size)
        for side, spec in walls.items():
            if spec[0] != kind:
                continue
            hit = sides == side
            active |= hit
            if kind == "velocity":
                any_hit = hit.any(axis=0)
                ux = np.where(any_hit, spec[1][0], ux)
                uy = np.where(any_hit, spec[1][1], uy)
                uz = np.where(any_hit, spec[1][2], uz)
        nodes = np.nonzero(active.any(axis=0))[0]
        if nodes.size == 0:
            continue
        qarr = np.where(active[:, nodes], q, -1.0).astype(np.float64)
        bc = BoundaryNodes(k=kf[nodes].astype(np.int32), q27=qarr.ravel())
        if kind == "velocity":
            bc.velocity = (ux[nodes], uy[nodes], uz[nodes])
        grid.bcs[kind] = bc
    return grid


def channel(n: tuple, inflow: float, q: float = 0.5) -> Grid:
    """Channel along x: velocity inflow at x-min, pressure (drho = 0) at x-max, no-slip y/z walls."""
    walls = {s: ("noslip",) for s in (d0P0, d0M0, d00P, d00M)}
    walls[dM00] = ("velocity", (inflow, 0.0, 0.0))
    walls[dP00] = ("pressure", 0.0)
    return walled_box(n, walls, q=q)


def cavity(n: int, lid_velocity: float = 0.05) -> Grid:
    """Lid-driven cavity: n^3 fluid nodes, lid moving in +x on the z-max side, no-slip elsewhere."""
    walls = {s: ("noslip",) for s in (dP00, dM00, d0P0, d0M0, d00M)}
    walls[d00P] = ("velocity", (lid_velocity, 0.0, 0.0))
    return walled_box((n, n, n), walls)


def sphere_channel(n: tuple, center: tuple, radius: float, inflow: float, wall_q: float = 0.5) -> Grid:
    """Channel along x (velocity inflow at x-min, pressure outlet at x-max, no-slip y/z walls)
    with a solid sphere; the sphere's links form the "geometry" BC list with the sub-grid
    distance q of each link from the exact ray-sphere intersection.

    n: fluid extent (fx, fy, fz); center/radius in fluid-node coordinates (node (0,0,0) is the
    first fluid node)."""
    g = channel(n, inflow, q=wall_q)
    x, y, z = g.coords()
    # fluid coordinates: the walled box pads one solid layer on each non-periodic side
    cx0, cy0, cz0 = (np.asarray(center, dtype=np.float64) + 1.0)
    X = np.stack([x, y, z]).astype(np.float64)
    C = np.array([cx0, cy0, cz0])
    inside = ((X - C[:, None]) ** 2).sum(0) <= radius * radius
    fluid = (g.geo == GEO_FLUID) & ~inside
    g.geo = np.where(fluid, GEO_FLUID, GEO_SOLID).astype(np.uint32)
    g.fluid = np.nonzero(fluid)[0].astype(np.uint32)
    # drop sphere nodes from the other BC lists (a sphere touching a wall is not supported)
    for bc in g.bcs.values():
        if np.any(inside[bc.k]):
            raise ValueError("the sphere overlaps a boundary layer")
    kf = np.nonzero(fluid)[0]
    q27 = np.full((27, kf.size), -1.0)
    for d in range(1, 27):
        c = np.array([CX[d], CY[d], CZ[d]], dtype=np.float64)
        xn = (x[kf] + CX[d]) % g.nx, (y[kf] + CY[d]) % g.ny, (z[kf] + CZ[d]) % g.nz
        hit = inside[g.index(*xn)]
        if not np.any(hit):
            continue
        rel = X[:, kf[hit]] - C[:, None]
        a = c @ c
        b = 2.0 * (c @ rel)
        cc = (rel * rel).sum(0) - radius * radius
        t = (-b - np.sqrt(np.maximum(b * b - 4 * a * cc, 0.0))) / (2 * a)
        q27[d, hit] = np.clip(t, 0.0, 1.0)
    nodes = np.nonzero((q27 >= 0).any(axis=0))[0]
    g.bcs["geometry"] = BoundaryNodes(k=kf[nodes].astype(np.int32), q27=q27[:, nodes].ravel())
    return g


def taylor_green_fields(grid: Grid, u0: float):
    """Taylor-Green vortex on a periodic box: drho and velocity per node, with node coordinates
    x_i = 2 pi i / nx (k = 2 pi / nx on every axis) and the pressure perturbation of
    apps/gpu/TaylorGreenVortex/TaylorGreenVortex.cpp:143-147:
        u = u0 sin(kx) cos(ky) cos(kz), v = -u0 cos(kx) sin(ky) cos(kz), w = 0,
        drho = 3 (u0^2 / 16) (cos 2kx + cos 2ky) (cos 2kz + 2)."""
    x, y, z = grid.coords()
    k = 2 * np.pi / grid.nx
    sx, cx_ = np.sin(k * x), np.cos(k * x)
    sy, cy_ = np.sin(k * y), np.cos(k * y)
    cz_ = np.cos(k * z)
    ux = u0 * sx * cy_ * cz_
    uy = -u0 * cx_ * sy * cz_
    uz = np.zeros_like(ux)
    drho = 3.0 * (u0 ** 2 / 16.0) * (np.cos(2 * k * x) + np.cos(2 * k * y)) * (np.cos(2 * k * z) + 2.0)
    return drho, ux, uy, uz


def taylor_green_omega(n: int, u0: float = 0.05, re: float = 1600.0) -> float:
    """Relaxation rate of the Taylor-Green case: nu = u0 (n / 2 pi) / Re
    (apps/gpu/TaylorGreenVortex/TaylorGreenVortex.cpp:97-100), omega = 1 / (3 nu + 1/2)."""
    nu = u0 * (n / (2 * np.pi)) / re
    return 1.0 / (3.0 * nu + 0.5)


def gaussian_random_field(shape: tuple, correlation_length: float, seed: int, low: float = 0.0,
                          high: float = 0.1) -> np.ndarray:
    """Seeded Gaussian random field on a periodic (nz, ny, nx) grid: white noise filtered in
    Fourier space with exp(-k^2 l^2 / 2), mapped linearly onto [low, high]; returned flattened
    in node order k = x + nx (y + ny z)."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    noise = rng.normal(0.0, 1.0, (nz, ny, nx))
    kz = np.fft.fftfreq(nz) * 2 * np.pi
    ky = np.fft.fftfreq(ny) * 2 * np.pi
    kx = np.fft.fftfreq(nx) * 2 * np.pi
    K2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    field = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.exp(-0.5 * K2 * correlation_length ** 2)))
    lo, hi = field.min(), field.max()
    return (low + (high - low) * (field - lo) / (hi - lo)).ravel()
