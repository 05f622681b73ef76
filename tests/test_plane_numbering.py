"""The plane-flat numbering of the one-node-per-thread box kernels takes y = t / nx as
trunc((t + 1/2) * (1/nx)) in double precision (vf_lattice.h, plane_row). IEEE double arithmetic
is the same on the host, so the claim that this is exact for every t < 2^32 is checked here on
the cases closest to failing: the first and last column of rows near 2^32, for every nx up to
2^16 and for large ones."""
import numpy as np


def plane_row(t, nx):
    return np.trunc((t.astype(np.float64) + 0.5) * (1.0 / np.float64(nx))).astype(np.int64)


def test_plane_row_is_exact_at_row_ends():
    rng = np.random.default_rng(5)
    for nx in list(range(1, 4097)) + list(rng.integers(4097, 1 << 16, 2000)) + [65535, 65536, 1 << 20, (1 << 30) - 1]:
        nx = int(nx)
        rows = np.concatenate([np.arange(0, 64), rng.integers(0, (1 << 32) // nx, 64), [(1 << 32) // nx - 1]])
        rows = rows[rows * nx + nx - 1 < (1 << 32)]
        for r in (0, nx - 1):
            t = rows * nx + r
            assert np.array_equal(plane_row(t, nx), rows), (nx, r)
