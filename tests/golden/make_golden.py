"""Write the golden vectors of the K17 path from the reference's own arithmetic.

Runs only where oracle/_ref/ was built from the reference tree's src/lbm headers
(`make -C oracle ref`). Writes tests/golden/k17_golden.npz: seeded near-equilibrium
shifted populations, including nodes whose third-order moments are large enough to
engage the quadric limiters, and the reference's post-collision populations and
macroscopic values in FP64 and FP32 for several relaxation rates and limiter settings.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

SEED = 20261202
CASES = [  # (omega, limiter)
    (1.2, 0.01),
    (1.0 / (3 * 0.032 + 0.5), 0.01),  # cavity, Re = 100 at 64^3
    (1.99, 1.0e6),                    # Taylor-Green setting: limiter switched off
    (1.8, 1.0e-4),                    # strongly limited
]


def near_equilibrium(rng, n, u_amp, noise):
    from virtualfluids_amd.grid import CX, CY, CZ, W

    drho = rng.normal(0.0, 0.01, n)
    u = rng.normal(0.0, u_amp, (n, 3))
    cu = u[:, :1] * CX + u[:, 1:2] * CY + u[:, 2:3] * CZ
    usq = (u ** 2).sum(1)[:, None]
    f = W * (drho[:, None] + (1 + drho[:, None]) * (3 * cu + 4.5 * cu ** 2 - 1.5 * usq))
    return f + W * rng.normal(0.0, noise, (n, 27))


def main():
    if oracle.ref("f64") is None or oracle.ref("f32") is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref needs the reference tree)")
    rng = np.random.default_rng(SEED)
    fin = np.concatenate([near_equilibrium(rng, 192, 0.03, 1e-3), near_equilibrium(rng, 64, 0.08, 2e-2)])
    out = {"fin": fin, "omega": np.array([c[0] for c in CASES]), "limiter": np.array([c[1] for c in CASES])}
    for i, (om, lim) in enumerate(CASES):
        lims = (lim, lim, lim)
        f64, m64 = oracle.ref_k17_nodes(fin, om, lims, precision="f64")
        f32, m32 = oracle.ref_k17_nodes(fin.astype(np.float32), om, lims, precision="f32")
        out[f"fout64_{i}"], out[f"macro64_{i}"] = f64, m64
        out[f"fout32_{i}"], out[f"macro32_{i}"] = f32, m32
    # refinement cells: 8 source nodes each, with and without wall offsets, CF at all 8 fine
    # positions and FC at the centre
    cells = near_equilibrium(rng, 8 * 24, 0.04, 3e-3).reshape(24, 8, 27)
    offs = np.array([(0.0, 0.0, 0.0)] * 12 + [tuple(rng.choice([-0.5, 0.0, 0.5], 3)) for _ in range(12)])
    cf_pos = np.array([((q & 1) - 0.5) / 2 for q in range(8)]), np.array([((q >> 1 & 1) - 0.5) / 2 for q in range(8)]), \
        np.array([((q >> 2 & 1) - 0.5) / 2 for q in range(8)])
    coef, fcf, ffc = [], [], []
    for i in range(24):
        row = []
        for q in range(8):
            c, f = oracle.ref_interp_cell(cells[i], 1.7, offs[i], 0, 1.85, 0.5, (cf_pos[0][q], cf_pos[1][q], cf_pos[2][q]))
            row.append(f)
        coef.append(c)
        fcf.append(row)
        c2, f2 = oracle.ref_interp_cell(cells[i], 1.85, offs[i], 1, 1.7, 2.0)
        ffc.append(f2)
    out["refine_cells"], out["refine_offsets"] = cells, offs
    out["refine_coeffs"], out["refine_cf"], out["refine_fc"] = np.array(coef), np.array(fcf), np.array(ffc)
    mom = np.stack([oracle.ref_moments(f) for f in fin[:32]])
    out["moments_in"] = fin[:32]
    out["moments64"] = mom
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "k17_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
