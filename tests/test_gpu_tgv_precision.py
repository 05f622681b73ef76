"""Taylor-Green vortex, Re 1600, u0 0.05: FP16 and BF16 shifted storage against FP32 storage
and against the FP64 CPU oracle (the precision study of the 512^3 storage comparison).

* Same resolution, 128^3, 500 steps: GPU FP32 / FP16 / BF16 velocity fields against the FP64
  oracle on the same lattice: the velocity error the storage type adds (relative L2 over all
  nodes, and the kinetic-energy decay curve).
* 512^3, 2000 steps (FP32, FP16, BF16 storage on the GPU) against the FP64 oracle at 128^3 and
  500 steps, the same convective time (u0 fixed, 2000 / 512 = 500 / 128): every 4th node of
  the 512^3 lattice sits on a 128^3 node. The FP32 run measures the resolution difference; the
  FP16 and BF16 runs must stay as close to FP64 as FP32 does, and close to FP32 itself.

Node coordinates x_i = 2 pi i / n, the quadric limiter off (as in the Taylor-Green app).
"""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U0 = 0.05
NO_LIMITER = (1e6, 1e6, 1e6)


def _oracle_tgv(orc, n, steps, ke_every):
    from virtualfluids_amd.grid import periodic_box, taylor_green_fields, taylor_green_omega

    grid = periodic_box(n, n, n)
    omega = taylor_green_omega(n, U0)
    L = orc.Lattice(grid.nbx, grid.nby, grid.nbz)
    L.init_eq(*taylor_green_fields(grid, U0))
    ke = []

    def energy():
        drho, vx, vy, vz, _ = L.macroscopic()
        return float((0.5 * (vx * vx + vy * vy + vz * vz) * (1.0 + drho)).sum() / grid.N)

    ke.append(energy())
    for s in range(1, steps + 1):
        orc.collide_box(L.f, n, n, n, omega, L.even, limiters=NO_LIMITER)
        L.swap()
        if s % ke_every == 0:
            ke.append(energy())
    _, vx, vy, vz, _ = L.macroscopic()
    return np.stack([vx, vy, vz]), np.array(ke)


def _gpu_tgv(n, storage, steps, ke_every, fields=None, sample=1, grid=None):
    import torch

    from virtualfluids_amd.grid import periodic_box, taylor_green_fields, taylor_green_omega
    from virtualfluids_amd.lattice import Lattice

    grid = grid if grid is not None else periodic_box(n, n, n)
    L = Lattice(grid, taylor_green_omega(n, U0), storage=storage, limiters=NO_LIMITER)
    L.init_equilibrium(*(fields if fields is not None else taylor_green_fields(grid, U0)))
    ke = [L.kinetic_energy()[0] / grid.N]
    for s in range(1, steps + 1):
        L.collide()
        L.swap()
        if s % ke_every == 0:
            ke.append(L.kinetic_energy()[0] / grid.N)
    L.kinetic_energy()  # leaves the macroscopic fields of the final state in L.vx, L.vy, L.vz
    u = torch.stack([t.view(n, n, n)[::sample, ::sample, ::sample].reshape(-1) for t in (L.vx, L.vy, L.vz)])
    out = u.double().cpu().numpy(), np.array(ke)
    del L
    torch.cuda.empty_cache()
    return out


def _rel_l2(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.mark.timeout(900)
def test_tgv_storage_precision_against_fp64(gpu, orc):
    from virtualfluids_amd.grid import periodic_box, taylor_green_fields

    report = {}
    # same resolution: the storage type's own velocity error against FP64
    u64, ke64 = _oracle_tgv(orc, 128, 500, 10)
    same = {}
    decay64 = 1.0 - ke64[-1] / ke64[0]
    for storage in ("fp32", "fp16", "bf16"):
        u, ke = _gpu_tgv(128, storage, 500, 10)
        same[storage] = {"u_rel_l2": _rel_l2(u, u64),
                         "ke_max_rel": float(np.abs(ke / ke64 - 1.0).max()),
                         "ke_decay": float(1.0 - ke[-1] / ke[0]),
                         "ke_decay_rel_err_vs_fp64": float(abs((1.0 - ke[-1] / ke[0]) / decay64 - 1.0))}
    same["fp64_ke_decay"] = float(decay64)
    report["128^3_500_steps_vs_fp64_128^3"] = same
    print(json.dumps(report), flush=True)
    # 512^3 at the same convective time
    grid = periodic_box(512, 512, 512)
    fields = [np.asarray(a, dtype=np.float32) for a in taylor_green_fields(grid, U0)]
    big = {}
    runs = {}
    for storage in ("fp32", "fp16", "bf16"):
        runs[storage] = _gpu_tgv(512, storage, 2000, 40, fields=fields, sample=4, grid=grid)
        print(f"512^3 {storage} done", flush=True)
    del grid, fields
    for storage, (u, ke) in runs.items():
        decay = 1.0 - ke[-1] / ke[0]
        big[storage] = {"u_rel_l2_vs_fp64_128": _rel_l2(u, u64),
                        "u_rel_l2_vs_fp32_512": _rel_l2(u, runs["fp32"][0]),
                        "ke_max_rel_vs_fp32_512": float(np.abs(ke / runs["fp32"][1] - 1.0).max()),
                        "ke_end_over_start": float(ke[-1] / ke[0]),
                        # the storage noise against the physical decay: resolved when the decay
                        # has the sign and half the size of the FP32 run's
                        "resolves_ke_decay": bool(decay > 0.5 * (1.0 - runs["fp32"][1][-1] / runs["fp32"][1][0]))}
    big["fp64_128_ke_end_over_start"] = float(ke64[-1] / ke64[0])
    report["512^3_2000_steps"] = big
    print(json.dumps(report), flush=True)

    # measured on MI355X: 128^3 / 500 steps against FP64, relative L2 of u: FP32 3.1e-5,
    # FP16 2.6e-3, BF16 5.3e-2; 512^3 / 2000 steps against FP32 storage: FP16 2.4e-2, BF16
    # 9.2e-2 (per-step storage rounding of 2^-11 / 2^-8 accumulating as a random walk), FP32
    # against FP64 at 128^3 8.6e-4 (resolution). The bounds leave a factor of 2-4.
    assert same["fp32"]["u_rel_l2"] < 1e-4
    assert same["fp16"]["u_rel_l2"] < 1e-2
    assert same["bf16"]["u_rel_l2"] < 1.5e-1
    assert same["fp32"]["ke_max_rel"] < 2e-4
    # at 128^3 FP16 storage resolves the kinetic-energy decay (0.46 % over 500 steps): the sign,
    # and the size within 10 % of FP64's (FP32 within 1 %); BF16 does not (its energy error of
    # ~5e-3 is the size of the decay itself)
    assert decay64 > 0
    assert same["fp32"]["ke_decay"] > 0 and same["fp32"]["ke_decay_rel_err_vs_fp64"] < 1e-2
    assert same["fp16"]["ke_decay"] > 0 and same["fp16"]["ke_decay_rel_err_vs_fp64"] < 1e-1
    # at 512^3 and 2000 steps the FP32 run resolves the same 0.46 % decay; FP16 storage does not
    # (its rounding noise adds energy and the curve ends above its start: reported as
    # resolves_ke_decay, not asserted); FP16 stays within 2 % of FP32 in kinetic energy
    assert big["fp32"]["resolves_ke_decay"] and abs(big["fp32"]["ke_end_over_start"] - ke64[-1] / ke64[0]) < 2e-3
    assert big["fp32"]["u_rel_l2_vs_fp64_128"] < 3e-3
    assert big["fp16"]["u_rel_l2_vs_fp32_512"] < 5e-2
    assert big["bf16"]["u_rel_l2_vs_fp32_512"] < 2e-1
    assert big["fp16"]["ke_max_rel_vs_fp32_512"] < 2e-2
    assert big["bf16"]["ke_max_rel_vs_fp32_512"] < 4e-2
