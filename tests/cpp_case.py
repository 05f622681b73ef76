"""Write a case directory for the C++ host-layer test program (tests/cpp/test_host run DIR STEPS).

Raw little-endian arrays per level L (L<L>_nbx.u32, ..._geo.u32, ..._fluid.u32, ..._mask.u8,
BC lists ..._velocity_k.i32/_q27.f32/_vx.f32..., ..._pressure_kN.i32/_rho.f32, interface cells
..._cf_coarse.u32/_cf_fine.u32/_fc_coarse.u32/_fc_fine.u32, initial fields ..._init_rho.f32 ...)
and config.txt with key=value lines. Outputs come back as out_L<L>_{rho,vx,vy,vz}.f32.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_TEST = os.path.join(ROOT, "tests", "cpp", "test_host")
STORAGE_CODE = {"fp32": 0, "fp16": 1, "bf16": 2}


def _w(path, arr, dtype):
    np.ascontiguousarray(arr, dtype=dtype).tofile(path)


def write_case(d, levels, viscosity, limiter=0.01, storage="fp32", velocity_bc=None, noslip_bc=None,
               pressure_bc=False, interfaces=()):
    """levels: [(grid, (drho, ux, uy, uz))]; interfaces: [(cf, fc)] InterfaceCells per level pair."""
    os.makedirs(d, exist_ok=True)
    cfg = {"levels": len(levels), "viscosity": repr(viscosity), "limiter": repr(limiter),
           "storage": STORAGE_CODE[storage]}
    if velocity_bc:
        cfg["velocityBC"] = velocity_bc
    if noslip_bc:
        cfg["noSlipBC"] = noslip_bc
    if pressure_bc:
        cfg["pressureBC"] = "PressureNonEquilibriumCompressible"
    for L, (g, init) in enumerate(levels):
        pre = os.path.join(d, f"L{L}_")
        cfg[f"L{L}_nx"], cfg[f"L{L}_ny"], cfg[f"L{L}_nz"] = g.nx, g.ny, g.nz
        for name, arr in (("nbx", g.nbx), ("nby", g.nby), ("nbz", g.nbz), ("geo", g.geo), ("fluid", g.fluid)):
            _w(pre + name + ".u32", arr, np.uint32)
        if not np.all(g.geo == 19):
            _w(pre + "mask.u8", g.fluid_mask, np.uint8)
        for name, arr in zip(("rho", "vx", "vy", "vz"), init):
            _w(pre + f"init_{name}.f32", arr, np.float32)
        for name, bc in g.bcs.items():
            if bc.n == 0:
                continue
            _w(pre + f"{name}_k.i32", bc.k, np.int32)
            _w(pre + f"{name}_q27.f32", bc.q27, np.float32)
            if bc.velocity is not None:
                for c, v in zip(("vx", "vy", "vz"), bc.velocity):
                    _w(pre + f"{name}_{c}.f32", v, np.float32)
            if bc.kN is not None:
                _w(pre + f"{name}_kN.i32", bc.kN, np.int32)
            if bc.rho_bc is not None:
                _w(pre + f"{name}_rho.f32", bc.rho_bc, np.float32)
            if name == "pressure":
                cfg[f"L{L}_pressure_direction"] = bc.direction
    for L, (cf, fc) in enumerate(interfaces):
        pre = os.path.join(d, f"L{L}_")
        _w(pre + "cf_coarse.u32", cf.coarse, np.uint32)
        _w(pre + "cf_fine.u32", cf.fine, np.uint32)
        _w(pre + "fc_coarse.u32", fc.coarse, np.uint32)
        _w(pre + "fc_fine.u32", fc.fine, np.uint32)
    with open(os.path.join(d, "config.txt"), "w") as f:
        for k, v in cfg.items():
            f.write(f"{k}={v}\n")


def run_case(d, steps, timeout=120, restart=()):
    """restart: () or ("checkpoint", T, FILE, "bin"|"txt") or ("restart", FILE, "bin"|"txt")."""
    r = subprocess.run([HOST_TEST, "run", d, str(steps), *map(str, restart)], capture_output=True, text=True,
                       timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"test_host run failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    return r.stdout


def read_output(d, level):
    pre = os.path.join(d, f"out_L{level}_")
    return tuple(np.fromfile(pre + f"{n}.f32", dtype=np.float32) for n in ("rho", "vx", "vy", "vz"))
