"""Two-level grid refinement: interface cells, the scaling factory and the recursive time step.

Mirrors, on the host:
  GridScalingFactory        src/gpu/core/GridScaling/GridScalingFactory.h:42-66 (enum GridScaling,
                            setScalingFactory, getGridScalingFC/CF; cpp:38-69)
  GridScalingKernelManager  src/gpu/core/GridScaling/GridScalingKernelManager.cpp:45-75 and the
                            check of GridScalingKernelManager.h:77-84
  UpdateGrid27::updateGrid  src/gpu/core/Calculation/UpdateGrid27.cpp:54-96: the finer level takes
                            two steps, then the coarser level steps and the interface is refreshed
                            (fine -> coarse, then coarse -> fine).

Interface geometry of a nested box (GridInterface::findInterfaceCF/FC,
src/gpu/GridGenerator/grid/GridInterface.cpp:60-167). Along each axis the fine level covers
the coarse cells [s, e]; fine node j sits at coarse coordinate s + 1/4 + j/2, j = 0..2(e-s)-1.
  CF cells  coarse base c in [s, e-1]^3 on the surface of that block; the fine cell at
            c + 1/4 (fine base j = 2 (c - s)). Their 8 coarse nodes become "CFC" nodes.
  FC cells  coarse nodes c in [s+2, e-2]^3 on the surface of that block (the non-CFC nodes
            next to a CFC node); the fine cell at c - 1/4 (fine base j = 2 (c - s) - 1).
  Coarse nodes strictly inside the FC shell are not computed (STOPPER/INVALID_COARSE_UNDER_FINE):
  they are marked solid so the collision skips them.
The two outer fine layers (the CF cells) are overwritten once per coarse step; after the fine
level's two sub-steps only they are affected by what lies outside the fine box, so the FC cells
read fine nodes from the third layer inwards.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass

import numpy as np
import torch

from . import capi
from .grid import GEO_FLUID, GEO_SOLID, Grid, periodic_box


@dataclass
class InterfaceCells:
    """InterpolationCells + ICellNeigh on the host: cell base indices on both levels and the
    per-cell offsets of the interpolation point (zero away from walls)."""

    coarse: np.ndarray
    fine: np.ndarray
    offset: tuple | None = None

    @property
    def n(self) -> int:
        return int(self.coarse.size)


@dataclass
class NestedBox:
    coarse: Grid
    fine: Grid
    cf: InterfaceCells
    fc: InterfaceCells
    lo: tuple
    hi: tuple

    def fine_coords(self):
        """Fine node coordinates in coarse lattice units."""
        x, y, z = self.fine.coords()
        return tuple(s + 0.25 + 0.5 * np.asarray(a) for s, a in zip(self.lo, (x, y, z)))

    def coarse_valid(self) -> np.ndarray:
        """Coarse nodes that hold the solution (not computed under the fine level: solid)."""
        return self.coarse.geo == GEO_FLUID


def _surface(lo, hi, coords):
    """Points of the block [lo, hi]^3 (inclusive) that lie on its surface."""
    inside = np.ones(coords[0].size, dtype=bool)
    on = np.zeros(coords[0].size, dtype=bool)
    for a in range(3):
        inside &= (coords[a] >= lo[a]) & (coords[a] <= hi[a])
        on |= (coords[a] == lo[a]) | (coords[a] == hi[a])
    return inside & on


def nested_box(nc: tuple, lo: tuple, hi: tuple, coarse: Grid | None = None) -> NestedBox:
    """A fine box covering the coarse cells [lo, hi] inside a coarse grid: a periodic box of nc
    nodes, or the given dense coarse grid (a channel with its boundary lists, for instance),
    whose nodes and boundary nodes around the fine box must all be fluid."""
    lo = tuple(int(v) for v in lo)
    hi = tuple(int(v) for v in hi)
    if coarse is not None:
        nc = (coarse.nx, coarse.ny, coarse.nz)
    nc = tuple(int(v) for v in nc)
    for a in range(3):
        if hi[a] - lo[a] < 4:
            raise ValueError("the fine box must span at least 4 coarse cells per axis")
        if lo[a] < 1 or hi[a] > nc[a] - 2:
            raise ValueError("the fine box must keep one coarse layer clear of the periodic wrap")
    if coarse is None:
        coarse = periodic_box(*nc)
    mf = tuple(2 * (hi[a] - lo[a]) for a in range(3))
    fine = periodic_box(*mf)
    cx, cy, cz = coarse.coords()
    cc = (cx, cy, cz)

    # the interface reads one coarse layer around [lo, hi]: it must be plain fluid
    near = np.ones(coarse.N, dtype=bool)
    for a in range(3):
        near &= (cc[a] >= lo[a] - 1) & (cc[a] <= hi[a] + 1)
    if np.any(coarse.geo[near] != GEO_FLUID):
        raise ValueError("the fine box and its surrounding coarse layer must be fluid")
    for name, bc in coarse.bcs.items():
        if bc.n and np.any(near[bc.k]):
            raise ValueError(f"the {name} boundary reaches the fine box")

    # coarse nodes under the fine level that no coarse step reads: not computed
    under = np.ones(coarse.N, dtype=bool)
    for a in range(3):
        under &= (cc[a] >= lo[a] + 3) & (cc[a] <= hi[a] - 3)
    coarse.geo = np.where(under, GEO_SOLID, coarse.geo).astype(np.uint32)
    coarse.fluid = np.nonzero(coarse.geo == GEO_FLUID)[0].astype(np.uint32)

    cf_sel = _surface(lo, tuple(h - 1 for h in hi), cc)
    kc = np.nonzero(cf_sel)[0]
    jf = [2 * (cc[a][kc] - lo[a]) for a in range(3)]
    cf = InterfaceCells(kc.astype(np.uint32), fine.index(*jf).astype(np.uint32))

    fc_sel = _surface(tuple(l + 2 for l in lo), tuple(h - 2 for h in hi), cc)
    kc = np.nonzero(fc_sel)[0]
    jf = [2 * (cc[a][kc] - lo[a]) - 1 for a in range(3)]
    fc = InterfaceCells(kc.astype(np.uint32), fine.index(*jf).astype(np.uint32))
    return NestedBox(coarse, fine, cf, fc, lo, hi)


class DeviceInterfaceCells:
    """Interface cells resident on the device, with the C-ABI structs that point at them."""

    def __init__(self, cells: InterfaceCells, device):
        def u32(a):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).to(device)

        self.n = cells.n
        self.coarse = u32(cells.coarse)
        self.fine = u32(cells.fine)
        self.struct = capi.InterpolationCells()
        self.struct.coarseCellIndices = ctypes.c_void_p(self.coarse.data_ptr())
        self.struct.fineCellIndices = ctypes.c_void_p(self.fine.data_ptr())
        self.struct.numberOfCells = self.n
        self.neigh = capi.InterpolationCellNeighbor()
        self.offsets = None
        if cells.offset is not None:
            self.offsets = [torch.from_numpy(np.asarray(o, dtype=np.float32)).to(device) for o in cells.offset]
            self.neigh.x, self.neigh.y, self.neigh.z = (ctypes.c_void_p(t.data_ptr()) for t in self.offsets)
        self.device = device
        self.host = cells
        self._sources = None

    def sources(self, grid):
        """vf_interpolation_sources of the coarse cells (corner q of cell i: q's bits step +x, +y,
        +z through the grid's neighbour arrays), built once: (struct, workspace floats)."""
        if self._sources is None:
            base = np.asarray(self.host.coarse, dtype=np.int64)
            corners = np.empty((self.n, 8), dtype=np.int64)
            for q in range(8):
                k = base.copy()
                if q & 1:
                    k = grid.nbx[k].astype(np.int64)
                if q & 2:
                    k = grid.nby[k].astype(np.int64)
                if q & 4:
                    k = grid.nbz[k].astype(np.int64)
                corners[:, q] = k
            lanes_flat = corners.reshape(-1)  # entry 8 i + q
            order = np.argsort(lanes_flat, kind="stable")
            nodes, first_idx = np.unique(lanes_flat[order], return_index=True)
            first = np.append(first_idx, order.size).astype(np.uint32)

            def u32(a):
                return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).to(self.device)

            t = (u32(nodes), u32(first), u32(order))
            st = capi.InterpolationSources()
            st.node, st.first, st.lanes = (ctypes.c_void_p(x.data_ptr()) for x in t)
            st.numberOfNodes = int(nodes.size)
            self._sources = (st, t, 9 * 8 * self.n)
        return self._sources[0], self._sources[2]


class GridScalingFactory:
    """Which interpolation the refinement interface uses (GridScalingFactory.h:42-66)."""

    class GridScaling(enum.Enum):
        ScaleCompressible = 0
        NotSpecified = 1

    def __init__(self):
        self.grid_scaling_type = self.GridScaling.NotSpecified

    def set_scaling_factory(self, grid_scaling_type: "GridScalingFactory.GridScaling"):
        self.grid_scaling_type = grid_scaling_type

    def _get(self, name: str, has_turbulent_viscosity: bool):
        if self.grid_scaling_type != self.GridScaling.ScaleCompressible:
            return None
        if has_turbulent_viscosity:
            raise NotImplementedError("scaling with turbulent viscosity is not part of the K17 compressible path")
        return name

    def get_grid_scaling_fc(self, has_turbulent_viscosity: bool = False):
        return self._get("vf_ScaleFineToCoarseCompressible", has_turbulent_viscosity)

    def get_grid_scaling_cf(self, has_turbulent_viscosity: bool = False):
        return self._get("vf_ScaleCoarseToFineCompressible", has_turbulent_viscosity)


class GridScalingKernelManager:
    """Checks and runs the interface kernels (GridScalingKernelManager.cpp:45-75)."""

    def __init__(self, number_of_levels: int, factory: GridScalingFactory | None, fc_cells: int = 0,
                 cf_cells: int = 0, use_turbulent_viscosity: bool = False):
        self.scaling_fine_to_coarse = None
        self.scaling_coarse_to_fine = None
        if number_of_levels > 1:
            if factory is None:
                raise RuntimeError("There is more than one level, but no scalingFactory was provided.")
            fc = factory.get_grid_scaling_fc(use_turbulent_viscosity)
            cf = factory.get_grid_scaling_cf(use_turbulent_viscosity)
            self._check(fc, fc_cells, "scalingFineToCoarse")
            self._check(cf, cf_cells, "scalingCoarseToFine")
            self.scaling_fine_to_coarse = fc
            self.scaling_coarse_to_fine = cf

    @staticmethod
    def _check(fn, cells: int, name: str):
        if fn is None and cells > 0:
            raise RuntimeError(f"The scaling function {name} was not set!")

    @staticmethod
    def _run(fn, coarse, fine, cells: DeviceInterfaceCells, stream):
        from .lattice import stream_handle

        neigh = ctypes.byref(cells.neigh) if cells.offsets is not None else None
        capi.call(fn, coarse.pref(), fine.pref(), ctypes.byref(cells.struct), neigh, stream_handle(stream))

    def run_fine_to_coarse(self, coarse, fine, cells: DeviceInterfaceCells, stream=None):
        if cells.n:
            self._run(self.scaling_fine_to_coarse, coarse, fine, cells, stream)

    def run_coarse_to_fine(self, coarse, fine, cells: DeviceInterfaceCells, stream=None):
        if cells.n:
            self._run(self.scaling_coarse_to_fine, coarse, fine, cells, stream)
