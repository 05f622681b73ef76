"""Device state of one grid level and the per-kernel calls of the time step.

A ``Lattice`` owns, in HBM (torch tensors used as plain device memory):
the single in-place population buffer (27*N, FP32/FP16/BF16), neighbour arrays,
node types, the fluid-node list and the macroscopic output arrays. It fills the
``vf_lbm_parameter`` struct of the C ABI and calls the HIP kernels through it.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import capi
from .grid import Grid

STORAGE = {
    "fp32": (capi.VF_STORAGE_FP32, torch.float32),
    "fp16": (capi.VF_STORAGE_FP16, torch.float16),
    "bf16": (capi.VF_STORAGE_BF16, torch.bfloat16),
}


def omega_from_viscosity(nu: float) -> float:
    """omega = 1 / (3 nu + 1/2) (src/gpu/core/Parameter/Parameter.cpp:400-418)."""
    return 1.0 / (3.0 * nu + 0.5)


def _u32(a: np.ndarray, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).to(device)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def nodes_along_rows(k) -> bool:
    """Whether consecutive BC nodes are, for the most part, neighbours in x (a wall normal to y or
    z): the layout hint vf_q_boundary_conditions.nodesAlongRows of the adjoint kernels."""
    k = np.asarray(k, dtype=np.int64)
    return k.size > 1 and 2 * int(np.count_nonzero(np.diff(k) == 1)) >= k.size - 1


class DeviceBC:
    """A boundary-condition node list resident on the device (QforBoundaryConditions)."""

    def __init__(self, nodes, device, number_of_nodes: int | None = None):
        k = np.asarray(nodes.k)
        if number_of_nodes is not None and k.size and (k.min() < 0 or k.max() >= number_of_nodes):
            raise ValueError("BC node index out of range")
        # the BC kernels write each (node, direction) cell from one thread: a node listed twice
        # would be written by two threads in the same launch
        if np.unique(k).size != k.size:
            raise ValueError("a BC node is listed twice")
        self.n = nodes.n
        self.k = torch.from_numpy(nodes.k.astype(np.int32)).to(device)
        self.q27 = torch.from_numpy(nodes.q27.astype(np.float32)).to(device)
        self.vel = None
        if nodes.velocity is not None:
            self.vel = [torch.from_numpy(np.asarray(v, dtype=np.float32)).to(device) for v in nodes.velocity]
        self.kN = None if nodes.kN is None else torch.from_numpy(nodes.kN.astype(np.int32)).to(device)
        # whether no two nodes share a neighbour kN (the pressure adjoint then needs no atomics)
        self.kN_distinct = nodes.kN is None or np.unique(np.asarray(nodes.kN)).size == np.asarray(nodes.kN).size
        self.rho_bc = None if nodes.rho_bc is None else torch.from_numpy(nodes.rho_bc.astype(np.float32)).to(device)
        self.direction = nodes.direction
        self.struct = capi.QBoundaryConditions()
        s = self.struct
        s.k = _ptr(self.k)
        s.kN = _ptr(self.kN)
        s.q27 = _ptr(self.q27)
        s.numberOfBCnodes = self.n
        if self.vel is not None:
            s.Vx, s.Vy, s.Vz = (_ptr(v) for v in self.vel)
        s.RhoBC = _ptr(self.rho_bc)
        s.direction = self.direction
        s.nodesAlongRows = int(nodes_along_rows(k))


class Lattice:
    """One grid level on the GPU."""

    def __init__(self, grid: Grid, omega: float, storage: str = "fp32", limiters=(0.01, 0.01, 0.01),
                 forces=(0.0, 0.0, 0.0), level: int = 0, box: bool | None = None, device="cuda"):
        if storage not in STORAGE:
            raise ValueError(f"storage must be one of {sorted(STORAGE)}")
        self.grid = grid
        self.device = torch.device(device)
        self.storage = storage
        code, dtype = STORAGE[storage]
        N = grid.N
        self.N = N
        self.f = torch.zeros(27 * N, dtype=dtype, device=self.device)
        self.nbx = _u32(grid.nbx, self.device)
        self.nby = _u32(grid.nby, self.device)
        self.nbz = _u32(grid.nbz, self.device)
        self.geo = _u32(grid.geo, self.device)
        self.fluid = _u32(grid.fluid, self.device)
        all_fluid = bool(np.all(grid.geo == 19))
        self.mask = None if all_fluid else torch.from_numpy(grid.fluid_mask).to(self.device)
        self.rho = torch.zeros(N, dtype=torch.float32, device=self.device)
        self.vx = torch.zeros_like(self.rho)
        self.vy = torch.zeros_like(self.rho)
        self.vz = torch.zeros_like(self.rho)
        self.press = torch.zeros_like(self.rho)
        self.sums = torch.zeros(5, dtype=torch.float64, device=self.device)
        # the dense-box path needs no neighbour loads; default to it for boxes
        self.box = True if box is None else box
        self.bcs = {name: DeviceBC(nodes, self.device, N) for name, nodes in grid.bcs.items()}

        p = capi.LBMParameter()
        p.distributions = _ptr(self.f)
        p.storage = code
        p.isEvenTimestep = 1
        p.numberOfNodes = N
        # a box lattice addresses every kernel (BCs, interface, reductions) by (x, y, z): its
        # neighbour arrays are the periodic box's, and recomputing them saves two dependent loads
        # per node; the indexed lattice hands the arrays to every kernel
        if not self.box:
            p.neighborX, p.neighborY, p.neighborZ = _ptr(self.nbx), _ptr(self.nby), _ptr(self.nbz)
        p.typeOfGridNode = _ptr(self.geo)
        p.nx, p.ny, p.nz = grid.nx, grid.ny, grid.nz
        p.fluidMask = _ptr(self.mask)
        p.omega = omega
        p.quadricLimiters[:] = list(limiters)
        p.forces[:] = list(forces)
        p.forceFactor = 1.0 / (2 ** level)
        p.rho, p.velocityX, p.velocityY, p.velocityZ = (_ptr(t) for t in (self.rho, self.vx, self.vy, self.vz))
        p.pressure = _ptr(self.press)
        p.fluidNodeIndices = _ptr(self.fluid)
        p.numberOfFluidNodes = int(grid.fluid.size)
        self.param = p

    def set_porosity(self, porosity):
        """Brinkman porosity per node (numpy or torch, N values; None removes it): a drag
        acceleration -porosity * u applied through the collision's velocity shift."""
        if porosity is None:
            self.porosity = None
            self.param.porosity = None
            return
        t = torch.as_tensor(porosity, dtype=torch.float32).reshape(-1)
        if t.numel() != self.N:
            raise ValueError("porosity needs one value per node")
        if getattr(self, "porosity", None) is None:
            self.porosity = torch.zeros(self.N, dtype=torch.float32, device=self.device)
        self.porosity.copy_(t)
        self.param.porosity = _ptr(self.porosity)

    @property
    def state_bytes(self) -> int:
        """Bytes of the population buffer (27 values per node)."""
        return self.f.numel() * self.f.element_size()

    def set_working_set(self, nbytes: int):
        """Bytes of lattice state the caller's time step cycles through between two collisions
        of this lattice (every level of a hierarchy; 0: this lattice's own). The dense-box
        collision takes the streaming row kernel when it exceeds the 256 MB infinity cache and
        the one-node-per-thread kernel otherwise."""
        self.param.workingSetBytes = int(nbytes)

    # ---- parity -------------------------------------------------------------------------
    @property
    def even(self) -> bool:
        return bool(self.param.isEvenTimestep)

    @even.setter
    def even(self, v: bool):
        self.param.isEvenTimestep = 1 if v else 0

    def swap(self):
        """swapBetweenEvenAndOddTimestep (src/gpu/core/Calculation/UpdateGrid27.cpp:265-269)."""
        self.even = not self.even

    @property
    def omega(self) -> float:
        return float(self.param.omega)

    # ---- kernels -------------------------------------------------------------------------
    def pref(self):
        return ctypes.byref(self.param)

    def set_body_force(self, bfx, bfy, bfz):
        """Per-node body force read (and reset to zero) by the ApplyBodyForce, AllFeatures and
        SubDomainBorder collision templates (numpy or torch, N values each; None removes it)."""
        if bfx is None:
            self.body_force = None
            self.param.bodyForceX = self.param.bodyForceY = self.param.bodyForceZ = None
            return
        self.body_force = [torch.as_tensor(np.asarray(b, dtype=np.float32) if not torch.is_tensor(b) else b,
                                           dtype=torch.float32).reshape(-1).to(self.device).clone()
                           for b in (bfx, bfy, bfz)]
        if any(b.numel() != self.N for b in self.body_force):
            raise ValueError("body force needs one value per node")
        self.param.bodyForceX, self.param.bodyForceY, self.param.bodyForceZ = (_ptr(b) for b in self.body_force)

    def collide(self, write_macro: bool = False, stream=None, template: int | None = None):
        """One collide + stream step. template: a CollisionTemplate (capi.COLLISION_*); the
        default is Default, or WriteMacroVars with write_macro. The body-force templates run on
        the fluid-node list (runOnIndices)."""
        tmpl = capi.COLLISION_WRITE_MACRO_VARS if write_macro else capi.COLLISION_DEFAULT
        if template is not None:
            tmpl = int(template)
        if tmpl in (capi.COLLISION_APPLY_BODY_FORCE, capi.COLLISION_ALL_FEATURES, capi.COLLISION_SUBDOMAIN_BORDER):
            if self.box:
                raise ValueError("the body-force templates need the indexed lattice (box=False)")
        if self.box:
            capi.call("vf_k17_run_box", self.pref(), tmpl, stream_handle(stream))
        else:
            capi.call("vf_k17_run_on_indices", self.pref(), _ptr(self.fluid), int(self.grid.fluid.size), tmpl,
                      stream_handle(stream))

    def apply_bc(self, fn_name: str, name: str, stream=None):
        bc = self.bcs.get(name)
        if bc is None or bc.n == 0:
            return
        capi.call(fn_name, self.pref(), ctypes.byref(bc.struct), stream_handle(stream))

    def init_equilibrium(self, rho, ux, uy, uz, stream=None):
        """Shifted equilibrium from drho and velocity (numpy or torch, N values each)."""
        for dst, src in ((self.rho, rho), (self.vx, ux), (self.vy, uy), (self.vz, uz)):
            dst.copy_(torch.as_tensor(np.asarray(src, dtype=np.float32) if not torch.is_tensor(src) else src,
                                      dtype=torch.float32))
        capi.call("vf_InitNavierStokesCompressible", self.pref(), stream_handle(stream))

    def init_non_equilibrium(self, rho, ux, uy, uz, stream=None):
        """Equilibrium plus the velocity-gradient non-equilibrium part (InitNeq), from drho and
        velocity (numpy or torch, N values each)."""
        for dst, src in ((self.rho, rho), (self.vx, ux), (self.vy, uy), (self.vz, uz)):
            dst.copy_(torch.as_tensor(np.asarray(src, dtype=np.float32) if not torch.is_tensor(src) else src,
                                      dtype=torch.float32))
        if not hasattr(self, "nbinv"):
            self.nbinv = _u32(self.grid.neighbor_inverse(), self.device)
        capi.call("vf_InitNavierStokesCompressibleNonEquilibrium", self.pref(), _ptr(self.nbinv),
                  stream_handle(stream))

    def macroscopic(self, stream=None):
        """drho, vx, vy, vz, p from the pre-collision state (host numpy arrays)."""
        capi.call("vf_CalculateMacroscopicQuantitiesCompressible", self.pref(), stream_handle(stream))
        torch.cuda.synchronize(self.device)
        return tuple(t.cpu().numpy() for t in (self.rho, self.vx, self.vy, self.vz, self.press))

    def moment_sums(self, stream=None) -> np.ndarray:
        """FP64 sums over fluid nodes: drho, jx, jy, jz, kinetic energy."""
        capi.call("vf_MomentSums", self.pref(), _ptr(self.sums), stream_handle(stream))
        torch.cuda.synchronize(self.device)
        return self.sums.cpu().numpy().copy()

    def drag_lift_half(self, name: str, post: bool, stream=None) -> np.ndarray:
        """Per-node momentum-exchange half on the BC list `name` (DragLiftPost27 after the
        collision, DragLiftPre27 after the BC kernels and the swap): host array (3, nbc)."""
        bc = self.bcs[name]
        out = torch.zeros((3, max(bc.n, 1)), dtype=torch.float64, device=self.device)
        fn = "vf_DragLiftPostD27" if post else "vf_DragLiftPreD27"
        capi.call(fn, self.pref(), ctypes.byref(bc.struct), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                  stream_handle(stream))
        torch.cuda.synchronize(self.device)
        return out[:, :bc.n].cpu().numpy()

    def drag_force(self, name: str, stream=None) -> np.ndarray:
        """Force on the BC list `name` by momentum exchange (sum of pre - post halves), from
        the state after a step: FP64 (3,)."""
        bc = self.bcs[name]
        out = torch.zeros(3, dtype=torch.float64, device=self.device)
        capi.call("vf_DragLiftSum", self.pref(), ctypes.byref(bc.struct), _ptr(out), stream_handle(stream))
        torch.cuda.synchronize(self.device)
        return out.cpu().numpy()

    def kinetic_energy(self, stream=None):
        """KineticEnergyAnalyzer: (sum of 1/2 |u|^2 (1 + drho), node count) over fluid nodes whose
        six face neighbours and (-1,-1,-1) neighbour are fluid, from the macroscopic fields of the
        current state."""
        capi.call("vf_CalculateMacroscopicQuantitiesCompressible", self.pref(), stream_handle(stream))
        if not hasattr(self, "nbinv"):
            self.nbinv = _u32(self.grid.neighbor_inverse(), self.device)
        out = torch.zeros(2, dtype=torch.float64, device=self.device)
        capi.call("vf_KineticEnergy", self.pref(), _ptr(self.nbinv), _ptr(out), stream_handle(stream))
        torch.cuda.synchronize(self.device)
        e, c = out.cpu().numpy()
        return float(e), int(round(c))

    def populations(self) -> np.ndarray:
        """The raw 27*N buffer as FP32 numpy (slot layout of the current parity)."""
        torch.cuda.synchronize(self.device)
        return self.f.float().cpu().numpy()

    def set_populations(self, host: np.ndarray):
        self.f.copy_(torch.from_numpy(np.ascontiguousarray(host, dtype=np.float32)).to(self.f.dtype))
