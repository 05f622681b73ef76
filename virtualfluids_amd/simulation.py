"""Time-step driver with the reference's kernel ordering and factory semantics.

``BoundaryConditionFactory`` mirrors src/gpu/core/BoundaryConditions/BoundaryConditionFactory.h:55-174
(same enum names and defaults: velocity/pressure NotSpecified -> no function, no-slip
NoSlipDelayBounceBack -> an empty function). ``Simulation.step`` mirrors
UpdateGrid27::updateGrid for one level (src/gpu/core/Calculation/UpdateGrid27.cpp:54-96):
collision -> post-collision BCs (velocity, no-slip, geometry) -> parity swap -> pre-collision
BCs (pressure).
"""
from __future__ import annotations

import enum

from .lattice import Lattice


class BoundaryConditionFactory:
    class VelocityBC(enum.Enum):
        VelocityBounceBack = 0
        VelocityInterpolatedIncompressible = 1
        VelocityInterpolatedCompressible = 2
        VelocityWithPressureInterpolatedCompressible = 3
        NotSpecified = 4

    class NoSlipBC(enum.Enum):
        NoSlipDelayBounceBack = 0
        NoSlipBounceBack = 1
        NoSlipInterpolatedIncompressible = 2
        NoSlipInterpolatedCompressible = 3

    class PressureBC(enum.Enum):
        PressureNonEquilibriumIncompressible = 0
        PressureNonEquilibriumCompressible = 1
        OutflowNonReflective = 2
        OutflowNonReflectivePressureCorrection = 3
        NotSpecified = 4

    # C entry point of each implemented BC (incompressible and non-reflecting variants are
    # outside the K17 compressible path and are not provided)
    _VELOCITY = {
        VelocityBC.VelocityBounceBack: "vf_VelocityBounceBack",
        VelocityBC.VelocityInterpolatedCompressible: "vf_VelocityInterpolatedCompressible",
        VelocityBC.VelocityWithPressureInterpolatedCompressible: "vf_VelocityWithPressureInterpolatedCompressible",
    }
    _NOSLIP = {
        NoSlipBC.NoSlipBounceBack: "vf_NoSlipBounceBack",
        NoSlipBC.NoSlipInterpolatedCompressible: "vf_NoSlipInterpolatedCompressible",
    }
    _PRESSURE = {PressureBC.PressureNonEquilibriumCompressible: "vf_PressureNonEquilibriumCompressible"}

    def __init__(self):
        self.velocity = self.VelocityBC.NotSpecified
        self.no_slip = self.NoSlipBC.NoSlipDelayBounceBack
        self.pressure = self.PressureBC.NotSpecified
        self.geometry = self.NoSlipBC.NoSlipDelayBounceBack

    def set_velocity_boundary_condition(self, bc: "BoundaryConditionFactory.VelocityBC"):
        self.velocity = bc

    def set_no_slip_boundary_condition(self, bc: "BoundaryConditionFactory.NoSlipBC"):
        self.no_slip = bc

    def set_pressure_boundary_condition(self, bc: "BoundaryConditionFactory.PressureBC"):
        self.pressure = bc

    def set_geometry_boundary_condition(self, bc):
        """A velocity or no-slip BC for the geometry's nodes (BoundaryConditionFactory.h:129-139);
        default NoSlipDelayBounceBack."""
        self.geometry = bc

    def _resolve(self, table, value, unsupported_ok=False):
        if value in table:
            return table[value]
        if value.name in ("NotSpecified",):
            return None
        if value.name == "NoSlipDelayBounceBack":
            return ""  # empty function: solid nodes never collide, the eso-twist reflects
        raise NotImplementedError(f"{value.name} is not part of the K17 compressible path")

    def get_velocity_boundary_condition_post(self):
        return self._resolve(self._VELOCITY, self.velocity)

    def get_no_slip_boundary_condition_post(self):
        return self._resolve(self._NOSLIP, self.no_slip)

    def get_pressure_boundary_condition_pre(self):
        return self._resolve(self._PRESSURE, self.pressure)

    def get_geometry_boundary_condition_post(self):
        table = self._VELOCITY if isinstance(self.geometry, self.VelocityBC) else self._NOSLIP
        return self._resolve(table, self.geometry)


# launchers whose lists carry wall velocities (and, in the adjoint, their gradient)
_VELOCITY_LAUNCHERS = ("vf_VelocityBounceBack", "vf_VelocityInterpolatedCompressible",
                       "vf_VelocityWithPressureInterpolatedCompressible")


def post_launch_groups(lattice: Lattice, lists):
    """The launches of the post-collision lists [(name, launcher)], in launch order, as
    [(name, launcher, DeviceBC)]. A list without wall velocities that follows another list with
    the same launcher and no node in common (the channel walls, then an obstacle as the
    geometry list) joins that list's launch, named "a+b": every node keeps its one thread and its
    links, and a link writes a cell of the node across the wall, which no node of either list
    reads, so every value is what the two launches leave, and the adjoint keeps one thread per
    node as well. A short launch costs a few microseconds of latency whatever its size; the
    reference launches the lists one by one (UpdateGrid27.cpp:69-77). A joined list is a copy of
    the grid's host lists made here (its own device arrays): callers build the groups once per
    run, after the lists are final."""
    import numpy as np

    from .grid import BoundaryNodes
    from .lattice import DeviceBC, nodes_along_rows

    out = []
    for name, fn in lists:
        bc = lattice.bcs.get(name)
        if not fn or bc is None or bc.n == 0:
            continue
        if out:
            pname, pfn, pbc = out[-1]
            host = [lattice.grid.bcs[part] for part in pname.split("+")] + [lattice.grid.bcs[name]]
            k = np.concatenate([np.asarray(h.k) for h in host])
            # the joined list takes the kernel form (a thread per node or a lane per direction)
            # that each list takes alone, so that it rounds as they do
            rows = {bool(pbc.struct.nodesAlongRows), bool(bc.struct.nodesAlongRows), nodes_along_rows(k)}
            if (fn == pfn and fn not in _VELOCITY_LAUNCHERS and len(rows) == 1
                    and np.unique(k).size == k.size):
                q = np.concatenate([np.asarray(h.q27).reshape(27, h.n) for h in host], axis=1).reshape(-1)
                out[-1] = (pname + "+" + name, fn, DeviceBC(BoundaryNodes(k=k, q27=q), lattice.device, lattice.N))
                continue
        out.append((name, fn, bc))
    return out


class BCEntryTables:
    """Per-node tables of post-collision BC lists for a kernel that applies the lists itself
    (the fused bounce-back collision): per node its entry or -1, and per entry
    and list the node's index in the list (or -1) and the bitmask of its links d with
    0 <= q_d <= 1. names: the lists in launch order (at most three). Raises ValueError where a link
    leads to a node the collision updates (a fused kernel would race with that node's own
    accesses), or on a lattice without a fluid mask."""

    def __init__(self, lattice: Lattice, names, what: str = "bounce-back"):
        import numpy as np
        import torch

        from .grid import CX, CY, CZ

        if not lattice.box or lattice.mask is None:
            raise ValueError(f"{what} needs a dense box with solid nodes")
        if len(names) > 3:
            raise ValueError(f"{what}: at most three lists")
        g = lattice.grid
        fluid = g.fluid_mask
        all_k = np.unique(np.concatenate([g.bcs[name].k.astype(np.int64) for name in names])) if names else \
            np.zeros(0, np.int64)
        entry = np.full(g.N, -1, dtype=np.int32)
        entry[all_k] = np.arange(all_k.size, dtype=np.int32)
        index = np.full((all_k.size, 3), -1, dtype=np.int32)
        links = np.zeros((all_k.size, 3), dtype=np.uint32)
        x, y, z = g.coords()
        for l, name in enumerate(names):
            bc = g.bcs[name]
            k = bc.k.astype(np.int64)
            q = np.asarray(bc.q27).reshape(27, bc.n)
            e = entry[k]
            index[e, l] = np.arange(bc.n, dtype=np.int32)
            for d in range(1, 27):
                act = (q[d] >= 0) & (q[d] <= 1)
                if not np.any(act):
                    continue
                kk = k[act]
                m = g.index((x[kk] + CX[d]) % g.nx, (y[kk] + CY[d]) % g.ny, (z[kk] + CZ[d]) % g.nz)
                if np.any(fluid[m]):
                    raise ValueError(f"a {what} link of list '{name}' leads to a fluid node")
                links[e[act], l] |= np.uint32(1 << d)
        self.nodes = all_k
        self.count = int(all_k.size)
        dev = lattice.device
        self.entry = torch.from_numpy(entry).to(dev)
        self.index = torch.from_numpy(index.reshape(-1)).to(dev)
        self.links = torch.from_numpy(links.reshape(-1).view(np.int32)).to(dev)

    def pointers(self):
        import ctypes

        return (ctypes.c_void_p(self.entry.data_ptr()), ctypes.c_void_p(self.index.data_ptr()),
                ctypes.c_void_p(self.links.data_ptr()))


class FusedBounceBack:
    """The bounce-back BC lists of a dense-box lattice in the form vf_k17_run_box_bounce_back
    applies them inside the collision launch: a per-node entry, and per entry and list the node's
    index and its bounce-back links. lists: (name, velocity) pairs in launch order, velocity True
    for VelocityBounceBack and False for NoSlipBounceBack. Raises ValueError where the lists do
    not allow it: a link that leads to a node the collision updates (the epilogue would race
    with that node's own writes), or a lattice without a fluid mask."""

    def __init__(self, lattice: Lattice, lists):
        import ctypes

        from . import capi

        used = [(name, bool(vel)) for name, vel in lists
                if lattice.bcs.get(name) is not None and lattice.bcs[name].n > 0]
        self.tables = BCEntryTables(lattice, [name for name, _ in used], "bounce-back")
        st = capi.PostBounceBack()
        for l, (name, vel) in enumerate(used):
            st.lists[l] = ctypes.pointer(lattice.bcs[name].struct)
            st.velocity[l] = 1 if vel else 0
        st.nodeEntry, st.entryIndex, st.entryLinks = self.tables.pointers()
        st.numberOfEntries = self.tables.count
        self.struct = st
        self.lattice = lattice

    def run(self, collision_template: int = 0, stream=None):
        """One collision of the lattice followed by the lists (vf_k17_run_box_bounce_back)."""
        import ctypes

        from . import capi
        from .lattice import stream_handle

        capi.call("vf_k17_run_box_bounce_back", self.lattice.pref(), int(collision_template),
                  ctypes.byref(self.struct), stream_handle(stream))


_BOUNCE_BACK = {"vf_VelocityBounceBack": True, "vf_NoSlipBounceBack": False}


class Simulation:
    """Single-level time stepping of one Lattice.

    On a dense box whose post-collision BCs are all bounce-back (NoSlipBounceBack,
    VelocityBounceBack) at solid walls, the collision launch applies them itself
    (FusedBounceBack, vf_k17_run_box_bounce_back: the same values bit for bit, no BC launches);
    fuse_bounce_back=False keeps the separate launches."""

    def __init__(self, lattice: Lattice, bc_factory: BoundaryConditionFactory | None = None,
                 fuse_bounce_back: bool = True):
        self.lattice = lattice
        f = bc_factory or BoundaryConditionFactory()
        self.velocity_fn = f.get_velocity_boundary_condition_post()
        self.noslip_fn = f.get_no_slip_boundary_condition_post()
        self.geometry_fn = f.get_geometry_boundary_condition_post()
        self.pressure_fn = f.get_pressure_boundary_condition_pre()
        # BoundaryConditionKernelManager: BC nodes without a function is an error
        # (src/gpu/core/BoundaryConditions/BoundaryConditionKernelManager.cpp:53-83)
        for name, fn, label in (("velocity", self.velocity_fn, "velocityBoundaryConditionPost"),
                                ("pressure", self.pressure_fn, "directionalPressureBoundaryConditionPre")):
            bc = lattice.bcs.get(name)
            if bc is not None and bc.n > 0 and fn is None:
                raise RuntimeError(f"The boundary condition {label} was not set!")
        self.t = 0
        self.fused = None
        self.groups = None
        post = [("velocity", self.velocity_fn), ("noslip", self.noslip_fn), ("geometry", self.geometry_fn)]
        post = [(name, fn) for name, fn in post if fn and lattice.bcs.get(name) is not None and lattice.bcs[name].n]
        if fuse_bounce_back and getattr(lattice, "box", False) and post and all(fn in _BOUNCE_BACK for _, fn in post):
            try:
                self.fused = FusedBounceBack(lattice, [(name, _BOUNCE_BACK[fn]) for name, fn in post])
            except ValueError:
                self.fused = None  # the separate launches serve these lists

    def _launch(self, stream=None):
        """The launches of one step (UpdateGrid27::updateGrid for one level); swaps the parity."""
        L = self.lattice
        if self.fused is not None:
            self.fused.run(0, stream)
        else:
            self._post_collision(stream)
        L.swap()
        if self.pressure_fn:
            L.apply_bc(self.pressure_fn, "pressure", stream)

    def _post_collision(self, stream):
        """Collision and the post-collision BCs as separate launches (post_launch_groups)."""
        import ctypes

        from . import capi
        from .lattice import stream_handle

        L = self.lattice
        L.collide(stream=stream)
        if self.groups is None:
            self.groups = post_launch_groups(L, [("velocity", self.velocity_fn), ("noslip", self.noslip_fn),
                                                 ("geometry", self.geometry_fn)])
        for _, fn, bc in self.groups:
            capi.call(fn, L.pref(), ctypes.byref(bc.struct), stream_handle(stream))

    def step(self, stream=None):
        self._launch(stream)
        self.t += 1

    def run(self, steps: int, stream=None, graphs: bool = True):
        """`steps` time steps. With graphs (and no explicit stream) the steps' launches are
        captured once, 16 steps to a graph, and replayed (virtualfluids_amd.graphs.StepGraphs):
        the same kernels with the same arguments, without a host round trip per launch;
        graphs=False launches every kernel from the host."""
        if not graphs or stream is not None or steps < 3:
            for _ in range(steps):
                self.step(stream)
            return
        from .graphs import StepGraphs

        L = self.lattice

        def launch(even, st):
            L.even = even
            self._launch(st)

        # the steps' parity swaps are host state, which a replay does not run: set it after
        run = StepGraphs(launch)
        L.even = run.run(steps, L.even)
        run.close()
        self.t += steps


class MultiLevelSimulation:
    """Time stepping of a level hierarchy (level 0 coarsest) with refinement interfaces.

    One ``step`` is one coarse time step, following UpdateGrid27::updateGrid
    (src/gpu/core/Calculation/UpdateGrid27.cpp:54-96): a level first lets the next finer level
    take two steps, then collides, applies its BCs and swaps parity, and finally refreshes the
    interface to the finer level: fine -> coarse, then coarse -> fine
    (Refinement_noExchange, src/gpu/core/GridScaling/RefinementStrategy.cpp:169-178). The interface kernels read and
    write the finer level with the even parity, so every finer level starts with it.
    """

    def __init__(self, lattices, interfaces, bc_factories=None, scaling_factory=None):
        from .refinement import DeviceInterfaceCells, GridScalingKernelManager

        if len(interfaces) != len(lattices) - 1:
            raise ValueError("one (coarse-to-fine, fine-to-coarse) cell pair is needed per level pair")
        facs = bc_factories or [None] * len(lattices)
        self.lattices = list(lattices)
        self.levels = [Simulation(L, f) for L, f in zip(self.lattices, facs)]
        # a coarse step cycles through every level's populations
        total = sum(L.state_bytes for L in self.lattices)
        for L in self.lattices:
            L.set_working_set(total)
        for L in self.lattices[1:]:
            if not L.even:
                raise ValueError("finer levels must start with the even parity")
        self.cells = [(DeviceInterfaceCells(cf, L.device), DeviceInterfaceCells(fc, L.device))
                      for (cf, fc), L in zip(interfaces, self.lattices)]
        n_fc = self.cells[0][1].n if self.cells else 0
        n_cf = self.cells[0][0].n if self.cells else 0
        self.manager = GridScalingKernelManager(len(self.lattices), scaling_factory, n_fc, n_cf)
        self.t = 0

    def _update(self, level: int, stream):
        fine = len(self.lattices) - 1
        if level != fine:
            self._update(level + 1, stream)
            self._update(level + 1, stream)
        self.levels[level].step(stream)
        if level != fine:
            cf, fc = self.cells[level]
            coarse_l, fine_l = self.lattices[level], self.lattices[level + 1]
            self.manager.run_fine_to_coarse(coarse_l, fine_l, fc, stream)
            self.manager.run_coarse_to_fine(coarse_l, fine_l, cf, stream)

    def step(self, stream=None):
        self._update(0, stream)
        self.t += 1

    def run(self, steps: int, stream=None, graphs: bool = True):
        """`steps` coarse steps. With graphs (and no explicit stream) pairs of coarse steps
        replay one HIP graph after two steps from the host (after a pair every level's parity
        is back where it started); graphs=False launches every kernel from the host. The
        fields are the same bit for bit."""
        if not graphs or stream is not None or steps < 4:
            for _ in range(steps):
                self.step(stream)
            return
        import torch

        from .graphs import capture

        for _ in range(2):
            self.step()
        steps -= 2
        pairs = steps // 2
        if pairs:
            saved_t = [lv.t for lv in self.levels]
            g = capture(lambda st: (self._update(0, st), self._update(0, st)), torch.cuda.Stream())
            # the capture ran the host side of a pair (parities are back); no step was taken
            for lv, t in zip(self.levels, saved_t):
                lv.t = t
            for _ in range(pairs):
                g.replay()
            torch.cuda.current_stream().synchronize()  # before the graph is released
            del g
            for level, lv in enumerate(self.levels):
                lv.t += 2 * pairs * 2 ** level
            self.t += 2 * pairs
        if steps % 2:
            self.step()
