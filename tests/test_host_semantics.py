"""Factory and manager semantics of the Python host layer (CPU only), mirroring the reference's
own unit tests: BoundaryConditionFactoryTest.cpp, BCKernelManagerTest.cpp and
GridScalingKernelManagerTest.cpp (tests/unit-tests/gpu/core/...)."""
import types

import numpy as np
import pytest

from virtualfluids_amd.grid import BoundaryNodes
from virtualfluids_amd.refinement import GridScalingFactory, GridScalingKernelManager
from virtualfluids_amd.simulation import BoundaryConditionFactory, Simulation

F = BoundaryConditionFactory


def test_bc_factory_defaults():
    f = F()
    assert f.get_velocity_boundary_condition_post() is None
    assert f.get_no_slip_boundary_condition_post() == ""  # NoSlipDelayBounceBack: nothing to launch
    assert f.get_pressure_boundary_condition_pre() is None


def test_bc_factory_maps_to_c_entry_points():
    f = F()
    f.set_velocity_boundary_condition(F.VelocityBC.VelocityInterpolatedCompressible)
    f.set_no_slip_boundary_condition(F.NoSlipBC.NoSlipBounceBack)
    f.set_pressure_boundary_condition(F.PressureBC.PressureNonEquilibriumCompressible)
    assert f.get_velocity_boundary_condition_post() == "vf_VelocityInterpolatedCompressible"
    assert f.get_no_slip_boundary_condition_post() == "vf_NoSlipBounceBack"
    assert f.get_pressure_boundary_condition_pre() == "vf_PressureNonEquilibriumCompressible"
    f.set_velocity_boundary_condition(F.VelocityBC.VelocityInterpolatedIncompressible)
    with pytest.raises(NotImplementedError):
        f.get_velocity_boundary_condition_post()


def _fake_lattice(**counts):
    bcs = {name: BoundaryNodes(k=np.zeros(n, dtype=np.int32), q27=np.zeros(27 * n)) for name, n in counts.items()}
    return types.SimpleNamespace(bcs=bcs)


def test_bc_manager_checks():
    Simulation(_fake_lattice(velocity=0, pressure=0), F())  # no nodes: no function needed
    Simulation(_fake_lattice(noslip=4), F())  # the no-slip default is specified
    with pytest.raises(RuntimeError, match="velocityBoundaryConditionPost was not set"):
        Simulation(_fake_lattice(velocity=3), F())
    with pytest.raises(RuntimeError, match="PressureBoundaryConditionPre was not set"):
        Simulation(_fake_lattice(pressure=3), F())


def test_grid_scaling_manager_checks():
    f = GridScalingFactory()
    GridScalingKernelManager(1, None)
    GridScalingKernelManager(1, f)
    g = GridScalingFactory()
    g.set_scaling_factory(GridScalingFactory.GridScaling.ScaleCompressible)
    GridScalingKernelManager(1, g)
    GridScalingKernelManager(3, g)
    with pytest.raises(RuntimeError, match="no scalingFactory"):
        GridScalingKernelManager(3, None)
    GridScalingKernelManager(3, f)  # no interface cells: no function needed
    with pytest.raises(RuntimeError, match="scalingFineToCoarse was not set"):
        GridScalingKernelManager(3, f, fc_cells=100)
    with pytest.raises(RuntimeError, match="scalingCoarseToFine was not set"):
        GridScalingKernelManager(3, f, cf_cells=100)
    with pytest.raises(NotImplementedError):
        g.get_grid_scaling_fc(True)


def test_pressure_list_records_whether_its_neighbours_are_distinct():
    """The pressure adjoint skips atomics only where no two nodes share a neighbour kN."""
    from virtualfluids_amd.lattice import DeviceBC

    q = np.full((27, 3), -1.0)
    one = DeviceBC(BoundaryNodes(k=np.array([5, 6, 7]), q27=q, kN=np.array([1, 2, 3]), rho_bc=np.zeros(3),
                                 direction=1), "cpu", 10)
    assert one.kN_distinct
    shared = DeviceBC(BoundaryNodes(k=np.array([5, 6, 7]), q27=q, kN=np.array([1, 2, 1]), rho_bc=np.zeros(3),
                                    direction=1), "cpu", 10)
    assert not shared.kN_distinct


def test_bc_list_layout_hint_follows_the_nodes():
    """nodesAlongRows (one thread per node in the adjoints) for lists along x rows, not for faces
    normal to x; a list given twice is refused before any hint is formed."""
    from virtualfluids_amd.lattice import DeviceBC, nodes_along_rows

    x, y, z = np.meshgrid(np.arange(8), np.arange(6), np.arange(5), indexing="ij")
    k = x + 8 * (y + 6 * z)
    wall = np.sort(k[:, 0, :].ravel())   # y = 0: consecutive in x
    face = np.sort(k[0, :, :].ravel())   # x = 0: a stride of 8 between neighbours
    assert nodes_along_rows(wall) and not nodes_along_rows(face)
    assert not nodes_along_rows(np.array([3])) and not nodes_along_rows(np.array([], dtype=np.int64))
    q = np.full((27, wall.size), -1.0)
    bc = DeviceBC(BoundaryNodes(k=wall, q27=q), "cpu", 240)
    assert bc.struct.nodesAlongRows == 1
    bc = DeviceBC(BoundaryNodes(k=face, q27=np.full((27, face.size), -1.0)), "cpu", 240)
    assert bc.struct.nodesAlongRows == 0
    with pytest.raises(ValueError, match="listed twice"):
        DeviceBC(BoundaryNodes(k=np.array([4, 5, 4]), q27=np.full((27, 3), -1.0)), "cpu", 240)


def test_post_launch_groups_join_lists_with_one_launcher_and_no_common_node():
    """post_launch_groups: the sphere channel's walls and sphere (one no-slip launcher, no common
    node) become one launch over both lists, in list order; the inflow (wall velocities) stays on
    its own; lists with different launchers, or sharing a node, keep their own launches."""
    from virtualfluids_amd.grid import sphere_channel
    from virtualfluids_amd.lattice import DeviceBC
    from virtualfluids_amd.simulation import post_launch_groups

    class Host:
        pass

    g = sphere_channel((24, 14, 14), (9, 7, 7), 3.3, 0.04)
    L = Host()
    L.grid, L.device, L.N = g, "cpu", g.N
    L.bcs = {name: DeviceBC(b, "cpu", g.N) for name, b in g.bcs.items()}
    ns, vel = "vf_NoSlipInterpolatedCompressible", "vf_VelocityInterpolatedCompressible"
    groups = post_launch_groups(L, [("velocity", vel), ("noslip", ns), ("geometry", ns)])
    assert [(name, fn) for name, fn, _ in groups] == [("velocity", vel), ("noslip+geometry", ns)]
    joined = groups[1][2]
    wall, sphere = g.bcs["noslip"], g.bcs["geometry"]
    assert joined.n == wall.n + sphere.n
    assert np.array_equal(joined.k.numpy(), np.concatenate([wall.k, sphere.k]).astype(np.int32))
    q = joined.q27.numpy().reshape(27, joined.n)
    assert np.array_equal(q[:, :wall.n], np.asarray(wall.q27, np.float32).reshape(27, wall.n))
    assert np.array_equal(q[:, wall.n:], np.asarray(sphere.q27, np.float32).reshape(27, sphere.n))
    groups = post_launch_groups(L, [("noslip", ns), ("geometry", "vf_NoSlipBounceBack")])
    assert [name for name, _, _ in groups] == ["noslip", "geometry"]
    g.bcs["twin"] = g.bcs["noslip"]
    L.bcs["twin"] = L.bcs["noslip"]
    groups = post_launch_groups(L, [("noslip", ns), ("twin", ns)])
    assert [name for name, _, _ in groups] == ["noslip", "twin"]


def test_interface_sources_list_every_cell_corner_once():
    """vf_interpolation_sources built on the host: every lane 8 i + q appears once, under the node
    that is corner q of coarse cell i (+x, +y, +z steps through the neighbour arrays)."""
    import torch  # noqa: F401  (DeviceInterfaceCells keeps its arrays as tensors)

    from virtualfluids_amd.refinement import DeviceInterfaceCells, nested_box

    nb = nested_box((12, 10, 14), (2, 1, 3), (9, 8, 10))
    dc = DeviceInterfaceCells(nb.cf, "cpu")
    dc.sources(nb.coarse)
    nodes, first, lanes = (t.numpy().view(np.uint32).astype(np.int64) for t in dc._sources[1])
    assert lanes.size == 8 * nb.cf.n and np.array_equal(np.sort(lanes), np.arange(8 * nb.cf.n))
    assert first[0] == 0 and first[-1] == lanes.size and np.all(np.diff(first) > 0)
    g = nb.coarse
    for u in range(nodes.size):
        for l in lanes[first[u]:first[u + 1]]:
            i, q = divmod(int(l), 8)
            k = int(nb.cf.coarse[i])
            if q & 1:
                k = int(g.nbx[k])
            if q & 2:
                k = int(g.nby[k])
            if q & 4:
                k = int(g.nbz[k])
            assert k == nodes[u]
    assert nodes.size < 8 * nb.cf.n  # neighbouring cells share corners
