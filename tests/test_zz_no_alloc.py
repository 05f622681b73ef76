"""The coarse-to-fine interface adjoint that allocates nothing
(vf_ScaleCoarseToFineCompressibleAdjointGatherNoAlloc: its workspace is the library's own device
array) against the same adjoint with a caller workspace (vf_ScaleCoarseToFineCompressibleAdjointGather):
identical results, on two calls in a row, the second with fewer cells, and a clear error beyond the
array's 65536 cells."""
import ctypes

import numpy as np
import pytest

from helpers import near_equilibrium, pack_pre

pytestmark = pytest.mark.gpu


def _case(nb, cells, storage, seed):
    import torch

    from virtualfluids_amd.lattice import Lattice
    from virtualfluids_amd.refinement import DeviceInterfaceCells

    rng = np.random.default_rng(seed)
    C = Lattice(nb.coarse, 1.7, storage=storage, level=0)
    F = Lattice(nb.fine, 1.5, storage=storage, level=1)
    C.set_populations(pack_pre(near_equilibrium(rng, nb.coarse.N, u_amp=0.04, noise=5e-3), nb.coarse, True))
    F.set_populations(pack_pre(near_equilibrium(rng, nb.fine.N, u_amp=0.04, noise=5e-3), nb.fine, True))
    lc = torch.from_numpy(rng.normal(size=27 * nb.coarse.N).astype(np.float32)).cuda()
    lf = torch.from_numpy(rng.normal(size=27 * nb.fine.N).astype(np.float32)).cuda()
    return C, F, lc, lf, DeviceInterfaceCells(cells, C.device)


def _run(fn, C, F, dc, nb, lc, lf):
    import torch

    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import stream_handle

    src, nwork = dc.sources(nb.coarse)
    lc, lf = lc.clone(), lf.clone()
    args = [C.pref(), F.pref(), ctypes.byref(dc.struct), None, ctypes.byref(src)]
    if fn.endswith("Gather"):
        work = torch.full((nwork,), float("nan"), dtype=torch.float32, device="cuda")
        args.append(ctypes.c_void_p(work.data_ptr()))
    capi.call(fn, *args, ctypes.c_void_p(lc.data_ptr()), ctypes.c_void_p(lf.data_ptr()), stream_handle())
    torch.cuda.synchronize()
    return lc.cpu().numpy(), lf.cpu().numpy()


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_no_alloc_gather_matches_workspace_gather(gpu, storage):
    from virtualfluids_amd.refinement import InterfaceCells, nested_box

    nb = nested_box((16, 14, 18), (3, 2, 4), (12, 11, 13))
    half = InterfaceCells(nb.cf.coarse[: nb.cf.n // 2], nb.cf.fine[: nb.cf.n // 2])
    for cells, seed in ((nb.cf, 11), (half, 12)):  # the second call has fewer cells
        C, F, lc, lf, dc = _case(nb, cells, storage, seed)
        print(f"no-alloc gather {storage} cells={cells.n}", flush=True)
        want = _run("vf_ScaleCoarseToFineCompressibleAdjointGather", C, F, dc, nb, lc, lf)
        got = _run("vf_ScaleCoarseToFineCompressibleAdjointGatherNoAlloc", C, F, dc, nb, lc, lf)
        moved = np.abs(want[0] - lc.cpu().numpy()).max()
        assert moved > 1e-3  # the call did something
        # the same kernels in the same order on the same inputs: bit for bit
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_no_alloc_gather_rejects_too_many_cells(gpu):
    from virtualfluids_amd import capi
    from virtualfluids_amd.lattice import stream_handle
    from virtualfluids_amd.refinement import nested_box

    nb = nested_box((16, 14, 18), (3, 2, 4), (12, 11, 13))
    C, F, lc, lf, dc = _case(nb, nb.cf, "fp32", 13)
    src, _ = dc.sources(nb.coarse)
    big = capi.InterpolationCells.from_buffer_copy(dc.struct)
    big.numberOfCells = (1 << 16) + 1
    rc = capi.load().vf_ScaleCoarseToFineCompressibleAdjointGatherNoAlloc(
        C.pref(), F.pref(), ctypes.byref(big), None, ctypes.byref(src), ctypes.c_void_p(lc.data_ptr()),
        ctypes.c_void_p(lf.data_ptr()), stream_handle())
    assert rc == capi.VF_ERR_INVALID_ARGUMENT
    assert b"65536" in capi.load().vf_get_last_error()
