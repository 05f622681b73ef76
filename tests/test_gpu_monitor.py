"""The conservation check on the GPU's FP64 sums (vf_MomentSums through the C ABI): a moving
periodic box passes over many steps and is flagged after a stale-parity step or a NaN."""
import numpy as np
import pytest

from helpers import equilibrium, near_equilibrium, pack_pre
from virtualfluids_amd.grid import periodic_box
from virtualfluids_amd.monitor import ConservationCheck, ConservationError

pytestmark = pytest.mark.gpu


def _lattice(storage="fp32"):
    from virtualfluids_amd.lattice import Lattice

    grid = periodic_box(32, 12, 10)
    rng = np.random.default_rng(21)
    noise = near_equilibrium(rng, grid.N, u_amp=0.01, noise=2e-3)
    f0 = equilibrium(np.zeros(grid.N), np.tile([0.03, -0.02, 0.01], (grid.N, 1))) + (noise - noise.mean(0))
    L = Lattice(grid, 1.7, storage=storage)
    L.set_populations(pack_pre(f0, grid, True))
    return L


def test_conservation_check_on_gpu_sums(gpu):
    """FP32 storage: the default tolerances sit above FP32 round-off (16-bit storage rounds every
    population each step and needs tolerances of its own)."""
    L = _lattice()
    chk = ConservationCheck.from_lattice(L)
    for _ in range(300):
        L.collide()
        L.swap()
    chk.check_lattice(L, 300)
    L.swap()  # stale parity for one step
    L.collide()
    L.swap()
    with pytest.raises(ConservationError, match="momentum"):
        chk.check_lattice(L, 301)


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
def test_nan_is_flagged_on_gpu(gpu, storage):
    L = _lattice(storage)
    chk = ConservationCheck.from_lattice(L)
    for _ in range(3):
        L.collide()
        L.swap()
    L.f[5 * L.N + 100] = float("nan")
    L.collide()
    L.swap()
    with pytest.raises(ConservationError, match="not finite"):
        chk.check_lattice(L, 4)
