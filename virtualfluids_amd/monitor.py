"""Conservation check of a running lattice (failure detection on the host side).

A closed lattice (periodic, bounce-back walls, no force or porosity) conserves its mass and
momentum to FP32 round-off per step. ``ConservationCheck`` records the FP64 sums of a state
(vf_MomentSums: sum of drho and of the momentum over the fluid nodes) and flags a later state
whose sums are not finite (a NaN or Inf anywhere in the populations) or have drifted beyond the
tolerances: for example after a step taken with a stale parity (a wrong isEvenTimestep), which
reads every population as its opposite direction and so turns a net momentum around.
The reference has no such check (its CUDA errors exit through getLastCudaError; its restart
files replace NaN by 0, src/gpu/core/Restart/RestartObject.cpp:67).
"""
from __future__ import annotations

import numpy as np


class ConservationError(RuntimeError):
    """A state failed the conservation check."""


class ConservationCheck:
    """Checks FP64 sums (drho, jx, jy, jz[, kinetic energy]) against a reference state.

    The tolerance on the drift of sum(drho) and of each momentum component after `steps` steps is
    fluid_nodes * (atol_per_node + per_node_step * steps). FP32 round-off is coherent where every
    node holds similar values (a uniform flow drifts its momentum by ~2e-9 per node and step), so
    the defaults (1e-7 and 2e-8) sit ten times above it; a flipped or lost population changes
    the sums by ~1e-2 per affected node."""

    def __init__(self, reference_sums, fluid_nodes: int, atol_per_node: float = 1e-7,
                 per_node_step: float = 2e-8):
        ref = np.asarray(reference_sums, dtype=np.float64)[:4]
        if not np.all(np.isfinite(ref)):
            raise ConservationError("the reference state's sums are not finite")
        self.reference = ref
        self.fluid_nodes = int(fluid_nodes)
        self.atol_per_node = float(atol_per_node)
        self.per_node_step = float(per_node_step)

    @classmethod
    def from_lattice(cls, lattice, **tolerances) -> "ConservationCheck":
        """Record the current state of a Lattice (its FP64 sums from the GPU reduction)."""
        return cls(lattice.moment_sums(), int(lattice.grid.fluid.size), **tolerances)

    def tolerance(self, steps: int) -> float:
        return self.fluid_nodes * (self.atol_per_node + self.per_node_step * max(int(steps), 0))

    def problems(self, sums, steps: int) -> list:
        """Descriptions of every failed condition `steps` steps after the reference state (empty
        when the state passes)."""
        s = np.asarray(sums, dtype=np.float64)[:4]
        if not np.all(np.isfinite(s)):
            return ["the FP64 sums are not finite (NaN or Inf in the populations)"]
        tol = self.tolerance(steps)
        out = []
        dm = s[0] - self.reference[0]
        if abs(dm) > tol:
            out.append(f"mass drifted by {dm:.3e} (tolerance {tol:.1e})")
        for i, name in enumerate("xyz"):
            dj = s[1 + i] - self.reference[1 + i]
            if abs(dj) > tol:
                out.append(f"momentum {name} drifted by {dj:.3e} (tolerance {tol:.1e})")
        return out

    def check(self, sums, steps: int) -> None:
        """Raise ConservationError when the sums fail."""
        p = self.problems(sums, steps)
        if p:
            raise ConservationError("; ".join(p))

    def check_lattice(self, lattice, steps: int) -> None:
        """Check the current state of a Lattice, `steps` steps after the reference state."""
        self.check(lattice.moment_sums(), steps)
