"""Checkpoint and restart files of the population buffers (Python mirror of the C++ layer's
RestartObject, which follows src/gpu/core/Restart/RestartObject.cpp:43-201).

One float vector of 27 * N values per level, in the slot layout of the device buffer, NaN
written as 0:

* ``BinaryRestartObject`` (``.bin``): int32 number of levels, then per level int32 count and
  count float32 values (exact);
* ``ASCIIRestartObject`` (``.txt``): one line per level, every value followed by a space and
  printed as C++ streams print floats by default (``%g``, 6 significant digits; lossy, as the
  reference).

The files are the ones the C++ classes write (the tests compare them byte for byte).
``deserialize`` leaves the lattices' parity as it is: checkpoints written after an even number
of coarse steps resume exactly from a freshly initialised (even-parity) lattice, as in the
reference.
"""
from __future__ import annotations

import os
import warnings

import numpy as np


def restart_file_name(fname: str, step: int, my_id: int = 0) -> str:
    """fname + "_Restart_" + processID + "_" + step (src/gpu/core/Calculation/Simulation.cpp:93-96)."""
    return f"{fname}_Restart_{my_id}_{step}"


class RestartObject:
    extension = ""

    def __init__(self):
        self.fs: list[np.ndarray] = []

    def serialize(self, filename: str, lattices) -> None:
        """Download every level's populations (FP16/BF16 widened to float) and write them."""
        self.fs = []
        for L in lattices:
            f = L.populations().astype(np.float32)
            self.fs.append(np.where(np.isnan(f), np.float32(0.0), f))
        self.serialize_internal(filename)

    def deserialize(self, filename: str, lattices) -> None:
        """Read the file and upload every level (round to nearest even into FP16/BF16 storage)."""
        self.fs = []
        self.deserialize_internal(filename)
        if len(self.fs) < len(lattices):
            raise ValueError(f"restart file {filename} holds {len(self.fs)} level(s), the simulation has "
                             f"{len(lattices)}")
        for level, L in enumerate(lattices):
            if self.fs[level].size != 27 * L.N:
                raise ValueError(f"restart file {filename}: level {level} holds {self.fs[level].size} values, "
                                 f"expected {27 * L.N}")
            L.set_populations(self.fs[level])

    def delete_restart_file(self, filename: str) -> None:
        path = filename + self.extension
        try:
            os.remove(path)
        except OSError:
            warnings.warn(f'Could not delete the restart object file "{path}".')

    def serialize_internal(self, filename: str) -> None:
        raise NotImplementedError

    def deserialize_internal(self, filename: str) -> None:
        raise NotImplementedError


class ASCIIRestartObject(RestartObject):
    extension = ".txt"

    def serialize_internal(self, filename: str) -> None:
        with open(filename + self.extension, "w") as out:
            for level in self.fs:
                vals = np.char.mod("%g", np.asarray(level, dtype=np.float32).astype(np.float64))
                out.write("".join(v + " " for v in vals.tolist()))
                out.write("\n")

    def deserialize_internal(self, filename: str) -> None:
        path = filename + self.extension
        if not os.path.exists(path):
            raise FileNotFoundError(f"couldn't open check point file {filename}")
        with open(path) as f:
            for line in f:
                self.fs.append(np.array(line.split(), dtype=np.float32))


class BinaryRestartObject(RestartObject):
    extension = ".bin"

    def serialize_internal(self, filename: str) -> None:
        with open(filename + self.extension, "wb") as out:
            out.write(np.int32(len(self.fs)).tobytes())
            for level in self.fs:
                a = np.ascontiguousarray(level, dtype="<f4")
                if a.size > np.iinfo(np.int32).max:
                    raise ValueError("a level exceeds 2^31 values")
                out.write(np.int32(a.size).tobytes())
                out.write(a.tobytes())

    def deserialize_internal(self, filename: str) -> None:
        path = filename + self.extension
        if not os.path.exists(path):
            raise FileNotFoundError(f"couldn't open check point file {filename}")
        with open(path, "rb") as f:
            head = f.read(4)
            if len(head) < 4:
                raise ValueError(f"restart: truncated file {filename}")
            for _ in range(int(np.frombuffer(head, "<i4")[0])):
                n = f.read(4)
                if len(n) < 4:
                    raise ValueError(f"restart: truncated file {filename}")
                count = int(np.frombuffer(n, "<i4")[0])
                data = f.read(4 * count)
                if len(data) < 4 * count:
                    raise ValueError(f"restart: truncated file {filename}")
                self.fs.append(np.frombuffer(data, "<f4").copy())
