"""HIP-graph replay of time steps that work in place on fixed buffers.

A long run of short kernels is bound by the host launching them; capturing the steps in HIP
graphs and replaying them takes the host out of the loop. No counterpart in the reference (its
time loop launches every kernel from the host).
"""
from __future__ import annotations

import torch


class StepGraphs:
    """Runs a time step that works in place on fixed buffers from replayed HIP graphs.

    step(even, stream) launches one step on `stream`. Calling the object runs one step; run(n,
    even) runs n, replaying graphs of `chunk` steps (an even number, so that the parity comes
    back) where it can. The first two steps (one of each parity) run eagerly, so that every lazy
    set-up of the library happens outside a capture; from then on each graph (one step of each
    parity, `chunk` steps from each parity) is captured once (torch.cuda.CUDAGraph on a side
    stream of default priority) and replayed on the current stream. The captured launches read their
    kernel arguments at capture time: the buffers, BC arrays and porosity fields they name must
    stay where they are while the graphs live (the owner drops them with the run). The step must
    not allocate or clear memory with memset nodes. (torch.cuda.graph would also collect
    garbage and empty the allocator's cache around every capture, which costs more than a
    thousand replayed steps save.)"""

    def __init__(self, step, enabled: bool = True, chunk: int = 16):
        self.step = step
        self.enabled = enabled
        self.chunk = chunk if chunk % 2 == 0 else chunk + 1  # whole pairs: the parity comes back
        self.graphs = {}
        self.chunks = {}
        self.eager_calls = 0
        self.side = None

    def close(self):
        """Wait for the replays in flight, then release the graphs (an executable graph is not
        released while the GPU may still run it)."""
        if self.graphs or self.chunks:
            torch.cuda.current_stream().synchronize()
        self.graphs.clear()
        self.chunks.clear()

    def _capture(self, launch):
        if self.side is None:
            self.side = torch.cuda.Stream()
        return capture(launch, self.side)

    def run(self, n: int, even: bool) -> bool:
        """n steps from parity `even`, in graphs of `chunk` steps (one replay, one host round
        trip, per chunk) and single steps for the rest; returns the parity after them."""
        while n > 0:
            if self.enabled and self.eager_calls >= 2 and self.chunk > 1 and n >= self.chunk:
                g = self.chunks.get(even)
                if g is None:
                    def launch(st, e0=even):
                        for j in range(self.chunk):
                            self.step(e0 ^ (j % 2 == 1), st)

                    g = self.chunks[even] = self._capture(launch)
                g.replay()
                n -= self.chunk
            else:
                self(even)
                even = not even
                n -= 1
        return even

    def __call__(self, even: bool):
        if not self.enabled or self.eager_calls < 2:
            self.eager_calls += 1
            self.step(even, torch.cuda.current_stream())
            return
        g = self.graphs.get(even)
        if g is None:
            g = self.graphs[even] = self._capture(lambda st: self.step(even, st))
        g.replay()


def capture(launch, side: torch.cuda.Stream) -> torch.cuda.CUDAGraph:
    """A HIP graph of the launches launch(stream) makes, captured on `side` (default priority)
    after the current stream's pending work; replay() then runs them on the current stream."""
    cur = torch.cuda.current_stream()
    side.wait_stream(cur)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        g.capture_begin()
        try:
            launch(side)
        finally:
            g.capture_end()
    cur.wait_stream(side)
    return g
