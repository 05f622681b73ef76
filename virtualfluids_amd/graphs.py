"""HIP-graph replay of time steps that work in place on fixed buffers.

A long run of short kernels is bound by the host launching them; capturing one step per
parity in a HIP graph and replaying it takes the host out of the loop. No counterpart in the
reference (its time loop launches every kernel from the host).
"""
from __future__ import annotations

import torch


class StepGraphs:
    """Runs a time step that works in place on fixed buffers, replaying one HIP graph per parity.

    step(even, stream) launches one step on `stream`. The first two calls (one of each parity)
    run eagerly, so that every lazy set-up of the library happens outside a capture; from then
    on each parity's launches are captured once (torch.cuda.CUDAGraph on a side stream of
    default priority) and replayed on the current stream. The captured launches read their
    kernel arguments at capture time: the buffers, BC arrays and porosity fields they name must
    stay where they are while the graphs live (the owner drops them with the run). The step must
    not allocate or clear memory with memset nodes. (torch.cuda.graph would also collect
    garbage and empty the allocator's cache around every capture, which costs more than a
    thousand replayed steps save.)"""

    def __init__(self, step, enabled: bool = True):
        self.step = step
        self.enabled = enabled
        self.graphs = {}
        self.eager_calls = 0
        self.side = None

    def __call__(self, even: bool):
        if not self.enabled or self.eager_calls < 2:
            self.eager_calls += 1
            self.step(even, torch.cuda.current_stream())
            return
        g = self.graphs.get(even)
        if g is None:
            cur = torch.cuda.current_stream()
            if self.side is None:
                self.side = torch.cuda.Stream()
            self.side.wait_stream(cur)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.stream(self.side):
                g.capture_begin()
                try:
                    self.step(even, self.side)
                finally:
                    g.capture_end()
            cur.wait_stream(self.side)
            self.graphs[even] = g
        g.replay()
