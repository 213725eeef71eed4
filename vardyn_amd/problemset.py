"""Many independent 4D-Var problems of uneven grid sizes and window lengths on
one GPU.

Each problem keeps its own device-resident engine (``FourDVar``) and its own
captured HIP graph -- its grid, window, observation times and checkpoint
schedule are its own.  ``evaluate`` replays all graphs at once, spread over a
few HIP streams (default priority) so that the GPU runs kernels of different
problems side by side: small grids that cannot fill the chip alone fill it
together.  Problems are assigned to streams longest-first, each stream taking
the next problem while it has the least work queued (by grid points x steps).

The reference runs such problem sets as one process per subdomain
(mapping/run_assimilation.py:27-350, datlas-ocean/VarDyn); here they share
one process and one GPU.
"""
from __future__ import annotations

import heapq

import torch

from .problem import build_problem


class ProblemSet:
    def __init__(self, cases, precision="float64", tape="auto", n_streams=4, device=None,
                 autotune=True):
        self.cases = list(cases)
        self.problems = []
        for c in self.cases:
            pb = build_problem(c, precision=precision, control="ic", tape=tape, device=device)
            if autotune:
                pb.plan.autotune(reps=3)
            self.problems.append(pb)
        dev = self.problems[0].plan.device if self.problems else None
        self.device = dev
        self.streams = [torch.cuda.Stream(dev) for _ in range(max(1, int(n_streams)))]
        # longest-processing-time-first assignment of problems to streams
        work = [c.ny * c.nx * c.n_steps for c in self.cases]
        heap = [(0, s) for s in range(len(self.streams))]
        self.assign = [[] for _ in self.streams]
        for i in sorted(range(len(work)), key=lambda k: -work[k]):
            load, s = heapq.heappop(heap)
            self.assign[s].append(i)
            heapq.heappush(heap, (load + work[i], s))

    def __len__(self):
        return len(self.problems)

    def capture(self):
        for pb in self.problems:
            pb.engine.capture()
        torch.cuda.synchronize(self.device)
        return self

    def evaluate(self, X):
        """Evaluate every problem at its control vector ``X[i]`` (float64 device
        tensors) concurrently; returns [(J, G)] and, per problem, the time in ms
        from the common start to its completion."""
        start = torch.cuda.Event(enable_timing=True)
        cur = torch.cuda.current_stream(self.device)
        start.record(cur)
        ends = [None] * len(self.problems)
        out = [None] * len(self.problems)
        for s, idx in zip(self.streams, self.assign):
            s.wait_event(start)
            with torch.cuda.stream(s):
                for i in idx:
                    out[i] = self.problems[i].engine.evaluate(X[i])
                    ends[i] = torch.cuda.Event(enable_timing=True)
                    ends[i].record(s)
        for s in self.streams:
            cur.wait_stream(s)
        torch.cuda.synchronize(self.device)
        done_ms = [start.elapsed_time(e) for e in ends]
        return out, done_ms

    def zeros(self):
        return [torch.zeros(pb.engine.nctrl, dtype=torch.float64, device=pb.plan.device)
                for pb in self.problems]
