"""Many independent 2D problems of mixed sizes: grouped by size, each group one batched solver.

Every kernel launch of a group's setup and cycle covers all problems of the group (the problem
index is the last grid dimension), so a set of problems costs one launch per kernel per level per
size group instead of one per problem.  With nrhs > 1 every operator is solved for nrhs
right-hand sides at once: the group's solver builds one hierarchy per operator and its cycle
kernels read each operator once for its nrhs problems.  Each problem stops at its own tolerance
and gets exactly the result of its individual solve.

The groups are independent: each solver runs on a stream of its own, and `solve` drives the
groups from one host thread each (the C calls release the GIL), so the latency-bound coarse levels
and the per-cycle norm reads of one group overlap the work of the others.
"""
import threading
from collections import OrderedDict

import torch

from .cdr2 import Solver


class BatchSet:
    """Solvers for a list of 2D operators (device tensors (nstncl, ny+2, nx+2), any mix of sizes).

    `solve(rhs)` takes, per operator (same order), one right-hand side (nrhs = 1) or a stack of
    nrhs of them, and returns the solutions (shaped like the right-hand sides) and the residual
    histories (one array, or a list of nrhs arrays, per operator).  `concurrent=False` solves the
    groups one after another on the current stream.  Other keyword arguments are those of
    `Solver`.
    """

    def __init__(self, operators, concurrent=True, nrhs=1, **kw):
        self.groups = OrderedDict()   # shape -> list of problem indices
        for k, so in enumerate(operators):
            self.groups.setdefault(tuple(so.shape), []).append(k)
        self.n = len(operators)
        self.nrhs = int(nrhs)
        self.concurrent = concurrent and len(self.groups) > 1
        self.solvers, self.streams = {}, {}
        cur = torch.cuda.current_stream()
        for shape, idx in self.groups.items():
            stacked = torch.stack([operators[k] for k in idx]).contiguous()
            st = torch.cuda.Stream() if self.concurrent else cur  # default priority
            st.wait_stream(cur)
            with torch.cuda.stream(st):
                self.solvers[shape] = Solver(stacked, nrhs=self.nrhs, **kw)
            self.streams[shape] = st
        for st in self.streams.values():
            cur.wait_stream(st)

    def setup_seconds(self):
        return sum(s.times()[0] for s in self.solvers.values())

    def solve(self, rhs):
        xs = [None] * self.n
        hs = [None] * self.n
        cur = torch.cuda.current_stream()
        m = self.nrhs
        # problem-major stacks: right-hand side r of operator k at k * nrhs + r
        bs = {shape: torch.stack([rhs[k] for k in idx]).reshape((len(idx) * m,) + shape[-2:])
              .contiguous() for shape, idx in self.groups.items()}
        out, errors = {}, []

        def run(shape):
            try:
                st = self.streams[shape]
                st.wait_stream(cur)
                with torch.cuda.stream(st):
                    out[shape] = self.solvers[shape].solve(bs[shape])
            except BaseException as exc:  # re-raised in the calling thread
                errors.append(exc)

        if self.concurrent:
            threads = [threading.Thread(target=run, args=(shape,)) for shape in self.groups]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
        else:
            for shape in self.groups:
                run(shape)
        for st in self.streams.values():
            cur.wait_stream(st)
        if errors:
            raise errors[0]
        for shape, idx in self.groups.items():
            x, h = out[shape]
            for t, k in enumerate(idx):
                if m == 1:
                    xs[k] = x[t]
                    hs[k] = h[t]
                else:
                    xs[k] = x[t * m:(t + 1) * m]
                    hs[k] = h[t * m:(t + 1) * m]
        return xs, hs
