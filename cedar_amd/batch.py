"""Many independent 2D problems of mixed sizes: grouped by size, each group one batched solver.

Every kernel launch of a group's setup and cycle covers all problems of the group (the problem
index is the last grid dimension), so a set of problems costs one launch per kernel per level per
size group instead of one per problem.  Each problem stops at its own tolerance and gets exactly
the result of its individual solve.
"""
from collections import OrderedDict

import torch

from .cdr2 import Solver


class BatchSet:
    """Solvers for a list of 2D operators (device tensors (nstncl, ny+2, nx+2), any mix of sizes).

    `solve(rhs)` takes one right-hand side per problem (same order) and returns the solutions and
    residual histories in that order.  Keyword arguments are those of `Solver`.
    """

    def __init__(self, operators, **kw):
        self.groups = OrderedDict()   # shape -> list of problem indices
        for k, so in enumerate(operators):
            self.groups.setdefault(tuple(so.shape), []).append(k)
        self.n = len(operators)
        self.solvers = {}
        for shape, idx in self.groups.items():
            stacked = torch.stack([operators[k] for k in idx]).contiguous()
            self.solvers[shape] = Solver(stacked, **kw)

    def setup_seconds(self):
        return sum(s.times()[0] for s in self.solvers.values())

    def solve(self, rhs):
        xs = [None] * self.n
        hs = [None] * self.n
        for shape, idx in self.groups.items():
            b = torch.stack([rhs[k] for k in idx]).contiguous()
            x, h = self.solvers[shape].solve(b)
            for t, k in enumerate(idx):
                xs[k] = x[t]
                hs[k] = h[t]
        return xs, hs
