"""The solve loops' stopping rule (cdr_hip_stop_rule: host code, no GPU).

With stall_cycles = 0 it must be Cedar's loop (include/cedar/multilevel.h:283-297): stop at the
first cycle whose rho < tol, else run every cycle given.  The added guards stop at the first
NaN/Inf rho (CDR_ERR_NONFINITE) and, with stall_cycles = k > 0, at the first cycle that ends k
cycles without a new smallest rho (CDR_ERR_STALLED).
"""
import ctypes
import math

import numpy as np

from cedar_amd import _abi

OK, ARG, NONFINITE, STALLED = 0, 1, 4, 7


def rule(rho, tol, k=0):
    arr = (ctypes.c_double * max(len(rho), 1))(*rho)
    n = ctypes.c_int(-1)
    st = _abi.load().cdr_hip_stop_rule(arr, len(rho), tol, k, ctypes.byref(n))
    return st, n.value


def test_converges_at_first_rho_below_tol():
    assert rule([0.5, 0.1, 2e-11, 1e-12], 1e-10) == (OK, 3)
    # rho == tol does not stop (strict <, as the reference's `rel_resid < tol`)
    assert rule([0.5, 1e-10, 9e-11], 1e-10) == (OK, 3)


def test_runs_every_cycle_without_convergence():
    assert rule([0.5, 0.6, 0.7, 0.8], 1e-10) == (OK, 4)
    assert rule([], 1e-10) == (OK, 0)


def test_nonfinite_stops_at_once():
    assert rule([0.5, float("nan"), 0.1], 1e-10) == (NONFINITE, 2)
    assert rule([float("inf")], 1e-10) == (NONFINITE, 1)


def test_stall_counts_cycles_without_a_new_minimum():
    # best 1 -> 0.5; 0.6 (1), 0.55 (2), 0.7 (3) -> stop at cycle 4
    assert rule([0.5, 0.6, 0.55, 0.7, 0.1], 1e-10, k=3) == (STALLED, 4)
    # a new minimum resets the count: 0.6 (1), 0.4 new, 0.45 (1), 0.46 (2)
    assert rule([0.5, 0.6, 0.4, 0.45, 0.46], 1e-10, k=2) == (STALLED, 5)
    # equal to the minimum is no decrease
    assert rule([0.5, 0.5], 1e-10, k=1) == (STALLED, 2)
    # rho_0 = 1: a first cycle that does not reduce the residual counts
    assert rule([1.0, 2.0], 1e-10, k=2) == (STALLED, 2)
    # convergence wins when the same cycle would also be a stall
    assert rule([0.5, 0.6, 1e-12], 1e-10, k=2) == (OK, 3)
    # stall off: the reference's loop
    assert rule([0.5, 0.6, 0.55, 0.7], 1e-10, k=0) == (OK, 4)


def reference_loop(rho, tol):
    """include/cedar/multilevel.h:283-297 on a given history: cycles run."""
    for i, r in enumerate(rho):
        if r < tol:
            return i + 1
    return len(rho)


def brute_stall(rho, tol, k):
    best, since = 1.0, 0
    for i, r in enumerate(rho):
        if not math.isfinite(r):
            return NONFINITE, i + 1
        if r < tol:
            return OK, i + 1
        if r < best:
            best, since = r, 0
        else:
            since += 1
            if k > 0 and since >= k:
                return STALLED, i + 1
    return OK, len(rho)


def test_random_histories_match_the_reference_loop_and_brute_force():
    rng = np.random.default_rng(11)
    for _ in range(300):
        n = int(rng.integers(0, 40))
        rho = list(np.exp(rng.normal(-5, 4, size=n)))
        tol = float(10.0 ** rng.integers(-12, -2))
        st, nc = rule(rho, tol, 0)
        assert st == OK and nc == reference_loop(rho, tol)
        k = int(rng.integers(1, 5))
        assert rule(rho, tol, k) == brute_stall(rho, tol, k)


def test_argument_errors():
    assert rule([0.5], -1.0)[0] == ARG
    assert rule([0.5], 1e-10, k=-1)[0] == ARG
    n = ctypes.c_int()
    assert _abi.load().cdr_hip_stop_rule(None, 3, 1e-10, 0, ctypes.byref(n)) == ARG
