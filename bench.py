"""Benchmark: time to a 1e-10 relative residual on one MI355X, FP32-V-cycle + FP64 iterative
refinement ("mixed") against plain FP64, plus achieved HBM bandwidth of the dominant kernel.

Workload (default): 2D 9-point diffusion -div(k grad u) = f on 4095 x 4095 interior nodes,
Dirichlet, k lognormal (sigma 2, 16x16-cell blocks, contrast <= 1e4, seed 2), f ~ U[-1,1];
V(1,1) cycles with alternating x/y line relaxation, 11 levels.  One "step" is one complete
solve from x0 = 0 to ||r|| / ||r0|| < 1e-10 with the hierarchy already built and all inputs
resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4095] [--cpu-runs 3]
                    [--levels-json PATH]

Prints one JSON line (rank 0), headline fields first.  `--workload c1` times the reference's
CPU-runnable case (511^2 5-point Poisson, point V(2,1)), `--workload c3` the 3D
plane-relaxation solve (255^3 anisotropic 7-point finite-volume problem, plane-xy V(1,1), FP64
and mixed), `--workload c4` the batched multi-size 2D solves and `--workload c5` the
differentiated solves (forward + adjoint gradient, 16 x 1025^2); each prints its own line.  A
record whose timed solves did not reach the tolerance is marked "valid": false and the run exits
with status 1.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def time_kernel(fn, reps, torch):
    """Average duration of fn() (enqueued on torch's current stream) with HIP events."""
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def measured_traffic():
    """HBM bytes per launch of the dominant kernel from the committed counter summary
    (two rocprofv3 --pmc runs, calibrated against a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "relax_xlines_f32_4095_counters.md")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return int(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def level_rates(solver, b, word, nstncl, torch):
    """Achieved bandwidth of every V-cycle phase on every level of one solve, from the solver's
    per-level device timers (cycles run eagerly while timed).  Algorithmic bytes per call, with
    N the level's points and w the word size: relaxation (one line-xy step: an x and a y sweep)
    2 (S+3) w N, residual (S+3) w N, restriction 3.25 w N_fine, interp_add 6.25 w N_fine."""
    import json as _json
    import tempfile
    solver.timers(True)
    solver.solve(b)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "t.json")
        solver.save_timers(path)
        levels = _json.load(open(path))["levels"]
    solver.timers(False)
    nl = solver.nlevels
    per = {"relaxation": 2 * (nstncl + 3), "residual": nstncl + 3, "restrict": 3.25,
           "interp-add": 6.25}
    out = []
    for k, lv in enumerate(levels):
        lvl = nl - 1 - k
        nx, ny = solver.level(lvl)[:2]
        npts = nx * ny
        row = {"level": lvl, "n": [nx, ny]}
        for label, words in per.items():
            if label not in lv:
                continue
            t, cnt = float(lv[label]["time"]), int(lv[label]["count"])
            byt = words * word * npts * cnt
            row[label] = {"s": t, "calls": cnt, "GBps": byt / t / 1e9,
                          "frac": byt / t / 1e9 / HBM_PEAK_GBS}
        out.append(row)
    return sorted(out, key=lambda r: r["level"])


def kernel_only(args, so, b):
    """Dominant kernel alone (the fine-level FP32 x-line kernel, 5 sweeps = 10 launches) plus a
    copy of known size whose counters calibrate FETCH_SIZE / WRITE_SIZE."""
    import torch

    import cedar_amd
    from cedar_amd import _abi
    s = cedar_amd.Solver(so, precision="mixed", relaxation="line-xy", nrelax_pre=1,
                         nrelax_post=1)
    ii, jj = so.shape[2], so.shape[1]
    _, _, _, _, sor0, _ = s.level(0)
    sor_x = torch.empty((2, jj, ii), dtype=torch.float64, device="cuda")
    cedar_amd.cdr2._hip().hipMemcpy(ctypes.c_void_p(sor_x.data_ptr()), ctypes.c_void_p(sor0),
                                    2 * ii * jj * 8, 3)
    A = so.to(torch.float32).contiguous()
    sx = sor_x.to(torch.float32).contiguous()
    f = b.to(torch.float32).contiguous()
    q = torch.zeros_like(f)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    fx = _abi.load().cdr_hip_relax_lines_x2d_f32
    torch.cuda.synchronize()
    for _ in range(5):
        fx(A.data_ptr(), so.shape[0], f.data_ptr(), q.data_ptr(), sx.data_ptr(), None, ii, jj, 0,
           stream)
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"kernel_only": True, "copy_bytes": src.numel() * 4,
                      "launches": 10, "points": (ii - 2) * (jj - 2)}))


def bench_c3(args):
    """3D: 7-point FV -div(K grad u) = f, (kx, ky, kz) = (1, 1, 1e-3) k, lognormal k (sigma 1.5)
    on 8^3 blocks, 27-point Galerkin coarse operators, plane-xy V(1,1) to 1e-10, FP64 and mixed."""
    import torch

    import cedar_amd
    from cedar_amd import gallery as G
    n = args.n3
    t0 = time.time()
    so_np, b_np, _ = G.heterogeneous_7pt(n, seed=3, block=8, sigma=1.5, aniso=(1.0, 1.0, 1e-3))
    gen_s = time.time() - t0
    so = torch.as_tensor(so_np).cuda()
    b = torch.as_tensor(b_np).cuda()
    del so_np
    res = {}
    for prec in ("f64", "mixed"):
        s = cedar_amd.Solver3(so, relaxation=args.relax3, nrelax_pre=1, nrelax_post=1,
                              maxiter=60, tol=1e-10, precision=prec)
        setup_s = s.times()[0]
        for _ in range(args.warmup):
            s.solve(b)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            x, h = s.solve(b)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        res[prec] = {"seconds": dt, "iters": int(len(h)), "rho": float(h[-1]),
                     "converged": converged(h, 1e-10, 60), "setup_s": setup_s,
                     "setup_plus_solve_s": setup_s + dt, "per_cycle_s": dt / len(h), "x": x}
        del s
    x64, x32 = res["f64"].pop("x"), res["mixed"].pop("x")
    rel = float((torch.linalg.norm(x32 - x64) / torch.linalg.norm(x64)).item())
    roof = plane_sweep_roofline(args, so, torch)
    cpu = None
    if args.cpu_runs > 0:
        cpu = cpu_cycles_baseline3(so.cpu().numpy(), b.cpu().numpy(), args.relax3,
                                   res["f64"]["iters"], args.cpu_cycles3)
    emit({
        "metric": "time-to-1e-10 rel. residual (s), 3D plane relaxation; mixed vs FP64",
        "value": res["mixed"]["seconds"], "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["mixed"]["seconds"] * 1e3,
        "higher_is_better": False, "dtype": "f32 cycle + f64 refinement",
        "data": "synthetic (seeded lognormal block coefficients, uniform RHS)",
        "config": {"workload": f"3D 7-point FV, {n}^3, (kx,ky,kz)=(1,1,1e-3)k, lognormal k "
                               f"sigma=1.5 on 8^3 blocks, {args.relax3} V(1,1), tol 1e-10"},
        "fp64": res["f64"], "mixed": dict(res["mixed"], rel_l2_vs_fp64=rel),
        "speedup_mixed_vs_fp64": res["f64"]["seconds"] / res["mixed"]["seconds"],
        "roofline": roof, "cpu_baseline": cpu,
        "gen_s": gen_s}, res["f64"]["converged"] and res["mixed"]["converged"])


def plane_sweep_roofline(args, so, torch):
    """Fine-level xy-plane line sweep of the mixed 3D cycle: the batched FP32 x-line kernel over
    the 128 odd xy-planes of the 255^3 grid (5-point plane operators {p, pw, ps} of the 7-point
    stencil), one colour per launch, live HIP-event timing on the launching stream.  Algorithmic
    bytes per launch: (S+3) w / 2 per point with S = 3 planes, w = 4: c/w/s read once, f read,
    q read and written, over all planes of the batch."""
    import cedar_amd
    from cedar_amd import _abi
    ns3, kk, jj, ii = so.shape
    planes = list(range(1, kk - 1, 2))
    so2 = so[0:3, planes].permute(1, 0, 2, 3).contiguous()           # (nb, 3, jj, ii) FP64
    nb = so2.shape[0]
    sor = torch.zeros((nb, 2, jj, ii), dtype=torch.float64, device="cuda")
    for t in range(nb):
        cedar_amd.cdr2.setup_lines(so2[t], sor[t], "x")
    A, sx = so2.float().contiguous(), sor.float().contiguous()
    g = torch.Generator(device="cpu").manual_seed(7)
    f = (torch.rand((nb, jj, ii), generator=g, dtype=torch.float32) * 2 - 1).cuda()
    q = torch.zeros_like(f)
    lib = _abi.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    fx = lib.cdr_hip_relax_lines2d_batched_f32
    t = time_kernel(lambda: fx(0, A.data_ptr(), 3, f.data_ptr(), q.data_ptr(), sx.data_ptr(), ii,
                               jj, nb, 0, stream), args.kernel_reps, torch)
    npts = (ii - 2) * (jj - 2) * nb
    byt = (3 + 3) * 4 * npts / 2
    return {"bound": "hbm", "kernel": f"relax_xlines_seg_k<float, 5-point> batched over {nb} "
            f"xy-planes of {ii - 2}^2 (the fine-level plane sweep of the mixed cycle), one colour "
            "per launch, shape picked by the autotuner", "achieved": byt / (t / 2) / 1e9,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": byt / (t / 2) / 1e9 / HBM_PEAK_GBS,
            "traffic": None, "algorithmic_bytes_per_launch": byt, "avg_launch_s": t / 2}


def cpu_cycles_baseline3(so_np, b_np, relax, gpu_iters, ncyc):
    """The 3D FP64 oracle (plain C, one thread) on the same problem: setup, then `ncyc` V-cycles
    from x0 = 0 timed one by one (a full solve takes minutes on one core); the value is the
    median cycle time times the GPU's cycle count."""
    import statistics

    from oracle import oracle3 as O3
    t = time.perf_counter()
    o = O3.Solver3(so_np, relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=60, tol=1e-10)
    setup = time.perf_counter() - t
    x = np.zeros_like(b_np)
    cyc = []
    for k in range(ncyc):
        t = time.perf_counter()
        x = o.vcycle(x, b_np)
        cyc.append(time.perf_counter() - t)
        print(f"cpu baseline 3D cycle {k + 1}/{ncyc}: {cyc[-1]:.2f} s", file=sys.stderr,
              flush=True)
    med = statistics.median(cyc)
    return {"value": med * gpu_iters,
            "unit": "s (solve from x0=0 to 1e-10 = median cycle time x the cycle count; setup "
                    "excluded, as in value)",
            "cores": 1, "kind": "port",
            "sample": f"FP64 C oracle, this {so_np.shape[1] - 2}^3 problem, {relax} V(1,1): setup "
                      f"then {ncyc} cycles timed one by one (median {med:.2f} s per cycle) x "
                      f"{gpu_iters} cycles; host '{cpu_model()}', nproc {os.cpu_count()}, 1 "
                      "thread",
            "setup_s": setup, "per_cycle_s": med, "cycles_timed": ncyc}


def bench_c4(args):
    """Batched solves: 96 independent 2D 9-point problems, sizes drawn uniformly from
    {257, 513, 1025, 2049}^2 (seed 4), each with its own lognormal block coefficients (sigma 2,
    16x16 blocks) and 4 uniform right-hand sides; mixed precision to 1e-10 each, line-xy V(1,1).
    Problems are grouped by size; each group is one solver over all its operators and their 4
    right-hand sides (one launch per kernel per level for the whole group, each operator read
    once per launch for its 4 right-hand sides).  `--c4-separate` solves the right-hand sides one
    after another instead (one solver per group with one right-hand side per operator)."""
    import torch

    from cedar_amd import gallery as G
    from cedar_amd.batch import BatchSet
    t0 = time.time()
    sizes, ops_np, rhs_np = G.batched_set(nprob=args.nprob)
    ops = [torch.as_tensor(so).cuda() for so in ops_np]
    nr = len(rhs_np)
    rhs = [[torch.as_tensor(b).cuda() for b in rr] for rr in rhs_np]       # [r][k]
    del ops_np, rhs_np
    gen_s = time.time() - t0
    kw = dict(relaxation="line-xy", nrelax_pre=1, nrelax_post=1, maxiter=100, tol=1e-10,
              precision="mixed")
    shared = not args.c4_separate
    bs = BatchSet(ops, nrhs=nr if shared else 1, **kw)
    setup_s = bs.setup_seconds()
    stacked = [torch.stack([rhs[r][k] for r in range(nr)]) for k in range(len(ops))]

    def step():
        if shared:
            _, hs = bs.solve(stacked)
            return [h for hk in hs for h in hk]
        out = []
        for r in range(nr):
            _, hs = bs.solve(rhs[r])
            out += hs
        return out

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    iters = []
    worst = 0.0
    ok = True
    for _ in range(args.steps):
        hs = step()
        iters += [len(h) for h in hs]
        worst = max(worst, max(float(h[-1]) for h in hs))
        ok = ok and all(converged(h, kw["tol"], kw["maxiter"]) for h in hs)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / args.steps
    nsolve = nr * len(ops)
    emit({
        "metric": "batched 2D solves to 1e-10 each: time for all problems and right-hand sides",
        "value": dt, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32 cycle + f64 refinement",
        "data": "synthetic (seeded lognormal block coefficients, uniform RHS)",
        "config": {"workload": f"{len(ops)} 2D 9-point problems, sizes uniform in "
                               "{257,513,1025,2049}^2, 4 RHS each, line-xy V(1,1), tol 1e-10",
                   "groups": {str(k[1] - 2): len(v) for k, v in bs.groups.items()},
                   "rhs_per_solve": nr if shared else 1},
        "solves": nsolve, "solves_per_s": nsolve / dt, "setup_s": setup_s,
        "cycles_min_max": [min(iters), max(iters)], "worst_rho": worst, "gen_s": gen_s}, ok)


def bench_c5(args):
    """Differentiated solves: 16 problems of 1025^2 nodes, k = exp(theta) per cell with theta a
    5-cell box-filtered N(0,1) field (seeded), f ~ U[-1,1]; L = 0.5 ||u - u_obs||^2 with u_obs
    from a second seeded theta plus 1e-3 Gaussian noise.  One step = forward (assembly, setup,
    mixed solve to 1e-10) + backward (adjoint solve on the same hierarchy + gradient kernel)
    for the whole batch.  The gradient is compared with an FP64 solve pair (tol 1e-12)."""
    import torch

    from cedar_amd import adjoint as A
    from cedar_amd import gallery as G
    nb, n = 16, args.n5
    t0 = time.time()
    theta, theta_obs, f, noise = G.differentiated_set(nb, n)
    gen_s = time.time() - t0
    dev = lambda a: torch.as_tensor(a, dtype=torch.float64).cuda()
    f_d = dev(f)
    with torch.no_grad():
        uobs = A.solve(dev(theta_obs), f_d, precision="f64", tol=1e-12)
    uobs = uobs + 1e-3 * dev(noise)
    uobs[..., 0, :] = 0.0
    uobs[..., -1, :] = 0.0
    uobs[..., :, 0] = 0.0
    uobs[..., :, -1] = 0.0

    def step(**kw):
        th = dev(theta).requires_grad_(True)
        u = A.solve(th, f_d, **kw)
        loss = 0.5 * ((u - uobs) ** 2)[..., 1:-1, 1:-1].sum()
        loss.backward()
        return float(loss.item()), th.grad

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        loss, g = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / args.steps
    loss64, g64 = step(precision="f64", tol=1e-12)
    rel = float((g - g64).abs().max() / g64.abs().max())
    print(json.dumps({
        "metric": "differentiated 2D solves: forward + adjoint gradient for the batch",
        "value": dt, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": False,
        "dtype": "f32 cycle + f64 refinement (forward and adjoint), f64 gradient kernel",
        "data": "synthetic (seeded smoothed Gaussian log-coefficients, uniform RHS)",
        "config": {"workload": f"{nb} x {n}^2 9-point Q1 diffusion, k = exp(theta), "
                               "L = 0.5||u - u_obs||^2, line-xy V(1,1), tol 1e-10"},
        "gradients_per_s": nb / dt, "loss": loss, "loss_fp64": loss64,
        "grad_rel_err_vs_fp64": rel, "gen_s": gen_s}))


def cpu_solve_baseline(make_solver, b_np, runs, what):
    """The FP64 oracle (plain C, one thread) on the same problem, run to convergence: one
    warm-up (setup and one cycle), then `runs` runs of setup followed by a solve from x0 = 0 to
    the tolerance; median setup and median solve times.  `value` is the solve alone, like the
    GPU's `value` (hierarchy prebuilt); the setup is reported beside it."""
    import statistics
    o = make_solver()
    o.vcycle(np.zeros_like(b_np), b_np)
    setups, solves = [], []
    hist = None
    for _ in range(runs):
        t = time.perf_counter()
        o = make_solver()
        setups.append(time.perf_counter() - t)
        t = time.perf_counter()
        _, hist = o.solve(b_np)
        solves.append(time.perf_counter() - t)
        print(f"cpu baseline run {len(solves)}/{runs}: setup {setups[-1]:.2f} s, solve "
              f"{solves[-1]:.2f} s ({len(hist)} cycles)", file=sys.stderr, flush=True)
    return {
        "value": statistics.median(solves),
        "unit": "s (solve from x0=0 to 1e-10; setup excluded, as in value)",
        "cores": 1,
        "kind": "port",
        "sample": (f"FP64 C oracle, {what}: warm-up (setup + 1 cycle), then median of {runs} "
                   f"runs of setup + full solve ({len(hist)} cycles to rho={hist[-1]:.2e}); host "
                   f"'{cpu_model()}', nproc {os.cpu_count()}, 1 thread"),
        "setup_s": statistics.median(setups),
        "setup_plus_solve_s": statistics.median(setups) + statistics.median(solves),
        "iters": int(len(hist)),
        "runs": runs,
    }


def converged(h, tol, maxiter):
    return bool(len(h) > 0 and h[-1] < tol and len(h) <= maxiter)


def emit(line, ok):
    """Prints the JSON line; a record whose solves did not reach the tolerance is marked
    invalid and the run exits with status 1."""
    if not ok:
        line["valid"] = False
        line["invalid_reason"] = "a timed solve did not reach the tolerance"
    print(json.dumps(line), flush=True)
    if not ok:
        sys.exit(1)


def bench_c1(args):
    """The reference's CPU-runnable case: 2D 5-point Poisson on 511 x 511 interior nodes
    (src/2d/gallery.cc:7-39), f ~ U[-1, 1] (seed 1), point red-black Gauss-Seidel V(2,1) cycles
    to 1e-10, 8 levels; FP64 and mixed."""
    import torch

    import cedar_amd
    from cedar_amd import _abi
    from cedar_amd import gallery as G
    n = 511
    so_np, b_np = G.poisson_uniform_rhs(n, seed=1)
    so = torch.as_tensor(so_np).cuda()
    b = torch.as_tensor(b_np).cuda()
    kw = dict(relaxation="point", nrelax_pre=2, nrelax_post=1, maxiter=50, tol=1e-10)
    res = {}
    for prec in ("f64", "mixed"):
        s = cedar_amd.Solver(so, precision=prec, **kw)
        setup_s = s.times()[0]
        for _ in range(args.warmup):
            s.solve(b)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            x, h = s.solve(b)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        res[prec] = {"seconds": dt, "iters": int(len(h)), "rho": float(h[-1]),
                     "converged": converged(h, kw["tol"], kw["maxiter"]), "setup_s": setup_s,
                     "setup_plus_solve_s": setup_s + dt}
    # dominant kernel: one colour of the fine-level point Gauss-Seidel (FP64), live timing
    lib = _abi.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    sor = torch.zeros((2,) + tuple(b.shape), dtype=torch.float64, device="cuda")
    _abi.call("cdr_hip_setup_recip2d_f64", ctypes.c_void_p(so.data_ptr()),
              ctypes.c_void_p(sor.data_ptr()), n + 2, n + 2, 3, stream)
    q = torch.zeros_like(b)
    tg = time_kernel(lambda: lib.cdr_hip_relax_gs2d_f64(so.data_ptr(), 3, b.data_ptr(),
                                                        q.data_ptr(), sor.data_ptr(), n + 2,
                                                        n + 2, 0, stream), args.kernel_reps,
                     torch)
    npts = n * n
    bytes_sweep = 6 * 8 * npts  # w, s, 1/c, f, q read, q write: (S+3) words per point
    achieved = bytes_sweep / tg / 1e9
    cpu = None
    if args.cpu_runs > 0:
        from oracle import oracle as O
        cpu = cpu_solve_baseline(lambda: O.Solver(so_np, **kw), b_np, args.cpu_runs,
                                 f"{n}x{n} Poisson, point V(2,1)")
    f64, mix = res["f64"], res["mixed"]
    emit({
        "metric": "time-to-1e-10 rel. residual (s), 2D 5-point Poisson; mixed vs FP64",
        "value": f64["seconds"], "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": f64["seconds"] * 1e3, "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Poisson operator, seeded uniform RHS)",
        "config": {"workload": f"2D 5-point Poisson {n}x{n}, f~U[-1,1] seed 1, point GS "
                               "V(2,1), tol 1e-10", "levels": cedar_amd.Solver(so, **kw).nlevels,
                   "step": "one FP64 solve from x0=0 to rho<1e-10, hierarchy prebuilt"},
        "fp64": f64, "mixed": mix, "speedup_mixed_vs_fp64": f64["seconds"] / mix["seconds"],
        "roofline": {"bound": "hbm", "kernel": "gs_colour_k<double, 5-point>, fine level, one "
                     "sweep = 2 colour launches", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": bytes_sweep / 2, "avg_launch_s": tg / 2,
                     "note": "the 511^2 level (2 MB per array) stays in the caches; each colour "
                             "launch is bound by its launch latency, not by HBM"},
        "cpu_baseline": cpu,
    }, f64["converged"] and mix["converged"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=4095)
    ap.add_argument("--cpu-runs", type=int, default=3,
                    help="full CPU oracle solves (median) for the CPU baseline; 0 disables it")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--kernel-only", action="store_true",
                    help="run only the dominant kernel (fine-level FP32 x-line sweep) and a plain "
                         "256 MiB device copy, for hardware-counter runs")
    ap.add_argument("--levels-json", default="",
                    help="write the per-level, per-phase bandwidth table (device timers) here")
    ap.add_argument("--workload", choices=("c1", "c2", "c3", "c4", "c5"), default="c2")
    ap.add_argument("--n5", type=int, default=1025)
    ap.add_argument("--nprob", type=int, default=96)
    ap.add_argument("--c4-separate", action="store_true",
                    help="C4: solve the 4 right-hand sides one after another")
    ap.add_argument("--n3", type=int, default=255)
    ap.add_argument("--relax3", default="plane-xy")
    ap.add_argument("--cpu-cycles3", type=int, default=3,
                    help="3D oracle cycles timed for the C3 CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU runs are not part of it "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload in ("c1", "c3", "c4", "c5"):
        {"c1": bench_c1, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5}[args.workload](args)
        return

    import torch

    import cedar_amd
    from cedar_amd import _abi
    from cedar_amd import gallery as G

    n = args.n
    t0 = time.time()
    so_np, b_np, _ = G.heterogeneous_9pt(n, seed=2, block=16, sigma=2.0, contrast=1e4)
    gen_s = time.time() - t0
    so = torch.as_tensor(so_np).cuda()
    b = torch.as_tensor(b_np).cuda()
    kw = dict(relaxation="line-xy", nrelax_pre=1, nrelax_post=1, maxiter=100, tol=1e-10)

    if args.kernel_only:
        kernel_only(args, so, b)
        return

    solvers = {}
    for prec in ("f64", "mixed"):
        s = cedar_amd.Solver(so, precision=prec, **kw)
        solvers[prec] = (s, s.times()[0])

    def solve(prec):
        x = torch.zeros_like(b)
        x, h = solvers[prec][0].solve(b, x)
        return x, h

    results = {}
    for prec in ("f64", "mixed"):
        for _ in range(args.warmup):
            solve(prec)
    for prec in ("f64", "mixed"):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            x, h = solve(prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        results[prec] = dict(seconds=dt, iters=int(len(h)), rho=float(h[-1]), x=x, hist=h,
                             converged=converged(h, kw["tol"], kw["maxiter"]))

    # ---- achieved bandwidth per level and phase (the solver's device timers) ---------------
    per_level = {prec: level_rates(solvers[prec][0], b, 8 if prec == "f64" else 4, so.shape[0],
                                   torch) for prec in ("f64", "mixed")}
    if args.levels_json:
        with open(args.levels_json, "w") as f:
            json.dump(per_level, f, indent=1)

    def level_summary(rows):
        """per level: all phases' algorithmic bytes / all phases' time, as a fraction of peak"""
        out = []
        for r in rows:
            ph = [v for k, v in r.items() if isinstance(v, dict)]
            if not ph:
                continue
            byt = sum(v["GBps"] * v["s"] for v in ph)
            t = sum(v["s"] for v in ph)
            out.append([r["level"], r["n"][0], round(byt / t / HBM_PEAK_GBS, 3)])
        return out

    x64 = results["f64"]["x"].cpu().numpy()
    x32 = results["mixed"]["x"].cpu().numpy()
    mixed_vs_f64 = float(np.linalg.norm((x32 - x64)[1:-1, 1:-1]) / np.linalg.norm(x64[1:-1, 1:-1]))

    # ---- dominant kernel: fine-level zebra line relaxation, live HIP-event timing ----------
    ns = so.shape[0]
    ii, jj = so.shape[2], so.shape[1]
    npts = (ii - 2) * (jj - 2)
    lib = _abi.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    kern = {}
    for prec, dt, sfx in (("mixed", torch.float32, "f32"), ("f64", torch.float64, "f64")):
        s = solvers[prec][0]
        _, _, _, _, sor0, sor1 = s.level(0)
        sor_x = torch.empty((2, jj, ii), dtype=torch.float64, device="cuda")
        sor_y = torch.empty((2, jj, ii), dtype=torch.float64, device="cuda")
        nbytes = 2 * ii * jj * 8
        hip = cedar_amd.cdr2._hip()
        hip.hipMemcpy(ctypes.c_void_p(sor_x.data_ptr()), ctypes.c_void_p(sor0), nbytes, 3)
        hip.hipMemcpy(ctypes.c_void_p(sor_y.data_ptr()), ctypes.c_void_p(sor1), nbytes, 3)
        A = so.to(dt).contiguous()
        sx, sy = sor_x.to(dt).contiguous(), sor_y.to(dt).contiguous()
        q = torch.zeros_like(b, dtype=dt)
        f = b.to(dt).contiguous()
        w = 4 if sfx == "f32" else 8
        bytes_sweep = (ns + 3) * w * npts  # per full sweep (both colours): (S+3) w per point
        fx = getattr(lib, f"cdr_hip_relax_lines_x2d_{sfx}")
        fy = getattr(lib, f"cdr_hip_relax_lines_y2d_{sfx}")
        tx = time_kernel(lambda: fx(A.data_ptr(), ns, f.data_ptr(), q.data_ptr(), sx.data_ptr(),
                                    None, ii, jj, 0, stream), args.kernel_reps, torch)
        ty = time_kernel(lambda: fy(A.data_ptr(), ns, f.data_ptr(), q.data_ptr(), sy.data_ptr(),
                                    None, ii, jj, 0, stream), args.kernel_reps, torch)
        kern[prec] = dict(x_sweep_s=tx, y_sweep_direct_s=ty, bytes_sweep=bytes_sweep,
                          x_gbs=round(bytes_sweep / tx / 1e9, 1),
                          y_direct_gbs=round(bytes_sweep / ty / 1e9, 1))
        del A, sx, sy, q, f, sor_x, sor_y

    # ---- CPU baseline: the FP64 oracle (single thread) on the same problem -----------------
    cpu = None
    if args.cpu_runs > 0:
        from oracle import oracle as O
        cpu = cpu_solve_baseline(lambda: O.Solver(so_np, **kw), b_np, args.cpu_runs,
                                 f"this {n}x{n} problem, line-xy V(1,1)")

    f64, mix = results["f64"], results["mixed"]
    km = kern["mixed"]
    shapes = {(r["dir"], r["fp"]): r["kernel"] for r in cedar_amd.cdr2.autotune_report()
              if r["ii"] == ii and r["jj"] == jj and r["nb"] == 1 and r["stencil"] == 9}
    # The x-line kernel is the dominant kernel of the mixed solve: it runs the x sweeps and, where
    # the autotuner chose the transposed grid ("t" entry), the y sweeps as well.
    y_transposed = shapes.get(("t", 32), "").startswith("transposed")
    achieved = km["bytes_sweep"] / 2 / (km["x_sweep_s"] / 2) / 1e9  # per launch: one colour
    traffic = measured_traffic() if n == 4095 else None
    ok = f64["converged"] and mix["converged"]
    line = {
        "metric": "time-to-1e-10 rel. residual (s) + HBM GB/s %peak; mixed vs FP64 on same GPU",
        "value": mix["seconds"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mix["seconds"] * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 V-cycle + f64 refinement",
        "data": "synthetic (seeded lognormal block coefficients, uniform RHS)",
        "config": {
            "workload": f"2D 9-point -div(k grad u)=f, {n}x{n}, lognormal k sigma=2 on 16x16 "
                        "blocks (contrast<=1e4), V(1,1) line-xy, tol 1e-10",
            "levels": solvers["f64"][0].nlevels,
            "step": "one solve from x0=0 to rho<1e-10, hierarchy prebuilt",
        },
        "fp64": {"seconds": f64["seconds"], "iters": f64["iters"], "rho": f64["rho"],
                 "converged": f64["converged"], "setup_s": solvers["f64"][1],
                 "setup_plus_solve_s": solvers["f64"][1] + f64["seconds"]},
        "mixed": {"seconds": mix["seconds"], "iters": mix["iters"], "rho": mix["rho"],
                  "converged": mix["converged"], "setup_s": solvers["mixed"][1],
                  "setup_plus_solve_s": solvers["mixed"][1] + mix["seconds"],
                  "rel_l2_vs_fp64": mixed_vs_f64},
        "speedup_mixed_vs_fp64": f64["seconds"] / mix["seconds"],
        "roofline": {
            "bound": "hbm",
            "kernel": f"{shapes.get(('x', 32), 'x-lines')} <float, 9-point>, fine level, one "
                      "colour per launch (shape picked by the autotuner); runs the x sweeps"
                      + (" and, on the transposed grid, the y sweeps" if y_transposed else ""),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": ("profiles/relax_xlines_f32_4095_counters.md: rocprofv3 --pmc "
                               "FETCH_SIZE and WRITE_SIZE runs of bench.py --kernel-only, "
                               "calibrated on a 256 MiB copy" if traffic else None),
            "algorithmic_bytes_per_launch": km["bytes_sweep"] / 2,
            "avg_launch_s": km["x_sweep_s"] / 2,
        },
        "cpu_baseline": cpu,
        "levels_frac": {p: level_summary(per_level[p]) for p in ("f64", "mixed")},
        "kernels": kern,
        "line_shapes": {f"{d}_f{fp}": k for (d, fp), k in shapes.items()},
        "gen_s": round(gen_s, 2),
    }
    emit(line, ok)


if __name__ == "__main__":
    main()
