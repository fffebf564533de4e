"""Benchmark of the 3D Helmholtz hot path on one MI355X.

Workload (one "step"): the mixed-precision solve of the 200^3 heterogeneous
problem -- complex64 factorisation of every kz plane, then flexible GMRES with
the complex64 FFT-in-z preconditioner and the FP64 7-point operator until the
true relative residual is below 1e-8.  Inputs are resident in HBM before the
timed region.  Also reported: the same solve with the FP64 preconditioner, the
error of the mixed solution against it, the live roofline of the dominant
kernel (the block-tridiagonal plane solve) and the CPU baseline (the
reference's CPU path with MKL PARDISO plane solves on the host cores, or the
FP64 oracle where MKL is absent: the factorisation and one restart cycle at
full size, extrapolated to the GPU run's iteration count).

The default run adds the other configurations of the metric as keys of the
same JSON line, each with its own small step count (--extra-steps,
--extra-warmup): "config2" (128^3 incl. 16-point PML, FP64 preconditioner,
with the CPU baseline at that size),
"batched" (four frequencies on 96^3..192^3 grids, 8 sources each, right-hand
sides per second) and "adjoint" (128^3 16-source misfit gradients per second,
checked against the FP64 gradient and central finite differences).

    python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--size 200] [--m 35] [--config all|mixed200|batched|adjoint]
"""
import argparse
import importlib
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=200, help="grid nodes per axis (200: the 200^3 configuration)")
    ap.add_argument("--npml", type=int, default=None)
    ap.add_argument("--m", type=int, default=35, help="FGMRES restart length")
    ap.add_argument("--beta", type=float, default=1.0, help="complex shift of the preconditioner")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iters", type=int, default=20000)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--hodlr-eps", type=float, default=1e-4,
                    help="also solve with the HODLR-compressed complex64 factors at this eps "
                         "(the reference's HODLR threshold, main.cpp:161); < 0 skips")
    ap.add_argument("--hodlr-small", type=int, default=7, help="HODLR leaf size (smallsize)")
    ap.add_argument("--hodlr-one-way", action="store_true",
                    help="compress the one-directional factors (the reference's recursion) instead of the "
                         "twisted ones (two workgroups per plane over half chains)")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--config", choices=["all", "mixed200", "batched", "adjoint"], default="all",
                    help="all (default): the 200^3 mixed-precision solve as the headline, with the 128^3 FP64, "
                         "batched and adjoint configurations as extra keys; mixed200: the headline only; "
                         "batched: four frequencies on 96^3..192^3 grids with 8 sources each, solved "
                         "concurrently; adjoint: the 128^3 16-source misfit gradient")
    ap.add_argument("--extra-steps", type=int, default=1, help="timed steps of each extra configuration")
    ap.add_argument("--extra-warmup", type=int, default=1, help="untimed steps of each extra configuration")
    ap.add_argument("--fd-dirs", type=int, default=2, help="adjoint: finite-difference directions checked")
    ap.add_argument("--kernel-only", action="store_true",
                    help="factor, then run a 1 GiB device copy (counter reference) and 5 plane solves; "
                         "for hardware-counter passes")
    return ap.parse_args()


def pml_strength(npml, h, c_ref=1.0, refl=1e-6):
    """d0 of d(x) = d0 x^2 giving normal-incidence reflection `refl` through a
    layer of npml*h at speed c_ref: d0 = 3 c ln(1/R) / (2 delta)."""
    delta = npml * h
    return 3.0 * c_ref * math.log(1.0 / refl) / (2.0 * delta)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: every box of this build has a single MI355X and the solver "
              "path has no multi-GPU collectives; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    hz = importlib.import_module("iterative-helmholtz-solver-with-pardiso_amd")
    pr = hz.problem
    A = hz.abi
    if a.config == "batched":
        print(json.dumps(bench_batched(a, hz, np, torch, a.steps, a.warmup)), flush=True)
        return
    if a.config == "adjoint":
        print(json.dumps(bench_adjoint(a, hz, np, torch, a.steps, a.warmup)), flush=True)
        return

    n = a.size
    npml = a.npml if a.npml is not None else max(4, n // 10)
    h = 1.0 / n
    d0 = pr.cube_geometry(n, npml, 1.0)
    c = pr.random_velocity(d0, seed=a.seed)
    omega = pr.omega_for_ppw(float(c.min()), h, 10.0)
    pml_c = pml_strength(npml, h)
    d = pr.cube_geometry(n, npml, omega, beta=a.beta, pml_c=pml_c)
    src = (n // 2 - 1,) * 3
    f_host = pr.point_source(d, src, amplitude=1.0 / h ** 3)
    f = torch.from_numpy(f_host).cuda()
    S = hz.Solver(d, c)
    stream = torch.cuda.ExternalStream(S.stream)
    if a.kernel_only:
        kernel_only(S, d, A, torch)
        return

    def solve(prec):
        S.factor(prec)
        u, st = S.fgmres(f, m=a.m, restarts=max(1, a.max_iters // a.m), tol=a.tol, prec=prec, autotune=True)
        return u, st

    note("grid %d^3, omega %.2f, pml_c %.1f" % (n, omega, pml_c))
    # launch choices for one right-hand side (not timed; the solves reuse them)
    S.factor(A.PREC_C64)
    tuning = S.autotune(A.PREC_C64, 1)
    note("tuned: %s" % tuning)
    for w in range(a.warmup):
        _, stw = solve(A.PREC_C64)
        note("warmup %d: %d iterations" % (w, stw.iters))
    torch.cuda.synchronize()
    times = []
    u32, st32 = None, None
    t_fac = []
    for _ in range(a.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S.factor(A.PREC_C64)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        u32, st32 = S.fgmres(f, m=a.m, restarts=max(1, a.max_iters // a.m), tol=a.tol, prec=A.PREC_C64,
                             autotune=True)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        times.append(t2 - t0)
        t_fac.append(t1 - t0)
        note("timed step: %.2f s, %d iterations" % (t2 - t0, st32.iters))
    t_mixed = float(np.mean(times))

    # ---- roofline of the dominant kernel: the complex64 plane solve, timed with
    # HIP events on the problem's stream.  Algorithmic bytes per launch: the
    # 2 ny - 1 block reads per plane (nx^2 complex64, or the packed upper block
    # triangle for the symmetric kernel; the twisted solve: 2 ny packed blocks
    # per factor slot, both phases) plus four passes over the plane vectors
    # (read b, write t, read t, write x).
    nx, ny, nz = d.n
    N = d.N
    buf = torch.randn(N, dtype=torch.complex64, device="cuda")
    reps = 10
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    # untimed launches first: after the host-side pause the first few launches
    # run up to 15% slower while the clocks ramp back up
    for _ in range(3):
        A.check(A.lib().hz_plane_solve(S.h, A.PREC_C64, 0, 1, buf.data_ptr()), "hz_plane_solve")
    e0.record(stream)
    for _ in range(reps):
        A.check(A.lib().hz_plane_solve(S.h, A.PREC_C64, 0, 1, buf.data_ptr()), "hz_plane_solve")
    e1.record(stream)
    torch.cuda.synchronize()
    ms_kernel = e0.elapsed_time(e1) / reps
    fam = tuning.get("solve_kernel", 0)
    bytes_kernel = plane_solve_bytes(d.n, fam, S.slots()[0])
    achieved = bytes_kernel / (ms_kernel * 1e-3) / 1e9

    out = {
        "metric": "3D Helmholtz time-to-rel-res 1e-8, mixed-precision preconditioner (%d^3 heterogeneous)" % n,
        "value": round(t_mixed, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * t_mixed, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 preconditioner + f64 Krylov",
        "data": "synthetic",
        "config": {
            "workload": "%d^3 incl. %d-point PML, seeded random c in [0.7,1.5] (corr. 0.1), 10 ppw at c_min, "
                        "1 centre source, FGMRES(%d) to %.0e" % (n, npml, a.m, a.tol),
            "grid": [nx, ny, nz], "npml": npml, "omega": omega, "beta": a.beta, "pml_c": pml_c,
            "seed": a.seed, "restart": a.m,
        },
        "iters_mixed": st32.iters,
        "rel_res_mixed": st32.rel_res,
        "factor_c64_s": round(float(np.mean(t_fac)), 4),
        "tuning": tuning,
        "roofline": {
            "kernel": "%s (block-tridiagonal plane solve, complex64)" % plane_solve_kernel(nx, tuning),
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(n, fam),
            "traffic_source": traffic_source(n, fam),
            "bytes_per_launch": bytes_kernel,
            "ms_per_launch": round(ms_kernel, 4),
        },
    }

    if a.hodlr_eps >= 0:
        out["hodlr"] = bench_hodlr(a, S, A, f, u32, t_mixed, stream, buf, torch, np)

    if not a.skip_fp64:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u64, st64 = solve(A.PREC_FP64)
        torch.cuda.synchronize()
        t_fp64 = time.perf_counter() - t0
        note("fp64 solve: %.2f s, %d iterations" % (t_fp64, st64.iters))
        err = (torch.linalg.vector_norm(u32 - u64) / torch.linalg.vector_norm(u64)).item()
        out.update({"fp64_gpu_s": round(t_fp64, 4), "iters_fp64": st64.iters,
                    "speedup_mixed_vs_fp64": round(t_fp64 / t_mixed, 3), "err_mixed_vs_fp64": err})
        del u64
        S.close()  # free the factors before the CPU leg
        torch.cuda.empty_cache()

    if not a.skip_cpu:
        note("cpu baseline sample")
        out["cpu_baseline"] = cpu_baseline(d, c, f_host, st32.iters, a.m, a.cpu_threads)
    del u32, f, buf
    S.close()
    torch.cuda.empty_cache()
    if a.config == "all":
        for key, fn in (("config2", bench_config2), ("batched", bench_batched), ("adjoint", bench_adjoint)):
            note("extra configuration: %s" % key)
            try:
                out[key] = fn(a, hz, np, torch, a.extra_steps, a.extra_warmup)
            except Exception as e:  # the headline stands on its own
                out[key] = {"error": "%s: %s" % (type(e).__name__, e)}
            torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


def bench_hodlr(a, S, A, f, u_dense, t_dense, stream, buf, torch, np):
    """The same solve with the complex64 block inverses HODLR-compressed
    (hz_precond_compress, the reference's HODLR build at its threshold; by
    default the twisted factors' records): factorisation, compression and
    FGMRES timed together; the compressed plane solve timed with HIP events
    like the dense one."""
    twisted = not a.hodlr_one_way
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S.factor(A.PREC_C64)
    try:
        info = S.compress(A.PREC_C64, eps=a.hodlr_eps, smallsize=a.hodlr_small, twisted=twisted)
    except A.HzError:
        if not twisted:
            raise
        # no twisted factors for this width: the one-directional records
        twisted = False
        info = S.compress(A.PREC_C64, eps=a.hodlr_eps, smallsize=a.hodlr_small)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    u, st = S.fgmres(f, m=a.m, restarts=max(1, a.max_iters // a.m), tol=a.tol, prec=A.PREC_C64, autotune=True)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    note("hodlr eps %g: %.2f s, %d iterations" % (a.hodlr_eps, t2 - t0, st.iters))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(3):
        A.check(A.lib().hz_plane_solve(S.h, A.PREC_C64, 0, 1, buf.data_ptr()), "hz_plane_solve")
    reps = 10
    e0.record(stream)
    for _ in range(reps):
        A.check(A.lib().hz_plane_solve(S.h, A.PREC_C64, 0, 1, buf.data_ptr()), "hz_plane_solve")
    e1.record(stream)
    torch.cuda.synchronize()
    err = (torch.linalg.vector_norm(u - u_dense) / torch.linalg.vector_norm(u_dense)).item()
    S.compress(A.PREC_C64, eps=-1.0)
    return {"eps": a.hodlr_eps, "smallsize": a.hodlr_small, "twisted": twisted, "time_s": round(t2 - t0, 4),
            "iters": st.iters,
            "rel_res": st.rel_res, "factor_compress_s": round(t1 - t0, 4),
            "plane_solve_ms": round(e0.elapsed_time(e1) / reps, 4), "levels": info["levels"],
            "max_rank": info["max_rank"], "mean_rank": round(info["mean_rank"], 3),
            "record_bytes": info["bytes"], "dense_factor_bytes": info["dense_bytes"],
            "err_vs_dense_mixed": err, "speedup_vs_dense_mixed": round(t_dense / (t2 - t0), 3)}


def plane_solve_kernel(nx, tuning):
    """Name of the complex64 single-RHS plane-solve kernel the tuned choice launches."""
    fam = tuning.get("solve_kernel", 0)
    if fam == 6:
        return "k_solve_twist_pw<%d, 0> + k_solve_twist_pw<%d, 1>" % (nx, nx)
    if fam == 5:
        return "k_solve_sym2<%d, false>" % nx
    if fam == 3:
        return "k_solve_dma<float,2,5,20,200>"
    if fam == 1:
        return "k_solve_mr<float,...,%d,1>" % nx
    return "k_solve_fast<float,2,8>"


def note(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def bench_batched(a, hz, np, torch, steps, warmup):
    """Four frequencies, each on its own grid scaled for 10 points per minimum
    wavelength (96^3, 128^3, 160^3, 192^3 incl. 12..24-point PML), 8 point
    sources each at seeded uniform-random interior nodes (32 right-hand sides),
    seeded random-field velocity, complex64 preconditioner, FP64 FGMRES to 1e-8.
    The four grids are factored and solved concurrently (one stream each); each
    grid advances its 8 sources in one lockstep batch with graph-replayed
    Arnoldi steps.  Value: right-hand sides solved per second, factorisation
    included."""
    pr, A = hz.problem, hz.abi
    grids = [(96, 12), (128, 16), (160, 20), (192, 24)]
    nsrc = 8
    solvers, fs, info = [], [], []
    for i, (n, npml) in enumerate(grids):
        h = 1.0 / n
        d0 = pr.cube_geometry(n, npml, 1.0)
        c = pr.random_velocity(d0, seed=a.seed + i)
        omega = pr.omega_for_ppw(float(c.min()), h, 10.0)
        d = pr.cube_geometry(n, npml, omega, beta=a.beta, pml_c=pml_strength(npml, h))
        rng = np.random.default_rng(a.seed + 100 + i)
        pts = [tuple(int(v) for v in rng.integers(npml, n - npml, size=3)) for _ in range(nsrc)]
        f = np.stack([pr.point_source(d, p, amplitude=1.0 / h ** 3) for p in pts])
        solvers.append(hz.Solver(d, c))
        fs.append(torch.from_numpy(f).cuda())
        info.append({"grid": n, "npml": npml, "omega": omega})
        note("grid %d^3 omega %.2f" % (n, omega))
    # launch choices per grid, each tuned alone on the idle GPU (not timed)
    hz.factor_many(solvers, A.PREC_C64)
    for S, inf in zip(solvers, info):
        inf["tuning"] = S.autotune(A.PREC_C64, nsrc)
        note("grid %d^3 tuned: %s" % (inf["grid"], inf["tuning"]))
    for w in range(warmup):
        hz.factor_many(solvers, A.PREC_C64)
        hz.fgmres_many(solvers, fs, m=a.m, restarts=max(1, a.max_iters // a.m), tol=a.tol, prec=A.PREC_C64,
                       batch=nsrc, graph=True)
        note("warmup %d done" % w)
    times = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hz.factor_many(solvers, A.PREC_C64)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = hz.fgmres_many(solvers, fs, m=a.m, restarts=max(1, a.max_iters // a.m), tol=a.tol,
                             prec=A.PREC_C64, batch=nsrc, graph=True)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        times.append((t1 - t0, t2 - t0))
        note("timed step: factor %.2f s, total %.2f s" % (t1 - t0, t2 - t0))
    t_fac = float(np.mean([t[0] for t in times]))
    t_tot = float(np.mean([t[1] for t in times]))
    conv = True
    for k, (S, (u, st)) in enumerate(zip(solvers, res)):
        info[k]["iters"] = [s.iters for s in st]
        info[k]["rel_res_max"] = max(s.rel_res for s in st)
        info[k]["solve_s"] = round(max(s.t_solve_s for s in st), 3)
        conv = conv and all(s.converged for s in st)
    total_rhs = nsrc * len(grids)
    out = {
        "metric": "batched uneven solves: right-hand sides to rel-res 1e-8 per second (4 grids x 8 sources)",
        "value": round(total_rhs / t_tot, 4), "unit": "rhs/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * t_tot, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 preconditioner + f64 Krylov", "data": "synthetic",
        "config": {"workload": "96^3/128^3/160^3/192^3 incl. 12-24-point PML, seeded random c in [0.7,1.5], "
                               "10 ppw at c_min, 8 seeded interior sources per grid, FGMRES(%d) to %.0e, lockstep "
                               "batches of 8, graph-replayed chunks of Arnoldi steps, grids concurrent from one "
                               "host thread" % (a.m, a.tol),
                   "grids": info},
        "factor_s": round(t_fac, 4), "converged": conv,
    }
    for S in solvers:
        S.close()
    return out


def bench_adjoint(a, hz, np, torch, steps, warmup):
    """Adjoint misfit gradient on 128^3 incl. 16-point PML.  True model c = 1 +
    3 seeded Gaussian blobs (+-10%, width 0.08); current model its periodic
    Gaussian smoothing (sigma 0.15).  16 sources on a regular 4x4 grid on the
    plane 4 nodes below the top-z layer interface (k = npml + 4), receivers at
    every physical node of that plane, omega for 10 points per wavelength at
    c = 1.  Observed data: forward solves on the true model.  Timed: one misfit
    forward (16 sources, fields stored) + backward (16 adjoint solves with the
    conjugate-transposed factors + fused gradient), complex64 preconditioner,
    tolerance 1e-8.  Checked against the FP64-preconditioner gradient and
    central finite differences of J along seeded smooth directions."""
    pr, A = hz.problem, hz.abi
    n, npml = 128, 16
    h = 1.0 / n
    omega = pr.omega_for_ppw(1.0, h, 10.0)
    d = pr.cube_geometry(n, npml, omega, beta=a.beta, pml_c=pml_strength(npml, h))
    rng = np.random.default_rng(a.seed)
    lo, hi = (npml + 1) * h, (n - npml) * h
    centres = [tuple(rng.uniform(lo + 0.1, hi - 0.1, size=3)) for _ in range(3)]
    amps = [0.1 * (1 if rng.random() < 0.5 else -1) for _ in range(3)]
    c_true = pr.gaussian_blobs(d, centres, amps, 0.08)
    c_cur = pr.gaussian_smooth(c_true, d, 0.15)
    kz = npml + 4
    xs = [npml + int((i + 0.5) * (n - 2 * npml) / 4) for i in range(4)]
    f = np.zeros((16, d.N), complex)
    for q, (i, j) in enumerate([(i, j) for j in xs for i in xs]):
        f[q, pr.flat_index(d, i, j, kz)] = 1.0 / h ** 3
    rec = np.array([pr.flat_index(d, i, j, kz) for j in range(npml, n - npml) for i in range(npml, n - npml)])
    fd = torch.from_numpy(f).cuda()
    kw = dict(m=a.m, restarts=max(1, a.max_iters // a.m), tol=a.tol)
    # observed data from the true model
    St = hz.Solver(d, c_true)
    St.factor(A.PREC_C64)
    ut, stt = St.fgmres(fd, prec=A.PREC_C64, batch=16, **kw)
    data = ut.cpu().numpy()[:, rec]
    St.close()
    del St, ut
    torch.cuda.empty_cache()
    note("data from the true model: %d iterations (max over sources)" % max(s.iters for s in stt))
    S = hz.Solver(d, c_cur)
    S.factor(A.PREC_C64)
    for w in range(warmup):
        S.misfit_forward(fd, rec, data, prec=A.PREC_C64, batch=16, **kw)
        S.misfit_backward()
    times = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = S.misfit_forward(fd, rec, data, prec=A.PREC_C64, batch=16, **kw)
        g = S.misfit_backward()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        note("timed gradient: %.2f s, J = %.6e" % (times[-1], J))
    t = float(np.mean(times))
    g32 = g.cpu().numpy()
    # FP64-preconditioner reference gradient
    S.factor(A.PREC_FP64)
    J64 = S.misfit_forward(fd, rec, data, prec=A.PREC_FP64, batch=16, **kw)
    g64 = S.misfit_backward().cpu().numpy()
    err64 = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    S.close()
    del S
    torch.cuda.empty_cache()
    # central finite differences of J(m) along seeded smooth directions, m = 1/c^2
    m0 = 1.0 / c_cur ** 2
    fd_rel = []
    for k in range(a.fd_dirs):
        dm = pr.gaussian_smooth(np.random.default_rng(a.seed + 500 + k).standard_normal(d.N), d, 0.05)
        dm *= 1e-6 * np.linalg.norm(m0) / np.linalg.norm(dm)
        Jpm = []
        for sgn in (1.0, -1.0):
            Sp = hz.Solver(d, 1.0 / np.sqrt(m0 + sgn * dm))
            Sp.factor(A.PREC_FP64)
            Jpm.append(Sp.misfit_forward(fd, rec, data, prec=A.PREC_FP64, batch=16, m=a.m,
                                         restarts=max(1, a.max_iters // a.m), tol=1e-11))
            Sp.close()
            del Sp
            torch.cuda.empty_cache()
        dJ_fd = (Jpm[0] - Jpm[1]) / 2.0
        dJ_adj = float(np.dot(g32, dm))
        fd_rel.append(abs(dJ_fd - dJ_adj) / abs(dJ_fd))
        note("fd direction %d: adjoint %.6e fd %.6e rel %.2e" % (k, dJ_adj, dJ_fd, fd_rel[-1]))
    out = {
        "metric": "adjoint misfit gradients per second (128^3, 16 sources, mixed precision, 1e-8)",
        "value": round(1.0 / t, 5), "unit": "grads/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * t, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 preconditioner + f64 Krylov", "data": "synthetic",
        "config": {"workload": "128^3 incl. 16-point PML, 3 Gaussian blobs +-10%% (w 0.08) smoothed by sigma 0.15, "
                               "16 sources on a 4x4 grid at k = npml + 4, %d receivers on that plane, 10 ppw at c=1, "
                               "FGMRES(%d) to %.0e, sources in one lockstep batch" % (rec.size, a.m, a.tol),
                   "omega": omega},
        "J": J, "J_fp64": J64, "grad_rel_err_vs_fp64": err64, "fd_rel_err": fd_rel,
    }
    return out


def bench_config2(a, hz, np, torch, steps, warmup):
    """128^3 incl. 16-point PML, seeded random c in [0.7, 1.5] (correlation 0.1),
    10 points per minimum wavelength, one centre source, FP64 preconditioner
    and FGMRES to relative residual 1e-8 (the value); the mixed-precision
    solve and its error against the FP64 one beside it.  The true residual of
    the FP64 solution is recomputed with the 7-point operator.  Unless
    --skip-cpu: the reference's CPU path at the same size (cpu_baseline, to
    the FP64 run's iteration count) and the GPU's speed-up over it."""
    pr, A = hz.problem, hz.abi
    n, npml = 128, 16
    h = 1.0 / n
    c = pr.random_velocity(pr.cube_geometry(n, npml, 1.0), seed=a.seed)
    omega = pr.omega_for_ppw(float(c.min()), h, 10.0)
    d = pr.cube_geometry(n, npml, omega, beta=a.beta, pml_c=pml_strength(npml, h))
    f_host = pr.point_source(d, (n // 2 - 1,) * 3, amplitude=1.0 / h ** 3)
    f = torch.from_numpy(f_host).cuda()
    S = hz.Solver(d, c)
    kw = dict(m=a.m, restarts=max(1, a.max_iters // a.m), tol=a.tol)
    res = {}
    for prec, name in ((A.PREC_FP64, "fp64"), (A.PREC_C64, "mixed")):
        S.factor(prec)
        S.autotune(prec, 1)
        for _ in range(warmup):
            S.fgmres(f, prec=prec, **kw)
        times = []
        for _ in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            S.factor(prec)
            u, st = S.fgmres(f, prec=prec, **kw)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        res[name] = (float(np.mean(times)), u, st)
        note("config2 %s: %.2f s, %d iterations" % (name, res[name][0], st.iters))
    t64, u64, st64 = res["fp64"]
    t32, u32, st32 = res["mixed"]
    _, rel = S.residual(f, u64)
    out = {
        "metric": "time-to-rel-res 1e-8, FP64 preconditioner (128^3 heterogeneous)",
        "value": round(t64, 4), "unit": "s", "steps": steps, "warmup": warmup, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "128^3 incl. 16-point PML, seeded random c in [0.7,1.5] (corr. 0.1), 10 ppw at "
                               "c_min, 1 centre source, FGMRES(%d) to %.0e" % (a.m, a.tol),
                   "omega": omega, "beta": a.beta, "seed": a.seed},
        "iters_fp64": st64.iters, "true_rel_res_fp64": float(rel[0]), "converged": bool(st64.converged),
        "mixed_s": round(t32, 4), "iters_mixed": st32.iters,
        "err_mixed_vs_fp64": (torch.linalg.vector_norm(u32 - u64) / torch.linalg.vector_norm(u64)).item(),
    }
    S.close()
    if not a.skip_cpu:
        note("config2 cpu baseline sample")
        cpu = cpu_baseline(d, c, f_host, st64.iters, a.m, a.cpu_threads)
        out["cpu_baseline"] = cpu
        out["speedup_fp64_gpu_vs_cpu"] = round(cpu["value"] / t64, 1)
    return out


def traffic_source(n, fam):
    """Where roofline.traffic comes from: it is not measured in this run (the
    counter passes need rocprofv3), it is read from the committed summary."""
    names = {3: "k_solve_dma_c64_200cube_pmc.md", 5: "k_solve_sym2_c64_200cube_pmc.md",
             6: "k_solve_twist_c64_200cube_pmc.md"}
    if n != 200 or fam not in names or not os.path.exists(os.path.join(ROOT, "profiles", names[fam])):
        return None
    return "committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary profiles/%s (not measured in this run)" % names[fam]


def plane_solve_bytes(dims, fam, nslot):
    """Algorithmic HBM bytes of one complex64 single-RHS plane solve: 2 ny - 1
    block reads per plane (the full nx^2 block, or for the symmetric kernel the
    upper block triangle of 8x8 tiles, 64 T (T+1)/2 entries with T = nx/8; the
    twisted solve reads 2 ny packed blocks per factor slot for both of its
    planes: ny - 1 in phase 0, ny + 1 in phase 1) and four passes over the
    nx ny nz vector."""
    nx, ny, nz = dims
    if fam in (5, 6):
        t = nx // 8
        blk = 64 * t * (t + 1) // 2
    else:
        blk = nx * nx
    blocks = 2 * ny * nslot if fam == 6 else (2 * ny - 1) * nz
    return blocks * blk * 8 + 4 * 8 * nx * ny * nz


def measured_traffic(n, fam):
    """HBM bytes per launch of the plane solve from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE passes (scaled by a 1 GiB copy), for the 200^3 grid."""
    names = {3: "k_solve_dma_c64_200cube_pmc.md", 5: "k_solve_sym2_c64_200cube_pmc.md",
             6: "k_solve_twist_c64_200cube_pmc.md"}
    if n != 200 or fam not in names:
        return None
    path = os.path.join(ROOT, "profiles", names[fam])
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def kernel_only(S, d, A, torch):
    S.factor(A.PREC_C64)
    src = torch.zeros(1 << 26, dtype=torch.complex128, device="cuda")  # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    buf = torch.randn(d.N, dtype=torch.complex64, device="cuda")
    # 5 launches of each single-RHS kernel that fits the grid: twisted packed
    # (family 6, two phase launches), symmetric packed (family 5), then the
    # full-block LDS ring (family 3, 200-wide planes)
    fams = [6, 5, 3] if d.n[0] == 200 else [6, 5]
    saved = os.environ.get("HZ_SOLVE_KERNEL")
    for fam in fams:
        os.environ["HZ_SOLVE_KERNEL"] = {6: "twist", 5: "sym", 3: "dma"}[fam]
        for _ in range(5):
            A.check(A.lib().hz_plane_solve(S.h, A.PREC_C64, 0, 1, buf.data_ptr()), "hz_plane_solve")
        A.check(A.lib().hz_problem_sync(S.h), "sync")
    if saved is None:
        os.environ.pop("HZ_SOLVE_KERNEL", None)
    else:
        os.environ["HZ_SOLVE_KERNEL"] = saved
    # then 5 launches of the HODLR-compressed solve (eps 1e-4, smallsize 7):
    # one-directional records, then the twisted ones
    for tw in (False, True):
        info = S.compress(A.PREC_C64, eps=1e-4, smallsize=7, twisted=tw)
        for _ in range(5):
            A.check(A.lib().hz_plane_solve(S.h, A.PREC_C64, 0, 1, buf.data_ptr()), "hz_plane_solve")
        A.check(A.lib().hz_problem_sync(S.h), "sync")
    print(json.dumps({"kernel_only": True, "copy_bytes": 2 * (1 << 30),
                      "plane_solve_bytes": {str(f): plane_solve_bytes(d.n, f, S.slots()[0]) for f in fams},
                      "hodlr_record_bytes": info["bytes"]}), flush=True)


def _cpu_threads(threads):
    if threads is None:
        try:
            threads = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            threads = min(16, os.cpu_count() or 1)
    return threads


def cpu_sample(orc, d, c, iters, m, threads):
    """Time-to-solution of the FP64 oracle (banded-LU plane solves, recursive
    mixed-radix z-FFT, 7-point operator, MGS) on `threads` host threads,
    extrapolated from a sample: the factorisation and one solve of one plane
    per thread, one z-FFT pair, one operator apply and one MGS step timed on
    the full-size problem, scaled to all planes and `iters` FGMRES(m)
    iterations."""
    import numpy as np
    P = orc.Problem(d, c)
    nz, n2, N = d.n[2], d.n[0] * d.n[1], d.N
    kp = min(threads, nz)
    t0 = time.perf_counter()
    P.factor_planes(0, kp)
    t_fac = time.perf_counter() - t0
    b = np.random.default_rng(0).standard_normal(kp * n2) + 0j
    t0 = time.perf_counter()
    P.solve_planes(0, kp, b)
    t_sol = time.perf_counter() - t0
    x = np.random.default_rng(1).standard_normal(N) + 1j
    t0 = time.perf_counter()
    y = P.fft_z(x, -1)
    P.fft_z(y, +1)
    t_fft = time.perf_counter() - t0
    t0 = time.perf_counter()
    P.op_apply(x)
    t_op = time.perf_counter() - t0
    V = [np.random.default_rng(s).standard_normal(N) + 0j for s in range(2)]
    t0 = time.perf_counter()
    h = np.vdot(V[0], V[1])
    V[1] -= h * V[0]
    t_mgs1 = time.perf_counter() - t0
    del V, x, y, b, P
    planes = nz / kp
    t_factor = t_fac * planes
    t_iter = t_sol * planes + t_fft + t_op + t_mgs1 * (m / 2.0 + 1)
    return {"total": t_factor + iters * t_iter, "factor": t_factor, "iter": t_iter, "kp": kp, "t_fac": t_fac}


def cpu_baseline(d, c, f_host, iters, m, threads):
    """The reference's CPU path on the host cores: MKL PARDISO plane solves
    (oracle/pardiso_cpu.py) when the MKL runtime is present, else the FP64
    oracle's banded LU (cpu_baseline_oracle).  Measured at full size: the
    PARDISO factorisation of every solved plane and one whole FGMRES(m)
    restart cycle from zero on the benchmark's right-hand side (preconditioner,
    operator, modified Gram-Schmidt, solution update and restart residual);
    value = factorisation + the GPU run's iteration count x the cycle's time
    per iteration."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pardiso_cpu
    sys.path.pop(0)
    if not pardiso_cpu.available():
        return cpu_baseline_oracle(d, c, iters, m, threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    sys.path.pop(0)
    threads = _cpu_threads(threads)
    orc.lib().or_set_threads(threads)
    pardiso_cpu.set_threads(threads)
    P = orc.Problem(d, c)
    t0 = time.perf_counter()
    M = pardiso_cpu.PardisoPrecond(P, threads=threads)
    t_fac = time.perf_counter() - t0
    M.apply(f_host)  # first applies carry one-time set-up (FFT plans, solve buffers)
    t0 = time.perf_counter()
    _, it, _ = pardiso_cpu.fgmres(P, M, f_host, m=m, tol=1e-30, max_iters=m)
    t_cycle = time.perf_counter() - t0
    M.release()
    del P, M
    t_iter = t_cycle / max(it, 1)
    return {
        "value": round(t_fac + iters * t_iter, 2),
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "solver": "MKL PARDISO (mtype 13, libmkl_rt) per plane, scipy z-FFTs, oracle 7-point operator, "
                  "numpy FGMRES(%d) with modified Gram-Schmidt" % m,
        "sample": "the reference's CPU path with its plane solver on %d host threads at full size: PARDISO "
                  "factorisation of every solved plane (%.2f s) and one FGMRES(%d) restart cycle from zero on the "
                  "same right-hand side (%d iterations, %.2f s), extrapolated to the GPU run's %d iterations"
                  % (threads, t_fac, m, it, t_cycle, iters),
        "factor_s": round(t_fac, 3),
        "iter_s": round(t_iter, 4),
        "cycle_iters": it,
    }


def cpu_baseline_oracle(d, c, iters, m, threads):
    """The FP64 oracle on the host cores.  The 200^3 time is an extrapolated
    sample (cpu_sample), calibrated against a measured end-to-end solve of the
    64^3 configuration (c = 1, 8-point PML, 10 points per wavelength, centre
    source, FGMRES(m) to 1e-8) on the same threads: value = sample estimate x
    (measured 64^3 time / sample estimate at 64^3)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    sys.path.pop(0)
    threads = _cpu_threads(threads)
    orc.lib().or_set_threads(threads)
    hz = importlib.import_module("iterative-helmholtz-solver-with-pardiso_amd")
    pr = hz.problem
    est = cpu_sample(orc, d, c, iters, m, threads)
    # calibration: the 64^3 configuration end to end
    n1, p1 = 64, 8
    h1 = 1.0 / n1
    d1 = pr.cube_geometry(n1, p1, pr.omega_for_ppw(1.0, h1, 10.0), beta=d.beta, pml_c=pml_strength(p1, h1))
    c1 = np.ones(d1.N)
    f1 = pr.point_source(d1, (n1 // 2 - 1,) * 3, amplitude=1.0 / h1 ** 3)
    P1 = orc.Problem(d1, c1)
    t0 = time.perf_counter()
    _, st1 = P1.fgmres(f1, m=m, rest=400, tol=1e-8)
    t64 = time.perf_counter() - t0
    del P1
    est64 = cpu_sample(orc, d1, c1, st1.iters, m, threads)
    ratio = t64 / est64["total"]
    return {
        "value": round(est["total"] * ratio, 2),
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": "FP64 oracle on %d host threads: banded-LU factor + solve of %d of %d planes, one z-FFT pair, "
                  "one 7-point apply, one MGS step at full size, extrapolated to all planes and %d FGMRES "
                  "iterations (%.1f s), times the calibration ratio %.3f = measured / extrapolated time of the "
                  "64^3 configuration solved end to end (%.2f s measured, %d iterations)"
                  % (threads, est["kp"], d.n[2], iters, est["total"], ratio, t64, st1.iters),
        "extrapolated_s": round(est["total"], 2),
        "factor_s_est": round(est["factor"], 2),
        "iter_s_est": round(est["iter"], 4),
        "calibration": {"config": "64^3 incl. 8-point PML, c = 1, 10 ppw, centre source, FGMRES(%d) to 1e-8" % m,
                        "measured_s": round(t64, 3), "iters": st1.iters, "extrapolated_s": round(est64["total"], 3),
                        "ratio": round(ratio, 4), "converged": bool(st1.converged)},
    }


if __name__ == "__main__":
    main()
