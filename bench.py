"""Benchmark of the 3D Helmholtz hot path on one MI355X.

Workload (one "step"): the mixed-precision solve of the 200^3 heterogeneous
problem -- complex64 factorisation of every kz plane, then flexible GMRES with
the complex64 FFT-in-z preconditioner and the FP64 7-point operator until the
true relative residual is below 1e-8.  Inputs are resident in HBM before the
timed region.  Also reported: the same solve with the FP64 preconditioner, the
error of the mixed solution against it, the live roofline of the dominant
kernel (the block-tridiagonal plane solve) and the CPU baseline (the FP64
oracle on the host cores, on a bounded sample of the same workload).

    python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--size 200] [--m 60]
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
    ap.add_argument("--m", type=int, default=30, help="FGMRES restart length")
    ap.add_argument("--beta", type=float, default=1.0, help="complex shift of the preconditioner")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iters", type=int, default=20000)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=None)
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
        u, st = S.fgmres(f, m=a.m, restarts=max(1, a.max_iters // a.m), tol=a.tol, prec=prec)
        return u, st

    def note(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    note("grid %d^3, omega %.2f, pml_c %.1f" % (n, omega, pml_c))
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
        u32, st32 = S.fgmres(f, m=a.m, restarts=max(1, a.max_iters // a.m), tol=a.tol, prec=A.PREC_C64)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        times.append(t2 - t0)
        t_fac.append(t1 - t0)
        note("timed step: %.2f s, %d iterations" % (t2 - t0, st32.iters))
    t_mixed = float(np.mean(times))

    # ---- roofline of the dominant kernel: the complex64 plane solve, timed with
    # HIP events on the problem's stream.  Algorithmic bytes per launch: the
    # 2 ny - 1 block reads of nx^2 complex64 per plane plus four passes over the
    # plane vectors (read b, write t, read t, write x).
    nx, ny, nz = d.n
    N = d.N
    buf = torch.randn(N, dtype=torch.complex64, device="cuda")
    reps = 10
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        A.check(A.lib().hz_plane_solve(S.h, A.PREC_C64, 0, 1, buf.data_ptr()), "hz_plane_solve")
    e1.record(stream)
    torch.cuda.synchronize()
    ms_kernel = e0.elapsed_time(e1) / reps
    bytes_kernel = (2 * ny - 1) * nx * nx * 8 * nz + 4 * 8 * N
    achieved = bytes_kernel / (ms_kernel * 1e-3) / 1e9

    out = {
        "metric": "3D Helmholtz time-to-rel-res 1e-8, mixed-precision preconditioner (200^3 heterogeneous)",
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
        "roofline": {
            "kernel": "%s (block-tridiagonal plane solve, complex64)" % plane_solve_kernel(nx),
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(n),
            "bytes_per_launch": bytes_kernel,
            "ms_per_launch": round(ms_kernel, 4),
        },
    }

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
        out["cpu_baseline"] = cpu_baseline(d, c, st32.iters, a.m, a.cpu_threads)
    print(json.dumps(out), flush=True)


def plane_solve_kernel(nx):
    """Name of the complex64 plane-solve kernel libhz.so launches for nx."""
    return "k_solve_dma<float,2,5,20,200>" if nx == 200 else "k_solve_fast<float,2,8>"


def measured_traffic(n):
    """HBM bytes per launch of the plane solve from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE passes (scaled by a 1 GiB copy), for the 200^3 grid."""
    if n != 200:
        return None
    path = os.path.join(ROOT, "profiles", "plane_solve_c64_200cube_pmc.md")
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
    for _ in range(5):
        A.check(A.lib().hz_plane_solve(S.h, A.PREC_C64, 0, 1, buf.data_ptr()), "hz_plane_solve")
    A.check(A.lib().hz_problem_sync(S.h), "sync")
    nx, ny, nz = d.n
    print(json.dumps({"kernel_only": True, "copy_bytes": 2 * (1 << 30),
                      "plane_solve_bytes": (2 * ny - 1) * nx * nx * 8 * nz + 4 * 8 * d.N}), flush=True)


def cpu_baseline(d, c, iters, m, threads):
    """The FP64 oracle (banded-LU plane solves, recursive mixed-radix z-FFT,
    7-point operator, MGS) on the host cores: the factorisation and one solve of
    a sample of planes (one per thread), one z-FFT pair, one operator apply and
    one MGS step are timed on the full-size problem; the time to solution is
    extrapolated to all planes and to the GPU run's iteration count."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    sys.path.pop(0)
    if threads is None:
        try:
            threads = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            threads = min(16, os.cpu_count() or 1)
    orc.lib().or_set_threads(threads)
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
    del V, x, y, b
    planes = nz / kp
    t_factor = t_fac * planes
    t_iter = t_sol * planes + t_fft + t_op + t_mgs1 * (m / 2.0 + 1)
    total = t_factor + iters * t_iter
    return {
        "value": round(total, 2),
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": "FP64 oracle on %d host threads: banded-LU factor + solve of %d of %d planes, one z-FFT pair, "
                  "one 7-point apply, one MGS step at full size; extrapolated to all planes and %d FGMRES "
                  "iterations (factor %.1fs measured for %d planes, %.3fs per iteration estimated)"
                  % (threads, kp, nz, iters, t_fac, kp, t_iter),
        "factor_s_est": round(t_factor, 2),
        "iter_s_est": round(t_iter, 4),
    }


if __name__ == "__main__":
    main()
