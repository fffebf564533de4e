"""Host-side mirror of the reference's hot-path interface over the C ABI.

One ``Solver`` owns one ``hz_problem`` (one frequency on one grid).  Vectors are
torch complex128 tensors on the GPU (shape ``(N,)`` or ``(nrhs, N)``); torch
only provides device memory here, every operation runs in libhz.so.

Method                   reference function (Source/...)
``factor``               PARDISO phases 11+22 per plane, FGMRES.cpp:176-298
``precond_apply``        Solve3DSparseUsingFT_CSR, functions.cpp:5837-6061
``L0_apply``             Multiply3DSparseUsingFT, functions.cpp:6247-6318
``op_pre_apply``         ApplyCoeffMatrixA_CSR, functions.cpp:5554-5575
``op_apply``             true operator (beta3D stencil / L0 - deltaL)
``fgmres``               FGMRES, FGMRES.cpp:6-725
``bicgstab``             BCGSTAB, BcGSTAB.cpp:6-666
"""
import ctypes as C

import numpy as np

from . import abi
from .abi import PREC_FP64, PREC_C64, DRIVER_FGMRES, DRIVER_GMRES_REF, ORTHO_REF_CONJ, ORTHO_MGS  # noqa: F401
from .abi import MASK_FULL, MASK_PHYS  # noqa: F401


def _torch():
    import torch
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class Stats(dict):
    __getattr__ = dict.__getitem__


def _stats(st, hist=None):
    s = Stats(iters=st.iters, restarts=st.restarts, converged=bool(st.converged), breakdown=bool(st.breakdown),
              rel_res=st.rel_res, res=st.res, diff_sol=st.diff_sol, t_solve_s=st.t_solve_s)
    if hist is not None:
        s["hist"] = hist[: min(st.n_hist, hist.size)].copy()
    return s


class Solver:
    def __init__(self, desc, c):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("the hz solver runs on the GPU; no device is visible")
        self.desc = desc
        self.N = desc.N
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1)
        if c.size != self.N:
            raise ValueError("sound speed has %d values, grid has %d" % (c.size, self.N))
        self._d = abi.make_desc(desc)
        h = C.c_void_p()
        torch.cuda.synchronize()
        abi.check(abi.lib().hz_problem_create(C.byref(self._d), c.ctypes.data_as(C.POINTER(C.c_double)),
                                              C.byref(h)), "hz_problem_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            abi.lib().hz_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ helpers
    def _vec(self, x):
        torch = _torch()
        if not (torch.is_tensor(x) and x.is_cuda and x.dtype == torch.complex128):
            raise TypeError("expected a complex128 CUDA tensor")
        x = x.contiguous()
        if x.numel() % self.N:
            raise ValueError("vector length %d is not a multiple of N=%d" % (x.numel(), self.N))
        return x, x.numel() // self.N

    def _run(self, name, *args):
        _torch().cuda.synchronize()
        abi.check(getattr(abi.lib(), name)(self.h, *args), name)
        abi.check(abi.lib().hz_problem_sync(self.h), "hz_problem_sync")

    @property
    def stream(self):
        return abi.lib().hz_problem_stream(self.h)

    def factor_bytes(self):
        dims = (C.c_int64 * 4)()
        b = C.c_double()
        abi.check(abi.lib().hz_problem_info(self.h, dims, C.byref(b)), "hz_problem_info")
        return b.value

    def slots(self):
        """(factor slots, planes skipped by SAVE_MEM) of the plane map."""
        ns, nk = C.c_int(), C.c_int()
        abi.check(abi.lib().hz_problem_slots(self.h, C.byref(ns), C.byref(nk)), "hz_problem_slots")
        return ns.value, nk.value

    # ---------------------------------------------------- preconditioner
    def factor(self, prec=PREC_FP64):
        self._run("hz_precond_factor", int(prec))

    def compress(self, prec=PREC_C64, eps=1e-5, smallsize=16, free_dense=False, twisted=False):
        """HODLR compression of the factors of `prec` (hz_precond_compress):
        later forward preconditioner applies use the compressed records;
        eps < 0 drops them.  twisted: compress the twisted complex64 factors
        (two-phase half-chain sweeps)."""
        flags = (abi.HODLR_FREE_DENSE if free_dense else 0) | (abi.HODLR_TWISTED if twisted else 0)
        self._run("hz_precond_compress", int(prec), float(eps), int(smallsize), flags)
        return self.hodlr_info(prec)

    def hodlr_info(self, prec=PREC_C64):
        info = abi.HodlrInfo()
        abi.check(abi.lib().hz_precond_hodlr_info(self.h, int(prec), C.byref(info)), "hz_precond_hodlr_info")
        return {f: getattr(info, f) for f, _ in abi.HodlrInfo._fields_}

    def hodlr_ranks(self, k, prec=PREC_C64):
        """(ny, nodes) ranks of plane k's compressed blocks."""
        info = self.hodlr_info(prec)
        ny = self.desc.n[1]
        out = np.zeros(ny * info["nodes"], np.int32)
        abi.check(abi.lib().hz_precond_hodlr_ranks(self.h, int(prec), int(k),
                                                   out.ctypes.data_as(C.POINTER(C.c_int32)), out.size),
                  "hz_precond_hodlr_ranks")
        return out.reshape(ny, info["nodes"])

    def autotune(self, prec=PREC_C64, nrhs=1):
        """Time the preconditioner's launch variants for nrhs right-hand sides
        and keep the fastest; returns the choice."""
        info = abi.TuneInfo()
        self._run("hz_autotune", int(prec), int(nrhs), C.byref(info))
        return dict(solve_kernel=info.solve_kernel, group_cap=info.group_cap, fft_lines=info.fft_lines,
                    from_cache=bool(info.from_cache), solve_ms=info.solve_ms, fft_ms=info.fft_ms)

    def precond_apply(self, x, prec=PREC_FP64, adjoint=False):
        x, nr = self._vec(x)
        out = _torch().empty_like(x)
        name = "hz_precond_apply_adjoint" if adjoint else "hz_precond_apply"
        self._run(name, int(prec), nr, _ptr(x), _ptr(out))
        return out

    def plane_solve(self, x, prec=PREC_FP64, adjoint=False):
        """In place on z-transformed data (complex128 for FP64, complex64 for C64)."""
        torch = _torch()
        want = torch.complex128 if prec == PREC_FP64 else torch.complex64
        if x.dtype != want or not x.is_contiguous():
            raise TypeError("plane_solve needs a contiguous %s tensor" % want)
        self._run("hz_plane_solve", int(prec), int(adjoint), x.numel() // self.N, _ptr(x))
        return x

    def fft_z(self, x, sign, prec=PREC_FP64):
        x, nr = self._vec(x)
        out = _torch().empty_like(x)
        self._run("hz_fft_z", int(prec), int(sign), nr, _ptr(x), _ptr(out))
        return out

    # --------------------------------------------------------- operators
    def L0_apply(self, x):
        x, nr = self._vec(x)
        out = _torch().empty_like(x)
        self._run("hz_L0_apply", nr, _ptr(x), _ptr(out))
        return out

    def op_apply(self, x, adjoint=False):
        x, nr = self._vec(x)
        out = _torch().empty_like(x)
        self._run("hz_op_apply_adjoint" if adjoint else "hz_op_apply", nr, _ptr(x), _ptr(out))
        return out

    def op_pre_apply(self, x, prec=PREC_FP64):
        x, nr = self._vec(x)
        out = _torch().empty_like(x)
        self._run("hz_op_apply_preconditioned", int(prec), nr, _ptr(x), _ptr(out))
        return out

    def residual(self, f, u, mask=MASK_FULL):
        f, nr = self._vec(f)
        u, _ = self._vec(u)
        r = _torch().empty_like(f)
        rel = (C.c_double * nr)()
        self._run("hz_residual", nr, _ptr(f), _ptr(u), _ptr(r), int(mask), rel)
        return r, np.array(rel[:])

    # ----------------------------------------------------------- drivers
    def fgmres(self, f, m=30, restarts=20, tol=1e-8, driver=DRIVER_FGMRES, prec=PREC_FP64, ortho=ORTHO_MGS,
               mask=MASK_FULL, monitor=0, breakdown=0.0, hist_cap=8192, check=False, batch=0, graph=False,
               autotune=False, chunk=0):
        """Solve for every row of f (shape (N,) or (nrhs, N)); FGMRES advances
        `batch` right-hand sides in lockstep (0: up to 16), `graph` replays each
        chunk of `chunk` Arnoldi steps (0: 1, or 5 with graphs) from a captured
        HIP graph."""
        f, nr = self._vec(f)
        u = _torch().zeros_like(f)
        hist = np.zeros(hist_cap)
        o = abi.KrylovOpts(driver=int(driver), m=int(m), max_restarts=int(restarts), ortho=int(ortho),
                           mask=int(mask), monitor=int(monitor), tol=float(tol), breakdown=float(breakdown),
                           hist=hist.ctypes.data_as(C.POINTER(C.c_double)), hist_cap=hist_cap, prec=int(prec),
                           batch=int(batch), graph=int(bool(graph)), autotune=int(bool(autotune)), chunk=int(chunk))
        st = (abi.KrylovStats * nr)()
        _torch().cuda.synchronize()
        rc = abi.lib().hz_fgmres(self.h, nr, _ptr(f), _ptr(u), C.byref(o), st)
        if rc not in (abi.HZ_OK, 5, 6) or (check and rc != abi.HZ_OK):
            abi.check(rc, "hz_fgmres")
        stats = [_stats(st[q], hist if q == 0 else None) for q in range(nr)]
        return u, (stats[0] if nr == 1 else stats)

    def bicgstab(self, f, prec=PREC_FP64, maxit=200, tol=1e-8, tol_abs=False, mask=MASK_FULL, hist_cap=8192):
        f, nr = self._vec(f)
        u = _torch().zeros_like(f)
        st = (abi.KrylovStats * nr)()
        hist = np.zeros(hist_cap)
        _torch().cuda.synchronize()
        rc = abi.lib().hz_bicgstab(self.h, int(prec), nr, _ptr(f), _ptr(u), int(maxit), float(tol), int(tol_abs),
                                   int(mask), hist.ctypes.data_as(C.POINTER(C.c_double)), hist_cap, st)
        if rc not in (abi.HZ_OK, 5):
            abi.check(rc, "hz_bicgstab")
        stats = [_stats(st[q], hist if q == 0 else None) for q in range(nr)]
        return u, (stats[0] if nr == 1 else stats)

    # ----------------------------------------------------------- adjoint
    def misfit_forward(self, f, rec, data, m=40, restarts=20, tol=1e-8, prec=PREC_C64, store_fields=True, batch=0,
                       graph=False):
        """J = 1/2 sum_s ||P u_s - d_s||^2 (hz_misfit_forward); store_fields True,
        False or "auto" (against the field budget)."""
        f, nsrc = self._vec(f)
        rec = np.ascontiguousarray(rec, dtype=np.int64)
        data = np.ascontiguousarray(data, dtype=np.complex128).reshape(nsrc, rec.size)
        o = abi.KrylovOpts(driver=DRIVER_FGMRES, m=int(m), max_restarts=int(restarts), ortho=ORTHO_MGS,
                           mask=MASK_FULL, monitor=0, tol=float(tol), breakdown=0.0, hist=None, hist_cap=0,
                           prec=int(prec), batch=int(batch), graph=int(bool(graph)))
        J = C.c_double()
        store = abi.FIELDS_AUTO if store_fields in ("auto", None) else int(bool(store_fields))
        self._run("hz_misfit_forward", nsrc, _ptr(f), rec.size, rec.ctypes.data_as(C.POINTER(C.c_int64)),
                  C.c_void_p(data.ctypes.data), C.byref(o), store, C.byref(J))
        return J.value

    def misfit_budget(self, nbytes):
        """Field budget of store_fields="auto" (0: half the free device memory)."""
        abi.check(abi.lib().hz_misfit_set_budget(self.h, float(nbytes)), "hz_misfit_set_budget")

    def misfit_fields_stored(self):
        v = C.c_int()
        abi.check(abi.lib().hz_misfit_fields_stored(self.h, C.byref(v)), "hz_misfit_fields_stored")
        return bool(v.value)

    def misfit_backward(self):
        torch = _torch()
        g = torch.empty(self.N, dtype=torch.float64, device="cuda")
        self._run("hz_misfit_backward", _ptr(g))
        return g


# ------------------------------------------------------- several problems
def factor_many(solvers, prec=PREC_C64):
    """Factor several problems (frequencies / grids) concurrently."""
    torch = _torch()
    hs = (C.c_void_p * len(solvers))(*[S.h.value for S in solvers])
    torch.cuda.synchronize()
    abi.check(abi.lib().hz_precond_factor_multi(len(solvers), hs, int(prec)), "hz_precond_factor_multi")


def fgmres_many(solvers, fs, m=30, restarts=20, tol=1e-8, prec=PREC_C64, ortho=ORTHO_MGS, mask=MASK_FULL,
                batch=0, graph=False, autotune=False, hist_cap=8192, chunk=0):
    """FGMRES on several problems at once: fs[i] (complex128 CUDA tensor of
    shape (N_i,) or (nrhs_i, N_i)) is solved on solvers[i]; every problem runs
    its lockstep batches on its own stream, all queued from the calling
    thread.  Returns [(u_i, [stats...])]."""
    torch = _torch()
    n = len(solvers)
    fv, nr, us, sts = [], [], [], []
    for S, f in zip(solvers, fs):
        f, k = S._vec(f)
        fv.append(f)
        nr.append(k)
        us.append(torch.zeros_like(f))
        sts.append((abi.KrylovStats * k)())
    hist = np.zeros(hist_cap)
    o = abi.KrylovOpts(driver=DRIVER_FGMRES, m=int(m), max_restarts=int(restarts), ortho=int(ortho), mask=int(mask),
                       monitor=0, tol=float(tol), breakdown=0.0, hist=hist.ctypes.data_as(C.POINTER(C.c_double)),
                       hist_cap=hist_cap, prec=int(prec), batch=int(batch), graph=int(bool(graph)),
                       autotune=int(bool(autotune)), chunk=int(chunk))
    hs = (C.c_void_p * n)(*[S.h.value for S in solvers])
    nra = (C.c_int * n)(*nr)
    fa = (C.c_void_p * n)(*[f.data_ptr() for f in fv])
    ua = (C.c_void_p * n)(*[u.data_ptr() for u in us])
    sa = (C.POINTER(abi.KrylovStats) * n)(*[C.cast(st, C.POINTER(abi.KrylovStats)) for st in sts])
    torch.cuda.synchronize()
    rc = abi.lib().hz_fgmres_multi(n, hs, nra, fa, ua, C.byref(o), sa)
    if rc not in (abi.HZ_OK, 5, 6):
        abi.check(rc, "hz_fgmres_multi")
    out = []
    for i in range(n):
        out.append((us[i], [_stats(sts[i][q], hist if (i == 0 and q == 0) else None) for q in range(nr[i])]))
    return out
