"""ctypes wrapper of the CPU FP64 oracle (liboracle_hz.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and by the
CPU-baseline leg of bench.py, always as the checker.  The product package never
imports this module.

All complex arrays are numpy complex128, layout l = i + nx*j + nx*ny*k.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_hz.so")


class OrDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int64 * 3),
        ("h", C.c_double * 3),
        ("l", C.c_double * 3),
        ("pml", C.c_int32 * 3),
        ("spg", C.c_int32 * 3),
        ("omega", C.c_double),
        ("beta", C.c_double),
        ("pml_c", C.c_double),
        ("mode", C.c_int32),
        ("kz_symbol", C.c_int32),
        ("c0_rule", C.c_int32),
        ("skip_ratio", C.c_int32),
        ("c_bg", C.c_double),
        ("skip_k0", C.c_double),
    ]


class OrStats(C.Structure):
    _fields_ = [
        ("iters", C.c_int32),
        ("restarts", C.c_int32),
        ("converged", C.c_int32),
        ("breakdown", C.c_int32),
        ("rel_res", C.c_double),
        ("res", C.c_double),
        ("diff_sol", C.c_double),
        ("n_hist", C.c_int32),
        ("reserved", C.c_int32),
    ]


def build():
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, dp, i64, i32, d = C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.c_int, C.c_double
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(OrDesc), dp]
        L.or_destroy.argtypes = [vp]
        L.or_set_threads.argtypes = [i32]
        L.or_alpha.argtypes = [i64, i32, d, d, d, dp]
        L.or_tables.argtypes = [vp, i32, dp, dp, dp]
        L.or_plane_solved.argtypes = [vp, C.POINTER(C.c_int)]
        for nm in ("or_c0", "or_kappa", "or_lamz", "or_deltaL", "or_factor"):
            getattr(L, nm).argtypes = [vp] + ([dp] if nm != "or_factor" else [])
        L.or_fft_line.argtypes = [i64, dp, dp, i32]
        L.or_dft_brute.argtypes = [i64, dp, dp, i32]
        L.or_fft_z.argtypes = [vp, dp, dp, i32]
        L.or_plane_matvec.argtypes = [vp, i64, dp, dp]
        L.or_plane_solve.argtypes = [vp, i64, dp, i32]
        L.or_factor_planes.argtypes = [vp, i64, i64]
        L.or_solve_planes.argtypes = [vp, i64, i64, dp]
        L.or_precond_apply.argtypes = [vp, dp, dp, i32]
        L.or_L0_apply.argtypes = [vp, dp, dp]
        L.or_op_apply.argtypes = [vp, dp, dp, i32]
        L.or_op_pre_apply.argtypes = [vp, dp, dp]
        L.or_masked_nrm2.restype = d
        L.or_masked_nrm2.argtypes = [vp, dp, i32]
        L.or_lsq.argtypes = [i64, i64, dp, i64, dp]
        L.or_gmres_ref.argtypes = [vp, dp, dp, i32, i32, d, d, i32, i32, i32, dp, i32, C.POINTER(OrStats)]
        L.or_fgmres.argtypes = [vp, dp, dp, i32, i32, d, i32, i32, dp, i32, C.POINTER(OrStats)]
        L.or_bicgstab.argtypes = [vp, dp, dp, i32, d, i32, i32, dp, i32, C.POINTER(OrStats)]
        L.or_misfit_grad.restype = i32
        L.or_misfit_grad.argtypes = [vp, i32, dp, i32, C.POINTER(C.c_int64), dp, i32, i32, d, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def make_desc(desc):
    """Build the oracle's struct from any object with the descriptor attributes."""
    o = OrDesc()
    for a in range(3):
        o.n[a] = int(desc.n[a])
        o.h[a] = float(desc.h[a])
        o.l[a] = float(desc.l[a])
        o.pml[a] = int(desc.pml[a])
        o.spg[a] = int(desc.spg[a])
    o.omega = float(desc.omega)
    o.beta = float(desc.beta)
    o.pml_c = float(desc.pml_c)
    o.mode = int(desc.mode)
    o.kz_symbol = int(desc.kz_symbol)
    o.c0_rule = int(desc.c0_rule)
    o.c_bg = float(desc.c_bg)
    o.skip_ratio = int(getattr(desc, "skip_ratio", 0))
    o.skip_k0 = float(getattr(desc, "skip_k0", 0.0))
    return o


def alpha(n, pml, omega, pml_c, s):
    out = np.zeros(2)
    lib().or_alpha(n, pml, omega, pml_c, s, _dp(out))
    return complex(out[0], out[1])


def fft_line(x, sign):
    x = _c128(x)
    out = np.zeros_like(x)
    lib().or_fft_line(x.size, _dp(x), _dp(out), sign)
    return out


def dft_brute(x, sign):
    x = _c128(x)
    out = np.zeros_like(x)
    lib().or_dft_brute(x.size, _dp(x), _dp(out), sign)
    return out


def lsq(H, b):
    """min ||b - H y|| by Householder QR; returns y (len = H.shape[1])."""
    H = np.asfortranarray(H, dtype=np.complex128).copy(order="F")
    b = _c128(b).copy()
    m, n = H.shape
    lib().or_lsq(m, n, _dp(H), m, _dp(b))
    return b[:n].copy()


class Stats:
    def __init__(self, st, hist):
        self.iters = st.iters
        self.restarts = st.restarts
        self.converged = bool(st.converged)
        self.breakdown = bool(st.breakdown)
        self.rel_res = st.rel_res
        self.res = st.res
        self.diff_sol = st.diff_sol
        self.hist = hist[: min(st.n_hist, hist.size)].copy()


class Problem:
    """One oracle problem: descriptor + sound speed c (N reals)."""

    def __init__(self, desc, c):
        self.desc = desc
        self.n = tuple(int(v) for v in desc.n)
        self.N = self.n[0] * self.n[1] * self.n[2]
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1)
        assert c.size == self.N
        self._d = make_desc(desc)
        self.h = lib().or_create(C.byref(self._d), _dp(c))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def tables(self, ax):
        n = self.n[ax]
        cm, cp, cd = (np.zeros(n, np.complex128) for _ in range(3))
        lib().or_tables(self.h, ax, _dp(cm), _dp(cp), _dp(cd))
        return cm, cp, cd

    def c0(self):
        out = np.zeros(self.n[0] * self.n[1], np.complex128)
        lib().or_c0(self.h, _dp(out))
        return out

    def kappa(self):
        out = np.zeros(self.n[0] * self.n[1], np.complex128)
        lib().or_kappa(self.h, _dp(out))
        return out

    def plane_solved(self):
        """1 for the planes the preconditioner solves, 0 for those SAVE_MEM skips."""
        out = np.zeros(self.n[2], dtype=np.int32)
        lib().or_plane_solved(self.h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def lamz(self):
        out = np.zeros(self.n[2])
        lib().or_lamz(self.h, _dp(out))
        return out

    def deltaL(self):
        out = np.zeros(self.N, np.complex128)
        lib().or_deltaL(self.h, _dp(out))
        return out

    def _unary(self, fn, x, *extra):
        x = _c128(x).reshape(-1)
        assert x.size == self.N, (x.size, self.N)
        out = np.zeros_like(x)
        fn(self.h, _dp(x), _dp(out), *extra)
        return out

    def fft_z(self, x, sign):
        return self._unary(lib().or_fft_z, x, sign)

    def plane_matvec(self, k, x):
        x = _c128(x).reshape(-1)
        out = np.zeros_like(x)
        lib().or_plane_matvec(self.h, k, _dp(x), _dp(out))
        return out

    def plane_solve(self, k, b, adjoint=False):
        b = _c128(b).reshape(-1).copy()
        lib().or_plane_solve(self.h, k, _dp(b), int(adjoint))
        return b

    def factor(self):
        lib().or_factor(self.h)

    def factor_planes(self, k0, k1):
        lib().or_factor_planes(self.h, k0, k1)

    def solve_planes(self, k0, k1, b):
        b = _c128(b).reshape(-1).copy()
        lib().or_solve_planes(self.h, k0, k1, _dp(b))
        return b

    def precond_apply(self, x, adjoint=False):
        return self._unary(lib().or_precond_apply, x, int(adjoint))

    def L0_apply(self, x):
        return self._unary(lib().or_L0_apply, x)

    def op_apply(self, x, adjoint=False):
        return self._unary(lib().or_op_apply, x, int(adjoint))

    def op_pre_apply(self, x):
        return self._unary(lib().or_op_pre_apply, x)

    def masked_nrm2(self, x, mask=1):
        x = _c128(x).reshape(-1)
        return lib().or_masked_nrm2(self.h, _dp(x), mask)

    def gmres_ref(self, f, m=10, rest=5, thresh=1e-4, tol=1e-4, mask=1, ortho=0, monitor=0, cap=4096):
        f = _c128(f).reshape(-1)
        assert f.size == self.N, (f.size, self.N)
        u = np.zeros_like(f)
        hist = np.zeros(cap)
        st = OrStats()
        lib().or_gmres_ref(self.h, _dp(f), _dp(u), m, rest, thresh, tol, mask, ortho, monitor,
                           _dp(hist), cap, C.byref(st))
        return u, Stats(st, hist)

    def fgmres(self, f, m=30, rest=20, tol=1e-8, mask=0, adjoint=False, cap=4096):
        f = _c128(f).reshape(-1)
        assert f.size == self.N, (f.size, self.N)
        u = np.zeros_like(f)
        hist = np.zeros(cap)
        st = OrStats()
        lib().or_fgmres(self.h, _dp(f), _dp(u), m, rest, tol, mask, int(adjoint), _dp(hist), cap, C.byref(st))
        return u, Stats(st, hist)

    def bicgstab(self, f, maxit=100, tol=1e-8, tol_abs=1, mask=1, cap=4096):
        f = _c128(f).reshape(-1)
        assert f.size == self.N, (f.size, self.N)
        u = np.zeros_like(f)
        hist = np.zeros(cap)
        st = OrStats()
        lib().or_bicgstab(self.h, _dp(f), _dp(u), maxit, tol, tol_abs, mask, _dp(hist), cap, C.byref(st))
        return u, Stats(st, hist)

    def misfit_grad(self, f, rec, data, m=40, rest=20, tol=1e-10):
        f = _c128(f)
        nsrc = f.shape[0]
        rec = np.ascontiguousarray(rec, dtype=np.int64)
        data = _c128(data)
        J = C.c_double(0.0)
        grad = np.zeros(self.N)
        u = np.zeros((nsrc, self.N), np.complex128)
        ok = lib().or_misfit_grad(self.h, nsrc, _dp(f), rec.size, rec.ctypes.data_as(C.POINTER(C.c_int64)),
                                  _dp(data), m, rest, tol, C.byref(J), _dp(grad), _dp(u))
        return J.value, grad, u, bool(ok)
