"""The reference's CPU path with its own plane solver, MKL PARDISO (test and
benchmark infrastructure: the CPU baseline of bench.py; the library never
loads it).

The reference factors every z-plane's 5-point PML operator with PARDISO
(complex unsymmetric, mtype 13, 1-based CSR; FGMRES.cpp:274-284) and solves
each plane per preconditioner apply (phase 33, functions.cpp:5959) between
batched z-DFTs (functions.cpp:5876-5881, 6041-6046).  This module does the
same on the host with the MKL runtime of the image (libmkl_rt, called
through ctypes with the documented PARDISO interface), numpy/scipy FFTs
along z, the oracle's 7-point operator and a right-preconditioned FGMRES
with modified Gram-Schmidt in numpy."""
import ctypes as C
import os

import numpy as np

MKL_PATHS = ("/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1", "libmkl_rt.so")
_lib = None
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C")


def lib():
    global _lib
    if _lib is None:
        # MKL's OpenMP threads sleep at once after a parallel region instead of
        # spinning against the oracle's and scipy's threads between calls
        os.environ.setdefault("KMP_BLOCKTIME", "0")
        last = None
        for p in MKL_PATHS:
            try:
                L = C.CDLL(p)
                break
            except OSError as e:
                last = e
        else:
            raise RuntimeError("MKL runtime not found: %s" % last)
        L.pardisoinit.argtypes = [C.c_void_p, C.POINTER(C.c_int), _i32]
        L.pardiso.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                              C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p, _i32, _i32, C.c_void_p,
                              C.POINTER(C.c_int), _i32, C.POINTER(C.c_int), C.c_void_p, C.c_void_p,
                              C.POINTER(C.c_int)]
        L.MKL_Set_Num_Threads.argtypes = [C.c_int]
        L.MKL_Set_Num_Threads_Local.argtypes = [C.c_int]
        L.MKL_Set_Num_Threads_Local.restype = C.c_int
        _lib = L
    return _lib


def available():
    try:
        lib()
        return True
    except RuntimeError:
        return False


def set_threads(n):
    lib().MKL_Set_Num_Threads(int(n))


_pools = {}


def pool(threads):
    """Host threads that each run PARDISO sequentially on their own planes
    (ctypes releases the interpreter lock during the call; PARDISO handles are
    independent): planes in parallel rather than MKL's threads inside each
    small plane, which wake up and sleep again for every call."""
    from concurrent.futures import ThreadPoolExecutor
    if threads not in _pools:
        _pools[threads] = ThreadPoolExecutor(threads, initializer=lambda: lib().MKL_Set_Num_Threads_Local(1))
    return _pools[threads]


def plane_csr(P, k):
    """1-based CSR of plane k's 5-point operator M_k (rows x fastest):
    columns (i, j-1), (i-1, j), (i, j), (i+1, j), (i, j+1) with the
    coefficients cm_y(j), cm_x(i), cd_x(i) + cd_y(j) + kappa(i, j) - lambda_k,
    cp_x(i), cp_y(j) (GenSparseMatrixOnline2DwithPMLFast, functions.cpp:3270-3370)."""
    nx, ny = P.n[0], P.n[1]
    cmx, cpx, cdx = P.tables(0)
    cmy, cpy, cdy = P.tables(1)
    kap = P.kappa().reshape(ny, nx)
    lam = P.lamz()[k]
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    I, J = I.reshape(-1), J.reshape(-1)
    l = I + nx * J
    parts = [
        (J > 0, l - nx, cmy[J]),
        (I > 0, l - 1, cmx[I]),
        (np.ones_like(I, bool), l, cdx[I] + cdy[J] + kap.reshape(-1) - lam),
        (I < nx - 1, l + 1, cpx[I]),
        (J < ny - 1, l + nx, cpy[J]),
    ]
    cnt = sum(m.astype(np.int32) for m, _, _ in parts)
    ia = np.zeros(nx * ny + 1, np.int32)
    ia[1:] = np.cumsum(cnt)
    ja = np.zeros(ia[-1], np.int32)
    a = np.zeros(ia[-1], np.complex128)
    pos = ia[:-1].copy()
    for m, col, val in parts:
        r = np.nonzero(m)[0]
        ja[pos[r]] = col[r]
        a[pos[r]] = val[r]
        pos[r] += 1
    return ia + 1, ja + 1, a


class PlaneSolver:
    """PARDISO handle of one plane (mtype 13, the default iparm of pardisoinit)."""

    def __init__(self, ia, ja, a):
        self.ia, self.ja, self.a = ia, ja, a
        self.n = ia.size - 1
        self.pt = (C.c_void_p * 64)()
        self.iparm = np.zeros(64, np.int32)
        self.mtype = C.c_int(13)
        lib().pardisoinit(self.pt, C.byref(self.mtype), self.iparm)
        self.alive = False

    def _call(self, phase, b=None, x=None):
        one, zero = C.c_int(1), C.c_int(0)
        err = C.c_int(0)
        n = C.c_int(self.n)
        ph = C.c_int(phase)
        bp = b.ctypes.data if b is not None else None
        xp = x.ctypes.data if x is not None else None
        lib().pardiso(self.pt, C.byref(one), C.byref(one), C.byref(self.mtype), C.byref(ph), C.byref(n),
                      self.a.ctypes.data, self.ia, self.ja, None, C.byref(one), self.iparm, C.byref(zero), bp, xp,
                      C.byref(err))
        if err.value != 0:
            raise RuntimeError("pardiso phase %d: error %d" % (phase, err.value))

    def factor(self):
        """Phases 11 + 22 (analysis and numerical factorisation)."""
        self._call(12)
        self.alive = True

    def solve(self, b):
        x = np.empty_like(b)
        self._call(33, np.ascontiguousarray(b), x)
        return x

    def release(self):
        if self.alive:
            self._call(-1)
            self.alive = False

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _fft_z(x, shape, sign, workers):
    import scipy.fft as sf
    X = x.reshape(shape)
    if sign < 0:
        return sf.fft(X, axis=0, workers=workers).reshape(-1)
    return sf.ifft(X, axis=0, workers=workers).reshape(-1)


class PardisoPrecond:
    """L0^{-1} with PARDISO plane solves (Solve3DSparseUsingFT_CSR): z-DFT,
    one phase-33 solve per solved plane (planes the SAVE_MEM rule skips are
    zero), inverse z-DFT.  planes: the plane indices to factor (all solved
    planes by default)."""

    def __init__(self, P, threads=8, planes=None):
        self.P = P
        self.threads = threads
        set_threads(threads)
        nz = P.n[2]
        solved = P.plane_solved()
        self.planes = {}
        ks = [k for k in range(nz) if solved[k]] if planes is None else list(planes)
        solvers = [PlaneSolver(*plane_csr(P, k)) for k in ks]
        list(pool(threads).map(lambda q: q.factor(), solvers))
        self.planes = dict(zip(ks, solvers))

    def apply(self, f):
        P = self.P
        nx, ny, nz = P.n
        n2 = nx * ny
        t = _fft_z(np.asarray(f, np.complex128), (nz, ny, nx), -1, self.threads)
        out = np.zeros_like(t)

        def one(item):
            k, s = item
            out[k * n2:(k + 1) * n2] = s.solve(t[k * n2:(k + 1) * n2])

        list(pool(self.threads).map(one, self.planes.items()))
        return _fft_z(out, (nz, ny, nx), +1, self.threads)

    def release(self):
        for s in self.planes.values():
            s.release()


def fgmres(P, M, f, m=35, tol=1e-8, max_iters=20000):
    """Right-preconditioned restarted FGMRES(m) from zero: Z_j = M^{-1} V_j,
    modified Gram-Schmidt, Givens rotations, stop at ||f - A u|| <= tol ||f||
    (the true residual at each restart and at the end).  Returns (u, iters,
    converged)."""
    from scipy.linalg.blas import dznrm2, zaxpy, zdotc
    f = np.asarray(f, np.complex128)
    nf = np.linalg.norm(f)
    u = np.zeros_like(f)
    it = 0
    while it < max_iters:
        r = f - P.op_apply(u)
        beta = np.linalg.norm(r)
        if beta <= tol * nf:
            return u, it, True
        V = [r / beta]
        Z = []
        H = np.zeros((m + 1, m), np.complex128)
        cs = np.zeros(m, np.complex128)
        sn = np.zeros(m, np.complex128)
        g = np.zeros(m + 1, np.complex128)
        g[0] = beta
        j_last = 0
        for j in range(m):
            z = M.apply(V[j])
            Z.append(z)
            w = P.op_apply(z)
            for i in range(j + 1):  # modified Gram-Schmidt with BLAS zdotc / zaxpy in place
                H[i, j] = zdotc(V[i], w)
                w = zaxpy(V[i], w, a=-H[i, j])
            hn = dznrm2(w)
            H[j + 1, j] = hn
            # rotations G_i = [[conj(c), conj(s)], [-s, c]] with c = a/r, s = b/r zero H[i+1, i]
            for i in range(j):
                t = np.conj(cs[i]) * H[i, j] + np.conj(sn[i]) * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            a, b = H[j, j], H[j + 1, j]
            den = np.sqrt(abs(a) ** 2 + abs(b) ** 2)
            cs[j] = a / den if den > 0 else 1.0
            sn[j] = b / den if den > 0 else 0.0
            H[j, j] = den
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = np.conj(cs[j]) * g[j]
            it += 1
            j_last = j + 1
            if abs(g[j + 1]) <= tol * nf or hn == 0.0 or it >= max_iters:
                break
            V.append(w / hn)
        y = np.linalg.solve(np.triu(H[:j_last, :j_last]), g[:j_last])
        for i in range(j_last):
            u = u + y[i] * Z[i]
    r = f - P.op_apply(u)
    return u, it, bool(np.linalg.norm(r) <= tol * nf)
