"""HODLR restatement of the compressed plane solve (test infrastructure: the
CPU reference that the GPU's compressed preconditioner is checked against;
nothing in the library calls it).

The reference's HODLR build replaces every dense block inverse of the
block-tridiagonal plane solve by a recursively compressed matrix:

* the tree (SymRecCompressStruct, HODLR/BinaryTreesHODLR.cpp:386-416): a node
  of order m is a dense leaf when m <= smallsize, otherwise it splits into
  n1 = m - ceil(m/2) leading and n2 = ceil(m/2) trailing rows; its lower
  off-diagonal block A21 = A[n1:, :n1] is compressed and the two diagonal
  blocks recurse;
* the compression (LowRankApproxStruct, :298-384): the SVD A21 = U S V^T keeps
  the leading singular triplets while S[j] / S[0] >= eps (the loop breaks at
  the first j below), folds S into U and stores U (n2 x p), VT (p x n1);
* the product (RecMultLStructWork2, :1095-1130, the SYMMETRY form in which
  A12 = A21^T): y1 = left x1 + VT^T (U^T x2), y2 = right x2 + U (VT x1);
* the sweeps (DirSolveFastDiagStruct, :2971-3018) with the Schur recursion of
  DirFactFastDiagStructOnline (:2870-2969) for distinct couplings.

The compressed matrices are K_j = G_j D_j, the block inverses scaled by the
diagonal D_j = diag(dx(i) dy(j)) with d(i) / d(i-1) = cm(i) / cp(i-1) per axis,
which makes them complex symmetric (the symmetric form the reference uses with
SYMMETRY); G_j v = K_j (D_j^{-1} v).  The blocks G_j come from the exact Schur
recursion in float64 (numpy inverses); the reference forms them in HODLR
arithmetic at the same eps, so the two agree to O(eps)."""
import math

import numpy as np


def tree(n, small):
    """Internal nodes (a, m, n1, n2, level) in breadth-first order and leaves
    (a, m, level) of the reference's recursion on order n."""
    nodes, leaves = [], []
    todo = [(0, n, 0)]
    while todo:
        nxt = []
        for a, m, lev in todo:
            if m <= small:
                leaves.append((a, m, lev))
                continue
            n2 = int(math.ceil(m / 2.0))
            n1 = m - n2
            nodes.append((a, m, n1, n2, lev))
            nxt.append((a, n1, lev + 1))
            nxt.append((a + n1, n2, lev + 1))
        todo = nxt
    leaves.sort()
    return nodes, leaves


def low_rank(A, eps):
    """LowRankApproxStruct: (U S, VT) of the truncated SVD of A."""
    U, S, VT = np.linalg.svd(A, full_matrices=False)
    p = 0
    for j in range(S.size):
        if not S[j] / S[0] >= eps:
            break
        p = j + 1
    return U[:, :p] * S[:p], VT[:p].copy(), S


class Hodlr:
    """Symmetric HODLR form of a complex symmetric matrix K."""

    def __init__(self, K, eps, small):
        n = K.shape[0]
        self.n, self.eps, self.small = n, eps, small
        self.nodes, self.leaves = tree(n, small)
        self.lr = []      # (U, VT) per internal node
        self.sv = []      # singular values of each A21
        for a, m, n1, n2, lev in self.nodes:
            U, VT, S = low_rank(K[a + n1:a + m, a:a + n1], eps)
            self.lr.append((U, VT))
            self.sv.append(S)
        self.dense = [K[a:a + m, a:a + m].copy() for a, m, _ in self.leaves]

    def ranks(self):
        return np.array([U.shape[1] for U, _ in self.lr], dtype=np.int64)

    def matvec(self, x):
        y = np.zeros(self.n, np.complex128)
        for (a, m, _), D in zip(self.leaves, self.dense):
            y[a:a + m] += D @ x[a:a + m]
        for (a, m, n1, n2, _), (U, VT) in zip(self.nodes, self.lr):
            x1, x2 = x[a:a + n1], x[a + n1:a + m]
            y[a:a + n1] += VT.T @ (U.T @ x2)
            y[a + n1:a + m] += U @ (VT @ x1)
        return y

    def to_dense(self):
        K = np.zeros((self.n, self.n), np.complex128)
        for (a, m, _), D in zip(self.leaves, self.dense):
            K[a:a + m, a:a + m] = D
        for (a, m, n1, n2, _), (U, VT) in zip(self.nodes, self.lr):
            B = U @ VT
            K[a + n1:a + m, a:a + n1] = B
            K[a:a + n1, a + n1:a + m] = B.T
        return K


def sym_scaling(cm, cp):
    """d(0) = 1, d(i) = d(i-1) cm(i) / cp(i-1) (d D^{-1} M symmetric)."""
    d = np.ones(cm.size, np.complex128)
    for i in range(1, cm.size):
        d[i] = d[i - 1] * cm[i] / cp[i - 1] if cp[i - 1] != 0 else d[i - 1]
    return d


def plane_blocks(P, k):
    """Diagonal blocks D_j (nx x nx) of plane k's operator and the couplings
    cm_y, cp_y (y-line j couples to j-1 by cm_y(j), to j+1 by cp_y(j))."""
    nx, ny = P.n[0], P.n[1]
    cmx, cpx, cdx = P.tables(0)
    cmy, cpy, cdy = P.tables(1)
    kap = P.kappa().reshape(ny, nx)
    lam = P.lamz()[k]
    Dx = np.diag(cdx) + np.diag(cmx[1:], -1) + np.diag(cpx[:-1], 1)
    return [Dx + np.diag(cdy[j] + kap[j] - lam) for j in range(ny)], cmy, cpy


def schur_inverses(P, k):
    """G_j = S_j^{-1}: S_0 = D_0, S_j = D_j - cm_y(j) cp_y(j-1) G_{j-1}
    (DirFactFastDiagStructOnline with distinct couplings)."""
    D, cmy, cpy = plane_blocks(P, k)
    G = []
    for j, Dj in enumerate(D):
        S = Dj if j == 0 else Dj - cmy[j] * cpy[j - 1] * G[-1]
        G.append(np.linalg.inv(S))
    return G


def compress_plane(P, k, eps, small):
    """The compressed K_j of plane k and the scalings dx, dy."""
    cmx, cpx, _ = P.tables(0)
    cmy, cpy, _ = P.tables(1)
    dx, dy = sym_scaling(cmx, cpx), sym_scaling(cmy, cpy)
    H = [Hodlr(G * (dx * dy[j])[None, :], eps, small) for j, G in enumerate(schur_inverses(P, k))]
    return H, dx, dy


def plane_solve(P, k, b, H, dx, dy):
    """DirSolveFastDiagStruct: forward tb_j = f_j - cm_y(j) G_{j-1} tb_{j-1},
    backward x_j = G_j (tb_j - cp_y(j) x_{j+1}), with G_j v = K_j (v / (dx dy_j))."""
    nx, ny = P.n[0], P.n[1]
    cmy, cpy, _ = P.tables(1)
    f = np.asarray(b, np.complex128).reshape(ny, nx)
    G = lambda j, v: H[j].matvec(v / (dx * dy[j]))
    tb = [f[0].copy()]
    for j in range(1, ny):
        tb.append(f[j] - cmy[j] * G(j - 1, tb[j - 1]))
    x = np.zeros((ny, nx), np.complex128)
    x[ny - 1] = G(ny - 1, tb[ny - 1])
    for j in range(ny - 2, -1, -1):
        x[j] = G(j, tb[j] - cpy[j] * x[j + 1])
    return x.reshape(-1)


class HodlrPrecond:
    """L0^{-1} with every solved plane's block inverses compressed
    (Solve3DSparseUsingFT_HODLR, functions.cpp:6064-6244): z-DFT, compressed
    plane solves (planes the SAVE_MEM rule skips are zero), inverse z-DFT."""

    def __init__(self, P, eps, small):
        self.P = P
        nz = P.n[2]
        solved = P.plane_solved()
        self.planes = {}
        for k in range(nz):
            if solved[k]:
                self.planes[k] = compress_plane(P, k, eps, small)

    def ranks(self, k):
        return np.stack([h.ranks() for h in self.planes[k][0]])

    def apply(self, f):
        P = self.P
        n2 = P.n[0] * P.n[1]
        t = P.fft_z(np.asarray(f, np.complex128), -1)
        out = np.zeros_like(t)
        for k in range(P.n[2]):
            if k in self.planes:
                H, dx, dy = self.planes[k]
                out[k * n2:(k + 1) * n2] = plane_solve(P, k, t[k * n2:(k + 1) * n2], H, dx, dy)
        return P.fft_z(out, +1)
