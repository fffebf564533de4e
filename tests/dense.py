"""Dense assemblies of the operators for tiny grids (numpy), used to pin the
oracle by brute force: Kronecker sums of the 1D PML second differences, the
z-DFT as an explicit matrix, and numpy.linalg solves."""
import numpy as np


def tridiag(cm, cd, cp):
    n = len(cd)
    D = np.diag(np.asarray(cd, dtype=np.complex128))
    for i in range(n):
        if i > 0:
            D[i, i - 1] = cm[i]
        if i < n - 1:
            D[i, i + 1] = cp[i]
    return D


def axis_matrix(tables):
    cm, cp, cd = tables
    return tridiag(cm, cd, cp)


def plane_matrix(P, k):
    nx, ny = P.n[0], P.n[1]
    Dx = axis_matrix(P.tables(0))
    Dy = axis_matrix(P.tables(1))
    M = np.kron(np.eye(ny), Dx) + np.kron(Dy, np.eye(nx))
    M += np.diag(P.kappa() - P.lamz()[k])
    return M


def dft_matrix(n):
    j = np.arange(n)
    return np.exp(-2j * np.pi * np.outer(j, j) / n)


def L0_matrix(P):
    nx, ny, nz = P.n
    n2 = nx * ny
    F = dft_matrix(nz)
    Fi = np.conj(F) / nz
    B = np.zeros((P.N, P.N), np.complex128)
    for k in range(nz):
        B[k * n2:(k + 1) * n2, k * n2:(k + 1) * n2] = plane_matrix(P, k)
    T = np.kron(F, np.eye(n2))
    Ti = np.kron(Fi, np.eye(n2))
    return Ti @ B @ T


def stencil_matrix(P, c):
    nx, ny, nz = P.n
    Dx = axis_matrix(P.tables(0))
    Dy = axis_matrix(P.tables(1))
    Dz = axis_matrix(P.tables(2))
    A = (np.kron(np.eye(nz * ny), Dx) + np.kron(np.eye(nz), np.kron(Dy, np.eye(nx)))
         + np.kron(Dz, np.eye(nx * ny)))
    A += np.diag(P.desc.omega ** 2 / np.asarray(c) ** 2)
    return A


def true_matrix(P, c):
    """REF: L0 - diag(deltaL); S7: 7-point stencil."""
    if P.desc.mode == 0:
        return L0_matrix(P) - np.diag(P.deltaL())
    return stencil_matrix(P, c)


def min_residuals(Aop, r0, nsteps):
    """||r_j|| of the minimal-residual iterate over K_j(Aop, r0), j = 1..nsteps,
    from an explicit orthonormal Krylov basis (numpy QR) and lstsq."""
    n = r0.size
    K = np.zeros((n, nsteps), np.complex128)
    v = r0.copy()
    out = []
    for j in range(nsteps):
        K[:, j] = v
        Q, _ = np.linalg.qr(K[:, :j + 1])
        W = Aop @ Q
        y, *_ = np.linalg.lstsq(W, r0, rcond=None)
        out.append(np.linalg.norm(r0 - W @ y))
        v = Aop @ Q[:, -1]
    return np.array(out)
