"""Pins the CPU FP64 oracle against closed forms, numpy/LAPACK routines, dense
brute-force assemblies and invariants the reference's own checks rely on.

None of these checks retypes the oracle's formulas: each compares with an
independent route (hand-evaluated values, numpy.fft, numpy.linalg, explicit
Kronecker assemblies, the GMRES minimal-residual property, an adjoint identity,
finite differences)."""
import math

import numpy as np
import pytest

from dense import (axis_matrix, dft_matrix, L0_matrix, min_residuals, plane_matrix,
                   stencil_matrix, true_matrix)


def _pkg():
    from conftest import load_pkg
    return load_pkg()


def tiny_ref(homo=True, beta=0.5, nxp=4, nyp=5, pml=2, spg=2, nzp=4):
    pr = _pkg().problem
    d = pr.ref_geometry(n_nopml=nxp, pml=pml, spg=spg, nz_nopml=nzp, length=(160.0, 160.0, 160.0),
                        beta=beta, homo=homo)
    # make nx != ny to catch axis mix-ups
    ny = nyp + 2 * pml
    ly = 160.0 + 2.0 * pml * 160.0 / nyp
    d = d.replace(n=(d.n[0], ny, d.n[2]), l=(d.l[0], ly, d.l[2]), h=(d.h[0], ly / ny, d.h[2]))
    c = pr.ref_sound_speed(d)
    return d, c


def tiny_s7(n=(8, 7, 6), npml=2, omega=None, beta=0.5, seed=3):
    pr = _pkg().problem
    h = 1.0 / max(n)
    om = omega if omega is not None else pr.omega_for_ppw(0.8, h, ppw=8)
    d = pr.cube_geometry(n, npml, om, beta=beta)
    rng = np.random.default_rng(seed)
    c = 0.8 + 0.4 * rng.random(d.N)
    return d, c


# ---------------------------------------------------------------- closed forms
def test_alpha_hand_values(oracle):
    w = 25.0
    # left layer: x = 1 - s/pml, d = 1000 x^2, alpha = -i w / (d - i w)
    for s, x in [(0.0, 1.0), (2.5, 1.0 - 2.5 / 3.0), (0.5, 1.0 - 0.5 / 3.0)]:
        dd = 1000.0 * x * x
        want = (-1j * w) / (dd - 1j * w)
        assert abs(oracle.alpha(10, 3, w, 1000.0, s) - want) < 1e-15
    # interior and the asymmetric right edge: s = n - pml - 1/2 is outside the layer
    assert oracle.alpha(10, 3, w, 1000.0, 3.0) == 1.0
    assert oracle.alpha(10, 3, w, 1000.0, 6.5) == 1.0
    # right layer: x = 1 + (s - n + 1)/pml
    for s in (7.0, 7.5, 9.0, 9.5):
        x = 1.0 + (s - 10 + 1) / 3.0
        want = (-1j * w) / (1000.0 * x * x - 1j * w)
        assert abs(oracle.alpha(10, 3, w, 1000.0, s) - want) < 1e-15
    assert oracle.alpha(10, 0, w, 1000.0, 0.0) == 1.0
    # alpha = 1 / (1 + i d / w)
    a = oracle.alpha(10, 3, w, 1000.0, 0.0)
    assert abs(a - 1.0 / (1.0 + 1j * 1000.0 / w)) < 1e-15


def test_tables_hand_values(oracle):
    d, c = tiny_s7()
    P = oracle.Problem(d, c)
    cm, cp, cd = P.tables(0)
    n, pml, w, h = d.n[0], d.pml[0], d.omega, d.h[0]
    al = lambda s: oracle.alpha(n, pml, w, 1000.0, s)
    for i in (0, 1, 3, n - 1):
        assert abs(cm[i] - al(i) * al(i - 0.5) / h ** 2) < 1e-9 * abs(cm[i])
        assert abs(cp[i] - al(i) * al(i + 0.5) / h ** 2) < 1e-9 * abs(cp[i])
    # away from the layers the stencil is the plain second difference
    i = n // 2
    assert abs(cm[i] - 1 / h ** 2) < 1e-9 / h ** 2 and abs(cd[i] + 2 / h ** 2) < 1e-9 / h ** 2


def test_plane_matrix_symmetric_after_row_scaling(oracle):
    """TestSymmSparseMatrixOnline2DwithPML (Tests.cpp:686-757): diag(1/(ax ay)) M
    is complex symmetric."""
    d, c = tiny_ref(homo=False)
    P = oracle.Problem(d, c)
    nx, ny = d.n[0], d.n[1]
    M = plane_matrix(P, 3)
    ax = np.array([oracle.alpha(nx, d.pml[0], d.omega, 1000.0, i) for i in range(nx)])
    ay = np.array([oracle.alpha(ny, d.pml[1], d.omega, 1000.0, j) for j in range(ny)])
    s = 1.0 / np.kron(ay, np.ones(nx)) / np.kron(np.ones(ny), ax)
    S = s[:, None] * M
    assert np.abs(S - S.T).max() < 1e-12 * np.abs(S).max()
    # the oracle's plane matvec is this matrix
    x = np.random.default_rng(0).standard_normal(nx * ny) + 1j
    assert np.allclose(P.plane_matvec(3, x), M @ x, rtol=1e-13, atol=0)


def test_kz_map_reference_quirk(oracle):
    """FGMRES.cpp:182-190: k < nz/2 - 1 -> k, else nz - k (so k = nz/2 - 1 maps
    to nz/2 + 1)."""
    d, c = tiny_ref(nzp=8)       # nz = 12
    P = oracle.Problem(d, c)
    nz, lz = d.n[2], d.l[2]
    got = P.lamz()
    kk = [0, 1, 2, 3, 4, 7, 6, 5, 4, 3, 2, 1]
    want = [4 * math.pi ** 2 * k * k / lz ** 2 for k in kk]
    assert np.allclose(got, want, rtol=1e-15)
    # corrected and FD symbols
    P1 = oracle.Problem(d.replace(kz_symbol=1), c)
    assert np.allclose(P1.lamz(), [4 * math.pi ** 2 * min(k, nz - k) ** 2 / lz ** 2 for k in range(nz)])
    P2 = oracle.Problem(d.replace(kz_symbol=2), c)
    hz = d.h[2]
    assert np.allclose(P2.lamz(), [(2 - 2 * math.cos(2 * math.pi * k / nz)) / hz ** 2 for k in range(nz)])


def test_c0_running_sum_closed_form(oracle):
    """SetSoundSpeed2D (functions.cpp:1716-1726) never resets its sum:
    S_m = c^-2 (1 - Nz^-m) / (1 - 1/Nz) for the m-th physical column."""
    d, c = tiny_ref(homo=True)
    P = oracle.Problem(d, c)
    c0 = P.c0()
    nx, ny = d.n[0], d.n[1]
    pml, Nz = d.pml[0], d.n[2] - 2 * d.spg[2]
    cz = d.meta["c_z"]
    m = 0
    for j in range(pml, ny - pml):
        for i in range(pml, nx - pml):
            m += 1
            S = cz ** -2 * (1 - Nz ** (-m)) / (1 - 1.0 / Nz)
            assert abs(c0[i + nx * j] - S ** -0.5) < 1e-12 * cz
    # the PML columns copy the edge: corner (0,0) equals the first physical column
    assert c0[0] == c0[pml + nx * pml]
    assert c0[nx - 1] == c0[(nx - pml - 1) + nx * pml]
    # the limit is c sqrt((N-1)/N)
    assert abs(c0[(nx - pml - 1) + nx * (ny - pml - 1)] - cz * math.sqrt((Nz - 1) / Nz)) < 1e-9 * cz


def test_deltaL_band_and_taper(oracle):
    """GenerateDeltaL (functions.cpp:1789-1823): band spg..nz-spg inclusive,
    Gaussian tapers with sigma 0.3 copied from the band edges."""
    d, c = tiny_ref(homo=False)
    P = oracle.Problem(d, c)
    dL = P.deltaL().reshape(d.shape_zyx)
    c0 = P.c0().reshape(d.n[1], d.n[0])
    cc = c.reshape(d.shape_zyx)
    w2, b = d.omega ** 2, d.beta
    nz, spg, hz = d.n[2], d.spg[2], d.h[2]
    for k in (spg, nz // 2, nz - spg):
        want = w2 * ((1 + 1j * b) / c0 ** 2 - 1 / cc[k] ** 2)
        assert np.allclose(dL[k], want, rtol=1e-13)
    z0, zN = spg * hz, (nz - spg) * hz
    for k in range(spg):
        t = math.exp(-((k * hz - z0) ** 2) / z0 ** 2 / 0.09)
        assert np.allclose(dL[k], dL[spg] * t, rtol=1e-13)
    for k in range(nz - spg + 1, nz):
        t = math.exp(-((k * hz - zN) ** 2) / (d.l[2] - zN) ** 2 / 0.09)
        assert np.allclose(dL[k], dL[nz - spg] * t, rtol=1e-13)


# ------------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 15, 16, 20, 25, 49, 64, 96, 150, 160, 192, 200])
def test_fft_line_vs_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(oracle.fft_line(x, -1), np.fft.fft(x), rtol=1e-13, atol=1e-12)
    assert np.allclose(oracle.fft_line(x, +1), np.fft.ifft(x), rtol=1e-13, atol=1e-13)
    if n <= 64:
        assert np.allclose(oracle.dft_brute(x, -1), np.fft.fft(x), rtol=1e-12, atol=1e-11)
        assert np.allclose(oracle.dft_brute(x, +1), np.fft.ifft(x), rtol=1e-12, atol=1e-12)


def test_fft_z_lines(oracle):
    d, c = tiny_s7(n=(5, 3, 12))
    P = oracle.Problem(d, c)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(d.N) + 1j * rng.standard_normal(d.N)
    X = x.reshape(d.shape_zyx)
    assert np.allclose(P.fft_z(x, -1).reshape(d.shape_zyx), np.fft.fft(X, axis=0), rtol=1e-13, atol=1e-12)
    assert np.allclose(P.fft_z(x, +1).reshape(d.shape_zyx), np.fft.ifft(X, axis=0), rtol=1e-13, atol=1e-13)


def test_lsq_vs_numpy(oracle):
    rng = np.random.default_rng(2)
    for m, n in [(2, 1), (5, 4), (11, 10)]:
        H = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        b = rng.standard_normal(m) + 1j * rng.standard_normal(m)
        y = oracle.lsq(H, b)
        want, *_ = np.linalg.lstsq(H, b, rcond=None)
        assert np.allclose(y, want, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------- plane solves
@pytest.mark.parametrize("mode", ["ref", "s7"])
def test_plane_solve_vs_dense(oracle, mode):
    d, c = tiny_ref(homo=False) if mode == "ref" else tiny_s7()
    P = oracle.Problem(d, c)
    rng = np.random.default_rng(5)
    n2 = d.n[0] * d.n[1]
    for k in range(d.n[2]):
        M = plane_matrix(P, k)
        b = rng.standard_normal(n2) + 1j * rng.standard_normal(n2)
        x = P.plane_solve(k, b)
        assert np.linalg.norm(x - np.linalg.solve(M, b)) < 1e-12 * np.linalg.norm(x)
        xa = P.plane_solve(k, b, adjoint=True)
        assert np.linalg.norm(xa - np.linalg.solve(M.conj().T, b)) < 1e-12 * np.linalg.norm(xa)


def test_plane_solve_needs_pivoting(oracle):
    """beta = 0 without layers gives an indefinite plane; banded LU with
    partial pivoting still matches the dense solve."""
    d, c = tiny_s7(n=(9, 8, 4), npml=0, beta=0.0, omega=30.0)
    P = oracle.Problem(d, np.ones(d.N))
    rng = np.random.default_rng(6)
    for k in range(d.n[2]):
        M = plane_matrix(P, k)
        b = rng.standard_normal(M.shape[0]) + 0j
        x = P.plane_solve(k, b)
        assert np.linalg.norm(M @ x - b) < 1e-10 * np.linalg.norm(b)


# ---------------------------------------------------------- 3D operators
@pytest.mark.parametrize("mode", ["ref_homo", "ref_hetero", "s7"])
def test_precond_and_operators_vs_dense(oracle, mode):
    if mode == "s7":
        d, c = tiny_s7()
    else:
        d, c = tiny_ref(homo=(mode == "ref_homo"))
    P = oracle.Problem(d, c)
    L0 = L0_matrix(P)
    A = true_matrix(P, c)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(d.N) + 1j * rng.standard_normal(d.N)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(P.L0_apply(x), L0 @ x) < 1e-12
    assert rel(P.precond_apply(x), np.linalg.solve(L0, x)) < 1e-11
    assert rel(P.precond_apply(x, adjoint=True), np.linalg.solve(L0.conj().T, x)) < 1e-11
    assert rel(P.op_apply(x), A @ x) < 1e-12
    assert rel(P.op_apply(x, adjoint=True), A.conj().T @ x) < 1e-12
    if d.mode == 0:
        want = x - P.deltaL() * np.linalg.solve(L0, x)
    else:
        want = A @ np.linalg.solve(L0, x)
    assert rel(P.op_pre_apply(x), want) < 1e-11
    # L0 L0^{-1} x = x (TEST_L0, FGMRES.cpp:606-621)
    assert rel(P.L0_apply(P.precond_apply(x)), x) < 1e-11


def test_s7_equals_ref_structure_without_z_layers(oracle):
    """With z PML off and the FD z symbol, L0 is exactly the 7-point stencil with
    the complex-shifted background: the S7 stencil with c = c0 and beta = 0 is L0."""
    d, c = tiny_s7(n=(6, 5, 8), npml=(2, 2, 0), beta=0.0)
    # z-invariant medium so that c0 = c
    c2 = np.broadcast_to(c.reshape(d.shape_zyx)[:1], d.shape_zyx).reshape(-1).copy()
    P = oracle.Problem(d, c2)
    x = np.random.default_rng(8).standard_normal(d.N) + 0.5j
    # periodic vs Dirichlet z differ only through the wrap-around couplings
    A = stencil_matrix(P, c2)
    n2, nz = d.n[0] * d.n[1], d.n[2]
    hz2 = d.h[2] ** 2
    A[:n2, (nz - 1) * n2:] += np.eye(n2) / hz2
    A[(nz - 1) * n2:, :n2] += np.eye(n2) / hz2
    assert np.linalg.norm(P.L0_apply(x) - A @ x) < 1e-12 * np.linalg.norm(A @ x)


# ------------------------------------------------------------ Krylov drivers
def test_fgmres_minimal_residual_and_solution(oracle):
    d, c = tiny_s7()
    P = oracle.Problem(d, c)
    A = stencil_matrix(P, c)
    f = np.zeros(d.N, complex)
    f[d.N // 2 + 3] = 1.0
    u, st = P.fgmres(f, m=150, rest=3, tol=1e-11)
    assert st.converged
    ud = np.linalg.solve(A, f)
    assert np.linalg.norm(u - ud) < 1e-9 * np.linalg.norm(ud)
    # the Givens estimates of the first cycle equal the minimal residual over
    # K_j(A L0^{-1}, f) (x0 = 0)
    Aop = A @ np.linalg.inv(L0_matrix(P))
    mr = min_residuals(Aop, f, 8) / np.linalg.norm(f)
    assert np.allclose(st.hist[:8], mr, rtol=1e-6, atol=1e-14)


@pytest.mark.parametrize("ortho", [0, 1])
def test_gmres_ref_converges_to_dense(oracle, ortho):
    d, c = tiny_ref(homo=False)
    P = oracle.Problem(d, c)
    A = true_matrix(P, c)
    f = np.zeros(d.N, complex)
    f[d.N // 2] = -1.0 / (d.h[0] * d.h[1] * d.h[2])
    u, st = P.gmres_ref(f, m=30, rest=6, thresh=1e-14, tol=1e-10, mask=0, ortho=ortho)
    ud = np.linalg.solve(A, f)
    assert st.converged
    assert np.linalg.norm(u - ud) < 1e-7 * np.linalg.norm(ud)


def test_gmres_ref_textbook_is_minimal_residual(oracle):
    d, c = tiny_ref(homo=False)
    P = oracle.Problem(d, c)
    f = np.zeros(d.N, complex)
    f[d.N // 2] = 1.0
    _, st = P.gmres_ref(f, m=10, rest=1, thresh=1e-14, tol=1e-30, mask=0, ortho=1)
    Apre = np.eye(d.N) - np.diag(P.deltaL()) @ np.linalg.inv(L0_matrix(P))
    mr = min_residuals(Apre, f, 10) / np.linalg.norm(f)
    assert np.allclose(st.hist[:10], mr, rtol=1e-7)


def test_gmres_ref_conj_quirk(oracle):
    """The Arnoldi coefficient w^H v_i (FGMRES.cpp:511) equals the textbook
    v_i^H w when every quantity is real (beta = 0, no layers, real source), and
    differs once the operator is complex."""
    d, c = tiny_ref(homo=False, beta=0.0, pml=0)
    P = oracle.Problem(d, c)
    f = np.zeros(d.N, complex)
    f[d.N // 2] = 1.0
    _, s0 = P.gmres_ref(f, m=8, rest=1, thresh=1e-14, tol=1e-30, mask=0, ortho=0)
    _, s1 = P.gmres_ref(f, m=8, rest=1, thresh=1e-14, tol=1e-30, mask=0, ortho=1)
    assert np.allclose(s0.hist[:8], s1.hist[:8], rtol=1e-9)
    d2, c2 = tiny_ref(homo=False)
    P2 = oracle.Problem(d2, c2)
    f = np.zeros(d2.N, complex)
    f[d2.N // 2] = 1.0
    _, t0 = P2.gmres_ref(f, m=8, rest=1, thresh=1e-14, tol=1e-30, mask=0, ortho=0)
    _, t1 = P2.gmres_ref(f, m=8, rest=1, thresh=1e-14, tol=1e-30, mask=0, ortho=1)
    assert np.abs(t0.hist[:8] - t1.hist[:8]).max() > 1e-6


@pytest.mark.parametrize("mode", ["ref", "s7"])
def test_bicgstab_converges_to_dense(oracle, mode):
    d, c = tiny_ref(homo=False) if mode == "ref" else tiny_s7()
    P = oracle.Problem(d, c)
    A = true_matrix(P, c)
    f = np.zeros(d.N, complex)
    f[d.N // 2] = 1.0
    u, st = P.bicgstab(f, maxit=200, tol=1e-11, tol_abs=0, mask=0)
    assert st.converged
    ud = np.linalg.solve(A, f)
    assert np.linalg.norm(u - ud) < 1e-8 * np.linalg.norm(ud)


# ------------------------------------------------------------------ adjoint
def test_misfit_gradient_vs_dense_and_fd(oracle):
    d, c = tiny_s7(n=(7, 6, 6), npml=1)
    nsrc = 2
    rng = np.random.default_rng(11)
    f = np.zeros((nsrc, d.N), complex)
    f[0, 2 + 7 * (2 + 6 * 3)] = 1.0
    f[1, 4 + 7 * (3 + 6 * 2)] = 1.0
    rec = np.array([1 + 7 * (1 + 6 * 4), 3 + 7 * (4 + 6 * 4), 5 + 7 * (2 + 6 * 4)])
    data = rng.standard_normal((nsrc, rec.size)) * 1e-2 + 1j * 1e-2
    P = oracle.Problem(d, c)
    J, g, u, ok = P.misfit_grad(f, rec, data, m=60, rest=5, tol=1e-13)
    assert ok
    w2 = d.omega ** 2

    def dense_J(cc):
        PP = oracle.Problem(d, cc)
        A = stencil_matrix(PP, cc)
        Jt, gt = 0.0, np.zeros(d.N)
        for s in range(nsrc):
            us = np.linalg.solve(A, f[s])
            r = us[rec] - data[s]
            Jt += 0.5 * np.vdot(r, r).real
            rhs = np.zeros(d.N, complex)
            rhs[rec] = r
            lam = np.linalg.solve(A.conj().T, rhs)
            gt += -w2 * (np.conj(us) * lam).real
        return Jt, gt

    Jd, gd = dense_J(c)
    assert abs(J - Jd) < 1e-9 * abs(Jd)
    assert np.linalg.norm(g - gd) < 1e-7 * np.linalg.norm(gd)
    # central finite differences in m = 1/c^2
    m0 = 1.0 / c ** 2
    for t in range(3):
        dm = rng.standard_normal(d.N) * m0
        eps = 1e-6
        Jp, _ = dense_J(1.0 / np.sqrt(m0 + eps * dm))
        Jm, _ = dense_J(1.0 / np.sqrt(m0 - eps * dm))
        fd = (Jp - Jm) / (2 * eps)
        assert abs(fd - g @ dm) < 1e-5 * abs(fd)


# ------------------------------------------------------- physics (Green's fn)
@pytest.mark.slow
def test_s7_point_source_matches_green_function(oracle):
    """Homogeneous c = 1, 10 points per wavelength: with f = +1/h^3 the solution
    approaches -e^{ikr}/(4 pi r) away from the source and the PML."""
    pr = _pkg().problem
    n, npml = 40, 8
    h = 1.0 / n
    om = pr.omega_for_ppw(1.0, h, 10)
    d = pr.cube_geometry(n, npml, om)
    c = np.ones(d.N)
    P = oracle.Problem(d, c)
    idx = (n // 2 - 1,) * 3
    f = pr.point_source(d, idx, amplitude=1.0 / h ** 3)
    u, st = P.fgmres(f, m=40, rest=10, tol=1e-8)
    assert st.converged
    G = pr.green_function(d, idx, c=1.0)
    X = np.arange(n)
    I, J, K = np.meshgrid(X, X, X, indexing="ij")
    r = np.sqrt((I - idx[0]) ** 2 + (J - idx[1]) ** 2 + (K - idx[2]) ** 2).transpose(2, 1, 0).reshape(-1) * h
    sel = d.phys_mask() & (r >= 3 * h) & (r <= 0.25)
    err = np.linalg.norm(u[sel] + G[sel]) / np.linalg.norm(G[sel])
    assert err < 0.1, err


def test_bicgstab_count_moves_under_roundoff_perturbations(hz, oracle):
    """Why the GPU BiCGStab test holds the iteration count to a band and not to
    +-1: on the 12x10x14 S7 case of that test, the oracle's own BiCGStab count
    moves by several iterations when the right-hand side is perturbed at the
    1e-16 relative level (seeded), while the solutions agree to the solver
    tolerance."""
    pr = hz.problem
    n = (12, 10, 14)
    h = 1.0 / max(n)
    d = pr.cube_geometry(n, 3, pr.omega_for_ppw(0.8, h, 8), beta=0.5, pml_c=30.0)
    c = pr.random_velocity(d, 3, corr_len=0.25, amp=0.1, cmin=0.8, cmax=1.2)
    P = oracle.Problem(d, c)
    f = pr.point_source(d, amplitude=1.0)
    u0, st0 = P.bicgstab(f, maxit=300, tol=1e-9, tol_abs=0, mask=0)
    rng = np.random.default_rng(0)
    its = []
    for _ in range(12):
        fp = f + 1e-16 * np.linalg.norm(f) * rng.standard_normal(f.size) / np.sqrt(f.size)
        u, st = P.bicgstab(fp, maxit=300, tol=1e-9, tol_abs=0, mask=0)
        assert st.converged
        assert np.linalg.norm(u - u0) / np.linalg.norm(u0) < 1e-7
        its.append(st.iters)
    print("bicgstab counts under 1e-16 perturbations: base %d, %s" % (st0.iters, its))
    assert max(its + [st0.iters]) - min(its + [st0.iters]) >= 3


def test_ref_operator_beta_cancels_in_band(hz, oracle):
    """Pins the shift (1 + i beta) of kappa in L0 and of delta-L together: with
    A_ref = L0 - diag(deltaL), both carry i beta omega^2 / c0^2, which cancels
    on the band planes spg..nz-spg (deltaL there is the full
    omega^2 ((1 + i beta)/c0^2 - 1/c^2)), while on the sponge planes the taper
    t < 1 leaves A(beta1) x - A(beta2) x = i (beta1 - beta2) omega^2/c0^2 (1 - t) x:
    zero on the band rows, purely imaginary with a positive factor on the
    sponge rows for beta1 > beta2 and real x.  A sign or a dropped i in either
    shift breaks one of the two."""
    pr = hz.problem
    d0 = pr.ref_geometry(n_nopml=8, pml=3, spg=4, nz_nopml=8, length=(240.0,) * 3, homo=False)
    c = pr.ref_sound_speed(d0)
    x = np.random.default_rng(5).standard_normal(d0.N) + 0j
    outs = []
    for beta in (1.0, 0.25):
        d = d0.replace(beta=beta)
        outs.append(oracle.Problem(d, c).op_apply(x))
    diff = (outs[0] - outs[1]).reshape(d0.n[2], -1)
    X = x.reshape(d0.n[2], -1).real
    nz, spg = d0.n[2], d0.spg[2]
    band = np.arange(spg, nz - spg + 1)
    sponge = np.setdiff1d(np.arange(nz), band)
    scale = np.abs(outs[0]).max()
    assert np.abs(diff[band]).max() < 1e-12 * scale
    q = diff[sponge] / X[sponge]
    assert np.abs(q.real).max() < 1e-9 * np.abs(q.imag).max()
    assert np.all(q.imag > 0)


@pytest.mark.parametrize("mode", ["ref", "s7"])
def test_save_mem_plane_skipping(hz, oracle, mode):
    """SAVE_MEM (FGMRES.cpp:196-206, functions.cpp:6023-6027, 6286-6295): the
    planes with kz^2 = lambda_k >= ratio k0^2 are skipped, and the preconditioner
    and L0 with skipping equal the full ones followed by zeroing those z
    wavenumbers: P_skip f = F^-1 mask F P f, with F the unscaled DFT along z
    (numpy.fft), so no skipped plane leaks into the result and no kept plane
    changes."""
    pr = hz.problem
    if mode == "ref":
        d = pr.ref_geometry(n_nopml=8, pml=3, spg=4, nz_nopml=12, length=(240.0,) * 3, nu=4.0, homo=True)
        k0 = 2 * np.pi * 4.0 / 1280.0
        d = d.replace(skip_ratio=5, skip_k0=k0)
        # L = 400 here: planes with (2 pi q / L)^2 >= 5 k0^2 have q >= 5^0.5 L / 320 = 2.8
        c = pr.ref_sound_speed(d)
    else:
        n = (10, 9, 16)
        d = pr.cube_geometry(n, 2, pr.omega_for_ppw(0.8, 1.0 / 16, 8), beta=0.5, pml_c=30.0)
        c = 0.8 + 0.4 * np.random.default_rng(4).random(d.N)
        d = d.replace(skip_ratio=2)
    full = oracle.Problem(d.replace(skip_ratio=0), c)
    skip = oracle.Problem(d, c)
    lam = full.lamz()
    if d.skip_k0 > 0:
        k0 = d.skip_k0
    else:
        k0 = d.omega / np.abs(full.c0()).min()
    keep = lam < d.skip_ratio * k0 ** 2
    assert np.array_equal(skip.plane_solved().astype(bool), keep)
    assert 0 < keep.sum() < keep.size
    nz = d.n[2]
    f = np.random.default_rng(9).standard_normal(d.N) + 1j * np.random.default_rng(10).standard_normal(d.N)

    def zmask(v):
        V = np.fft.fft(v.reshape(nz, -1), axis=0)
        V[~keep] = 0
        return np.fft.ifft(V, axis=0).reshape(-1)
    for ps, pf in ((skip.precond_apply(f), full.precond_apply(f)), (skip.L0_apply(f), full.L0_apply(f))):
        assert np.linalg.norm(ps - zmask(pf)) <= 1e-12 * np.linalg.norm(pf)
