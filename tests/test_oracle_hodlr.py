"""Pins of the HODLR restatement (oracle/hodlr.py): the reference's row tree,
its truncation rule against the singular value decomposition (Eckart-Young),
the symmetric product against the dense matrix, and the compressed plane
solve at eps = 0 against the oracle's banded-LU preconditioner (an independent
factorisation)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import hodlr  # noqa: E402

from conftest import load_pkg  # noqa: E402


def load_problem_module():
    return load_pkg().problem


def test_tree_follows_the_reference_split():
    """n2 = ceil(m/2) trailing rows, n1 = m - n2 leading rows, leaves at
    m <= smallsize (SymRecCompressStruct, BinaryTreesHODLR.cpp:386-416)."""
    nodes, leaves = hodlr.tree(200, 50)
    assert nodes == [(0, 200, 100, 100, 0), (0, 100, 50, 50, 1), (100, 100, 50, 50, 1)]
    assert leaves == [(0, 50, 2), (50, 50, 2), (100, 50, 2), (150, 50, 2)]
    nodes, leaves = hodlr.tree(13, 3)
    assert nodes == [(0, 13, 6, 7, 0), (0, 6, 3, 3, 1), (6, 7, 3, 4, 1), (9, 4, 2, 2, 2)]
    assert leaves == [(0, 3, 2), (3, 3, 2), (6, 3, 2), (9, 2, 3), (11, 2, 3)]
    # every row in exactly one leaf, for odd and even orders
    for n, s in [(200, 13), (97, 7), (24, 3), (64, 8)]:
        _, lv = hodlr.tree(n, s)
        rows = np.concatenate([np.arange(a, a + m) for a, m, _ in lv])
        assert np.array_equal(np.sort(rows), np.arange(n)) and all(m <= s for _, m, _ in lv)


def test_truncation_is_the_optimal_rank_rule():
    """The kept rank p has S[p-1]/S[0] >= eps > S[p]/S[0], and the spectral
    error of U VT is S[p] (the best rank-p error)."""
    rng = np.random.default_rng(5)
    for m, n, eps in [(30, 20, 1e-3), (20, 31, 1e-6), (16, 16, 0.0)]:
        U0 = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))[0]
        V0 = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
        s = 10.0 ** -np.arange(min(m, n))
        A = (U0[:, :s.size] * s) @ V0[:s.size]
        Ul, VT, S = hodlr.low_rank(A, eps)
        p = Ul.shape[1]
        assert np.allclose(S, s, rtol=1e-8, atol=1e-15 * s[0])
        assert S[p - 1] / S[0] >= eps and (p == S.size or S[p] / S[0] < eps)
        err = np.linalg.norm(A - Ul @ VT, 2)
        assert abs(err - (S[p] if p < S.size else 0.0)) <= 1e-12 * S[0]


def test_symmetric_product_and_dense_form():
    """On a complex symmetric matrix: to_dense rebuilds the upper blocks as the
    transposes of the compressed lower ones, and matvec equals to_dense @ x;
    at eps = 0 the form is exact."""
    rng = np.random.default_rng(6)
    n = 37
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    K = B + B.T
    H = hodlr.Hodlr(K, 0.0, 4)
    assert np.abs(H.to_dense() - K).max() <= 1e-12 * np.abs(K).max()
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(H.matvec(x) - K @ x).max() <= 1e-12 * np.abs(K @ x).max()
    H = hodlr.Hodlr(K, 0.3, 4)
    assert np.abs(H.matvec(x) - H.to_dense() @ x).max() <= 1e-12 * np.abs(K @ x).max()
    Kd = H.to_dense()
    assert np.abs(Kd - Kd.T).max() == 0.0


def _cube(oracle, n=16, npml=3, seed=3, beta=1.0):
    pr = load_problem_module()
    h = 1.0 / n
    c = pr.random_velocity(pr.cube_geometry(n, npml, 1.0), seed=seed)
    om = pr.omega_for_ppw(float(c.min()), h, 10.0)
    d = pr.cube_geometry(n, npml, om, beta=beta, pml_c=3.0 * np.log(1e6) / (2 * npml * h))
    return oracle.Problem(d, c)


def test_scaled_block_inverses_are_symmetric(oracle):
    """K_j = G_j diag(dx dy_j) is complex symmetric (the scaling makes the
    plane operator symmetric)."""
    P = _cube(oracle)
    cmx, cpx, _ = P.tables(0)
    cmy, cpy, _ = P.tables(1)
    dx, dy = hodlr.sym_scaling(cmx, cpx), hodlr.sym_scaling(cmy, cpy)
    for j, G in enumerate(hodlr.schur_inverses(P, 2)):
        K = G * (dx * dy[j])[None, :]
        assert np.abs(K - K.T).max() <= 1e-13 * np.abs(K).max()


@pytest.mark.parametrize("mode", ["s7", "ref"])
def test_exact_compression_equals_banded_lu(oracle, mode):
    """eps = 0: the compressed sweeps give the banded-LU preconditioner (an
    independent factorisation of the same planes) to 1e-12; eps > 0 stays
    within a small multiple of eps."""
    if mode == "s7":
        P = _cube(oracle)
    else:
        from ref_driver import ref_desc
        pr = load_problem_module()
        d = ref_desc(0)
        P = oracle.Problem(d, pr.ref_sound_speed(d))
    rng = np.random.default_rng(7)
    f = rng.standard_normal(P.N) + 1j * rng.standard_normal(P.N)
    ref = P.precond_apply(f)
    u = hodlr.HodlrPrecond(P, 0.0, 3).apply(f)
    assert np.linalg.norm(u - ref) <= 1e-12 * np.linalg.norm(ref)
    for eps in (1e-6, 1e-3):
        u = hodlr.HodlrPrecond(P, eps, 3).apply(f)
        assert np.linalg.norm(u - ref) <= 20 * eps * np.linalg.norm(ref)
