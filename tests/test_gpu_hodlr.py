"""HODLR-compressed plane factors on the GPU (hz_precond_compress and the
compressed plane solve, the reference's HODLR build: SymRecCompressStruct /
LowRankApproxStruct, HODLR/BinaryTreesHODLR.cpp:298-416, applied by
DirSolveFastDiagStruct :2971-3018 in Solve3DSparseUsingFT_HODLR,
functions.cpp:6064-6244) against the oracle's restatement (oracle/hodlr.py)
and the dense factors."""
import os
import sys

import numpy as np
import pytest

from conftest import load_pkg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import hodlr as ohodlr  # noqa: E402


@pytest.fixture(scope="module")
def hz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return load_pkg()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).cuda()


def _cube(hz, n, npml, seed=3, beta=1.0):
    pr = hz.problem
    nn = n if np.isscalar(n) else n[0]
    p0 = npml if np.isscalar(npml) else npml[0]
    h = 1.0 / nn
    d0 = pr.cube_geometry(n, npml, 1.0)
    c = pr.random_velocity(d0, seed=seed)
    om = pr.omega_for_ppw(float(c.min()), h, 10.0)
    d = pr.cube_geometry(n, npml, om, beta=beta, pml_c=3.0 * np.log(1e6) / (2 * p0 * h))
    return d, c


def _case(hz, kind):
    if kind == "s7":
        return _cube(hz, 24, 4)
    from ref_driver import ref_desc
    d = ref_desc(0)
    return d, hz.problem.ref_sound_speed(d)


def _rhs(N, seed=11):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(N) + 1j * rng.standard_normal(N)


def test_exact_compression_equals_dense(hz, oracle):
    """eps = 0 (FP64): every singular value above the FP64 floor is kept, the
    compressed preconditioner equals the dense one and the oracle's (1e-10),
    and each stored rank covers the oracle's singular values above 1e-12 S[0]."""
    print("case hodlr eps=0 fp64 s7 24^3", flush=True)
    d, c = _case(hz, "s7")
    S, P = hz.Solver(d, c), oracle.Problem(d, c)
    f = _rhs(d.N)
    S.factor(hz.abi.PREC_FP64)
    u_dense = S.precond_apply(_dev(f), hz.abi.PREC_FP64).cpu().numpy()
    info = S.compress(hz.abi.PREC_FP64, eps=0.0, smallsize=3)
    assert info["active"] == 1 and info["levels"] == 3 and info["nodes"] == 7
    u = S.precond_apply(_dev(f), hz.abi.PREC_FP64).cpu().numpy()
    assert _rel(u, u_dense) < 1e-11
    assert _rel(u, P.precond_apply(f)) < 1e-10
    H = ohodlr.HodlrPrecond(P, 0.0, 3)
    for k in (0, 5, 12):
        gr = S.hodlr_ranks(k, hz.abi.PREC_FP64)
        for j in range(d.n[1]):
            for q, sv in enumerate(H.planes[k][0][j].sv):
                need = int(np.sum(sv / sv[0] >= 1e-12))
                assert need <= gr[j, q] <= sv.size, (k, j, q, gr[j, q], need)
    S.close()


@pytest.mark.parametrize("kind,prec,eps", [("s7", 0, 1e-8), ("s7", 1, 1e-5), ("s7", 1, 1e-3), ("ref", 1, 1e-4),
                                           ("ref", 0, 1e-6)])
def test_ranks_and_apply_vs_oracle(hz, oracle, kind, prec, eps):
    """Ranks equal the oracle's truncation of the exact singular values except
    at nodes whose deciding singular value lies within the input precision of
    eps (at most 1 apart there); the compressed preconditioner equals the
    oracle's compressed one within eps (plus complex64 rounding) and the dense
    one within 20 eps."""
    print("case hodlr %s prec %d eps %g" % (kind, prec, eps), flush=True)
    d, c = _case(hz, kind)
    S, P = hz.Solver(d, c), oracle.Problem(d, c)
    S.factor(prec)
    info = S.compress(prec, eps=eps, smallsize=3)
    assert info["active"] == 1 and info["max_rank"] >= 1
    H = ohodlr.HodlrPrecond(P, eps, 3)
    band = 1e-6 if prec == 0 else 0.3
    solved = P.plane_solved()
    total = off = 0
    for k in range(d.n[2]):
        if not solved[k]:
            continue
        gr = S.hodlr_ranks(k, prec)
        orr = H.ranks(k)
        assert gr.shape == orr.shape
        for j, q in zip(*np.nonzero(gr != orr)):
            sv = H.planes[k][0][j].sv[q]
            lo = min(gr[j, q], orr[j, q])
            assert abs(int(gr[j, q]) - int(orr[j, q])) <= 1, (k, j, q)
            assert abs(sv[lo] / sv[0] / eps - 1.0) < band, (k, j, q, sv[lo] / sv[0])
            off += 1
        total += gr.size
    assert off <= max(1, total // 100)
    f = _rhs(d.N)
    u = S.precond_apply(_dev(f), prec).cpu().numpy()
    floor = 1e-12 if prec == 0 else 2e-5
    assert _rel(u, H.apply(f)) < 2 * eps + floor
    assert _rel(u, P.precond_apply(f)) < 20 * eps + floor
    S.close()


def test_compressed_fgmres_matches_oracle(hz, oracle):
    """Mixed FGMRES with the compressed complex64 preconditioner (eps 1e-5,
    REF mode, point source) reaches 1e-8 and the oracle's solution, in at most
    10% more iterations than with the dense complex64 factors."""
    print("case hodlr fgmres ref", flush=True)
    d, c = _case(hz, "ref")
    S, P = hz.Solver(d, c), oracle.Problem(d, c)
    f = hz.problem.point_source(d)
    S.factor(hz.abi.PREC_C64)
    _, st_d = S.fgmres(_dev(f), m=20, restarts=50, tol=1e-8, prec=hz.abi.PREC_C64)
    S.compress(hz.abi.PREC_C64, eps=1e-5, smallsize=3)
    u, st = S.fgmres(_dev(f), m=20, restarts=50, tol=1e-8, prec=hz.abi.PREC_C64)
    u_o, st_o = P.fgmres(f, m=20, rest=50, tol=1e-10)
    print("iterations dense %d compressed %d oracle %d" % (st_d.iters, st.iters, st_o.iters))
    assert st.converged and st.iters <= int(1.1 * st_d.iters) + 1
    assert _rel(u.cpu().numpy(), u_o) < 1e-6
    S.close()


def test_free_dense_refactor_and_adjoint(hz):
    """HZ_HODLR_FREE_DENSE releases the dense factors (the compressed apply is
    unchanged, the adjoint apply reports HZ_ERR_STATE); eps < 0 and a new
    factorisation drop the compressed form."""
    print("case hodlr free dense", flush=True)
    d, c = _case(hz, "s7")
    S = hz.Solver(d, c)
    f = _dev(_rhs(d.N, 5))
    S.factor(hz.abi.PREC_C64)
    u0 = S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy()
    S.compress(hz.abi.PREC_C64, eps=1e-4, smallsize=3)
    u1 = S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy()
    ua = S.precond_apply(f, hz.abi.PREC_C64, adjoint=True)  # dense factors kept: works
    assert torch.isfinite(torch.view_as_real(ua)).all()
    info = S.compress(hz.abi.PREC_C64, eps=1e-4, smallsize=3, free_dense=True)
    assert info["dense_bytes"] == 0 and info["bytes"] > 0
    assert S.factor_bytes() == pytest.approx(info["bytes"])
    u2 = S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy()
    assert _rel(u2, u1) < 1e-6
    with pytest.raises(hz.abi.HzError) as e:
        S.precond_apply(f, hz.abi.PREC_C64, adjoint=True)
    assert e.value.code == 3
    S.factor(hz.abi.PREC_C64)
    assert S.hodlr_info(hz.abi.PREC_C64)["active"] == 0
    assert _rel(S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy(), u0) < 1e-6
    S.compress(hz.abi.PREC_C64, eps=1e-4, smallsize=3)
    S.compress(hz.abi.PREC_C64, eps=-1.0, smallsize=3)
    assert S.hodlr_info(hz.abi.PREC_C64)["active"] == 0
    with pytest.raises(hz.abi.HzError):
        S.compress(hz.abi.PREC_C64, eps=1e-4, smallsize=24)  # nothing to split
    S.close()


def test_wide_planes_against_dense(hz):
    """200-wide planes (a 200 x 200 x 16 grid, leaves of 12-13 rows, depth 4):
    the compressed complex64 preconditioner stays within 20 eps of the dense
    one, and the records take less memory than the packed dense factors."""
    print("case hodlr 200x200x16", flush=True)
    d, c = _cube(hz, (200, 200, 16), (20, 20, 4))
    S = hz.Solver(d, c)
    f = _dev(_rhs(d.N, 6))
    S.factor(hz.abi.PREC_C64)
    u0 = S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy()
    for eps in (1e-4, 1e-5):
        info = S.compress(hz.abi.PREC_C64, eps=eps, smallsize=13)
        assert info["levels"] == 4 and info["leaf_max"] == 13
        u = S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy()
        err = _rel(u, u0)
        print("eps %g: max rank %d mean %.2f, %.1f MB vs dense %.1f MB, rel diff %.2e" % (
            eps, info["max_rank"], info["mean_rank"], info["bytes"] / 1e6, info["dense_bytes"] / 1e6, err))
        assert err < 20 * eps + 2e-5
        assert info["bytes"] < 0.5 * info["dense_bytes"]
    S.close()


def test_twisted_compression_against_dense(hz):
    """HODLR records of the twisted complex64 factors (HZ_HODLR_TWISTED, read
    from their packed form): the two-phase compressed sweep stays within 20 eps
    of the dense twisted solve, as close as the one-directional records are to
    theirs; several right-hand sides give each one's single solve; the adjoint
    apply after it completes and uses the dense one-directional factors; with
    the dense factors released the forward apply is unchanged."""
    print("case hodlr twisted 200x60x16", flush=True)
    d, c = _cube(hz, (200, 60, 16), (20, 6, 4))
    S = hz.Solver(d, c)
    f = _dev(_rhs(d.N, 8))
    S.factor(hz.abi.PREC_C64)
    u0 = S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy()
    ua0 = S.precond_apply(f, hz.abi.PREC_C64, adjoint=True).cpu().numpy()
    for eps in (1e-4, 1e-5):
        S.compress(hz.abi.PREC_C64, eps=eps, smallsize=13)
        err1 = _rel(S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy(), u0)
        info = S.compress(hz.abi.PREC_C64, eps=eps, smallsize=13, twisted=True)
        u = S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy()
        err = _rel(u, u0)
        print("eps %g: twisted max rank %d mean %.2f, %.1f MB; rel diff %.2e (one-directional %.2e)" % (
            eps, info["max_rank"], info["mean_rank"], info["bytes"] / 1e6, err, err1))
        assert err < 20 * eps + 2e-5 and err < 3 * err1 + 1e-5
        assert info["bytes"] < 0.5 * info["dense_bytes"]
    f3 = torch.stack([f, 2.0 * f, _dev(_rhs(d.N, 9))])
    u3 = S.precond_apply(f3, hz.abi.PREC_C64).cpu().numpy()
    for q in range(3):
        assert _rel(u3[q], S.precond_apply(f3[q], hz.abi.PREC_C64).cpu().numpy()) < 1e-6
    assert _rel(S.precond_apply(f, hz.abi.PREC_C64, adjoint=True).cpu().numpy(), ua0) < 1e-12
    S.compress(hz.abi.PREC_C64, eps=1e-4, smallsize=13, twisted=True, free_dense=True)
    u2 = S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy()
    assert _rel(u2, S.precond_apply(f, hz.abi.PREC_C64).cpu().numpy()) < 1e-12
    with pytest.raises(hz.abi.HzError) as e:
        S.precond_apply(f, hz.abi.PREC_C64, adjoint=True)
    assert e.value.code == 3
    S.close()


def test_twisted_compressed_fgmres(hz, oracle):
    """Mixed FGMRES with the twisted HODLR preconditioner (eps 1e-5) reaches
    1e-8 and the oracle's solution, within 10% of the dense twisted solve's
    iterations."""
    print("case hodlr twisted fgmres", flush=True)
    pr = hz.problem
    d = pr.cube_geometry((128, 24, 16), (8, 4, 4), pr.omega_for_ppw(0.8, 1.0 / 16, 8), beta=0.5, pml_c=30.0)
    c = pr.random_velocity(d, 3, corr_len=0.25, amp=0.1, cmin=0.8, cmax=1.2)
    S, P = hz.Solver(d, c), oracle.Problem(d, c)
    f = hz.problem.point_source(d)
    S.factor(hz.abi.PREC_C64)
    _, st_d = S.fgmres(_dev(f), m=20, restarts=80, tol=1e-8, prec=hz.abi.PREC_C64)
    S.compress(hz.abi.PREC_C64, eps=1e-5, smallsize=8, twisted=True)
    u, st = S.fgmres(_dev(f), m=20, restarts=80, tol=1e-8, prec=hz.abi.PREC_C64)
    u_o, st_o = P.fgmres(f, m=20, rest=80, tol=1e-10)
    print("iterations dense %d twisted compressed %d oracle %d" % (st_d.iters, st.iters, st_o.iters))
    assert st.converged and st.iters <= int(1.1 * st_d.iters) + 1
    assert _rel(u.cpu().numpy(), u_o) < 1e-6
    S.close()
