"""GPU parity of the twisted two-plane packed plane solve (HZ_SOLVE_KERNEL=twist).

The twisted factors split every factor slot's block-tridiagonal chain at
m = ny/2: top-down inverses G_j below m, the middle inverse T_m and bottom-up
inverses B_j above m.  One workgroup per (slot, half) walks half the chain
for both planes that share the slot (k and nz - k).  It is the same linear
operator as the one-directional block LU (the reference's plane solve), so:

  plane solve vs the FP64 plane solve     within 2x the one-directional packed
                                          solve's error (+1e-6)
  preconditioner vs the FP64 oracle       within 2x the full-block error
"""
import numpy as np
import pytest

from conftest import load_pkg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def hz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return load_pkg()


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).cuda()


def _host(x):
    return x.cpu().numpy()


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _case(hz, mode, nx, ny, nz, seed):
    pr = hz.problem
    if mode == "s7":
        py = 3 if ny >= 8 else 0
        d = pr.cube_geometry((nx, ny, nz), (8, py, 2), pr.omega_for_ppw(0.8, 1.0 / nx, 8), beta=0.5, pml_c=30.0)
        c = 0.8 + 0.4 * np.random.default_rng(seed).random(d.N)
    else:
        d = pr.ref_geometry(n_nopml=nx - 16, pml=8, spg=2, nz_nopml=max(nz - 4, 2), length=(240.0,) * 3, homo=False)
        d = d.replace(n=(nx, ny, d.n[2]), l=(d.l[0], d.l[0] * ny / nx, d.l[2]), h=(d.h[0], d.h[0], d.h[2]))
        c = pr.ref_sound_speed(d)
    return d, c


def _cmp(hz, monkeypatch, S, N, nrhs, seed):
    """Largest relative difference between the twisted and the one-directional
    packed solves of the same complex64 data, and the two solves' errors against
    the FP64 plane solve (the factors of both precisions must be built)."""
    xd = _rand((nrhs, N), seed)
    x = torch.from_numpy(xd).to(torch.complex64).cuda()
    want = S.plane_solve(torch.from_numpy(xd.astype(np.complex64).astype(np.complex128)).cuda(), hz.abi.PREC_FP64)
    monkeypatch.setenv("HZ_SOLVE_KERNEL", "sym")
    ref = S.plane_solve(x.clone(), hz.abi.PREC_C64)
    monkeypatch.setenv("HZ_SOLVE_KERNEL", "twist")
    got = S.plane_solve(x.clone(), hz.abi.PREC_C64)
    torch.cuda.synchronize()
    nrm = lambda a, b: (torch.linalg.vector_norm(a.to(torch.complex128) - b) / torch.linalg.vector_norm(b)).item()
    d = max(nrm(got[q], ref[q].to(torch.complex128)) for q in range(nrhs))
    e_tw = max(nrm(got[q], want[q]) for q in range(nrhs))
    e_sym = max(nrm(ref[q], want[q]) for q in range(nrhs))
    return d, e_tw, e_sym


@pytest.mark.parametrize("mode,nx,ny,nz", [
    ("s7", 128, 12, 6), ("s7", 160, 13, 7), ("s7", 192, 12, 8), ("s7", 200, 15, 6),
    ("ref", 128, 20, 8), ("ref", 200, 20, 9),
    ("s7", 128, 2, 6), ("s7", 128, 3, 5), ("s7", 200, 4, 6),
])
def test_twist_matches_packed_solve(hz, oracle, monkeypatch, mode, nx, ny, nz):
    """Twisted vs one-directional packed plane solve for 1 and 3 right-hand
    sides (even and odd ny down to 2, even and odd nz: single-plane slots at
    k = 0 and k = nz/2), then the whole complex64 preconditioner against the
    FP64 oracle."""
    print("case twist %s nx=%d ny=%d nz=%d" % (mode, nx, ny, nz), flush=True)
    d, c = _case(hz, mode, nx, ny, nz, nx + ny)
    S = hz.Solver(d, c)
    S.factor(hz.abi.PREC_FP64)
    S.factor(hz.abi.PREC_C64)
    for nrhs in (1, 3):
        e, e_tw, e_sym = _cmp(hz, monkeypatch, S, d.N, nrhs, 7 * nx + ny + nrhs)
        print("  nrhs=%d twisted vs packed %.2e; vs FP64 twisted %.2e, packed %.2e" % (nrhs, e, e_tw, e_sym))
        # the two complex64 factorisations eliminate in different orders: they
        # differ at the complex64 factors' own accuracy (up to 2e-4 against
        # FP64 in the reference mode's strong layers)
        assert e_tw < 2.0 * e_sym + 1e-6, (nrhs, e_tw, e_sym)
        assert e < 2.0 * e_sym + 1e-6, (nrhs, e, e_sym)
    xb = _rand(d.N, 11 + nx)
    want = oracle.Problem(d, c).precond_apply(xb)
    monkeypatch.setenv("HZ_SOLVE_KERNEL", "twist")
    e_tw = _rel(_host(S.precond_apply(_dev(xb), hz.abi.PREC_C64)), want)
    monkeypatch.setenv("HZ_SOLVE_KERNEL", "mr")
    e_full = _rel(_host(S.precond_apply(_dev(xb), hz.abi.PREC_C64)), want)
    print("  preconditioner error twisted %.2e, full %.2e" % (e_tw, e_full))
    assert e_tw < 2.0 * e_full + 1e-7


def test_twist_save_mem_and_many_rhs(hz, oracle, monkeypatch):
    """Skipped planes (SAVE_MEM, zero output) and slots with one or two planes;
    66 right-hand sides run as two launch pairs (64 + 2)."""
    print("case twist save_mem", flush=True)
    pr = hz.problem
    d = pr.cube_geometry((128, 6, 16), (8, 2, 2), pr.omega_for_ppw(0.8, 1.0 / 16, 8), beta=0.5, pml_c=30.0)
    c = 0.8 + 0.4 * np.random.default_rng(7).random(d.N)
    d = d.replace(skip_ratio=2)
    P = oracle.Problem(d, c)
    keep = P.plane_solved().astype(bool)
    assert 0 < keep.sum() < keep.size
    S = hz.Solver(d, c)
    S.factor(hz.abi.PREC_FP64)
    S.factor(hz.abi.PREC_C64)
    e, e_tw, e_sym = _cmp(hz, monkeypatch, S, d.N, 66, 5)
    print("  66 rhs twisted vs packed %.2e; vs FP64 twisted %.2e, packed %.2e" % (e, e_tw, e_sym))
    assert e < 1e-4 and e_tw < 2.0 * e_sym + 1e-6
    f = _rand(d.N, 21)
    monkeypatch.setenv("HZ_SOLVE_KERNEL", "twist")
    assert _rel(_host(S.precond_apply(_dev(f), hz.abi.PREC_C64)), P.precond_apply(f)) < 1e-4


def test_twist_fgmres_iterations(hz, oracle, monkeypatch):
    """Mixed-precision FGMRES with the twisted preconditioner: the same
    iteration count (+-1) and solution (1e-7) as with the one-directional
    packed solve and as the oracle's FP64 solve."""
    print("case twist fgmres", flush=True)
    pr = hz.problem
    d = pr.cube_geometry((128, 24, 16), (8, 4, 4), pr.omega_for_ppw(0.8, 1.0 / 16, 8), beta=0.5, pml_c=30.0)
    c = pr.random_velocity(d, 3, corr_len=0.25, amp=0.1, cmin=0.8, cmax=1.2)
    src = pr.point_source(d)
    S = hz.Solver(d, c)
    S.factor(hz.abi.PREC_C64)
    out = {}
    for k in ("sym", "twist"):
        monkeypatch.setenv("HZ_SOLVE_KERNEL", k)
        u, st = S.fgmres(_dev(src), m=20, restarts=40, tol=1e-9, prec=hz.abi.PREC_C64)
        out[k] = (_host(u), st)
        print("  %s: %d iterations, converged %s" % (k, st.iters, st.converged))
    u_o, st_o = oracle.Problem(d, c).fgmres(src, m=20, rest=40, tol=1e-9)
    print("  oracle: %d iterations" % st_o.iters)
    assert st_o.converged
    for k, (u, st) in out.items():
        assert st.converged
        assert abs(st.iters - out["sym"][1].iters) <= 1
        assert _rel(u, u_o) < 1e-7, (k, _rel(u, u_o))


def test_twist_lazy_top_chain(hz, monkeypatch):
    """With the twisted factors the factorisation stops the top-down chain at
    ny/2 and completes it on the first use of a one-directional kernel (packed,
    full-block, adjoint, HODLR): every such result equals the one from a
    factorisation that ran the whole chain at once (HZ_FACTOR_FULL), bit for
    bit, and the twisted solve is the same either way."""
    print("case twist lazy", flush=True)
    d, c = _case(hz, "ref", 160, 20, 9, 5)
    monkeypatch.delenv("HZ_SOLVE_KERNEL", raising=False)
    monkeypatch.delenv("HZ_FACTOR_FULL", raising=False)
    lazy = hz.Solver(d, c)
    lazy.factor(hz.abi.PREC_C64)
    monkeypatch.setenv("HZ_FACTOR_FULL", "1")
    full = hz.Solver(d, c)
    full.factor(hz.abi.PREC_C64)
    monkeypatch.delenv("HZ_FACTOR_FULL")
    assert lazy.factor_bytes() == full.factor_bytes()
    x = torch.from_numpy(_rand((3, d.N), 17)).to(torch.complex64).cuda()

    def run(S, kernel, nrhs=1, adjoint=False):
        if kernel:
            monkeypatch.setenv("HZ_SOLVE_KERNEL", kernel)
        else:
            monkeypatch.delenv("HZ_SOLVE_KERNEL", raising=False)
        y = S.plane_solve(x[:nrhs].clone(), hz.abi.PREC_C64, adjoint=adjoint)
        torch.cuda.synchronize()
        return y.cpu().numpy()

    # the twisted solve first (the lazy factors stay incomplete), then every
    # one-directional family
    assert np.array_equal(run(lazy, ""), run(full, ""))
    for kernel, nrhs, adjoint in (("sym", 1, False), ("", 3, False), ("mfma", 3, False), ("", 1, True)):
        a, b = run(lazy, kernel, nrhs, adjoint), run(full, kernel, nrhs, adjoint)
        print("  %s nrhs=%d adjoint=%s: max |diff| %.3g" % (kernel or "default", nrhs, adjoint, np.abs(a - b).max()))
        assert np.array_equal(a, b), (kernel, nrhs, adjoint)
    # HODLR compression of a lazily factored problem
    monkeypatch.delenv("HZ_SOLVE_KERNEL", raising=False)
    lazy2 = hz.Solver(d, c)
    lazy2.factor(hz.abi.PREC_C64)
    i1 = lazy2.compress(hz.abi.PREC_C64, eps=1e-4, smallsize=7)
    i2 = full.compress(hz.abi.PREC_C64, eps=1e-4, smallsize=7)
    assert i1["max_rank"] == i2["max_rank"] and i1["bytes"] == i2["bytes"]
    for k in (0, 3):
        assert np.array_equal(lazy2.hodlr_ranks(k, hz.abi.PREC_C64), full.hodlr_ranks(k, hz.abi.PREC_C64))
