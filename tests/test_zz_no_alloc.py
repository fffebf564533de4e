"""The allocation-free preconditioner apply (hz_precond_apply_noalloc): same
result as hz_precond_apply on two calls in a row, the second with a smaller
input on another problem, in both precisions; no device memory taken by a
call; requests beyond its fixed array rejected with an error."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_pkg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def hz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return load_pkg()


def _problem(hz, shape, npml, seed):
    pr = hz.problem
    h = 1.0 / max(shape)
    d = pr.cube_geometry(shape, npml, pr.omega_for_ppw(0.8, h, 8), beta=0.5, pml_c=30.0)
    c = 0.8 + 0.4 * np.random.default_rng(seed).random(d.N)
    S = hz.Solver(d, c)
    S.factor(hz.abi.PREC_FP64)
    S.factor(hz.abi.PREC_C64)
    return d, S


def _rand(N, nrhs, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nrhs, N)) + 1j * rng.standard_normal((nrhs, N))
    return torch.from_numpy(x).cuda()


def _noalloc(hz, S, prec, x):
    out = torch.empty_like(x)
    torch.cuda.synchronize()
    rc = hz.abi.lib().hz_precond_apply_noalloc(S.h, int(prec), x.numel() // S.N, C.c_void_p(x.data_ptr()),
                                               C.c_void_p(out.data_ptr()))
    hz.abi.check(rc, "hz_precond_apply_noalloc")
    hz.abi.check(hz.abi.lib().hz_problem_sync(S.h), "hz_problem_sync")
    return out


def test_noalloc_matches_precond_apply(hz):
    # a 128-wide grid (packed symmetric plane solve) with 3 right-hand sides,
    # then a smaller grid with one
    cases = [((128, 10, 12), (8, 2, 2), 3), ((24, 20, 16), 4, 1)]
    for prec in (hz.abi.PREC_FP64, hz.abi.PREC_C64):
        for k, (shape, npml, nrhs) in enumerate(cases):
            print("case noalloc prec %d shape %s nrhs %d" % (prec, shape, nrhs), flush=True)
            d, S = _problem(hz, shape, npml, 7 + k)
            x = _rand(d.N, nrhs, 100 + k)
            ref = S.precond_apply(x, prec)
            got = _noalloc(hz, S, prec, x)
            err = (torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref)).item()
            print("  rel diff %.3e, max abs diff %.3e" % (err, (got - ref).abs().max().item()))
            assert err <= 1e-14, err
            # a repeated call takes no device memory
            free0 = torch.cuda.mem_get_info()[0]
            got2 = _noalloc(hz, S, prec, x)
            assert torch.cuda.mem_get_info()[0] == free0
            assert torch.equal(got2, got)
            S.close()


def test_noalloc_rejects_oversized_request(hz):
    print("case noalloc oversized", flush=True)
    d, S = _problem(hz, (24, 20, 16), 4, 9)
    # 64 MiB of complex128 plane values hold 4 Mi points: 547 vectors of 7680 exceed it
    nrhs = (64 << 20) // (16 * d.N) + 1
    x = torch.zeros(nrhs * d.N, dtype=torch.complex128, device="cuda")
    out = torch.empty_like(x)
    torch.cuda.synchronize()
    rc = hz.abi.lib().hz_precond_apply_noalloc(S.h, hz.abi.PREC_FP64, nrhs, C.c_void_p(x.data_ptr()),
                                               C.c_void_p(out.data_ptr()))
    assert rc == 1  # HZ_ERR_ARG
    assert b"fixed" in hz.abi.lib().hz_last_error()
    # the complex64 form of the same request fits
    rc = hz.abi.lib().hz_precond_apply_noalloc(S.h, hz.abi.PREC_C64, nrhs // 2, C.c_void_p(x.data_ptr()),
                                               C.c_void_p(out.data_ptr()))
    hz.abi.check(rc, "hz_precond_apply_noalloc")
    hz.abi.check(hz.abi.lib().hz_problem_sync(S.h), "hz_problem_sync")
    S.close()
