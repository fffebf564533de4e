"""The PARDISO CPU path of the baseline (oracle/pardiso_cpu.py): its plane
CSR, preconditioner and FGMRES against the oracle.  Skipped where the MKL
runtime is absent."""
import os
import sys

import numpy as np
import pytest

from conftest import load_pkg

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import pardiso_cpu  # noqa: E402

pytestmark = pytest.mark.skipif(not pardiso_cpu.available(), reason="MKL runtime (libmkl_rt) absent")


def _cube(oracle, n=16, npml=3):
    pr = load_pkg().problem
    h = 1.0 / n
    c = pr.random_velocity(pr.cube_geometry(n, npml, 1.0), seed=3)
    om = pr.omega_for_ppw(float(c.min()), h, 10.0)
    d = pr.cube_geometry(n, npml, om, beta=1.0, pml_c=3.0 * np.log(1e6) / (2 * npml * h))
    return d, c, oracle.Problem(d, c)


def test_plane_csr_is_the_plane_operator(oracle):
    """The CSR times x equals the oracle's plane matvec (1e-14), columns sorted
    and 1-based."""
    d, c, P = _cube(oracle)
    n2 = d.n[0] * d.n[1]
    x = np.random.default_rng(1).standard_normal(n2) + 1j * np.random.default_rng(2).standard_normal(n2)
    for k in (0, 3, 9):
        ia, ja, a = pardiso_cpu.plane_csr(P, k)
        assert ia[0] == 1 and ia[-1] == a.size + 1
        y = np.zeros(n2, complex)
        for r in range(n2):
            cols = ja[ia[r] - 1:ia[r + 1] - 1] - 1
            assert np.all(np.diff(cols) > 0)
            y[r] = np.dot(a[ia[r] - 1:ia[r + 1] - 1], x[cols])
        ref = P.plane_matvec(k, x)
        assert np.abs(y - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("mode", ["s7", "ref"])
def test_pardiso_preconditioner_and_fgmres(oracle, mode):
    """PARDISO plane solves give the oracle's preconditioner (1e-10); the numpy
    FGMRES with it reaches 1e-8 and the oracle's solution (1e-6)."""
    if mode == "s7":
        d, c, P = _cube(oracle)
    else:
        from ref_driver import ref_desc
        d = ref_desc(0)
        c = load_pkg().problem.ref_sound_speed(d)
        P = oracle.Problem(d, c)
    rng = np.random.default_rng(4)
    f = rng.standard_normal(d.N) + 1j * rng.standard_normal(d.N)
    M = pardiso_cpu.PardisoPrecond(P, threads=2)
    ref = P.precond_apply(f)
    assert np.linalg.norm(M.apply(f) - ref) <= 1e-10 * np.linalg.norm(ref)
    u, it, ok = pardiso_cpu.fgmres(P, M, f, m=20, tol=1e-8, max_iters=2000)
    M.release()
    u_o, st = P.fgmres(f, m=20, rest=100, tol=1e-10)
    assert ok and np.linalg.norm(u - u_o) <= 1e-6 * np.linalg.norm(u_o)
