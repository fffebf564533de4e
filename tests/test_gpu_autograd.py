"""The misfit gradient through torch.autograd (tests/misfit_autograd.py, a
thin test-only wrapper of hz_misfit_forward / hz_misfit_backward): equal to
the direct C-ABI gradient and the oracle's, scaled by the chain rule, checked
by central differences; and the field-storage policy HZ_FIELDS_AUTO against
the field budget."""
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


def _case(hz):
    pr = hz.problem
    n, npml = (14, 12, 16), 3
    h = 1.0 / n[0]
    om = pr.omega_for_ppw(1.0, h, 10.0)
    d = pr.cube_geometry(n, npml, om, beta=1.0, pml_c=3.0 * np.log(1e6) / (2.0 * npml * h))
    c = pr.gaussian_smooth(1.0 + 0.2 * np.random.default_rng(2).standard_normal(d.N), d, 0.15)
    f = np.zeros((2, d.N), complex)
    f[0, pr.flat_index(d, 5, 5, 6)] = 1.0
    f[1, pr.flat_index(d, 8, 6, 9)] = 1.0
    rec = np.array([pr.flat_index(d, i, j, 12) for i in range(3, 11, 2) for j in range(3, 9, 2)])
    rng = np.random.default_rng(9)
    data = (rng.standard_normal((2, rec.size)) + 1j * rng.standard_normal((2, rec.size))) * 1e-3
    return d, c, torch.from_numpy(f).cuda(), rec, data


def test_autograd_gradient(hz, oracle):
    """m.grad from J.backward() equals the direct hz_misfit_backward gradient
    (1e-12) and the oracle's dense adjoint gradient (1e-7, FP64 solves); (3 J)
    gives 3 m.grad; the directional derivative matches central differences of
    J along two smooth directions (1e-5)."""
    from misfit_autograd import MisfitFunction
    print("case autograd misfit 14x12x16", flush=True)
    d, c, f, rec, data = _case(hz)
    P = oracle.Problem(d, c)
    J_o, g_o, _, ok = P.misfit_grad(f.cpu().numpy(), rec, data, m=60, rest=10, tol=1e-12)
    assert ok
    FP64 = hz.abi.PREC_FP64
    m = torch.tensor(1.0 / c ** 2, dtype=torch.float64, device="cuda", requires_grad=True)
    J = MisfitFunction.apply(m, hz, d, f, rec, data, FP64, 1e-12, True)
    assert abs(J.item() - J_o) < 1e-9 * abs(J_o)
    J.backward()
    g = m.grad.detach().cpu().numpy()
    assert np.linalg.norm(g - g_o) < 1e-7 * np.linalg.norm(g_o)
    S = hz.Solver(d, c)
    S.factor(FP64)
    S.misfit_forward(f, rec, data, m=40, restarts=50, tol=1e-12, prec=FP64)
    g_direct = S.misfit_backward().cpu().numpy()
    S.close()
    assert np.linalg.norm(g - g_direct) < 1e-12 * np.linalg.norm(g_direct)
    m.grad = None
    (3.0 * MisfitFunction.apply(m, hz, d, f, rec, data, FP64, 1e-12, True)).backward()
    assert np.linalg.norm(m.grad.cpu().numpy() - 3.0 * g) < 1e-10 * np.linalg.norm(3.0 * g)
    m0 = (1.0 / c ** 2)
    for k in range(2):
        dm = hz.problem.gaussian_smooth(np.random.default_rng(40 + k).standard_normal(d.N), d, 0.1)
        dm *= 1e-6 * np.linalg.norm(m0) / np.linalg.norm(dm)
        Jpm = [MisfitFunction.apply(torch.tensor(m0 + s * dm, device="cuda"), hz, d, f, rec, data, FP64, 1e-13,
                                    True).item() for s in (1.0, -1.0)]
        fd = (Jpm[0] - Jpm[1]) / 2.0
        assert abs(float(np.dot(g, dm)) - fd) < 1e-5 * abs(fd)


def test_field_policy_auto(hz):
    """HZ_FIELDS_AUTO keeps the forward fields when they fit the budget and
    recomputes them otherwise; both give the stored-fields gradient."""
    print("case misfit field budget", flush=True)
    d, c, f, rec, data = _case(hz)
    C64 = hz.abi.PREC_C64
    S = hz.Solver(d, c)
    S.factor(C64)
    S.misfit_forward(f, rec, data, tol=1e-10, prec=C64, store_fields=True)
    assert S.misfit_fields_stored()
    g_ref = S.misfit_backward().cpu().numpy()
    S.misfit_forward(f, rec, data, tol=1e-10, prec=C64, store_fields="auto")  # default budget: half the free HBM
    assert S.misfit_fields_stored()
    S.misfit_budget(2 * d.N * 16 - 1)  # one byte short of the two fields
    S.misfit_forward(f, rec, data, tol=1e-10, prec=C64, store_fields="auto")
    assert not S.misfit_fields_stored()
    g = S.misfit_backward().cpu().numpy()
    assert np.linalg.norm(g - g_ref) < 1e-6 * np.linalg.norm(g_ref)
    S.misfit_budget(2 * d.N * 16)
    S.misfit_forward(f, rec, data, tol=1e-10, prec=C64, store_fields="auto")
    assert S.misfit_fields_stored()
    S.close()
