"""The file-based rows of the reference's path driven end to end on the GPU:
a velocity model read from an OverThrust-format file into a device solve
(SetSoundSpeed3DFromFile feeding FGMRES, FGMRES.cpp:78-82), and GPU solutions
written as sol3D_N%d.dat files and passed through the reference's accuracy
studies (main.cpp:827-834; Runge and Beta3D, functions.cpp:932-1239), each
against the oracle or the numpy restatement on the same data."""
import os
import sys

import numpy as np
import pytest

from conftest import load_pkg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def hz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return load_pkg()


@pytest.fixture(scope="module")
def ref():
    sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
    import files_ref
    sys.path.pop(0)
    return files_ref


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).cuda()


def test_velocity_file_model_through_gpu_solves(hz, oracle, ref, tmp_path):
    """OVERTRUST-style run on the reference grid (x/y PML 4, z sponge 4): the
    model comes from a float32 [y][x][z] file over the physical box (seeded
    smooth field, 1500-2500 m/s); the GPU problem built from hz_velocity_from_file
    and the oracle built from the numpy restatement of the reader give the same
    reference GMRES solve (iterations +-1, solution 1e-10) and the same FP64
    preconditioner (1e-10)."""
    print("case velocity file -> gpu solve", flush=True)
    pr = hz.problem
    d = pr.ref_geometry(n_nopml=16, pml=4, spg=4, nz_nopml=12, length=(1600.0, 1600.0, 1200.0), nu=4.0,
                        beta=0.5, homo=False)
    nx, ny, nz = d.n
    n1, n2, n3 = nx - 8, ny - 8, nz - 8
    rng = np.random.default_rng(31)
    g = rng.standard_normal((n2, n1, n3))
    for ax in range(3):  # a little smoothing along every axis
        g = (g + np.roll(g, 1, ax) + np.roll(g, -1, ax)) / 3.0
    box = (2000.0 + 500.0 * g / np.abs(g).max()).astype(np.float32)
    path = tmp_path / hz.files.velocity_file_name(100.0)
    box.tofile(path)
    c = hz.files.velocity_from_file(path, d)
    c_ref = ref.sound_from_box(d, box)
    assert np.array_equal(c, c_ref)
    S, P = hz.Solver(d, c), oracle.Problem(d, c_ref)
    S.factor(hz.abi.PREC_FP64)
    f = pr.point_source(d)
    assert _rel(S.precond_apply(_dev(f), hz.abi.PREC_FP64).cpu().numpy(), P.precond_apply(f)) < 1e-10
    u_o, st_o = P.gmres_ref(f, m=10, rest=30, thresh=1e-12, tol=1e-8, mask=1, ortho=0)
    u, st = S.fgmres(_dev(f), m=10, restarts=30, tol=1e-8, driver=hz.abi.DRIVER_GMRES_REF,
                     ortho=hz.abi.ORTHO_REF_CONJ, mask=hz.abi.MASK_PHYS, breakdown=1e-12)
    err = _rel(u.cpu().numpy(), u_o)
    print("file model: iterations gpu %d oracle %d, u err %.2e" % (st.iters, st_o.iters, err))
    assert st.converged and st_o.converged and abs(st.iters - st_o.iters) <= 1
    assert err < (1e-10 if st.iters == st_o.iters else 1e-7)


def test_gpu_solutions_through_accuracy_studies(hz, ref, tmp_path):
    """Point-source solves on the GPU (FP64 FGMRES to 1e-10) with physical boxes
    of 15, 31 and 63 nodes per side, written by hz_write_solution as
    sol3D_N%d.dat and read by the accuracy studies: Runge's ratio of the three
    and Beta3D's shell norms of the 31/63 pair equal the restatement's on the
    same files (1e-12); the written text is the restatement's; every norm is
    finite and the solution is nonzero."""
    print("case gpu solutions -> accuracy studies", flush=True)
    pr = hz.problem
    npml = 9  # grids of 33, 49 and 81 nodes: z lengths with prime factors the z-FFT takes (<= 64)
    sols, paths, descs = [], [], []
    for nphys in (15, 31, 63):
        n = nphys + 2 * npml
        h = 1.0 / n
        d = pr.cube_geometry(n, npml, pr.omega_for_ppw(1.0, 1.0 / 24, 10.0), beta=0.5,
                             pml_c=3.0 * np.log(1e6) / (2.0 * npml * h))
        S = hz.Solver(d, np.ones(d.N))
        S.factor(hz.abi.PREC_FP64)
        f = pr.point_source(d, (n // 2,) * 3, amplitude=1.0 / h ** 3)
        u, st = S.fgmres(_dev(f), m=30, restarts=200, tol=1e-10, prec=hz.abi.PREC_FP64)
        assert st.converged
        uh = u.cpu().numpy()
        path = tmp_path / ("sol3D_N%d.dat" % nphys)
        hz.files.write_solution(path, d, uh)
        assert path.read_text() == ref.solution_text(d, uh)
        sols.append(np.loadtxt(path).view(np.complex128).reshape(-1))
        paths.append(path)
        descs.append(d)
        S.close()
    h = 1.0 / 64
    got = hz.files.runge_ratio(*paths, 15, 31, 63, h)
    want = ref.runge_ratio(*sols, 15, 31, 63, h)
    print("Runge: f1 %.6e f2 %.6e ratio %.4f" % tuple(got))
    assert np.all(np.isfinite(got)) and got[0] > 0
    assert np.allclose(got, want, rtol=1e-12, atol=0)
    # Beta3D over the 31 / 63 pair: coarse spacing 1/32, the source on coarse
    # node 15 (every shell 2h <= r <= e h holds nodes)
    src, hc = (0.5, 0.5, 0.5), 1.0 / 32
    radius, norm = hz.files.beta3d_norms(paths[1], paths[2], 32, hc, src)
    rr, rn = ref.beta3d_norms(sols[1], sols[2], 32, hc, src)
    assert np.array_equal(radius, rr)
    assert np.all(np.isfinite(norm))
    assert np.allclose(norm, rn, rtol=1e-12, atol=0, equal_nan=True)
