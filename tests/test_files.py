"""Velocity-model ingestion and the reference's solution / convergence files
(host entry points of libhz.so, no GPU) against the numpy restatement in
oracle/files_ref.py and against a convergence log the reference itself wrote
(tests/golden/convergence_N49_..._FGMRES.dat, copied from its HETERO/0203 runs)."""
import os

import numpy as np
import pytest

from conftest import load_pkg

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "convergence_N49_PML20_Lx3200_FREQ4_SPG  1920_BETA0.500_FGMRES.dat")


@pytest.fixture(scope="module")
def hz():
    return load_pkg()


@pytest.fixture(scope="module")
def ref():
    import sys
    sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
    import files_ref
    sys.path.pop(0)
    return files_ref


def _ref_geometry(hz, nxp=6, nyp=5, pml=3, spg=2, nzp=4):
    pr = hz.problem
    d = pr.ref_geometry(n_nopml=nxp, pml=pml, spg=spg, nz_nopml=nzp, length=(240.0,) * 3)
    ny = nyp + 2 * pml
    return d.replace(n=(d.n[0], ny, d.n[2]), l=(d.l[0], d.h[0] * ny, d.l[2]), h=(d.h[0], d.h[0], d.h[2]))


@pytest.mark.parametrize("kind", ["ref", "s7"])
def test_velocity_from_file_matches_restatement(hz, ref, tmp_path, kind):
    if kind == "ref":
        d = _ref_geometry(hz)
    else:
        d = hz.problem.cube_geometry((9, 7, 8), (2, 1, 2), 30.0)
    lx, ly, lz = (int(d.pml[a]) + int(d.spg[a]) for a in range(3))
    n1, n2, n3 = d.n[0] - 2 * lx, d.n[1] - 2 * ly, d.n[2] - 2 * lz
    box = (1000.0 + 800.0 * np.random.default_rng(3).random((n2, n1, n3))).astype(np.float32)
    path = tmp_path / hz.files.velocity_file_name(90.0)
    assert path.name == "OT_small_at_90m.bin"
    box.tofile(path)
    c = hz.files.velocity_from_file(path, d)
    want = ref.sound_from_box(d, box)
    assert np.array_equal(c, want)
    # the physical box holds the file's values; layers repeat the boundary nodes
    c3 = c.reshape(d.shape_zyx)
    assert np.array_equal(c3[lz:-lz, ly:-ly, lx:-lx], box.transpose(2, 0, 1).astype(np.float64))
    assert np.array_equal(c3[:, :, 0], c3[:, :, lx])
    assert np.array_equal(c3[:, 0, :], c3[:, ly, :])
    assert np.array_equal(c3[0], c3[lz])


def test_velocity_from_file_errors(hz, tmp_path):
    d = _ref_geometry(hz)
    lx, ly, lz = (int(d.pml[a]) + int(d.spg[a]) for a in range(3))
    count = (d.n[0] - 2 * lx) * (d.n[1] - 2 * ly) * (d.n[2] - 2 * lz)
    short = tmp_path / "short.bin"
    np.ones(count - 1, dtype=np.float32).tofile(short)
    with pytest.raises(hz.abi.HzError, match="float32 values"):
        hz.files.velocity_from_file(short, d)
    bad = tmp_path / "bad.bin"
    v = np.ones(count, dtype=np.float32)
    v[5] = -1.0
    v.tofile(bad)
    with pytest.raises(hz.abi.HzError, match="speed"):
        hz.files.velocity_from_file(bad, d)
    with pytest.raises(hz.abi.HzError, match="cannot open"):
        hz.files.velocity_from_file(tmp_path / "absent.bin", d)


def test_solution_file_format_and_round_trip(hz, ref, tmp_path):
    d = _ref_geometry(hz)
    rng = np.random.default_rng(11)
    u = rng.standard_normal(d.N) * 10.0 ** rng.integers(-3, 3, d.N) + 1j * rng.standard_normal(d.N)
    path = tmp_path / ("sol3D_N%d.dat" % (d.n[0] - 2 * (d.pml[0] + d.spg[0])))
    hz.files.write_solution(path, d, u)
    text = path.read_text()
    assert text == ref.solution_text(d, u)
    back = hz.files.read_solution(path, d)
    phys = ref.reduce_pml(d, np.arange(d.N))
    mask = np.zeros(d.N, dtype=bool)
    mask[phys] = True
    assert np.all(back[~mask] == 0)
    assert np.max(np.abs(back[mask] - u[mask]) / np.maximum(1.0, np.abs(u[mask]))) < 1e-15
    # one value short / one value extra are rejected
    lines = text.splitlines(True)
    (tmp_path / "short.dat").write_text("".join(lines[:-1]))
    with pytest.raises(hz.abi.HzError, match="missing"):
        hz.files.read_solution(tmp_path / "short.dat", d)
    (tmp_path / "long.dat").write_text(text + "1.0 2.0\n")
    with pytest.raises(hz.abi.HzError, match="more values"):
        hz.files.read_solution(tmp_path / "long.dat", d)


def test_convergence_log_reproduces_reference_file(hz, ref, tmp_path):
    """Parsing the reference's own log and writing it back gives the same bytes."""
    raw = open(GOLDEN, "rb").read()
    rows = [line.split() for line in raw.decode().splitlines()]
    it = [int(r[0]) for r in rows]
    res, rel, diff = ([float(r[c]) for r in rows] for c in (1, 2, 3))
    out = tmp_path / "conv.dat"
    hz.files.write_convergence(out, it, res, rel, diff)
    assert out.read_bytes() == raw
    assert out.read_text() == ref.convergence_text(it, res, rel, diff)


def test_convergence_file_name_matches_reference(hz, ref):
    # N = 49 physical nodes, 20 PML nodes, Lx = 3200, nu = 4, sponge 15 nodes of
    # h_z = 64 on each side, beta = 0.5 (the reference's HETERO/0203 run)
    pr = hz.problem
    d = pr.ref_geometry(n_nopml=49, pml=20, spg=15, nz_nopml=20, length=(3200.0, 3200.0, 1280.0))
    d = d.replace(h=(d.h[0], d.h[1], 64.0))
    name = hz.files.convergence_file_name(d, 3200.0, 4.0, 0.5)
    assert name == os.path.basename(GOLDEN)
    assert name == ref.convergence_name(49, 20, 3200.0, 4.0, 64.0 * 2 * 15, 0.5)


def _write_box(path, u):
    path.write_text("".join("%18.16f %18.16f\n" % (v.real, v.imag) for v in np.asarray(u).reshape(-1)))


def _box_index(n, s):
    """Fine-grid index s (i + 1) - 1 of every node of an n-per-side box, [k][j][i] order."""
    m = s * (np.arange(n) + 1) - 1
    k, j, i = np.meshgrid(m, m, m, indexing="ij")
    return i.reshape(-1), j.reshape(-1), k.reshape(-1)


def test_runge_ratio_closed_form(hz, ref, tmp_path):
    """Three nested boxes (4, 9, 19 per side: steps 4 and 2, as the reference's
    49, 99, 199) that agree up to constants at coinciding nodes: every
    difference is that constant, so f = h n^3 |c_fine - c| exactly.  A wrong
    restriction index, axis order or step makes the differences vary."""
    n1, n2, n3, h = 4, 9, 19, 0.5

    def field(i, j, k):  # integer-valued, distinct along every axis
        return (3 * i + 5 * j * j - 7 * k) + 1j * (i * k - 2 * j)

    c1, c2, c3 = 0.25 - 0.5j, 0.625 + 0.0j, 1.0 + 0.5j  # |c3 - c1| = 1.25, |c3 - c2| = 0.625
    u1 = field(*_box_index(n1, 4)) + c1
    u2 = field(*_box_index(n2, 2)) + c2
    u3 = field(*_box_index(n3, 1)) + c3
    paths = [tmp_path / ("sol3D_N%d.dat" % (n + 1)) for n in (n1, n2, n3)]
    for p, u in zip(paths, (u1, u2, u3)):
        _write_box(p, u)
    f1, f2, ratio = hz.files.runge_ratio(*paths, n1, n2, n3, h)
    assert f1 == h * n1 ** 3 * abs(c3 - c1) == 40.0
    assert f2 == h * n2 ** 3 * abs(c3 - c2) == 227.8125
    assert ratio == f1 / f2
    # random boxes: the restatement on the same file contents
    rng = np.random.default_rng(5)
    us = []
    for p, n in zip(paths, (n1, n2, n3)):
        _write_box(p, rng.standard_normal(n ** 3) + 1j * rng.standard_normal(n ** 3))
        us.append(np.loadtxt(p).view(np.complex128).reshape(-1))
    got = hz.files.runge_ratio(*paths, n1, n2, n3, h)
    want = ref.runge_ratio(*us, n1, n2, n3, h)
    assert np.allclose(got, want, rtol=1e-13, atol=0)


def test_runge_ratio_errors(hz, tmp_path):
    paths = [tmp_path / ("b%d.dat" % n) for n in (4, 9, 19)]
    for p, n in zip(paths, (4, 9, 19)):
        _write_box(p, np.ones(n ** 3, dtype=complex))
    with pytest.raises(hz.abi.HzError, match="not nested"):
        hz.files.runge_ratio(paths[0], paths[1], paths[2], 4, 8, 19, 1.0)
    with pytest.raises(hz.abi.HzError, match="more values"):
        hz.files.runge_ratio(paths[0], paths[1], paths[2], 3, 9, 19, 1.0)
    with pytest.raises(hz.abi.HzError, match="64 of 125 values"):
        hz.files.runge_ratio(paths[0], paths[1], paths[2], 5, 9, 29, 1.0)


def test_beta3d_norms_lattice_counts(hz, ref, tmp_path):
    """h = 2^-6 and the source on a node make every offset, radius and
    comparison exact, so shell membership is the integer test 4 <= |v|^2 <= e^2.
    The coarse solution is 1 + eps on the nodes with |v|^2 <= R2 and 1 elsewhere,
    the fine one is 1 at the coarse nodes and 1e3 between them: the norm of
    shell e is eps sqrt(#{4 <= |v|^2 <= min(e^2, R2)} / #{4 <= |v|^2 <= e^2})."""
    grid, h, eps, R2 = 12, 2.0 ** -6, 0.5, 10
    src = (grid // 2 * h,) * 3
    n1, n2 = grid - 1, 2 * grid - 1
    off = np.arange(1, grid) - grid // 2  # (i + 1) - grid / 2
    vk, vj, vi = np.meshgrid(off, off, off, indexing="ij")
    q = (vi * vi + vj * vj + vk * vk).reshape(-1)
    uc = 1.0 + eps * (q <= R2) + 0j
    fi, fj, fk = _box_index(n2, 1)
    on = ((fi + 1) % 2 == 0) & ((fj + 1) % 2 == 0) & ((fk + 1) % 2 == 0)
    uf = np.where(on, 1.0, 1e3) + 0j
    pc, pf = tmp_path / "sol3D_N12.dat", tmp_path / "sol3D_N24.dat"
    _write_box(pc, uc)
    _write_box(pf, uf)
    radius, norm = hz.files.beta3d_norms(pc, pf, grid, h, src)
    assert len(norm) == grid - 3
    for e in range(2, grid - 1):
        inside = (q >= 4) & (q <= e * e)
        want = eps * np.sqrt(np.count_nonzero(inside & (q <= R2)) / np.count_nonzero(inside))
        assert radius[e - 2] == e * h
        assert abs(norm[e - 2] - want) <= 1e-14 * want, (e, norm[e - 2], want)
    # random fields and an off-node source: the shell-by-shell restatement
    rng = np.random.default_rng(8)
    grid, h, src = 10, 0.1, (0.47, 0.52, 0.5)
    n1, n2 = grid - 1, 2 * grid - 1
    _write_box(pc, rng.standard_normal(n1 ** 3) + 1j * rng.standard_normal(n1 ** 3))
    _write_box(pf, rng.standard_normal(n2 ** 3) + 1j * rng.standard_normal(n2 ** 3))
    radius, norm = hz.files.beta3d_norms(pc, pf, grid, h, src)
    rr, rn = ref.beta3d_norms(np.loadtxt(pc).view(np.complex128).reshape(-1),
                              np.loadtxt(pf).view(np.complex128).reshape(-1), grid, h, src)
    assert np.array_equal(radius, rr)
    assert np.isnan(norm[0]) and np.isnan(rn[0])  # no node in 2h <= r <= 2h: 0 / 0, as the reference
    assert np.allclose(norm, rn, rtol=1e-12, atol=0, equal_nan=True)


def test_beta_norms_file(hz, ref, tmp_path):
    radius, norm = np.arange(2, 9) * 0.02, np.linspace(0.01, 0.3, 7)
    name = hz.files.beta_norm_file_name(49, 5)
    assert name == "beta_3D_norm2_Nx49Ny49_SPG5.dat" == ref.beta_norm_name(49, 5)
    hz.files.write_beta_norms(tmp_path / name, radius, norm)
    assert (tmp_path / name).read_text() == ref.beta_norm_text(radius, norm)
    with pytest.raises(hz.abi.HzError, match="shells"):
        hz.files.beta3d_norms(tmp_path / name, tmp_path / name, 3, 1.0, (0.0, 0.0, 0.0))
