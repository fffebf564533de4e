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
