"""Runs tests/ref_shim_driver.cpp (a C++ caller written against the
reference's signatures, linked against libhz.so) on the reference grid of the
tests, with the medium given as built-in, a caller's array or an OverThrust
file, and loads what it wrote."""
import os
import subprocess

import numpy as np
import pytest

from conftest import load_pkg

DRV = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "iterative-helmholtz-solver-with-pardiso_amd", "_build", "ref_shim_driver")

N_NOPML, NZ_NOPML, PML, SPG, LENGTH, NU, BETA, C_Z = 8, 8, 4, 3, 240.0, 4.0, 0.5, 1280.0


def ref_desc(homo):
    hz = load_pkg()
    return hz.problem.ref_geometry(n_nopml=N_NOPML, pml=PML, spg=SPG, nz_nopml=NZ_NOPML, length=(LENGTH,) * 3,
                                   nu=NU, beta=BETA, c_z=C_Z, homo=bool(homo))


def medium(kind, tmp_path):
    """(homo flag, driver medium argument, sound speed over the grid)."""
    hz = load_pkg()
    if kind in ("homo", "linear"):
        homo = 1 if kind == "homo" else 0
        return homo, "builtin", hz.problem.ref_sound_speed(ref_desc(homo))
    d = ref_desc(0)
    nx, ny, nz = d.n
    if kind == "array":
        # a seeded field of the caller's, already extended into the x/y layers
        box = 1200.0 + 300.0 * np.random.default_rng(11).random((nz, ny - 2 * PML, nx - 2 * PML))
        c = np.pad(box, ((0, 0), (PML, PML), (PML, PML)), mode="edge").reshape(-1)
        path = tmp_path / "c.bin"
        c.astype(np.float64).tofile(path)
        return 0, "array:%s" % path, c
    # an OverThrust-format file: float32, [y][x][z], physical box
    data = (1500.0 + 500.0 * np.random.default_rng(12).random((N_NOPML, N_NOPML, NZ_NOPML))).astype(np.float32)
    path = tmp_path / "OT_test.bin"
    data.tofile(path)
    import files_ref
    return 0, "file:%s" % path, files_ref.sound_from_box(d, data)


def run_driver(tmp_path, homo, medium_arg="builtin", f=None, m=10, rest=5, rel_res=1e-8):
    if not os.path.exists(DRV):
        pytest.fail("the reference-signature test driver is not built (run __graft_entry__.build())")
    hz = load_pkg()
    d = ref_desc(homo)
    if f is None:
        f = hz.problem.point_source(d)
    fin = tmp_path / "f.bin"
    np.asarray(f, dtype=np.complex128).tofile(fin)
    out = str(tmp_path / "res")
    r = subprocess.run([DRV, str(N_NOPML), str(NZ_NOPML), str(PML), str(SPG), repr(LENGTH), repr(NU), repr(BETA),
                        str(homo), repr(C_Z), str(m), str(rest), repr(rel_res), str(fin), out, medium_arg],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)

    def load(name, dtype=np.complex128):
        return np.fromfile(out + "_%s.bin" % name, dtype=dtype)
    return d, load, r
