"""CPU checks of the C ABI: the library builds, loads, and exports every entry
point the public headers declare (no device calls)."""
import ctypes
import os
import re

from conftest import ROOT, load_pkg


def _declared(header, pattern):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    return sorted(set(re.findall(pattern, txt)))


def test_library_exports_every_declared_symbol():
    hz = load_pkg()
    L = hz.abi.lib()
    names = _declared("hz.h", r"\b(hz_[a-z0-9_]+)\s*\(")
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # every declared entry point has a ctypes signature in the binding
    assert set(names) <= set(hz.abi.SIGNATURES) | {"hz_ref_configure", "hz_ref_release"}


def test_reference_shim_symbols():
    hz = load_pkg()
    path = hz.abi.LIB_PATH
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    for fn in ("FGMRES", "BCGSTAB", "Solve3DSparseUsingFT_CSR", "ApplyCoeffMatrixA_CSR", "Multiply3DSparseUsingFT",
               "alpha", "beta3D", "beta2D_pml", "GenSparseMatrixOnline2DwithPMLFast", "SetSoundSpeed3D",
               "SetSoundSpeed3DFromFile", "SetSoundSpeed2D", "GenerateDeltaL", "HeteroSoundSpeed3DExtensionToPML",
               "HeteroSoundSpeed2DExtensionToPML"):
        assert re.search(r"_Z\d+%s" % fn, out), fn
    L = hz.abi.lib()
    for fn in ("hz_ref_configure", "hz_ref_release", "hz_ref_set_sound_speed", "hz_ref_last_status"):
        assert hasattr(L, fn), fn


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the ABI structs have the C compiler's size and offsets."""
    import subprocess
    hz = load_pkg()
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stddef.h>
#include <stdio.h>
#include "hz.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(hz_desc), sizeof(hz_krylov_opts),
         sizeof(hz_krylov_stats), offsetof(hz_desc, c_bg), offsetof(hz_krylov_opts, hist),
         offsetof(hz_krylov_stats, t_solve_s), offsetof(hz_krylov_opts, batch), offsetof(hz_krylov_opts, autotune),
         sizeof(hz_tune_info), offsetof(hz_tune_info, fft_ms), offsetof(hz_desc, skip_ratio),
         offsetof(hz_desc, skip_k0), offsetof(hz_krylov_opts, chunk));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    A = hz.abi
    want = [ctypes.sizeof(A.HzDesc), ctypes.sizeof(A.KrylovOpts), ctypes.sizeof(A.KrylovStats),
            A.HzDesc.c_bg.offset, A.KrylovOpts.hist.offset, A.KrylovStats.t_solve_s.offset,
            A.KrylovOpts.batch.offset, A.KrylovOpts.autotune.offset, ctypes.sizeof(A.TuneInfo),
            A.TuneInfo.fft_ms.offset, A.HzDesc.skip_ratio.offset, A.HzDesc.skip_k0.offset, A.KrylovOpts.chunk.offset]
    assert got == want


def test_package_fails_loudly_without_library(monkeypatch):
    hz = load_pkg()
    monkeypatch.setattr(hz.abi, "LIB_PATH", "/nonexistent/libhz.so")
    monkeypatch.setattr(hz.abi, "_lib", None)
    try:
        hz.abi.lib()
    except RuntimeError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("expected a loud failure")
