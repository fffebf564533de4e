// ref_shim_driver.cpp -- a caller written against the reference's own
// signatures (include/hz_ref.h), the way the reference calls them, linked
// against libhz.so instead of the MKL build: the grid set-up of main.cpp:122-211,
// the medium and delta-L set-up of FGMRES.cpp:77-96 (SetSoundSpeed3D or
// SetSoundSpeed3DFromFile, the two layer extensions, SetSoundSpeed2D,
// GenerateDeltaL), the plane assembly of FGMRES.cpp:140-141 and 256
// (GenSparseMatrixOnline2DwithPMLFast), the coefficient functions alpha,
// beta2D_pml and beta3D, then Solve3DSparseUsingFT_CSR, ApplyCoeffMatrixA_CSR
// (with the delta-L it computed), Multiply3DSparseUsingFT and FGMRES (main.cpp:470).
// Test program: reads the right-hand side from a file and writes every result
// to a file.
//
//   ref_shim_driver n_nopml nz_nopml pml spg length nu beta homo c_z m rest rel_res in.bin out_prefix [medium]
//
// medium: "builtin" (default: HOMO c_z or the linear model), "array:<path>"
// (N doubles over the whole grid, given to hz_ref_set_sound_speed) or
// "file:<path>" (an OverThrust float32 file, hz_ref_config.velocity_file).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hz_ref.h"

static size_m axis(int n_nopml, int layer, double length, bool pml) {
  size_m a{};
  a.n_nopml = n_nopml;
  a.n = n_nopml + 2 * layer;
  a.l = length + 2.0 * layer * length / n_nopml;
  a.h = a.l / a.n;
  a.pml_pts = pml ? layer : 0;
  a.spg_pts = pml ? 0 : layer;
  return a;
}

template <class T>
static bool read_all(const char* path, std::vector<T>& v) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  const size_t got = fread(v.data(), sizeof(T), v.size(), f);
  fclose(f);
  return got == v.size();
}

template <class T>
static bool write_all(const std::string& path, const std::vector<T>& v) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  const size_t put = fwrite(v.data(), sizeof(T), v.size(), f);
  fclose(f);
  return put == v.size();
}

int main(int argc, char** argv) {
  if (argc != 15 && argc != 16) {
    fprintf(stderr,
            "usage: %s n_nopml nz_nopml pml spg length nu beta homo c_z m rest rel_res in.bin out_prefix [medium]\n",
            argv[0]);
    return 2;
  }
  const int n_nopml = atoi(argv[1]), nz_nopml = atoi(argv[2]), pml = atoi(argv[3]), spg = atoi(argv[4]);
  const double length = atof(argv[5]), nu = atof(argv[6]), beta = atof(argv[7]);
  const int homo = atoi(argv[8]);
  const double c_z = atof(argv[9]);
  const int m = atoi(argv[10]), rest = atoi(argv[11]);
  const double rel_res = atof(argv[12]);
  const std::string out = argv[14];
  const std::string med = argc == 16 ? argv[15] : "builtin";

  size_m x = axis(n_nopml, pml, length, true), y = axis(n_nopml, pml, length, true);
  size_m z = axis(nz_nopml, spg, length, false);
  const size_t N = (size_t)x.n * y.n * z.n, n2 = (size_t)x.n * y.n;
  std::string vfile;
  if (med.rfind("file:", 0) == 0) vfile = med.substr(5);
  hz_ref_config cfg = {nu, c_z, homo, rel_res, 1e-8, 1000.0, HZ_PREC_FP64, vfile.empty() ? nullptr : vfile.c_str()};
  if (hz_ref_configure(&cfg) != HZ_OK) return 3;
  const point source = {x.l / 2, y.l / 2, z.l / 2};

  // ---- medium and delta-L (FGMRES.cpp:77-96)
  std::vector<dtype> sound3D(N, dtype(0.0)), sound2D(n2, dtype(0.0)), deltaL(N, dtype(0.0));
  if (med.rfind("array:", 0) == 0) {
    std::vector<double> c(N);
    if (!read_all(med.substr(6).c_str(), c)) return 4;
    if (hz_ref_set_sound_speed(c.data(), (int64_t)N) != HZ_OK) return 3;
    for (size_t l = 0; l < N; ++l) sound3D[l] = c[l];
  } else if (!vfile.empty()) {
    SetSoundSpeed3DFromFile(x, y, z, sound3D.data(), source);
    if (hz_ref_last_status() != HZ_OK) return 6;
    // z sponge planes: the nearest physical plane (the reference leaves them unset)
    for (int k = 0; k < z.n; ++k) {
      const int kk = k < z.spg_pts ? z.spg_pts : (k >= z.n - z.spg_pts ? z.n - z.spg_pts - 1 : k);
      if (kk != k) memcpy(&sound3D[(size_t)k * n2], &sound3D[(size_t)kk * n2], n2 * sizeof(dtype));
    }
  } else {
    SetSoundSpeed3D(x, y, z, sound3D.data(), source);
  }
  HeteroSoundSpeed3DExtensionToPML(x, y, z, sound3D.data());
  SetSoundSpeed2D(x, y, z, sound3D.data(), sound2D.data(), source);
  HeteroSoundSpeed2DExtensionToPML(x, y, sound2D.data());
  GenerateDeltaL(x, y, z, sound3D.data(), sound2D.data(), deltaL.data(), beta);

  // ---- one plane's CSR (FGMRES.cpp:140-141, 256): pattern with kwave 0, then a shift
  const int nnz = 5 * x.n * y.n - 2 * x.n - 2 * y.n;
  std::vector<int> ia(n2 + 1), ja(nnz), freqs(n2);
  std::vector<dtype> vals(nnz);
  zcsr A;
  A.ia = ia.data();
  A.ja = ja.data();
  A.values = vals.data();
  A.non_zeros = nnz;
  GenSparseMatrixOnline2DwithPMLFast(-1, x, y, &A, dtype(0.0), freqs.data());
  const dtype kwave(0.75e-3, 0.2e-3);
  GenSparseMatrixOnline2DwithPMLFast(3, x, y, &A, kwave, freqs.data());
  if (hz_ref_last_status() != HZ_OK) return 7;

  // ---- the 9-point planes (DIAG_PATTERN 9): the reference's weights, and sigma = 0.25
  //      (FGMRES.cpp:140, 144)
  const int nnz9 = 9 * x.n * y.n - 6 * x.n - 6 * y.n + 4;
  std::vector<int> ia9(n2 + 1), ja9(nnz9), fr9(n2), ia9s(n2 + 1), ja9s(nnz9), fr9s(n2);
  std::vector<dtype> v9(nnz9), v9s(nnz9);
  zcsr A9, A9s;
  A9.ia = ia9.data(); A9.ja = ja9.data(); A9.values = v9.data(); A9.non_zeros = nnz9;
  A9s.ia = ia9s.data(); A9s.ja = ja9s.data(); A9s.values = v9s.data(); A9s.non_zeros = nnz9;
  GenSparseMatrixOnline2DwithPMLand9Pts(-1, x, y, &A9, dtype(0.0), fr9.data());
  GenSparseMatrixOnline2DwithPMLand9Pts(3, x, y, &A9, kwave, fr9.data());
  GenSparseMatrixOnline2DwithPMLand9Points(-1, x, y, z, &A9s, kwave, fr9s.data(), 0.25);
  if (hz_ref_last_status() != HZ_OK) return 8;

  // ---- coefficient functions at every (half-)node of x and a set of 3D nodes
  std::vector<dtype> alph, b2, b3;
  for (int t = -1; t <= 2 * x.n; ++t) alph.push_back(alpha(x, 0.5 * t));
  for (int j = 0; j < y.n; ++j)
    for (int i = 0; i < x.n; ++i)
      for (int c = -2; c <= 2; ++c) b2.push_back(beta2D_pml(x, y, c, kwave, i, j));
  for (int k = 0; k < z.n; k += 3)
    for (int j = 0; j < y.n; j += 2)
      for (int i = 0; i < x.n; ++i)
        for (int c = -3; c <= 3; ++c) b3.push_back(beta3D(x, y, z, c, i, j, k));

  // ---- the hot path
  std::vector<dtype> f(N), u(N), g(N), lf(N), x_orig(N);
  if (!read_all(argv[13], f)) return 4;
  int* iparm = nullptr;
  int* perm = nullptr;
  size_t* pt = nullptr;
  zcsr** D2csr = nullptr;
  Solve3DSparseUsingFT_CSR(x, y, z, iparm, perm, pt, D2csr, f.data(), u.data(), beta, 1e-12);
  ApplyCoeffMatrixA_CSR(x, y, z, iparm, perm, pt, D2csr, f.data(), deltaL.data(), g.data(), beta, 1e-12);
  Multiply3DSparseUsingFT(x, y, z, iparm, perm, pt, D2csr, f.data(), lf.data(), 1e-12);
  std::vector<dtype> sol(N);
  double diff_sol = -1.0;
  FGMRES(x, y, z, m, rest, source, sol.data(), x_orig.data(), f.data(), 1e-12, diff_sol, beta);
  const int st = hz_ref_last_status();
  hz_ref_release();
  if (!write_all(out + "_precond.bin", u) || !write_all(out + "_applyA.bin", g) || !write_all(out + "_L0.bin", lf) ||
      !write_all(out + "_dl.bin", deltaL) || !write_all(out + "_fgmres.bin", sol) ||
      !write_all(out + "_sound3d.bin", sound3D) || !write_all(out + "_sound2d.bin", sound2D) ||
      !write_all(out + "_csr_ia.bin", ia) || !write_all(out + "_csr_ja.bin", ja) ||
      !write_all(out + "_csr_val.bin", vals) || !write_all(out + "_alpha.bin", alph) ||
      !write_all(out + "_beta2d.bin", b2) || !write_all(out + "_beta3d.bin", b3) ||
      !write_all(out + "_csr9_ia.bin", ia9) || !write_all(out + "_csr9_ja.bin", ja9) ||
      !write_all(out + "_csr9_val.bin", v9) || !write_all(out + "_csr9s_ia.bin", ia9s) ||
      !write_all(out + "_csr9s_ja.bin", ja9s) || !write_all(out + "_csr9s_val.bin", v9s))
    return 5;
  printf("diff_sol %.17g status %d\n", diff_sol, st);
  return 0;
}
