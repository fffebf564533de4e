// ref_shim_driver.cpp -- a caller written against the reference's own
// signatures (include/hz_ref.h), the way the reference's main() calls them
// (main.cpp:122-211 grid set-up, main.cpp:470 FGMRES call), linked against
// libhz.so instead of the MKL build.  Test program: reads the right-hand side
// from a file, calls Solve3DSparseUsingFT_CSR, ApplyCoeffMatrixA_CSR,
// Multiply3DSparseUsingFT and FGMRES, and writes each result to a file.
//
//   ref_shim_driver n_nopml nz_nopml pml spg length nu beta homo c_z m rest rel_res in.bin out_prefix
#include <cstdio>
#include <cstdlib>
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

static bool read_all(const char* path, std::vector<dtype>& v) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  const size_t got = fread(v.data(), sizeof(dtype), v.size(), f);
  fclose(f);
  return got == v.size();
}

static bool write_all(const std::string& path, const std::vector<dtype>& v) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  const size_t put = fwrite(v.data(), sizeof(dtype), v.size(), f);
  fclose(f);
  return put == v.size();
}

int main(int argc, char** argv) {
  if (argc != 15) {
    fprintf(stderr, "usage: %s n_nopml nz_nopml pml spg length nu beta homo c_z m rest rel_res in.bin out_prefix\n",
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

  size_m x = axis(n_nopml, pml, length, true), y = axis(n_nopml, pml, length, true);
  size_m z = axis(nz_nopml, spg, length, false);
  const size_t N = (size_t)x.n * y.n * z.n;
  hz_ref_config cfg = {nu, c_z, homo, rel_res, 1e-8, 1000.0, HZ_PREC_FP64};
  if (hz_ref_configure(&cfg) != HZ_OK) return 3;

  std::vector<dtype> f(N), u(N), g(N), dl(N), lf(N), x_orig(N);
  if (!read_all(argv[13], f)) return 4;
  for (size_t i = 0; i < N; ++i) dl[i] = dtype(1e-3 * (double)(i % 7), -2e-3 * (double)(i % 5));
  int* iparm = nullptr;
  int* perm = nullptr;
  size_t* pt = nullptr;
  zcsr** D2csr = nullptr;
  const point source = {x.l / 2, y.l / 2, z.l / 2};

  Solve3DSparseUsingFT_CSR(x, y, z, iparm, perm, pt, D2csr, f.data(), u.data(), beta, 1e-12);
  ApplyCoeffMatrixA_CSR(x, y, z, iparm, perm, pt, D2csr, f.data(), dl.data(), g.data(), beta, 1e-12);
  Multiply3DSparseUsingFT(x, y, z, iparm, perm, pt, D2csr, f.data(), lf.data(), 1e-12);
  std::vector<dtype> sol(N);
  double diff_sol = -1.0;
  FGMRES(x, y, z, m, rest, source, sol.data(), x_orig.data(), f.data(), 1e-12, diff_sol, beta);
  hz_ref_release();
  if (!write_all(out + "_precond.bin", u) || !write_all(out + "_applyA.bin", g) || !write_all(out + "_L0.bin", lf) ||
      !write_all(out + "_dl.bin", dl) || !write_all(out + "_fgmres.bin", sol))
    return 5;
  printf("diff_sol %.17g\n", diff_sol);
  return 0;
}
