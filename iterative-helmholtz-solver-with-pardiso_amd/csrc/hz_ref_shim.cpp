// hz_ref_shim.cpp -- the reference's hot-path functions (include/hz_ref.h)
// implemented on the hz C ABI: host arrays in and out, device work inside.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/hz_ref.h"
#include "hz_internal.h"

namespace {

hz_ref_config g_cfg = {4.0, 1280.0, 1, 1e-4, 1e-8, 1000.0, HZ_PREC_FP64};

struct Cached {
  hz_problem* p = nullptr;
  hz_desc d;
  double2* buf = nullptr;  // 3 N device scratch
  size_t N = 0;
};
Cached g_cache;

void fail(const char* where, int rc) {
  fprintf(stderr, "hz_ref %s: error %d: %s\n", where, rc, hz_last_error());
}

// Grid of the reference (main.cpp:122-211): x/y PML, z sponge, spectral z.
hz_desc make_desc(const size_m& x, const size_m& y, const size_m& z, double beta) {
  hz_desc d;
  memset(&d, 0, sizeof(d));
  const size_m* ax[3] = {&x, &y, &z};
  for (int a = 0; a < 3; ++a) {
    d.n[a] = ax[a]->n;
    d.h[a] = ax[a]->h;
    d.l[a] = ax[a]->l;
  }
  d.pml[0] = x.pml_pts;
  d.pml[1] = y.pml_pts;
  d.pml[2] = 0;
  d.spg[2] = z.spg_pts;
  d.omega = 2.0 * 3.141592653589793238462643 * g_cfg.nu;
  d.beta = beta;
  d.pml_c = g_cfg.pml_c;
  d.mode = HZ_MODE_REF;
  d.kz_symbol = HZ_KZ_REFERENCE;
  d.c0_rule = HZ_C0_REFERENCE;
  d.c_bg = g_cfg.homo ? g_cfg.c_z : 0.0;
  return d;
}

// SetSoundSpeed3D + HeteroSoundSpeed3DExtensionToPML (functions.cpp:1362-1430,
// 2675-2716).
std::vector<double> sound_speed(const size_m& x, const size_m& y, const size_m& z) {
  const int nx = x.n, ny = y.n, nz = z.n;
  std::vector<double> c((size_t)nx * ny * nz, g_cfg.c_z);
  if (g_cfg.homo) return c;
  const double c3 = 600.0;
  const double c2 = (1500.0 - c3) / y.l;
  const double c1 = (1800.0 - c3 - c2 * y.l) / x.l;
  const int px = x.pml_pts, py = y.pml_pts;
  std::vector<double> c2d((size_t)nx * ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int ii = i < px ? px : (i >= nx - px ? nx - px - 1 : i);
      const int jj = j < py ? py : (j >= ny - py ? ny - py - 1 : j);
      c2d[i + (size_t)nx * j] = c1 * ((double)ii + 1 - px) * x.h + c2 * ((double)jj + 1 - py) * y.h + c3;
    }
  for (int k = 0; k < nz; ++k) memcpy(&c[(size_t)nx * ny * k], c2d.data(), c2d.size() * sizeof(double));
  return c;
}

bool same(const hz_desc& a, const hz_desc& b) { return memcmp(&a, &b, sizeof(hz_desc)) == 0; }

hz_problem* problem_for(const size_m& x, const size_m& y, const size_m& z, double beta) {
  hz_desc d = make_desc(x, y, z, beta);
  if (g_cache.p && same(d, g_cache.d)) return g_cache.p;
  hz_ref_release();
  std::vector<double> c = sound_speed(x, y, z);
  hz_problem* p = nullptr;
  int rc = hz_problem_create(&d, c.data(), &p);
  if (rc) { fail("hz_problem_create", rc); return nullptr; }
  rc = hz_precond_factor(p, g_cfg.prec);
  if (rc) { fail("hz_precond_factor", rc); hz_problem_destroy(p); return nullptr; }
  g_cache.p = p;
  g_cache.d = d;
  g_cache.N = (size_t)d.n[0] * d.n[1] * d.n[2];
  if (hipMalloc((void**)&g_cache.buf, 3 * g_cache.N * sizeof(double2)) != hipSuccess) {
    fprintf(stderr, "hz_ref: device allocation failed\n");
    hz_ref_release();
    return nullptr;
  }
  return p;
}

void up(const dtype* h, double2* d, size_t n) { (void)hipMemcpy(d, h, n * sizeof(double2), hipMemcpyHostToDevice); }
void down(const double2* d, dtype* h, size_t n) { (void)hipMemcpy(h, d, n * sizeof(double2), hipMemcpyDeviceToHost); }

}  // namespace

extern "C" int hz_ref_configure(const hz_ref_config* cfg) {
  if (!cfg) return HZ_ERR_ARG;
  g_cfg = *cfg;
  hz_ref_release();
  return HZ_OK;
}

extern "C" void hz_ref_release(void) {
  if (g_cache.buf) (void)hipFree(g_cache.buf);
  if (g_cache.p) hz_problem_destroy(g_cache.p);
  g_cache = Cached();
}

void FGMRES(size_m x, size_m y, size_m z, int m, int rest, const point source, dtype* x_sol, dtype* x_orig,
            const dtype* f, double thresh, double& diff_sol, double beta_eq) {
  (void)source;
  (void)x_orig;
  hz_problem* p = problem_for(x, y, z, beta_eq);
  if (!p) return;
  const size_t N = g_cache.N;
  double2* df = g_cache.buf;
  double2* du = g_cache.buf + N;
  up(f, df, N);
  hz_krylov_opts o;
  memset(&o, 0, sizeof(o));
  o.driver = HZ_DRIVER_GMRES_REF;
  o.m = m;
  o.max_restarts = rest;
  o.ortho = HZ_ORTHO_REF_CONJ;
  o.mask = HZ_MASK_PHYS;
  o.monitor = 1;
  o.tol = g_cfg.rel_res_exit;
  o.breakdown = thresh;
  o.prec = g_cfg.prec;
  hz_krylov_stats st;
  int rc = hz_fgmres(p, 1, (const hz_c128*)df, (hz_c128*)du, &o, &st);
  if (rc && rc != HZ_ERR_NOT_CONVERGED && rc != HZ_ERR_BREAKDOWN) fail("FGMRES", rc);
  down(du, x_sol, N);
  diff_sol = st.diff_sol;
}

void BCGSTAB(size_m x, size_m y, size_m z, int m, const point source, dtype* x_sol, dtype* x_orig, const dtype* f,
             double thresh, double& diff_sol, double beta_eq) {
  (void)source;
  (void)x_orig;
  (void)thresh;
  hz_problem* p = problem_for(x, y, z, beta_eq);
  if (!p) return;
  const size_t N = g_cache.N;
  double2* df = g_cache.buf;
  double2* du = g_cache.buf + N;
  up(f, df, N);
  hz_krylov_stats st;
  int rc = hz_bicgstab(p, g_cfg.prec, 1, (const hz_c128*)df, (hz_c128*)du, m, g_cfg.res_exit, 1, HZ_MASK_PHYS,
                       nullptr, 0, &st);
  if (rc && rc != HZ_ERR_NOT_CONVERGED) fail("BCGSTAB", rc);
  down(du, x_sol, N);
  diff_sol = 0.0;
}

void Solve3DSparseUsingFT_CSR(size_m x, size_m y, size_m z, int* iparm, int* perm, size_t* pt, zcsr**& D2csr,
                              const dtype* f, dtype* x_sol, double beta_eq, double thresh) {
  (void)iparm; (void)perm; (void)pt; (void)D2csr; (void)thresh;
  hz_problem* p = problem_for(x, y, z, beta_eq);
  if (!p) return;
  const size_t N = g_cache.N;
  up(f, g_cache.buf, N);
  int rc = hz_precond_apply(p, g_cfg.prec, 1, (const hz_c128*)g_cache.buf, (hz_c128*)(g_cache.buf + N));
  if (rc) { fail("Solve3DSparseUsingFT_CSR", rc); return; }
  hz_problem_sync(p);
  down(g_cache.buf + N, x_sol, N);
}

void ApplyCoeffMatrixA_CSR(size_m x, size_m y, size_m z, int* iparm, int* perm, size_t* pt, zcsr**& D2csr,
                           const dtype* w, const dtype* deltaL, dtype* g, double beta_eq, double thresh) {
  // g = w - deltaL * L0^{-1} w with the caller's deltaL (functions.cpp:5567-5574).
  (void)iparm; (void)perm; (void)pt; (void)D2csr; (void)thresh;
  hz_problem* p = problem_for(x, y, z, beta_eq);
  if (!p) return;
  const size_t N = g_cache.N;
  double2* dw = g_cache.buf;
  double2* dt = g_cache.buf + N;
  double2* dl = g_cache.buf + 2 * N;
  up(w, dw, N);
  up(deltaL, dl, N);
  int rc = hz_precond_apply(p, g_cfg.prec, 1, (const hz_c128*)dw, (hz_c128*)dt);
  if (rc) { fail("ApplyCoeffMatrixA_CSR", rc); return; }
  hz::launch_w_minus_dl((int64_t)N, 1, dw, dl, dt, dt, p->stream);
  hz_problem_sync(p);
  down(dt, g, N);
}

void Multiply3DSparseUsingFT(size_m x, size_m y, size_m z, int* iparm, int* perm, size_t* pt, zcsr**& D2csr,
                             const dtype* u, dtype* f_sol, double thresh) {
  (void)iparm; (void)perm; (void)pt; (void)D2csr; (void)thresh;
  // The plane matrices of Multiply3DSparseUsingFT carry the shift of the last
  // factorisation (FGMRES.cpp:233): reuse the cached problem's shift.
  hz_problem* p = problem_for(x, y, z, g_cache.p ? g_cache.d.beta : 0.5);
  if (!p) return;
  const size_t N = g_cache.N;
  up(u, g_cache.buf, N);
  int rc = hz_L0_apply(p, 1, (const hz_c128*)g_cache.buf, (hz_c128*)(g_cache.buf + N));
  if (rc) { fail("Multiply3DSparseUsingFT", rc); return; }
  hz_problem_sync(p);
  down(g_cache.buf + N, f_sol, N);
}
