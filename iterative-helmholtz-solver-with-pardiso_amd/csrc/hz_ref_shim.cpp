// hz_ref_shim.cpp -- the reference's hot-path functions (include/hz_ref.h)
// implemented on the hz C ABI: host arrays in and out, device work inside.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/hz_ref.h"
#include "hz_internal.h"

namespace {

hz_ref_config g_cfg = {4.0, 1280.0, 1, 1e-4, 1e-8, 1000.0, HZ_PREC_FP64, nullptr};
std::string g_vfile;              // copy of g_cfg.velocity_file
std::vector<double> g_sound;      // hz_ref_set_sound_speed medium
thread_local int g_status = HZ_OK;
const double kPi = 3.141592653589793238462643;
double omega() { return 2.0 * kPi * g_cfg.nu; }

struct Cached {
  hz_problem* p = nullptr;
  hz_desc d;
  double2* buf = nullptr;  // 3 N device scratch
  size_t N = 0;
};
Cached g_cache;

void fail(const char* where, int rc) {
  g_status = rc;
  fprintf(stderr, "hz_ref %s: error %d: %s\n", where, rc, hz_last_error());
}

// a failed drop-in call leaves NaN in its output, so that no caller goes on
// with a stale or partial result
void poison(dtype* out, size_t n) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (size_t i = 0; i < n; ++i) out[i] = dtype(nan, nan);
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

// the medium of the drop-in solvers on this grid (precedence: the array of
// hz_ref_set_sound_speed, the configured OverThrust file, the built-in model)
bool medium(const size_m& x, const size_m& y, const size_m& z, const hz_desc& d, std::vector<double>& c) {
  const size_t N = (size_t)x.n * y.n * z.n;
  if (!g_sound.empty()) {
    if (g_sound.size() != N) {
      hz::set_error("hz_ref_set_sound_speed gave %zu values, the grid has %zu nodes", g_sound.size(), N);
      fail("medium", HZ_ERR_ARG);
      return false;
    }
    c = g_sound;
    return true;
  }
  if (!g_vfile.empty()) {
    c.assign(N, 0.0);
    const int rc = hz_velocity_from_file(g_vfile.c_str(), &d, c.data());
    if (rc) {
      fail("medium", rc);
      return false;
    }
    return true;
  }
  c = sound_speed(x, y, z);
  return true;
}

hz_problem* problem_for(const size_m& x, const size_m& y, const size_m& z, double beta) {
  g_status = HZ_OK;
  hz_desc d = make_desc(x, y, z, beta);
  if (g_cache.p && same(d, g_cache.d)) return g_cache.p;
  hz_ref_release();
  std::vector<double> c;
  if (!medium(x, y, z, d, c)) return nullptr;
  hz_problem* p = nullptr;
  int rc = hz_problem_create(&d, c.data(), &p);
  if (rc) { fail("hz_problem_create", rc); return nullptr; }
  rc = hz_precond_factor(p, g_cfg.prec);
  if (rc) { fail("hz_precond_factor", rc); hz_problem_destroy(p); return nullptr; }
  g_cache.p = p;
  g_cache.d = d;
  g_cache.N = (size_t)d.n[0] * d.n[1] * d.n[2];
  if (hipMalloc((void**)&g_cache.buf, 3 * g_cache.N * sizeof(double2)) != hipSuccess) {
    hz::set_error("device allocation of %zu bytes failed", 3 * g_cache.N * sizeof(double2));
    fail("problem_for", HZ_ERR_NOMEM);
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
  g_vfile = cfg->velocity_file ? cfg->velocity_file : "";
  g_cfg.velocity_file = nullptr;
  hz_ref_release();
  return HZ_OK;
}

extern "C" int hz_ref_set_sound_speed(const double* c, int64_t n) {
  if (c && n < 1) return HZ_ERR_ARG;
  if (c) {
    for (int64_t i = 0; i < n; ++i)
      if (!(c[i] > 0.0) || !std::isfinite(c[i])) {
        hz::set_error("sound speed %lld is %g", (long long)i, c[i]);
        return HZ_ERR_ARG;
      }
    g_sound.assign(c, c + n);
  } else {
    g_sound.clear();
  }
  hz_ref_release();
  return HZ_OK;
}

extern "C" int hz_ref_last_status(void) { return g_status; }

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
  if (!p) return poison(x_sol, (size_t)x.n * y.n * z.n);
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
  if (rc && rc != HZ_ERR_NOT_CONVERGED && rc != HZ_ERR_BREAKDOWN) {
    fail("FGMRES", rc);
    return poison(x_sol, N);
  }
  g_status = rc;
  down(du, x_sol, N);
  diff_sol = st.diff_sol;
}

void BCGSTAB(size_m x, size_m y, size_m z, int m, const point source, dtype* x_sol, dtype* x_orig, const dtype* f,
             double thresh, double& diff_sol, double beta_eq) {
  (void)source;
  (void)x_orig;
  (void)thresh;
  hz_problem* p = problem_for(x, y, z, beta_eq);
  if (!p) return poison(x_sol, (size_t)x.n * y.n * z.n);
  const size_t N = g_cache.N;
  double2* df = g_cache.buf;
  double2* du = g_cache.buf + N;
  up(f, df, N);
  hz_krylov_stats st;
  int rc = hz_bicgstab(p, g_cfg.prec, 1, (const hz_c128*)df, (hz_c128*)du, m, g_cfg.res_exit, 1, HZ_MASK_PHYS,
                       nullptr, 0, &st);
  if (rc && rc != HZ_ERR_NOT_CONVERGED) {
    fail("BCGSTAB", rc);
    return poison(x_sol, N);
  }
  g_status = rc;
  down(du, x_sol, N);
  diff_sol = 0.0;
}

void Solve3DSparseUsingFT_CSR(size_m x, size_m y, size_m z, int* iparm, int* perm, size_t* pt, zcsr**& D2csr,
                              const dtype* f, dtype* x_sol, double beta_eq, double thresh) {
  (void)iparm; (void)perm; (void)pt; (void)D2csr; (void)thresh;
  hz_problem* p = problem_for(x, y, z, beta_eq);
  if (!p) return poison(x_sol, (size_t)x.n * y.n * z.n);
  const size_t N = g_cache.N;
  up(f, g_cache.buf, N);
  int rc = hz_precond_apply(p, g_cfg.prec, 1, (const hz_c128*)g_cache.buf, (hz_c128*)(g_cache.buf + N));
  if (rc) { fail("Solve3DSparseUsingFT_CSR", rc); return poison(x_sol, N); }
  hz_problem_sync(p);
  down(g_cache.buf + N, x_sol, N);
}

void ApplyCoeffMatrixA_CSR(size_m x, size_m y, size_m z, int* iparm, int* perm, size_t* pt, zcsr**& D2csr,
                           const dtype* w, const dtype* deltaL, dtype* g, double beta_eq, double thresh) {
  // g = w - deltaL * L0^{-1} w with the caller's deltaL (functions.cpp:5567-5574).
  (void)iparm; (void)perm; (void)pt; (void)D2csr; (void)thresh;
  hz_problem* p = problem_for(x, y, z, beta_eq);
  if (!p) return poison(g, (size_t)x.n * y.n * z.n);
  const size_t N = g_cache.N;
  double2* dw = g_cache.buf;
  double2* dt = g_cache.buf + N;
  double2* dl = g_cache.buf + 2 * N;
  up(w, dw, N);
  up(deltaL, dl, N);
  int rc = hz_precond_apply(p, g_cfg.prec, 1, (const hz_c128*)dw, (hz_c128*)dt);
  if (rc) { fail("ApplyCoeffMatrixA_CSR", rc); return poison(g, N); }
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
  if (!p) return poison(f_sol, (size_t)x.n * y.n * z.n);
  const size_t N = g_cache.N;
  up(u, g_cache.buf, N);
  int rc = hz_L0_apply(p, 1, (const hz_c128*)g_cache.buf, (hz_c128*)(g_cache.buf + N));
  if (rc) { fail("Multiply3DSparseUsingFT", rc); return poison(f_sol, N); }
  hz_problem_sync(p);
  down(g_cache.buf + N, f_sol, N);
}

// ------------------------------------------------------------ set-up functions

dtype alpha(size_m xyz, double i) { return hz::stretch(xyz.n, xyz.pml_pts, omega(), g_cfg.pml_c, i); }

static dtype second_diff(const size_m& a, int i, int side) {
  const double h2 = a.h * a.h;
  const dtype ai = alpha(a, i);
  if (side == 0) return -ai * (alpha(a, i + 0.5) + alpha(a, i - 0.5)) / h2;
  return ai * alpha(a, side > 0 ? i + 0.5 : i - 0.5) / h2;
}

dtype beta2D_pml(size_m x, size_m y, int diag_case, dtype kwave_beta2, int i, int j) {
  switch (diag_case) {
    case 0: return second_diff(x, i, 0) + second_diff(y, j, 0) + kwave_beta2;
    case 1: return second_diff(x, i, 1);
    case -1: return second_diff(x, i, -1);
    case 2: return second_diff(y, j, 1);
    case -2: return second_diff(y, j, -1);
    default: return 0.0;
  }
}

dtype beta3D(size_m x, size_m y, size_m z, int diag_case, int i, int j, int k) {
  switch (diag_case) {
    case 0: return second_diff(x, i, 0) + second_diff(y, j, 0) + second_diff(z, k, 0);
    case 1: return second_diff(x, i, 1);
    case -1: return second_diff(x, i, -1);
    case 2: return second_diff(y, j, 1);
    case -2: return second_diff(y, j, -1);
    case 3: return second_diff(z, k, 1);
    case -3: return second_diff(z, k, -1);
    default: return 0.0;
  }
}

void GenSparseMatrixOnline2DwithPMLFast(int w, size_m x, size_m y, zcsr* Acsr, dtype kwave_beta2, int* freqs) {
  g_status = HZ_OK;
  const int nx = x.n, size = x.n * y.n;
  const int nnz = 5 * size - 2 * nx - 2 * y.n;
  if (!Acsr || !Acsr->values || !freqs || (w == -1 && (!Acsr->ia || !Acsr->ja)) || Acsr->non_zeros != nnz) {
    hz::set_error("plane CSR: non_zeros %d (pattern has %d) or a null array", Acsr ? Acsr->non_zeros : -1, nnz);
    fail("GenSparseMatrixOnline2DwithPMLFast", HZ_ERR_ARG);
    return;
  }
  if (w != -1) {
    for (int l = 0; l < size; ++l) Acsr->values[freqs[l]] += kwave_beta2;
    return;
  }
  // rows in order; per row the columns l - nx, l - 1, l, l + 1, l + nx
  int cnt = 0;
  for (int l = 0; l < size; ++l) {
    const int i = l % nx, j = l / nx;
    Acsr->ia[l] = cnt + 1;
    const int col[5] = {l - nx, l - 1, l, l + 1, l + nx};
    const int dc[5] = {-2, -1, 0, 1, 2};
    const bool has[5] = {l - nx >= 0, i > 0, true, i < nx - 1, l + nx < size};
    for (int e = 0; e < 5; ++e) {
      if (!has[e]) continue;
      Acsr->ja[cnt] = col[e] + 1;
      if (dc[e] == 0) freqs[l] = cnt;
      Acsr->values[cnt++] = beta2D_pml(x, y, dc[e], kwave_beta2, i, j);
    }
  }
  Acsr->ia[size] = cnt + 1;
}

// d2(a, i, side): the stretched second difference of axis a at node i: side 0
// the centre -a(i)(a(i+1/2) + a(i-1/2))/h^2, +1 / -1 the coupling a(i) a(i +- 1/2)/h^2
dtype beta2D_pml_9pts(size_m x, size_m y, int diag_case, dtype kwave_beta2, int i, int j, double sigma) {
  switch (diag_case) {
    case 0: return second_diff(y, j, 0) + (1.0 - 2.0 * sigma) * second_diff(x, i, 0) + kwave_beta2;
    case 1: return (1.0 - 2.0 * sigma) * second_diff(x, i, 1);
    case -1: return (1.0 - 2.0 * sigma) * second_diff(x, i, -1);
    case 3: return sigma * second_diff(x, i, 0) + second_diff(y, j, 1);
    case -3: return sigma * second_diff(x, i, 0) + second_diff(y, j, -1);
    case 2: case 4: return sigma * second_diff(x, i, 1);
    case -2: case -4: return sigma * second_diff(x, i, -1);
    default: return 0.0;
  }
}

dtype beta2D_pml_9pts_opt(size_m x, size_m y, int diag_case, dtype kwave_beta2, int i, int j, double a, double c,
                          double d) {
  const dtype corner = (1.0 - c - 4.0 * d) / 4.0 * kwave_beta2;
  switch (diag_case) {
    case 0:
      return a * (second_diff(y, j, 0) + second_diff(x, i, 0)) +
             (1.0 - a) * 0.5 * (second_diff(y, j, 0) + second_diff(x, i, 0)) + kwave_beta2 * c;
    case 1: return a * second_diff(x, i, 1) + kwave_beta2 * d;
    case -1: return a * second_diff(x, i, -1) + kwave_beta2 * d;
    case 3: return a * second_diff(y, j, 1) + kwave_beta2 * d;
    case -3: return a * second_diff(y, j, -1) + kwave_beta2 * d;
    case 2: return (1.0 - a) * 0.5 * second_diff(y, j, 1) + corner;
    case -2: return (1.0 - a) * 0.5 * second_diff(y, j, -1) + corner;
    case 4: return (1.0 - a) * 0.5 * second_diff(x, i, 1) + corner;
    case -4: return (1.0 - a) * 0.5 * second_diff(x, i, -1) + corner;
    default: return 0.0;
  }
}

// 9-point CSR with coefficient function coef(diag_case, i, j) (columns ascending,
// the neighbour classification of CheckDiag9Pts, functions.cpp:3373-3412)
template <class F>
static void csr9(int w, const size_m& x, const size_m& y, zcsr* Acsr, dtype kwave_beta2, int* freqs, F coef,
                 const char* where) {
  g_status = HZ_OK;
  const int nx = x.n, size = x.n * y.n;
  const int nnz = size + 2 * (size - y.n) + 2 * (size - nx) + 4 * (size - nx - y.n + 1);
  if (!Acsr || !Acsr->values || !freqs || (w == -1 && (!Acsr->ia || !Acsr->ja)) || Acsr->non_zeros != nnz) {
    hz::set_error("9-point plane CSR: non_zeros %d (pattern has %d) or a null array", Acsr ? Acsr->non_zeros : -1,
                  nnz);
    fail(where, HZ_ERR_ARG);
    return;
  }
  if (w != -1) {
    for (int l = 0; l < size; ++l) Acsr->values[freqs[l]] += kwave_beta2;
    return;
  }
  int cnt = 0;
  for (int l = 0; l < size; ++l) {
    const int i = l % nx, j = l / nx;
    const bool L = i > 0, R = i < nx - 1, D = l - nx >= 0, U = l + nx < size;
    // ascending columns: (x-1,y-1) (x,y-1) (x+1,y-1) (x-1,y) (x,y) (x+1,y) (x-1,y+1) (x,y+1) (x+1,y+1)
    const int col[9] = {l - nx - 1, l - nx, l - nx + 1, l - 1, l, l + 1, l + nx - 1, l + nx, l + nx + 1};
    const int dc[9] = {-4, -3, -2, -1, 0, 1, 2, 3, 4};
    const bool has[9] = {D && L, D, D && R, L, true, R, U && L, U, U && R};
    Acsr->ia[l] = cnt + 1;
    for (int e = 0; e < 9; ++e) {
      if (!has[e]) continue;
      Acsr->ja[cnt] = col[e] + 1;
      if (dc[e] == 0) freqs[l] = cnt;
      Acsr->values[cnt++] = coef(dc[e], i, j);
    }
  }
  Acsr->ia[size] = cnt + 1;
}

void GenSparseMatrixOnline2DwithPMLand9Pts(int w, size_m x, size_m y, zcsr* Acsr, dtype kwave_beta2, int* freqs) {
  // the reference's weights (functions.cpp:3553-3555)
  const double a = 1.0, c = 1.0, d = 0.0;
  csr9(w, x, y, Acsr, kwave_beta2, freqs,
       [&](int dc, int i, int j) { return beta2D_pml_9pts_opt(x, y, dc, kwave_beta2, i, j, a, c, d); },
       "GenSparseMatrixOnline2DwithPMLand9Pts");
}

void GenSparseMatrixOnline2DwithPMLand9Points(int w, size_m x, size_m y, size_m z, zcsr* Acsr, dtype kwave_beta2,
                                              int* freqs, double sigma) {
  (void)z;
  csr9(w, x, y, Acsr, kwave_beta2, freqs,
       [&](int dc, int i, int j) { return beta2D_pml_9pts(x, y, dc, kwave_beta2, i, j, sigma); },
       "GenSparseMatrixOnline2DwithPMLand9Points");
}

void SetSoundSpeed3D(size_m x, size_m y, size_m z, dtype* sound3D, point source) {
  (void)source;
  const int nx = x.n, ny = y.n, n2 = x.n * y.n;
  const double c3 = 600.0, c2 = (1500.0 - c3) / y.l, c1 = (1800.0 - c3 - c2 * y.l) / x.l;
  for (int k = 0; k < z.n; ++k)
    for (int j = y.pml_pts; j < ny - y.pml_pts; ++j)
      for (int i = x.pml_pts; i < nx - x.pml_pts; ++i) {
        const double px = ((double)i + 1 - x.pml_pts) * x.h, py = ((double)j + 1 - y.pml_pts) * y.h;
        sound3D[(size_t)k * n2 + (size_t)j * nx + i] = g_cfg.homo ? dtype(g_cfg.c_z) : dtype(c1 * px + c2 * py + c3);
      }
}

void SetSoundSpeed3DFromFile(size_m x, size_m y, size_m z, dtype* sound3D, point source) {
  (void)source;
  g_status = HZ_OK;
  hz_desc d = make_desc(x, y, z, 0.0);
  std::string path = g_vfile;
  if (path.empty()) {
    char name[64];
    hz_velocity_file_name(x.h, name, sizeof(name));
    path = name;
  }
  const size_t N = (size_t)x.n * y.n * z.n;
  std::vector<double> c(N);
  const int rc = hz_velocity_from_file(path.c_str(), &d, c.data());
  if (rc) {
    fail("SetSoundSpeed3DFromFile", rc);
    return poison(sound3D, N);
  }
  const int nx = x.n, ny = y.n, n2 = x.n * y.n;
  for (int k = z.spg_pts; k < z.n - z.spg_pts; ++k)
    for (int j = y.pml_pts; j < ny - y.pml_pts; ++j)
      for (int i = x.pml_pts; i < nx - x.pml_pts; ++i) {
        const size_t l = (size_t)k * n2 + (size_t)j * nx + i;
        sound3D[l] = c[l];
      }
}

// x/y layers of one plane copy the nearest physical node: rows first, then columns
static void extend_plane(const size_m& x, const size_m& y, dtype* p) {
  const int nx = x.n, ny = y.n, px = x.pml_pts, py = y.pml_pts;
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < px; ++i) p[i + (size_t)nx * j] = p[px + (size_t)nx * j];
    for (int i = nx - px; i < nx; ++i) p[i + (size_t)nx * j] = p[nx - px - 1 + (size_t)nx * j];
  }
  for (int i = 0; i < nx; ++i) {
    for (int j = 0; j < py; ++j) p[i + (size_t)nx * j] = p[i + (size_t)nx * py];
    for (int j = ny - py; j < ny; ++j) p[i + (size_t)nx * j] = p[i + (size_t)nx * (ny - py - 1)];
  }
}

void HeteroSoundSpeed3DExtensionToPML(size_m x, size_m y, size_m z, dtype* sound3D) {
  for (int k = 0; k < z.n; ++k) extend_plane(x, y, sound3D + (size_t)k * x.n * y.n);
}

void HeteroSoundSpeed2DExtensionToPML(size_m x, size_m y, dtype* sound2D) { extend_plane(x, y, sound2D); }

void SetSoundSpeed2D(size_m x, size_m y, size_m z, dtype* sound3D, dtype* sound2D, point source) {
  (void)source;
  const int nx = x.n, ny = y.n, n2 = x.n * y.n;
  const int k0 = z.spg_pts, k1 = z.n - z.spg_pts;
  // one accumulator for every column, divided by the z count after each
  // column and never cleared (functions.cpp:1716-1727)
  dtype acc = 0.0;
  for (int j = y.pml_pts; j < ny - y.pml_pts; ++j)
    for (int i = x.pml_pts; i < nx - x.pml_pts; ++i) {
      for (int k = k0; k < k1; ++k) {
        const dtype v = sound3D[(size_t)k * n2 + (size_t)j * nx + i];
        acc += 1.0 / (v * v);
      }
      acc /= (double)(k1 - k0);
      sound2D[(size_t)j * nx + i] = std::sqrt(1.0 / acc);
    }
}

void GenerateDeltaL(size_m x, size_m y, size_m z, dtype* sound3D, dtype* sound2D, dtype* deltaL, double beta_eq) {
  const int nx = x.n, ny = y.n, nz = z.n, spg = z.spg_pts;
  const size_t n2 = (size_t)nx * ny;
  const double w2 = omega() * omega();
  // band spg .. nz - spg (both ends), clipped to the grid
  const int k_last = nz - spg < nz - 1 ? nz - spg : nz - 1;
  for (int k = spg; k <= k_last; ++k)
    for (size_t ij = 0; ij < n2; ++ij) {
      const dtype b = sound2D[ij], v = sound3D[ij + n2 * k];
      deltaL[ij + n2 * k] = w2 * (dtype(1.0, beta_eq) / (b * b) - 1.0 / (v * v));
    }
  const double sig = 0.3, z0 = spg * z.h, zN = (nz - spg) * z.h, L0 = z0, Ln = z.l - zN;
  for (int k = 0; k < spg; ++k) {
    const double t = (k * z.h - z0) / L0 / sig;
    const double g = std::exp(-t * t);
    for (size_t ij = 0; ij < n2; ++ij) deltaL[ij + n2 * k] = deltaL[ij + n2 * spg] * g;
  }
  for (int k = nz - spg + 1; k < nz; ++k) {
    const double t = (k * z.h - zN) / Ln / sig;
    const double g = std::exp(-t * t);
    for (size_t ij = 0; ij < n2; ++ij) deltaL[ij + n2 * k] = deltaL[ij + n2 * (nz - spg)] * g;
  }
}
