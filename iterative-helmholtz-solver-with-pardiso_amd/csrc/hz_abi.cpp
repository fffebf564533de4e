// hz_abi.cpp -- the extern "C" entry points of include/hz.h and the Krylov
// drivers that sequence the device kernels.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "hz_internal.h"

namespace hz {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

void DevBuf::alloc(size_t b) {
  release();
  if (b == 0) return;
  hipError_t e = hipMalloc(&ptr, b);
  if (e != hipSuccess) {
    ptr = nullptr;
    set_error("hipMalloc(%zu bytes): %s", b, hipGetErrorString(e));
    throw Error{HZ_ERR_NOMEM};
  }
  bytes = b;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

void PinBuf::reserve(size_t b) {
  if (b <= bytes) return;
  release();
  hipError_t e = hipHostMalloc(&ptr, b, hipHostMallocDefault);
  if (e != hipSuccess) {
    ptr = nullptr;
    set_error("hipHostMalloc(%zu bytes): %s", b, hipGetErrorString(e));
    throw Error{HZ_ERR_NOMEM};
  }
  bytes = b;
}

void PinBuf::release() {
  if (ptr) (void)hipHostFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

typedef std::complex<double> zc;

static inline double2 d2(zc v) { return make_double2(v.real(), v.imag()); }
static inline zc zz(double2 v) { return zc(v.x, v.y); }

void grow(DevBuf& b, size_t bytes) {
  if (b.bytes < bytes) b.alloc(bytes);
}

void ensure_work(hz_problem* p, int nrhs) {
  const size_t b = (size_t)p->N * nrhs * sizeof(double2);
  grow(p->w_a, b);
  grow(p->w_b, b);
  grow(p->w_c, b);
  grow(p->scal, 4096 * sizeof(double2));
}

static void sync(hz_problem* p) { HZ_HIP(hipStreamSynchronize(p->stream)); }

// ------------------------------------------------------------ operators
// out = L0^{-1} in (or its adjoint), with `work` (nrhs N complex of the
// factor precision) holding the z-transformed planes.  epi: REF
// right-preconditioned operator out = in - deltaL * L0^{-1} in fused into the
// backward transform.
void precond_apply_ws(hz_problem* p, int prec, int nrhs, const double2* in, double2* out, bool adjoint, bool epi,
                      void* work) {
  HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
  HZ_CHECK(p->fac[prec].ready, HZ_ERR_STATE, "hz_precond_factor(prec=%d) has not run", prec);
  const double2* dL = p->d_dL.as<double2>();
  const int wp = prec == HZ_PREC_FP64 ? 0 : 1;  // work precision: double2 or float2
  launch_fft_z(p, prec, -1, nrhs, in, 0, work, wp, 0, nullptr, nullptr, p->stream);
  launch_plane_solve(p, prec, nrhs, work, adjoint, p->stream);
  launch_fft_z(p, prec, +1, nrhs, work, wp, out, 0, epi ? 1 : 0, in, dL, p->stream);
}

void precond_apply(hz_problem* p, int prec, int nrhs, const double2* in, double2* out, bool adjoint, bool epi) {
  HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
  HZ_CHECK(p->fac[prec].ready, HZ_ERR_STATE, "hz_precond_factor(prec=%d) has not run", prec);
  ensure_work(p, nrhs);
  precond_apply_ws(p, prec, nrhs, in, out, adjoint, epi, p->w_a.ptr);
}

static void L0_apply(hz_problem* p, int nrhs, const double2* in, double2* out, bool adjoint) {
  ensure_work(p, nrhs);
  launch_fft_z(p, HZ_PREC_FP64, -1, nrhs, in, 0, p->w_a.ptr, 0, 0, nullptr, nullptr, p->stream);
  launch_plane_matvec(p, nrhs, p->w_a.as<double2>(), p->w_b.as<double2>(), adjoint, p->stream);
  launch_fft_z(p, HZ_PREC_FP64, +1, nrhs, p->w_b.ptr, 0, out, 0, 0, nullptr, nullptr, p->stream);
}

// out = A in (true operator)
void op_apply(hz_problem* p, int nrhs, const double2* in, double2* out, bool adjoint) {
  if (p->d.mode == HZ_MODE_S7) {
    launch_stencil(p, nrhs, in, out, adjoint, nullptr, p->stream);
    return;
  }
  L0_apply(p, nrhs, in, out, adjoint);
  // out -= dL .* in  (conj(dL) for the adjoint)
  const double2* dL = adjoint ? p->d_dLc.as<double2>() : p->d_dL.as<double2>();
  launch_w_minus_dl(p->N, nrhs, out, dL, in, out, p->stream);
}

// r = f - A u
void residual(hz_problem* p, int nrhs, const double2* f, const double2* u, double2* r, bool adjoint) {
  if (p->d.mode == HZ_MODE_S7) {
    launch_stencil(p, nrhs, u, r, adjoint, f, p->stream);
    return;
  }
  op_apply(p, nrhs, u, r, adjoint);
  launch_sub((int64_t)p->N * nrhs, f, r, r, p->stream);
}

// out = A L0^{-1} in
static void op_pre_apply(hz_problem* p, int prec, int nrhs, const double2* in, double2* out) {
  if (p->d.mode == HZ_MODE_REF) {
    precond_apply(p, prec, nrhs, in, out, false, true);
    return;
  }
  ensure_work(p, nrhs);
  precond_apply(p, prec, nrhs, in, p->w_c.as<double2>(), false, false);
  launch_stencil(p, nrhs, p->w_c.as<double2>(), out, false, nullptr, p->stream);
}

// ------------------------------------------------------------ reductions
double2* dscal(hz_problem* p, int i) { return p->scal.as<double2>() + i; }

static zc dot_host(hz_problem* p, const double2* x, const double2* y, int mask = HZ_MASK_FULL) {
  launch_dot(p, p->N, x, y, mask, dscal(p, 0), p->stream);
  double2 h;
  HZ_HIP(hipMemcpyAsync(&h, dscal(p, 0), sizeof(double2), hipMemcpyDeviceToHost, p->stream));
  sync(p);
  return zz(h);
}

static double nrm_host(hz_problem* p, const double2* x, int mask = HZ_MASK_FULL) {
  launch_dot(p, p->N, x, nullptr, mask, dscal(p, 0), p->stream);
  double2 h;
  HZ_HIP(hipMemcpyAsync(&h, dscal(p, 0), sizeof(double2), hipMemcpyDeviceToHost, p->stream));
  sync(p);
  return h.x;
}

// Householder least squares min ||b - H y|| for an m x n (m >= n) column-major
// H with leading dimension ld (the zgels step of FGMRES.cpp:563-569).
static void lsq(int m, int n, std::vector<zc>& H, int ld, std::vector<zc>& b) {
  for (int j = 0; j < n; ++j) {
    double s2 = 0.0;
    for (int i = j; i < m; ++i) s2 += std::norm(H[i + ld * j]);
    double nr = std::sqrt(s2);
    if (nr == 0.0) continue;
    zc x0 = H[j + ld * j];
    zc ph = std::abs(x0) == 0.0 ? zc(1.0, 0.0) : x0 / std::abs(x0);
    std::vector<zc> v(m - j);
    for (int i = j; i < m; ++i) v[i - j] = H[i + ld * j];
    v[0] += ph * nr;  // v = x - alpha e1 with alpha = -ph * nr
    double vn = 0.0;
    for (auto& e : v) vn += std::norm(e);
    if (vn == 0.0) continue;
    for (int c = j; c < n; ++c) {
      zc s(0, 0);
      for (int i = j; i < m; ++i) s += std::conj(v[i - j]) * H[i + ld * c];
      s *= 2.0 / vn;
      for (int i = j; i < m; ++i) H[i + ld * c] -= s * v[i - j];
    }
    zc s(0, 0);
    for (int i = j; i < m; ++i) s += std::conj(v[i - j]) * b[i];
    s *= 2.0 / vn;
    for (int i = j; i < m; ++i) b[i] -= s * v[i - j];
  }
  for (int j = n - 1; j >= 0; --j) {
    zc s = b[j];
    for (int c = j + 1; c < n; ++c) s -= H[j + ld * c] * b[c];
    b[j] = s / H[j + ld * j];
  }
}

struct HistW {
  double* buf;
  int cap;
  int n = 0;
  void push(double v) {
    if (buf && n < cap) buf[n] = v;
    ++n;
  }
};

// The reference's restarted GMRES on (I - deltaL L0^{-1}) w = f
// (FGMRES.cpp:458-661), u = L0^{-1} w at the end.
static void gmres_ref_one(hz_problem* p, const double2* f, double2* u, const hz_krylov_opts& o,
                          hz_krylov_stats& st) {
  const int m = o.m;
  const int64_t N = p->N;
  ensure_krylov(p, m, 1);
  double2* V = p->kV.as<double2>();
  double2* w = p->kw.as<double2>();
  double2* x0 = p->kx.as<double2>();
  double2* xi = p->kaux.as<double2>();
  double2* xs = p->kZ.as<double2>();   // monitor u_k (slot 0), u_{k-1} (slot 1) when m >= 2
  double2* dH = dscal(p, 64);
  const int ldh = m + 1;
  double2* dY = dH + (size_t)ldh * m;
  HistW hist{o.hist, o.hist_cap};
  memset(&st, 0, sizeof(st));
  auto t0 = std::chrono::steady_clock::now();
  launch_fill(N, make_double2(0, 0), xi, p->stream);
  launch_fill(N, make_double2(0, 0), x0, p->stream);
  const double norm_f = (nrm_host(p, f, HZ_MASK_FULL));
  std::vector<zc> H((size_t)ldh * m, zc(0, 0)), Hg, eb(m + 1);
  std::vector<double2> hcol(m + 2), yd(m);
  double Res = 0.0, RelRes = 0.0, diff = 0.0;
  bool finish = false, have_prev = false;
  int total = 0;
  for (int rs = 0; rs < o.max_restarts && !finish; ++rs) {
    st.restarts++;
    HZ_HIP(hipMemcpyAsync(x0, xi, N * sizeof(double2), hipMemcpyDeviceToDevice, p->stream));
    op_pre_apply(p, o.prec, 1, x0, w);
    launch_sub(N, f, w, w, p->stream);  // r0 = f - A x0
    launch_dot(p, N, w, nullptr, HZ_MASK_FULL, dscal(p, 1), p->stream);
    double2 hb;
    HZ_HIP(hipMemcpyAsync(&hb, dscal(p, 1), sizeof(double2), hipMemcpyDeviceToHost, p->stream));
    sync(p);
    const double beta = hb.x;
    launch_scale_inv_dev(N, dscal(p, 1), w, V, p->stream);
    for (int j = 0; j < m; ++j) {
      op_pre_apply(p, o.prec, 1, V + (size_t)N * j, w);
      total++;
      for (int i = 0; i <= j; ++i) {
        double2* hij = dH + i + (size_t)ldh * j;
        if (o.ortho == HZ_ORTHO_REF_CONJ)
          launch_dot(p, N, w, V + (size_t)N * i, HZ_MASK_FULL, hij, p->stream);
        else
          launch_dot(p, N, V + (size_t)N * i, w, HZ_MASK_FULL, hij, p->stream);
        launch_axpy_dev(N, hij, true, false, V + (size_t)N * i, w, p->stream);
      }
      double2* hn_d = dH + (j + 1) + (size_t)ldh * j;
      launch_dot(p, N, w, nullptr, HZ_MASK_FULL, hn_d, p->stream);
      HZ_HIP(hipMemcpyAsync(hcol.data(), dH + (size_t)ldh * j, (j + 2) * sizeof(double2), hipMemcpyDeviceToHost,
                            p->stream));
      sync(p);
      for (int i = 0; i <= j + 1; ++i) H[i + (size_t)ldh * j] = zz(hcol[i]);
      const double hn = hcol[j + 1].x;
      if (std::fabs(hn) < o.breakdown) {
        st.breakdown = 1;
        break;
      }
      launch_scale_inv_dev(N, hn_d, w, V + (size_t)N * (j + 1), p->stream);
      std::fill(eb.begin(), eb.end(), zc(0, 0));
      eb[0] = beta;
      Hg = H;
      lsq(j + 2, j + 1, Hg, ldh, eb);
      for (int i = 0; i <= j; ++i) yd[i] = d2(eb[i]);
      HZ_HIP(hipMemcpyAsync(dY, yd.data(), (j + 1) * sizeof(double2), hipMemcpyHostToDevice, p->stream));
      HZ_HIP(hipMemcpyAsync(x0, xi, N * sizeof(double2), hipMemcpyDeviceToDevice, p->stream));
      launch_lincomb(N, j + 1, dY, V, N, x0, true, p->stream);
      op_pre_apply(p, o.prec, 1, x0, w);
      launch_sub(N, w, f, w, p->stream);
      Res = (nrm_host(p, w, o.mask));
      RelRes = Res / norm_f;
      hist.push(RelRes);
      if (o.monitor && m >= 2) {
        double2* cur = xs + (size_t)N * (total % 2);
        double2* prev = xs + (size_t)N * ((total + 1) % 2);
        precond_apply(p, o.prec, 1, x0, cur, false, false);
        if (have_prev) {
          launch_sub(N, cur, prev, w, p->stream);
          const double a = (nrm_host(p, w, HZ_MASK_PHYS));
          const double b = (nrm_host(p, prev, HZ_MASK_PHYS));
          diff = a / b;
        }
        have_prev = true;
      }
      if (!std::isfinite(RelRes)) {
        set_error("non-finite residual in GMRES");
        throw Error{HZ_ERR_NONFINITE};
      }
      if (RelRes < o.tol) {
        finish = true;
        break;
      }
    }
    HZ_HIP(hipMemcpyAsync(xi, x0, N * sizeof(double2), hipMemcpyDeviceToDevice, p->stream));
  }
  precond_apply(p, o.prec, 1, x0, u, false, false);
  sync(p);
  st.iters = total;
  st.converged = finish;
  st.rel_res = RelRes;
  st.res = Res;
  st.diff_sol = diff;
  st.n_hist = hist.n;
  st.t_solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// BiCGStab of BcGSTAB.cpp:397-568 (w-space recurrences on A L0^{-1}, x in
// u-space).  The residual uses the true operator A x - f directly.
static void bicgstab_one(hz_problem* p, int prec, const double2* f, double2* x, int maxit, double tol, int tol_abs,
                         int mask, HistW& hist, hz_krylov_stats& st) {
  const int64_t N = p->N;
  ensure_krylov(p, 8, 1);
  double2* V = p->kV.as<double2>();
  double2 *p0 = V, *s0 = V + N, *yn = V + 2 * N, *zn = V + 3 * N, *Ayn = V + 4 * N, *Azn = V + 5 * N,
          *r0 = V + 6 * N, *rst = V + 7 * N, *w = p->kw.as<double2>();
  memset(&st, 0, sizeof(st));
  st.restarts = 1;
  auto t0 = std::chrono::steady_clock::now();
  const double norm_f = (nrm_host(p, f, HZ_MASK_FULL));
  launch_fill(N, make_double2(0, 0), x, p->stream);
  HZ_HIP(hipMemcpyAsync(r0, f, N * sizeof(double2), hipMemcpyDeviceToDevice, p->stream));  // f - A_pre(0)
  HZ_HIP(hipMemcpyAsync(p0, r0, N * sizeof(double2), hipMemcpyDeviceToDevice, p->stream));
  launch_fill(N, make_double2(0, 0), rst, p->stream);
  launch_axpby(N, make_double2(2.0, 0), r0, make_double2(1.0, 0), rst, p->stream);
  const bool ref = p->d.mode == HZ_MODE_REF;
  auto apply_pre = [&](const double2* in, double2* pre, double2* out) {
    precond_apply(p, prec, 1, in, pre, false, false);
    if (ref) launch_w_minus_dl(N, 1, in, p->d_dL.as<double2>(), pre, out, p->stream);
    else launch_stencil(p, 1, pre, out, false, nullptr, p->stream);
  };
  double Res = 0, RelRes = 0;
  int it = 0;
  bool conv = false;
  for (int j = 0; j < maxit; ++j) {
    apply_pre(p0, yn, Ayn);
    zc alpha = -(dot_host(p, rst, r0) / dot_host(p, rst, Ayn));
    launch_axpbypcz(N, r0, d2(alpha), Ayn, make_double2(0, 0), Ayn, s0, p->stream);
    apply_pre(s0, zn, Azn);
    zc om = -(dot_host(p, Azn, s0) / dot_host(p, Azn, Azn));
    // r_{n+1} = s0 + om Azn  (into w)
    launch_axpbypcz(N, s0, d2(om), Azn, make_double2(0, 0), Azn, w, p->stream);
    alpha = -alpha;
    om = -om;
    launch_axpbypcz(N, x, d2(alpha), yn, d2(om), zn, x, p->stream);
    zc beta = (alpha / om) * (dot_host(p, rst, w) / dot_host(p, rst, r0));
    om = -om;
    // p = rn + beta (p + om Ayn)
    launch_axpby(N, d2(om * beta), Ayn, d2(beta), p0, p->stream);
    launch_axpby(N, make_double2(1, 0), w, make_double2(1, 0), p0, p->stream);
    HZ_HIP(hipMemcpyAsync(r0, w, N * sizeof(double2), hipMemcpyDeviceToDevice, p->stream));
    it++;
    residual(p, 1, f, x, w, false);
    Res = (nrm_host(p, w, mask));
    RelRes = Res / norm_f;
    hist.push(RelRes);
    if (!std::isfinite(RelRes)) {
      set_error("non-finite residual in BiCGStab");
      throw Error{HZ_ERR_NONFINITE};
    }
    if ((tol_abs ? Res : RelRes) < tol) {
      conv = true;
      break;
    }
  }
  sync(p);
  st.iters = it;
  st.converged = conv;
  st.res = Res;
  st.rel_res = RelRes;
  st.n_hist = hist.n;
  st.t_solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace hz

namespace hz {
// A lockstep FGMRES job over the ABI's arguments: batches of opts.batch
// right-hand sides (0: all of them, at most 16).
static LockJob lock_job(hz_problem* p, int nrhs, const hz_c128* f, hz_c128* u, const hz_krylov_opts& o,
                        hz_krylov_stats* st, bool adjoint) {
  LockJob J;
  J.p = p;
  J.nrhs = nrhs;
  J.f = (const double2*)f;
  J.u = (double2*)u;
  J.o = o;
  J.o.driver = HZ_DRIVER_FGMRES;
  J.st = st;
  J.adjoint = adjoint;
  J.batch = std::max(1, std::min(o.batch > 0 ? o.batch : 16, nrhs));
  return J;
}
}  // namespace hz

using namespace hz;

#define HZ_API_BEGIN try {
#define HZ_API_END                                    \
  }                                                   \
  catch (const hz::Error& e) {                        \
    return e.code;                                    \
  }                                                   \
  catch (const std::exception& e) {                   \
    hz::set_error("%s", e.what());                    \
    return HZ_ERR_ARG;                                \
  }                                                   \
  return HZ_OK;

extern "C" {

const char* hz_last_error(void) { return g_err.c_str(); }
int hz_version(void) { return 1; }

int hz_problem_create(const hz_desc* d, const double* c_host, hz_problem** out) {
  HZ_API_BEGIN
  HZ_CHECK(d && c_host && out, HZ_ERR_ARG, "null argument");
  for (int a = 0; a < 3; ++a) {
    HZ_CHECK(d->n[a] >= 1 && d->h[a] > 0.0, HZ_ERR_ARG, "axis %d: n=%lld h=%g", a, (long long)d->n[a], d->h[a]);
    HZ_CHECK(d->pml[a] >= 0 && d->spg[a] >= 0 && 2 * (d->pml[a] + d->spg[a]) < d->n[a], HZ_ERR_ARG,
             "axis %d: layers (pml %d, sponge %d) do not fit n=%lld", a, d->pml[a], d->spg[a], (long long)d->n[a]);
  }
  HZ_CHECK(d->mode == HZ_MODE_REF || d->mode == HZ_MODE_S7, HZ_ERR_ARG, "bad mode %d", d->mode);
  HZ_CHECK(d->n[0] <= 512, HZ_ERR_ARG, "nx=%lld exceeds the 512-node block limit", (long long)d->n[0]);
  HZ_CHECK(d->omega > 0.0, HZ_ERR_ARG, "omega must be positive");
  if (d->mode == HZ_MODE_REF) HZ_CHECK(d->spg[2] >= 1, HZ_ERR_ARG, "REF mode needs a z sponge (spg[2] >= 1)");
  hz_problem* p = new hz_problem();
  try {
    p->d = *d;
    p->nx = d->n[0]; p->ny = d->n[1]; p->nz = d->n[2];
    p->n2 = p->nx * p->ny;
    p->N = p->n2 * p->nz;
    HZ_HIP(hipGetDevice(&p->device));
    HZ_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    p->own_stream = true;
    setup_tables(p, c_host);
    if (d->mode == HZ_MODE_REF) {
      // conjugated delta-L for the adjoint operator
      std::vector<double2> dl(p->N);
      HZ_HIP(hipMemcpy(dl.data(), p->d_dL.ptr, p->N * sizeof(double2), hipMemcpyDeviceToHost));
      for (auto& e : dl) e.y = -e.y;
      p->d_dLc.alloc(p->N * sizeof(double2));
      HZ_HIP(hipMemcpy(p->d_dLc.ptr, dl.data(), p->N * sizeof(double2), hipMemcpyHostToDevice));
    }
    grow(p->scal, 4096 * sizeof(double2));
  } catch (...) {
    if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    throw;
  }
  *out = p;
  HZ_API_END
}

int hz_problem_destroy(hz_problem* p) {
  if (!p) return HZ_OK;
  (void)hipStreamSynchronize(p->stream);
  if (p->own_stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return HZ_OK;
}

int hz_problem_set_stream(hz_problem* p, void* s) {
  HZ_API_BEGIN
  HZ_CHECK(p, HZ_ERR_ARG, "null problem");
  HZ_HIP(hipStreamSynchronize(p->stream));
  if (p->own_stream) HZ_HIP(hipStreamDestroy(p->stream));
  p->own_stream = false;
  p->stream = (hipStream_t)s;
  if (!s) {
    HZ_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    p->own_stream = true;
  }
  HZ_API_END
}

int hz_problem_sync(hz_problem* p) {
  HZ_API_BEGIN
  HZ_CHECK(p, HZ_ERR_ARG, "null problem");
  sync(p);
  HZ_API_END
}

void* hz_problem_stream(const hz_problem* p) { return p ? (void*)p->stream : nullptr; }

int hz_plane_solve(hz_problem* p, int prec, int adjoint, int nrhs, void* x) {
  HZ_API_BEGIN
  HZ_CHECK(p && x && nrhs >= 1, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
  launch_plane_solve(p, prec, nrhs, x, adjoint != 0, p->stream);
  HZ_API_END
}

int hz_problem_slots(const hz_problem* p, int* nslot, int* nskip) {
  HZ_API_BEGIN
  HZ_CHECK(p, HZ_ERR_ARG, "null problem");
  if (nslot) *nslot = p->nslot;
  if (nskip) *nskip = p->nskip;
  HZ_API_END
}

int hz_problem_info(const hz_problem* p, int64_t* dims, double* factor_bytes) {
  HZ_API_BEGIN
  HZ_CHECK(p, HZ_ERR_ARG, "null problem");
  if (dims) { dims[0] = p->nx; dims[1] = p->ny; dims[2] = p->nz; dims[3] = p->N; }
  if (factor_bytes) {
    double b = 0;
    for (int q = 0; q < 2; ++q)
      b += (double)(p->fac[q].G.bytes + p->fac[q].Gsym.bytes + p->fac[q].Gtw.bytes + p->fac[q].hodlr.pool.bytes);
    *factor_bytes = b;
  }
  HZ_API_END
}

int hz_precond_factor(hz_problem* p, int prec) {
  HZ_API_BEGIN
  HZ_CHECK(p, HZ_ERR_ARG, "null problem");
  HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
  launch_factor(p, prec, p->stream);
  HZ_API_END
}

int hz_precond_apply(hz_problem* p, int prec, int nrhs, const hz_c128* in, hz_c128* out) {
  HZ_API_BEGIN
  HZ_CHECK(p && in && out && nrhs >= 1, HZ_ERR_ARG, "bad arguments");
  precond_apply(p, prec, nrhs, (const double2*)in, (double2*)out, false, false);
  HZ_API_END
}

int hz_precond_apply_adjoint(hz_problem* p, int prec, int nrhs, const hz_c128* in, hz_c128* out) {
  HZ_API_BEGIN
  HZ_CHECK(p && in && out && nrhs >= 1, HZ_ERR_ARG, "bad arguments");
  precond_apply(p, prec, nrhs, (const double2*)in, (double2*)out, true, false);
  HZ_API_END
}

int hz_precond_compress(hz_problem* p, int prec, double eps, int smallsize, int flags) {
  HZ_API_BEGIN
  HZ_CHECK(p, HZ_ERR_ARG, "null problem");
  HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
  HZ_CHECK((flags & ~(HZ_HODLR_FREE_DENSE | HZ_HODLR_TWISTED)) == 0, HZ_ERR_ARG, "unknown flags 0x%x", flags);
  hodlr_compress(p, prec, eps, smallsize, flags);
  HZ_API_END
}

int hz_precond_hodlr_info(const hz_problem* p, int prec, hz_hodlr_info* out) {
  HZ_API_BEGIN
  HZ_CHECK(p && out, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
  const Factors& f = p->fac[prec];
  const Hodlr& h = f.hodlr;
  memset(out, 0, sizeof(*out));
  out->active = h.active ? 1 : 0;
  out->smallsize = h.small;
  out->levels = h.lev;
  out->nodes = h.nn;
  out->leaf_max = h.mleaf;
  out->max_rank = h.pmax;
  out->eps = h.eps;
  out->mean_rank = h.mean_rank;
  out->bytes = h.bytes;
  out->dense_bytes = (double)(f.G.bytes + f.Gsym.bytes + f.Gtw.bytes);
  HZ_API_END
}

int hz_precond_hodlr_ranks(const hz_problem* p, int prec, int64_t k, int32_t* out, int64_t cap) {
  HZ_API_BEGIN
  HZ_CHECK(p && out, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
  const Hodlr& h = p->fac[prec].hodlr;
  HZ_CHECK(h.active, HZ_ERR_STATE, "no compressed factors of precision %d", prec);
  HZ_CHECK(k >= 0 && k < p->nz, HZ_ERR_ARG, "plane %lld outside 0..%lld", (long long)k, (long long)p->nz - 1);
  const int slot = p->fmap[k];
  HZ_CHECK(slot >= 0, HZ_ERR_ARG, "plane %lld is skipped (no factor)", (long long)k);
  const int64_t cnt = p->ny * (int64_t)h.nn;
  HZ_CHECK(cap >= cnt, HZ_ERR_ARG, "rank buffer of %lld entries, %lld needed", (long long)cap, (long long)cnt);
  const unsigned char* r = h.ranks.data() + (size_t)slot * cnt;
  for (int64_t e = 0; e < cnt; ++e) out[e] = r[e];
  HZ_API_END
}

int hz_L0_apply(hz_problem* p, int nrhs, const hz_c128* in, hz_c128* out) {
  HZ_API_BEGIN
  HZ_CHECK(p && in && out && nrhs >= 1, HZ_ERR_ARG, "bad arguments");
  L0_apply(p, nrhs, (const double2*)in, (double2*)out, false);
  HZ_API_END
}

int hz_op_apply(hz_problem* p, int nrhs, const hz_c128* in, hz_c128* out) {
  HZ_API_BEGIN
  HZ_CHECK(p && in && out && nrhs >= 1, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(in != out, HZ_ERR_ARG, "hz_op_apply is out of place");
  op_apply(p, nrhs, (const double2*)in, (double2*)out, false);
  HZ_API_END
}

int hz_op_apply_adjoint(hz_problem* p, int nrhs, const hz_c128* in, hz_c128* out) {
  HZ_API_BEGIN
  HZ_CHECK(p && in && out && nrhs >= 1, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(in != out, HZ_ERR_ARG, "hz_op_apply_adjoint is out of place");
  op_apply(p, nrhs, (const double2*)in, (double2*)out, true);
  HZ_API_END
}

int hz_op_apply_preconditioned(hz_problem* p, int prec, int nrhs, const hz_c128* in, hz_c128* out) {
  HZ_API_BEGIN
  HZ_CHECK(p && in && out && nrhs >= 1, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(in != out, HZ_ERR_ARG, "hz_op_apply_preconditioned is out of place");
  op_pre_apply(p, prec, nrhs, (const double2*)in, (double2*)out);
  HZ_API_END
}

int hz_residual(hz_problem* p, int nrhs, const hz_c128* f, const hz_c128* u, hz_c128* r, int mask, double* rel) {
  HZ_API_BEGIN
  HZ_CHECK(p && f && u && r && nrhs >= 1, HZ_ERR_ARG, "bad arguments");
  residual(p, nrhs, (const double2*)f, (const double2*)u, (double2*)r, false);
  if (rel) {
    for (int q = 0; q < nrhs; ++q) {
      const double2* fq = (const double2*)f + (size_t)q * p->N;
      const double2* rq = (const double2*)r + (size_t)q * p->N;
      const double a = (nrm_host(p, rq, mask));
      const double b = (nrm_host(p, fq, HZ_MASK_FULL));
      rel[q] = a / b;
    }
  }
  HZ_API_END
}

int hz_fgmres(hz_problem* p, int nrhs, const hz_c128* f, hz_c128* u, const hz_krylov_opts* o, hz_krylov_stats* st) {
  HZ_API_BEGIN
  HZ_CHECK(p && f && u && o && st && nrhs >= 1, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(o->m >= 1 && o->m <= 2000 && o->max_restarts >= 1, HZ_ERR_ARG, "bad restart parameters");
  int rc = HZ_OK;
  if (o->driver != HZ_DRIVER_GMRES_REF) {
    std::vector<LockJob> jobs(1);
    jobs[0] = lock_job(p, nrhs, f, u, *o, st, false);
    run_lockstep(jobs);
    if (jobs[0].rc != HZ_OK) {
      set_error("%s", jobs[0].msg.c_str());
      return jobs[0].rc;
    }
  }
  for (int q = 0; q < nrhs; ++q) {
    if (o->driver == HZ_DRIVER_GMRES_REF) {
      hz_krylov_opts oq = *o;
      if (q > 0) { oq.hist = nullptr; oq.hist_cap = 0; }
      gmres_ref_one(p, (const double2*)f + (size_t)q * p->N, (double2*)u + (size_t)q * p->N, oq, st[q]);
    }
    if (!st[q].converged && rc == HZ_OK) {
      rc = st[q].breakdown ? HZ_ERR_BREAKDOWN : HZ_ERR_NOT_CONVERGED;
      set_error("Krylov solve %d stopped at relative residual %.3e after %d iterations", q, st[q].rel_res,
                st[q].iters);
    }
  }
  return rc;
  HZ_API_END
}

int hz_bicgstab(hz_problem* p, int prec, int nrhs, const hz_c128* f, hz_c128* u, int maxit, double tol, int tol_abs,
                int mask, double* hist, int hist_cap, hz_krylov_stats* st) {
  HZ_API_BEGIN
  HZ_CHECK(p && f && u && st && nrhs >= 1 && maxit >= 1, HZ_ERR_ARG, "bad arguments");
  int rc = HZ_OK;
  for (int q = 0; q < nrhs; ++q) {
    HistW hw{q == 0 ? hist : nullptr, q == 0 ? hist_cap : 0};
    bicgstab_one(p, prec, (const double2*)f + (size_t)q * p->N, (double2*)u + (size_t)q * p->N, maxit, tol, tol_abs,
                 mask, hw, st[q]);
    if (!st[q].converged && rc == HZ_OK) {
      rc = HZ_ERR_NOT_CONVERGED;
      set_error("BiCGStab %d stopped at residual %.3e after %d iterations", q, st[q].res, st[q].iters);
    }
  }
  return rc;
  HZ_API_END
}

int hz_fft_z(hz_problem* p, int prec, int sign, int nrhs, const hz_c128* in, hz_c128* out) {
  HZ_API_BEGIN
  HZ_CHECK(p && in && out && nrhs >= 1 && (sign == 1 || sign == -1), HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
  launch_fft_z(p, prec, sign, nrhs, in, 0, out, 0, 0, nullptr, nullptr, p->stream);
  HZ_API_END
}

// ------------------------------------------------------- several problems
// Every problem (one frequency on its own grid) runs on its own stream; one
// host thread queues the work of all of them, so the GPU runs the problems'
// kernels side by side (their plane counts add up to fill the CUs).

// ------------------------------------------------------------- misfit
// Sources run in lockstep batches (opts.batch, default min(nsrc, 16)); the
// receiver restriction P and its transpose are device gathers/scatters.
static int misfit_batch(const hz_problem* p) {
  const int b = p->mopts.batch > 0 ? p->mopts.batch : 16;
  return std::max(1, std::min(b, p->nsrc));
}

// forward fields of sources [s0, s0 + nb) into us (stride N)
static void solve_lockstep(hz_problem* p, int nb, const double2* f, double2* u, const hz_krylov_opts& o,
                           hz_krylov_stats* st, bool adjoint) {
  std::vector<LockJob> jobs(1);
  jobs[0] = lock_job(p, nb, (const hz_c128*)f, (hz_c128*)u, o, st, adjoint);
  jobs[0].batch = nb;
  run_lockstep(jobs);
  if (jobs[0].rc != HZ_OK) {
    set_error("%s", jobs[0].msg.c_str());
    throw Error{jobs[0].rc};
  }
}

static void misfit_solve_forward(hz_problem* p, int s0, int nb, double2* us) {
  std::vector<hz_krylov_stats> st(nb);
  solve_lockstep(p, nb, p->src_f.as<double2>() + (size_t)p->N * s0, us, p->mopts, st.data(), false);
  for (int q = 0; q < nb; ++q)
    HZ_CHECK(st[q].converged, HZ_ERR_NOT_CONVERGED, "forward solve %d: relative residual %.3e", s0 + q,
             st[q].rel_res);
}

int hz_autotune(hz_problem* p, int prec, int nrhs, hz_tune_info* info) {
  HZ_API_BEGIN
  HZ_CHECK(p, HZ_ERR_ARG, "null problem");
  autotune(p, prec, nrhs, info);
  HZ_API_END
}

int hz_precond_factor_multi(int nprob, hz_problem* const* probs, int prec) {
  HZ_API_BEGIN
  HZ_CHECK(nprob >= 1 && probs, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
  for (int i = 0; i < nprob; ++i) HZ_CHECK(probs[i], HZ_ERR_ARG, "null problem %d", i);
  for (int i = 0; i < nprob; ++i)
    for (int k = 0; k < i; ++k) HZ_CHECK(probs[i] != probs[k], HZ_ERR_ARG, "problem %d listed twice", i);
  // all factorisations queued first (one stream each), then each finished
  int rc = HZ_OK;
  std::string msg;
  std::vector<bool> begun(nprob, false);
  for (int i = 0; i < nprob && rc == HZ_OK; ++i) {
    try {
      factor_begin(probs[i], prec, probs[i]->stream);
      begun[i] = true;
    } catch (const Error& e) {
      rc = e.code;
      msg = "problem " + std::to_string(i) + ": " + g_err;
    }
  }
  for (int i = 0; i < nprob; ++i) {
    if (!begun[i]) continue;
    try {
      factor_end(probs[i], prec, probs[i]->stream);
    } catch (const Error& e) {
      if (rc == HZ_OK) {
        rc = e.code;
        msg = "problem " + std::to_string(i) + ": " + g_err;
      }
    }
  }
  if (rc != HZ_OK) set_error("%s", msg.c_str());
  return rc;
  HZ_API_END
}

int hz_fgmres_multi(int nprob, hz_problem* const* probs, const int* nrhs, const hz_c128* const* f, hz_c128* const* u,
                    const hz_krylov_opts* opts, hz_krylov_stats* const* stats) {
  HZ_API_BEGIN
  HZ_CHECK(nprob >= 1 && probs && nrhs && f && u && opts && stats, HZ_ERR_ARG, "bad arguments");
  for (int i = 0; i < nprob; ++i) HZ_CHECK(probs[i] && nrhs[i] >= 1 && f[i] && u[i] && stats[i], HZ_ERR_ARG,
                                           "bad arguments for problem %d", i);
  for (int i = 0; i < nprob; ++i)
    for (int k = 0; k < i; ++k) HZ_CHECK(probs[i] != probs[k], HZ_ERR_ARG, "problem %d listed twice", i);
  HZ_CHECK(opts->driver == HZ_DRIVER_FGMRES, HZ_ERR_ARG, "hz_fgmres_multi runs the FGMRES driver");
  HZ_CHECK(opts->m >= 1 && opts->m <= 2000 && opts->max_restarts >= 1, HZ_ERR_ARG, "bad restart parameters");
  std::vector<LockJob> jobs(nprob);
  for (int i = 0; i < nprob; ++i) {
    hz_krylov_opts o = *opts;
    if (i > 0) { o.hist = nullptr; o.hist_cap = 0; }
    jobs[i] = lock_job(probs[i], nrhs[i], f[i], u[i], o, stats[i], false);
  }
  run_lockstep(jobs);
  for (int i = 0; i < nprob; ++i)
    if (jobs[i].rc != HZ_OK) {
      set_error("problem %d: %s", i, jobs[i].msg.c_str());
      return jobs[i].rc;
    }
  for (int i = 0; i < nprob; ++i)
    for (int q = 0; q < nrhs[i]; ++q) {
      const hz_krylov_stats& s = stats[i][q];
      if (!s.converged) {
        set_error("problem %d: Krylov solve %d stopped at relative residual %.3e after %d iterations", i, q,
                  s.rel_res, s.iters);
        return s.breakdown ? HZ_ERR_BREAKDOWN : HZ_ERR_NOT_CONVERGED;
      }
    }
  HZ_API_END
}

int hz_misfit_forward(hz_problem* p, int nsrc, const hz_c128* f, int nrec, const int64_t* rec, const hz_c128* data,
                      const hz_krylov_opts* o, int store_fields, double* J) {
  HZ_API_BEGIN
  HZ_CHECK(p && f && rec && data && o && J && nsrc >= 1 && nrec >= 1, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(p->d.mode == HZ_MODE_S7, HZ_ERR_STATE, "the misfit gradient is defined for HZ_MODE_S7");
  std::vector<int64_t> r(rec, rec + nrec);
  for (auto v : r) HZ_CHECK(v >= 0 && v < p->N, HZ_ERR_ARG, "receiver index %lld out of range", (long long)v);
  std::vector<int64_t> sorted = r;
  std::sort(sorted.begin(), sorted.end());
  HZ_CHECK(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(), HZ_ERR_ARG, "duplicate receivers");
  p->misfit_ready = false;
  p->rec = r;
  p->nsrc = nsrc;
  p->mopts = *o;
  p->mopts.driver = HZ_DRIVER_FGMRES;
  p->mopts.hist = nullptr;
  p->mopts.hist_cap = 0;
  const int B = misfit_batch(p);
  const size_t vb = (size_t)p->N * sizeof(double2);
  p->d_rec.alloc(sizeof(int64_t) * nrec);
  p->d_data.alloc(sizeof(double2) * (size_t)nsrc * nrec);
  p->d_res.alloc(sizeof(double2) * (size_t)B * nrec);
  HZ_HIP(hipMemcpyAsync(p->d_rec.ptr, r.data(), sizeof(int64_t) * nrec, hipMemcpyHostToDevice, p->stream));
  HZ_HIP(hipMemcpyAsync(p->d_data.ptr, data, sizeof(double2) * (size_t)nsrc * nrec, hipMemcpyHostToDevice,
                        p->stream));
  p->src_f.alloc(vb * nsrc);
  HZ_HIP(hipMemcpyAsync(p->src_f.ptr, f, vb * nsrc, hipMemcpyDeviceToDevice, p->stream));
  if (store_fields == HZ_FIELDS_AUTO) {
    // keep the fields when they fit the budget (the caller's, or half of the
    // device memory free now, after the problem's own buffers)
    double budget = p->field_budget;
    if (!(budget > 0.0)) {
      size_t free_b = 0, total_b = 0;
      HZ_HIP(hipMemGetInfo(&free_b, &total_b));
      budget = 0.5 * (double)free_b;
    }
    p->fields_stored = (double)vb * nsrc <= budget;
  } else {
    HZ_CHECK(store_fields == 0 || store_fields == 1, HZ_ERR_ARG, "store_fields=%d (0, 1 or HZ_FIELDS_AUTO)",
             store_fields);
    p->fields_stored = store_fields != 0;
  }
  if (p->fields_stored) p->fields_u.alloc(vb * nsrc);
  else grow(p->fields_u, vb * B);
  double Jt = 0.0;
  std::vector<double2> res((size_t)B * nrec);
  for (int s0 = 0; s0 < nsrc; s0 += B) {
    const int nb = std::min(B, nsrc - s0);
    double2* us = p->fields_u.as<double2>() + (p->fields_stored ? (size_t)p->N * s0 : 0);
    misfit_solve_forward(p, s0, nb, us);
    launch_rec_residual(nrec, nb, p->d_rec.as<int64_t>(), us, p->N, p->d_data.as<double2>() + (size_t)s0 * nrec,
                        p->d_res.as<double2>(), p->stream);
    HZ_HIP(hipMemcpyAsync(res.data(), p->d_res.ptr, sizeof(double2) * (size_t)nb * nrec, hipMemcpyDeviceToHost,
                          p->stream));
    sync(p);
    for (size_t i = 0; i < (size_t)nb * nrec; ++i) Jt += 0.5 * (res[i].x * res[i].x + res[i].y * res[i].y);
  }
  p->misfit_ready = true;
  *J = Jt;
  HZ_API_END
}

int hz_misfit_set_budget(hz_problem* p, double bytes) {
  HZ_API_BEGIN
  HZ_CHECK(p && bytes >= 0.0, HZ_ERR_ARG, "bad arguments");
  p->field_budget = bytes;
  HZ_API_END
}

int hz_misfit_fields_stored(const hz_problem* p, int* stored) {
  HZ_API_BEGIN
  HZ_CHECK(p && stored, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(p->misfit_ready, HZ_ERR_STATE, "hz_misfit_forward has not run");
  *stored = p->fields_stored ? 1 : 0;
  HZ_API_END
}

int hz_misfit_backward(hz_problem* p, double* grad) {
  HZ_API_BEGIN
  HZ_CHECK(p && grad, HZ_ERR_ARG, "bad arguments");
  HZ_CHECK(p->misfit_ready, HZ_ERR_STATE, "hz_misfit_forward has not run");
  const int64_t N = p->N;
  const int nrec = (int)p->rec.size();
  const double w2 = p->d.omega * p->d.omega;
  const int B = misfit_batch(p);
  grow(p->m_adj, (size_t)N * sizeof(double2) * 2 * B);
  double2* rhs = p->m_adj.as<double2>();
  double2* lam = rhs + (size_t)B * N;
  launch_fill_real(N, 0.0, grad, p->stream);
  std::vector<hz_krylov_stats> st(B);
  for (int s0 = 0; s0 < p->nsrc; s0 += B) {
    const int nb = std::min(B, p->nsrc - s0);
    double2* us = p->fields_u.as<double2>() + (p->fields_stored ? (size_t)N * s0 : 0);
    if (!p->fields_stored) misfit_solve_forward(p, s0, nb, us);
    // adjoint sources P^T (P u_s - d_s)
    launch_rec_residual(nrec, nb, p->d_rec.as<int64_t>(), us, N, p->d_data.as<double2>() + (size_t)s0 * nrec,
                        p->d_res.as<double2>(), p->stream);
    launch_fill((int64_t)nb * N, make_double2(0, 0), rhs, p->stream);
    launch_rec_scatter(nrec, nb, p->d_rec.as<int64_t>(), p->d_res.as<double2>(), rhs, N, p->stream);
    solve_lockstep(p, nb, rhs, lam, p->mopts, st.data(), true);
    for (int q = 0; q < nb; ++q)
      HZ_CHECK(st[q].converged, HZ_ERR_NOT_CONVERGED, "adjoint solve %d: relative residual %.3e", s0 + q,
               st[q].rel_res);
    // grad += -w^2 sum_s Re(conj(u_s) lambda_s)
    launch_grad_acc_multi(N, nb, -w2, us, N, lam, N, grad, p->stream);
  }
  sync(p);
  HZ_API_END
}

}  // extern "C"
