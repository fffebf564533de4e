// hz_setup.cpp -- host-side set-up of one frequency: PML tables, background
// medium, delta-L, z symbols and the z-FFT plan.  Everything here is O(N) or
// smaller and runs once per frequency before the device work.
#include <cmath>
#include <complex>

#include <algorithm>
#include <map>
#include <mutex>
#include <unordered_map>

#include "hz_internal.h"

namespace hz {

typedef std::complex<double> zc;
static inline double2 d2(zc v) { return make_double2(v.real(), v.imag()); }
static inline zc zz(double2 v) { return zc(v.x, v.y); }

// Stretching factor at (possibly half-integer) node s of one axis.  Inside a
// layer of `pml` nodes the layer coordinate runs 1 -> 1/pml from the outer edge
// inward: x = 1 - s/pml on the left, x = 1 + (s - n + 1)/pml on the right; the
// damping is pml_c x^2 and the factor 1 / (1 + i d / omega).  Half-integer
// points just inside n - pml on the right count as interior.
// (alpha and d, functions.cpp:1839-1878, 893-906.)
zc stretch(int64_t n, int pml, double omega, double pml_c, double s) {
  if (pml <= 0) return zc(1.0, 0.0);
  bool left = s < (double)pml;
  bool right = s >= (double)(n - pml);
  if (!left && !right) return zc(1.0, 0.0);
  double x = left ? 1.0 - s / (double)pml : 1.0 + (s - (double)n + 1.0) / (double)pml;
  double damp = pml_c * x * x;
  return zc(0.0, -omega) / zc(damp, -omega);
}

// Axis tables of the stretched second difference (beta3D / beta2D_pml,
// functions.cpp:2242-2303, 2347-2402): couplings a(i) a(i -/+ 1/2) / h^2 and the
// diagonal -a(i) (a(i + 1/2) + a(i - 1/2)) / h^2.  Adjoint couplings seen from
// node i: conj of the neighbour's coupling toward i.
static void build_axis(const hz_desc& d, int ax, std::vector<double2>& cm, std::vector<double2>& cp,
                       std::vector<double2>& cdg) {
  int64_t n = d.n[ax];
  double inv_h2 = 1.0 / (d.h[ax] * d.h[ax]);
  std::vector<zc> half(n + 1);  // a(i - 1/2) for i = 0..n
  for (int64_t i = 0; i <= n; ++i) half[i] = stretch(n, d.pml[ax], d.omega, d.pml_c, (double)i - 0.5);
  cm.resize(n); cp.resize(n); cdg.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    zc a = stretch(n, d.pml[ax], d.omega, d.pml_c, (double)i);
    zc lo = half[i], hi = half[i + 1];
    cm[i] = d2(a * lo * inv_h2);
    cp[i] = d2(a * hi * inv_h2);
    cdg[i] = d2(-(a * (hi + lo)) * inv_h2);
  }
}

// z symbol of plane k of the preconditioner.
static double z_symbol(const hz_desc& d, int64_t k) {
  const double pi = 3.141592653589793238462643;
  int64_t nz = d.n[2];
  switch (d.kz_symbol) {
    case HZ_KZ_REFERENCE: {  // FGMRES.cpp:182-190
      double q = (k < nz / 2 - 1) ? (double)k : (double)(nz - k);
      return (2.0 * pi * q / d.l[2]) * (2.0 * pi * q / d.l[2]);
    }
    case HZ_KZ_SPECTRAL: {
      double q = (double)(k < nz - k ? k : nz - k);
      return (2.0 * pi * q / d.l[2]) * (2.0 * pi * q / d.l[2]);
    }
    default: {
      // with q = min(k, nz - k) the symbol of k and nz - k is the same double
      const int64_t q = k < nz - k ? k : nz - k;
      double s = std::sin(pi * (double)q / (double)nz);
      return 4.0 * s * s / (d.h[2] * d.h[2]);
    }
  }
}

// Background speed per (x,y) column over the physical z range.
static void build_c0(hz_problem* p, const double* c) {
  const hz_desc& d = p->d;
  int64_t nx = p->nx, ny = p->ny, n2 = p->n2;
  int64_t z0 = d.pml[2] + d.spg[2], z1 = p->nz - d.pml[2] - d.spg[2];
  double nzp = (double)(z1 - z0);
  p->c0.assign(n2, make_double2(0, 0));
  if (d.c0_rule == HZ_C0_REFERENCE) {
    // SetSoundSpeed2D (functions.cpp:1716-1728): a single accumulator over the
    // physical columns (y outer, x inner) that is divided by the physical z
    // count after each column and never cleared; then the layer columns copy
    // the nearest physical edge, x first then y (functions.cpp:2636-2673).
    int64_t px = d.pml[0], py = d.pml[1];
    zc acc(0.0, 0.0);
    for (int64_t j = py; j < ny - py; ++j)
      for (int64_t i = px; i < nx - px; ++i) {
        for (int64_t k = z0; k < z1; ++k) {
          zc v(c[i + nx * j + n2 * k], 0.0);
          acc += 1.0 / (v * v);
        }
        acc /= nzp;
        p->c0[i + nx * j] = d2(std::sqrt(1.0 / acc));
      }
    for (int64_t j = 0; j < ny; ++j) {
      for (int64_t i = 0; i < px; ++i) p->c0[i + nx * j] = p->c0[px + nx * j];
      for (int64_t i = nx - px; i < nx; ++i) p->c0[i + nx * j] = p->c0[nx - px - 1 + nx * j];
    }
    for (int64_t i = 0; i < nx; ++i) {
      for (int64_t j = 0; j < py; ++j) p->c0[i + nx * j] = p->c0[i + nx * py];
      for (int64_t j = ny - py; j < ny; ++j) p->c0[i + nx * j] = p->c0[i + nx * (ny - py - 1)];
    }
  } else {
    std::vector<double> s(n2, 0.0);
    for (int64_t k = z0; k < z1; ++k)
      for (int64_t ij = 0; ij < n2; ++ij) {
        double v = c[ij + n2 * k];
        s[ij] += 1.0 / (v * v);
      }
    for (int64_t ij = 0; ij < n2; ++ij) p->c0[ij] = make_double2(std::sqrt(nzp / s[ij]), 0.0);
  }
}

// delta-L of the reference's implicit operator (GenerateDeltaL,
// functions.cpp:1789-1823): w^2 ((1 + i beta)/c0^2 - 1/c^2) on the planes
// spg..nz-spg (both ends included); the sponge planes take the nearest band
// plane times exp(-((z - z_edge)/(L_edge sigma))^2) with sigma = 0.3.
static void build_deltaL(hz_problem* p, const double* c, std::vector<double2>& dL) {
  const hz_desc& d = p->d;
  int64_t nz = p->nz, n2 = p->n2, spg = d.spg[2];
  double w2 = d.omega * d.omega, hz_ = d.h[2];
  dL.assign((size_t)p->N, make_double2(0, 0));
  int64_t k_last = (nz - spg < nz - 1) ? nz - spg : nz - 1;
  for (int64_t k = spg; k <= k_last; ++k)
    for (int64_t ij = 0; ij < n2; ++ij) {
      zc b = zz(p->c0[ij]);
      double v = c[ij + n2 * k];
      dL[ij + n2 * k] = d2(w2 * (zc(1.0, d.beta) / (b * b) - 1.0 / (v * v)));
    }
  const double sig = 0.3;
  double zl = spg * hz_, zr = (nz - spg) * hz_;
  double ll = zl, lr = d.l[2] - zr;
  for (int64_t k = 0; k < spg; ++k) {
    double t = (k * hz_ - zl) / (ll * sig);
    double g = std::exp(-t * t);
    for (int64_t ij = 0; ij < n2; ++ij) {
      double2 e = dL[ij + n2 * spg];
      dL[ij + n2 * k] = make_double2(e.x * g, e.y * g);
    }
  }
  for (int64_t k = nz - spg + 1; k < nz; ++k) {
    double t = (k * hz_ - zr) / (lr * sig);
    double g = std::exp(-t * t);
    for (int64_t ij = 0; ij < n2; ++ij) {
      double2 e = dL[ij + n2 * (nz - spg)];
      dL[ij + n2 * k] = make_double2(e.x * g, e.y * g);
    }
  }
}

// Factor slots of the planes: one per distinct z symbol among the planes the
// SAVE_MEM rule keeps (FGMRES.cpp:196-206: a plane whose kz^2 = lambda_k is at
// least skip_ratio k0^2 is neither factored nor solved).
static void build_plane_map(hz_problem* p) {
  const hz_desc& d = p->d;
  double k0 = d.skip_k0;
  if (k0 <= 0.0) {
    if (d.c_bg > 0.0) {
      k0 = d.omega / d.c_bg;
    } else {
      double cmin = 0.0;
      for (const double2& v : p->c0) {
        const double a = std::abs(zz(v));
        if (cmin == 0.0 || a < cmin) cmin = a;
      }
      k0 = d.omega / cmin;
    }
  }
  const double limit = (double)d.skip_ratio * k0 * k0;
  p->fmap.assign(p->nz, -1);
  p->splane.clear();
  std::map<double, int> slot_of;
  for (int64_t k = 0; k < p->nz; ++k) {
    const double lam = p->lamz[k];
    if (d.skip_ratio > 0 && lam >= limit) continue;
    auto it = slot_of.find(lam);
    if (it != slot_of.end()) {
      p->fmap[k] = it->second;
    } else {
      const int s = (int)p->splane.size();
      slot_of[lam] = s;
      p->splane.push_back((int)k);
      p->fmap[k] = s;
    }
  }
  p->nslot = (int)p->splane.size();
  p->nskip = (int)std::count(p->fmap.begin(), p->fmap.end(), -1);
  p->d_fmap.alloc(p->nz * sizeof(int));
  HZ_HIP(hipMemcpy(p->d_fmap.ptr, p->fmap.data(), p->nz * sizeof(int), hipMemcpyHostToDevice));
  if (p->nslot > 0) {
    p->d_splane.alloc(p->nslot * sizeof(int));
    HZ_HIP(hipMemcpy(p->d_splane.ptr, p->splane.data(), p->nslot * sizeof(int), hipMemcpyHostToDevice));
    std::vector<int2> sp(p->nslot, int2{-1, -1});
    p->slot_pairs = true;
    for (int64_t k = 0; k < p->nz; ++k) {
      const int s = p->fmap[k];
      if (s < 0) continue;
      if (sp[s].x < 0) sp[s].x = (int)k;
      else if (sp[s].y < 0) sp[s].y = (int)k;
      else p->slot_pairs = false;
    }
    p->d_slotp.alloc(p->nslot * sizeof(int2));
    HZ_HIP(hipMemcpy(p->d_slotp.ptr, sp.data(), p->nslot * sizeof(int2), hipMemcpyHostToDevice));
  } else {
    p->slot_pairs = false;
  }
}

static void build_fft_plan(hz_problem* p) {
  FftPlan& f = p->fft;
  int n = (int)p->nz;
  f.n = n;
  f.nstages = 0;
  int r = n;
  while (r % 4 == 0) { f.radix[f.nstages++] = 4; r /= 4; }
  while (r % 2 == 0) { f.radix[f.nstages++] = 2; r /= 2; }
  while (r % 3 == 0) { f.radix[f.nstages++] = 3; r /= 3; }
  while (r % 5 == 0) { f.radix[f.nstages++] = 5; r /= 5; }
  for (int q = 7; r > 1; q += 2)
    while (r % q == 0) {
      HZ_CHECK(f.nstages < 24, HZ_ERR_ARG, "nz=%d has too many factors", n);
      HZ_CHECK(q <= 64, HZ_ERR_ARG, "nz=%d has a prime factor %d > 64", n, q);
      f.radix[f.nstages++] = q;
      r /= q;
    }
  std::vector<double2> t64(n);
  std::vector<float2> t32(n);
  for (int t = 0; t < n; ++t) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
    double c = (double)cosl(a), s = (double)sinl(a);
    t64[t] = make_double2(c, s);
    t32[t] = make_float2((float)c, (float)s);
  }
  f.tw64.alloc(n * sizeof(double2));
  f.tw32.alloc(n * sizeof(float2));
  HZ_HIP(hipMemcpy(f.tw64.ptr, t64.data(), n * sizeof(double2), hipMemcpyHostToDevice));
  HZ_HIP(hipMemcpy(f.tw32.ptr, t32.data(), n * sizeof(float2), hipMemcpyHostToDevice));
}

void setup_tables(hz_problem* p, const double* c) {
  const hz_desc& d = p->d;
  for (int a = 0; a < 3; ++a) build_axis(d, a, p->cm[a], p->cp[a], p->cdg[a]);
  // device coefficient block: per axis [cm, cp, cdg, adj_cm, adj_cp]
  std::vector<double2> tab;
  for (int a = 0; a < 3; ++a) {
    int64_t n = d.n[a];
    std::vector<double2> acm(n), acp(n);
    for (int64_t i = 0; i < n; ++i) {
      acm[i] = i > 0 ? make_double2(p->cp[a][i - 1].x, -p->cp[a][i - 1].y) : make_double2(0, 0);
      acp[i] = i < n - 1 ? make_double2(p->cm[a][i + 1].x, -p->cm[a][i + 1].y) : make_double2(0, 0);
    }
    tab.insert(tab.end(), p->cm[a].begin(), p->cm[a].end());
    tab.insert(tab.end(), p->cp[a].begin(), p->cp[a].end());
    tab.insert(tab.end(), p->cdg[a].begin(), p->cdg[a].end());
    tab.insert(tab.end(), acm.begin(), acm.end());
    tab.insert(tab.end(), acp.begin(), acp.end());
  }
  p->d_tab.alloc(tab.size() * sizeof(double2));
  HZ_HIP(hipMemcpy(p->d_tab.ptr, tab.data(), tab.size() * sizeof(double2), hipMemcpyHostToDevice));

  build_c0(p, c);
  p->kappa.resize(p->n2);
  for (int64_t ij = 0; ij < p->n2; ++ij) {
    zc kw = d.c_bg > 0.0 ? zc(d.omega / d.c_bg, 0.0) : d.omega / zz(p->c0[ij]);
    p->kappa[ij] = d2(kw * kw * zc(1.0, d.beta));
  }
  p->lamz.resize(p->nz);
  for (int64_t k = 0; k < p->nz; ++k) p->lamz[k] = z_symbol(d, k);
  p->d_kappa.alloc(p->n2 * sizeof(double2));
  HZ_HIP(hipMemcpy(p->d_kappa.ptr, p->kappa.data(), p->n2 * sizeof(double2), hipMemcpyHostToDevice));
  p->d_lamz.alloc(p->nz * sizeof(double));
  HZ_HIP(hipMemcpy(p->d_lamz.ptr, p->lamz.data(), p->nz * sizeof(double), hipMemcpyHostToDevice));
  build_plane_map(p);

  // Schur coupling of y-line j to j-1: cm_y(j) * cp_y(j-1)
  std::vector<double2> sig(p->ny, make_double2(0, 0));
  for (int64_t j = 1; j < p->ny; ++j) sig[j] = d2(zz(p->cm[1][j]) * zz(p->cp[1][j - 1]));
  p->d_sig.alloc(p->ny * sizeof(double2));
  HZ_HIP(hipMemcpy(p->d_sig.ptr, sig.data(), p->ny * sizeof(double2), hipMemcpyHostToDevice));

  // Diagonal scalings D = diag(dx(i) dy(j)) with D^{-1} M_k complex symmetric:
  // d(i) / d(i-1) = cm(i) / cp(i-1) per axis (the reference's symmetry check,
  // Tests.cpp:686-757), so the block inverses G_j D_j are symmetric.
  {
    std::vector<double2> sd;
    for (int a = 0; a < 2; ++a) {
      const int64_t n = d.n[a];
      std::vector<zc> dd(n, zc(1.0, 0.0));
      for (int64_t i = 1; i < n; ++i) {
        const zc den = zz(p->cp[a][i - 1]);
        dd[i] = std::abs(den) > 0.0 ? dd[i - 1] * zz(p->cm[a][i]) / den : dd[i - 1];
      }
      p->symd[a].resize(n);
      for (int64_t i = 0; i < n; ++i) p->symd[a][i] = d2(dd[i]);
    }
    for (int a = 0; a < 2; ++a) sd.insert(sd.end(), p->symd[a].begin(), p->symd[a].end());
    for (int a = 0; a < 2; ++a)
      for (const double2& v : p->symd[a]) sd.push_back(d2(1.0 / zz(v)));
    p->d_symd.alloc(sd.size() * sizeof(double2));
    HZ_HIP(hipMemcpy(p->d_symd.ptr, sd.data(), sd.size() * sizeof(double2), hipMemcpyHostToDevice));
  }

  if (d.mode == HZ_MODE_S7) {
    std::vector<double> mass(p->N);
    double w2 = d.omega * d.omega;
    for (int64_t l = 0; l < p->N; ++l) mass[l] = w2 / (c[l] * c[l]);
    p->d_mass.alloc(p->N * sizeof(double));
    HZ_HIP(hipMemcpy(p->d_mass.ptr, mass.data(), p->N * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<double2> dL;
    build_deltaL(p, c, dL);
    p->d_dL.alloc(p->N * sizeof(double2));
    HZ_HIP(hipMemcpy(p->d_dL.ptr, dL.data(), p->N * sizeof(double2), hipMemcpyHostToDevice));
  }
  build_fft_plan(p);
}

void set_max_lds(const void* fn, size_t bytes) {
  static std::mutex mu;
  static std::unordered_map<const void*, size_t> done;
  std::lock_guard<std::mutex> lk(mu);
  size_t& have = done[fn];
  if (bytes <= have) return;
  HZ_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  have = bytes;
}

}  // namespace hz
