// hz_oracle.cpp -- CPU FP64 restatement of the reference's 3D Helmholtz hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke entry and
// the CPU-baseline leg of bench.py may load it; the product library never links,
// loads or calls it.  Every function cites the reference file:line it restates
// (paths relative to "Iterative Helholtz solver with PARDISO/Source/").
//
// Algorithms are deliberately independent of the GPU implementation where that
// is possible: the per-kz plane solver here is a banded LU with partial pivoting
// (LAPACK zgbtrf/zgbtrs form, the same role as PARDISO phase 22/33 in
// FGMRES.cpp:274-284 / functions.cpp:5959), the z transform is a recursive
// mixed-radix FFT, and the Krylov drivers follow the reference control flow.
//
// Complex data is interleaved (re, im) double, index l = i + nx*j + nx*ny*k
// (functions.cpp:225-230).
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;

extern "C" {

// Same field layout as hz_desc in include/hz.h (kept separate on purpose: the
// oracle shares no code with the product).
typedef struct {
  int64_t n[3];
  double h[3];
  double l[3];
  int32_t pml[3];
  int32_t spg[3];
  double omega;
  double beta;
  double pml_c;
  int32_t mode;       // 0 = REF (spectral z + delta-L), 1 = S7 (7-point PML stencil)
  int32_t kz_symbol;  // 0 = reference kz map, 1 = corrected spectral, 2 = periodic 2nd difference
  int32_t c0_rule;    // 0 = reference running sum, 1 = per-column harmonic mean
  int32_t skip_ratio; // > 0: SAVE_MEM plane skipping (FGMRES.cpp:196-206)
  double c_bg;        // > 0: constant background speed for kappa (reference HOMO)
  double skip_k0;     // wavenumber of the skip test (0: omega / c_bg, or omega / min c0)
} or_desc;

typedef struct {
  int32_t iters;        // total inner iterations performed
  int32_t restarts;     // restart cycles started
  int32_t converged;
  int32_t breakdown;
  double rel_res;       // last relative residual (definition depends on the driver)
  double res;           // last absolute residual
  double diff_sol;      // reference monitor ||u_k - u_{k-1}|| / ||u_{k-1}|| (phys box)
  int32_t n_hist;
  int32_t reserved;
} or_stats;

}  // extern "C"

static const double OR_PI = 3.141592653589793238462643;

// ---------------------------------------------------------------- PML profile
// d(x) = C x^2, functions.cpp:893-906 (C = 1000 in the reference).
static double pml_d(double c, double x) { return c * x * x; }

// alpha(xyz, i) = -i w / (d(x) - i w) inside the layer, 1 outside; the layer
// coordinate is x = 1 - i/pml on the left and 1 + (i - n + 1)/pml on the right.
// functions.cpp:1839-1878.
static cd alpha_at(int64_t n, int pml, double omega, double pml_c, double s) {
  if (pml == 0) return cd(1.0, 0.0);
  double hp = 1.0 / pml;
  double x;
  if (s < pml) {
    x = 1.0 - s * hp;
  } else if (s >= (double)(n - pml)) {
    x = 1.0 + (s - (double)n + 1.0) * hp;
  } else {
    return cd(1.0, 0.0);
  }
  return cd(0.0, -omega) / cd(pml_d(pml_c, x), -omega);
}

// ------------------------------------------------------------------- problem
struct OrProblem {
  or_desc d;
  int64_t nx, ny, nz, n2, N;
  std::vector<cd> cm[3], cp[3], cdg[3];   // per-axis stencil coefficient tables
  std::vector<double> c;                  // sound speed (N)
  std::vector<cd> c0;                     // background speed (nx*ny), complex as in the reference
  std::vector<double> lamz;               // z symbol per plane
  std::vector<cd> kappa;                  // (nx*ny) background term without the z symbol
  std::vector<cd> dL;                     // delta-L (N), REF mode
  std::vector<char> solve;                // per plane: 1 = factored and solved, 0 = skipped (SAVE_MEM)
  // plane factors (banded LU)
  bool factored = false;
  int64_t kl = 0, ku = 0, ldab = 0;
  int64_t k_off = 0;                      // first plane held in ab/ipiv
  std::vector<cd> ab;                     // nz * ldab * n2
  std::vector<int64_t> ipiv;              // nz * n2
};

// Coefficients of the 1D second-difference with PML stretching along one axis:
// beta2D_pml / beta3D, functions.cpp:2242-2303, 2347-2402:
//   off-diagonals alpha(i) alpha(i +- 1/2) / h^2,
//   diagonal     -alpha(i) (alpha(i + 1/2) + alpha(i - 1/2)) / h^2.
static void axis_tables(const or_desc& d, int ax, std::vector<cd>& cm, std::vector<cd>& cp,
                        std::vector<cd>& cdg) {
  int64_t n = d.n[ax];
  double h2 = d.h[ax] * d.h[ax];
  cm.resize(n); cp.resize(n); cdg.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    cd a = alpha_at(n, d.pml[ax], d.omega, d.pml_c, (double)i);
    cd ap = alpha_at(n, d.pml[ax], d.omega, d.pml_c, (double)i + 0.5);
    cd am = alpha_at(n, d.pml[ax], d.omega, d.pml_c, (double)i - 0.5);
    cm[i] = a * am / h2;
    cp[i] = a * ap / h2;
    cdg[i] = -a * (ap + am) / h2;
  }
}

// kz symbol for plane k.  Reference map (FGMRES.cpp:182-190): k < nz/2 - 1 uses
// k, else nz - k, symbol (2 pi k~ / Lz)^2.
static double z_symbol(const or_desc& d, int64_t k) {
  int64_t nz = d.n[2];
  if (d.kz_symbol == 0) {
    double kk = (k < nz / 2 - 1) ? (double)k : (double)(nz - k);
    return 4.0 * OR_PI * OR_PI * kk * kk / (d.l[2] * d.l[2]);
  } else if (d.kz_symbol == 1) {
    double kk = (double)std::min<int64_t>(k, nz - k);
    return 4.0 * OR_PI * OR_PI * kk * kk / (d.l[2] * d.l[2]);
  } else {
    double q = (double)std::min<int64_t>(k, nz - k);
    double s = std::sin(OR_PI * q / (double)nz);
    return 4.0 * s * s / (d.h[2] * d.h[2]);
  }
}

// SAVE_MEM (FGMRES.cpp:196-206): plane k is factored and solved only while its
// kz^2 = lambda_k is below ratio * k2, k2 = (omega / c_z)^2 (here skip_k0^2, or
// (omega / c_bg)^2, or omega^2 over the smallest |c0|^2); the preconditioner is
// zero on the other planes (functions.cpp:6023-6027), and so is L0
// (functions.cpp:6286-6295).
static void plane_flags(OrProblem& p) {
  const or_desc& d = p.d;
  double k0 = d.skip_k0;
  if (k0 <= 0.0) {
    if (d.c_bg > 0.0) {
      k0 = d.omega / d.c_bg;
    } else {
      double cmin = 0.0;
      for (const cd& v : p.c0)
        if (cmin == 0.0 || std::abs(v) < cmin) cmin = std::abs(v);
      k0 = d.omega / cmin;
    }
  }
  p.solve.assign(p.nz, 1);
  if (d.skip_ratio > 0)
    for (int64_t k = 0; k < p.nz; ++k)
      if (p.lamz[k] >= (double)d.skip_ratio * k0 * k0) p.solve[k] = 0;
}

// Background speed c0(i,j).
// rule 0: SetSoundSpeed2D (functions.cpp:1716-1728) -- one running sum over the
//   physical columns in j-outer / i-inner order, never reset, divided by the
//   physical z count after each column; then HeteroSoundSpeed2DExtensionToPML
//   (functions.cpp:2636-2673) copies the edge values into the x then y layers.
// rule 1: harmonic mean of c over the physical z range, for every column.
static void compute_c0(OrProblem& p) {
  const or_desc& d = p.d;
  int64_t nx = p.nx, ny = p.ny, nz = p.nz;
  int64_t zlo = d.pml[2] + d.spg[2], zhi = nz - d.pml[2] - d.spg[2];
  p.c0.assign(nx * ny, cd(0, 0));
  if (d.c0_rule == 0) {
    int px = d.pml[0], py = d.pml[1];
    cd S(0, 0);
    for (int64_t j = py; j < ny - py; ++j)
      for (int64_t i = px; i < nx - px; ++i) {
        for (int64_t k = zlo; k < zhi; ++k) {
          cd cc(p.c[i + nx * j + nx * ny * k], 0.0);
          S += 1.0 / (cc * cc);
        }
        S /= (double)(zhi - zlo);
        p.c0[i + nx * j] = std::sqrt(1.0 / S);
      }
    for (int64_t j = 0; j < ny; ++j) {
      for (int64_t i = 0; i < px; ++i) p.c0[i + nx * j] = p.c0[px + nx * j];
      for (int64_t i = nx - px; i < nx; ++i) p.c0[i + nx * j] = p.c0[(nx - px - 1) + nx * j];
    }
    for (int64_t i = 0; i < nx; ++i) {
      for (int64_t j = 0; j < py; ++j) p.c0[i + nx * j] = p.c0[i + nx * py];
      for (int64_t j = ny - py; j < ny; ++j) p.c0[i + nx * j] = p.c0[i + nx * (ny - py - 1)];
    }
  } else {
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        double s = 0.0;
        for (int64_t k = zlo; k < zhi; ++k) {
          double cc = p.c[i + nx * j + nx * ny * k];
          s += 1.0 / (cc * cc);
        }
        s /= (double)(zhi - zlo);
        p.c0[i + nx * j] = cd(std::sqrt(1.0 / s), 0.0);
      }
  }
}

// delta-L = w^2 [(1 + i beta)/c0^2 - 1/c^2] on planes spg..nz-spg inclusive, then
// a Gaussian taper (sigma = 0.3) copied from the first/last band plane into the
// sponge.  GenerateDeltaL, functions.cpp:1789-1823.
static void compute_deltaL(OrProblem& p) {
  const or_desc& d = p.d;
  int64_t nx = p.nx, ny = p.ny, nz = p.nz, n2 = p.n2;
  int64_t spg = d.spg[2];
  double w2 = d.omega * d.omega;
  p.dL.assign(p.N, cd(0, 0));
  double sigma = 0.3, hz = d.h[2];
  double z0 = spg * hz, zN = (nz - spg) * hz;
  double L0 = z0, Ln = d.l[2] - zN;
  int64_t khi = std::min<int64_t>(nz - spg, nz - 1);
  for (int64_t k = spg; k <= khi; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        int64_t ij = i + nx * j;
        cd c0 = p.c0[ij];
        cd cc(p.c[ij + n2 * k], 0.0);
        p.dL[ij + n2 * k] = w2 * (cd(1.0, d.beta) / (c0 * c0) - 1.0 / (cc * cc));
      }
  for (int64_t k = 0; k < spg; ++k) {
    double zp = k * hz;
    double t = std::exp(-(zp - z0) * (zp - z0) / (L0 * L0) / (sigma * sigma));
    for (int64_t ij = 0; ij < n2; ++ij) p.dL[ij + n2 * k] = p.dL[ij + n2 * spg] * t;
  }
  for (int64_t k = nz - spg + 1; k < nz; ++k) {
    double zp = k * hz;
    double t = std::exp(-(zp - zN) * (zp - zN) / (Ln * Ln) / (sigma * sigma));
    for (int64_t ij = 0; ij < n2; ++ij) p.dL[ij + n2 * k] = p.dL[ij + n2 * (nz - spg)] * t;
  }
}

// ---------------------------------------------------------------------- FFT
// Recursive mixed-radix decimation in time; forward sign e^{-2 pi i jk/n},
// unscaled (the DFTI convention, functions.cpp:5864-5869 and FFT.cpp:141-149).
static void fft_rec(const cd* x, int64_t stride, int64_t n, cd* out, int sign,
                    std::vector<cd>& scratch) {
  if (n == 1) { out[0] = x[0]; return; }
  int64_t p = 0;
  for (int64_t f = 2; f * f <= n; ++f) if (n % f == 0) { p = f; break; }
  if (p == 0) {  // prime length: direct sum
    for (int64_t k = 0; k < n; ++k) {
      cd s(0, 0);
      for (int64_t j = 0; j < n; ++j) {
        int64_t t = (j * k) % n;
        double a = 2.0 * OR_PI * (double)t / (double)n;
        s += x[j * stride] * cd(std::cos(a), sign * std::sin(a));
      }
      out[k] = s;
    }
    return;
  }
  int64_t m = n / p;
  std::vector<cd> sub(n);
  for (int64_t r = 0; r < p; ++r) fft_rec(x + r * stride, stride * p, m, &sub[r * m], sign, scratch);
  for (int64_t k = 0; k < n; ++k) {
    cd s(0, 0);
    for (int64_t r = 0; r < p; ++r) {
      int64_t t = (r * k) % n;
      double a = 2.0 * OR_PI * (double)t / (double)n;
      s += sub[r * m + (k % m)] * cd(std::cos(a), sign * std::sin(a));
    }
    out[k] = s;
  }
}

// Brute-force DFT, FFT.cpp:141-162 (forward unscaled, backward scaled by 1/n).
static void dft_brute(const cd* x, int64_t n, cd* out, int sign) {
  for (int64_t k = 0; k < n; ++k) {
    cd s(0, 0);
    for (int64_t j = 0; j < n; ++j) {
      double a = 2.0 * OR_PI * (double)((j * k) % n) / (double)n;
      s += x[j] * cd(std::cos(a), sign * std::sin(a));
    }
    out[k] = (sign > 0) ? s / (double)n : s;
  }
}

// z transform of every (x,y) line, stride nx*ny: the per-line DftiCompute loops
// at functions.cpp:5876-5881 (forward) and 6041-6046 (backward, scale 1/nz).
static void fft_z(const OrProblem& p, const cd* in, cd* out, int sign) {
  int64_t n2 = p.n2, nz = p.nz;
#pragma omp parallel
  {
    std::vector<cd> line(nz), res(nz), scratch;
#pragma omp for schedule(static)
    for (int64_t w = 0; w < n2; ++w) {
      for (int64_t k = 0; k < nz; ++k) line[k] = in[w + n2 * k];
      fft_rec(line.data(), 1, nz, res.data(), sign, scratch);
      double sc = (sign > 0) ? 1.0 / (double)nz : 1.0;
      for (int64_t k = 0; k < nz; ++k) out[w + n2 * k] = res[k] * sc;
    }
  }
}

// ------------------------------------------------------------ plane operator
// kappa_k(i,j) = (w/cb)^2 (1 + i beta) - kz_k^2 (FGMRES.cpp:233, 261-266):
// cb = c_bg (reference HOMO) or c0(i,j).
static inline cd plane_kappa(const OrProblem& p, int64_t k, int64_t ij) {
  return p.kappa[ij] - p.lamz[k];
}

// y = M_k x for one plane (5-point; GenSparseMatrixOnline2DwithPMLFast,
// functions.cpp:3296-3345 with the diagonal shift of FGMRES.cpp:254-267).
static void plane_matvec(const OrProblem& p, int64_t k, const cd* x, cd* y) {
  int64_t nx = p.nx, ny = p.ny;
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      int64_t l = i + nx * j;
      cd diag = (p.cdg[0][i] + p.cdg[1][j]) + plane_kappa(p, k, l);
      cd s = diag * x[l];
      if (j > 0) s += p.cm[1][j] * x[l - nx];
      if (i > 0) s += p.cm[0][i] * x[l - 1];
      if (i < nx - 1) s += p.cp[0][i] * x[l + 1];
      if (j < ny - 1) s += p.cp[1][j] * x[l + nx];
      y[l] = s;
    }
}

// Band storage, LAPACK convention (0-based): A(i,c) at ab[(kl+ku+i-c) + c*ldab].
static inline cd& AB(std::vector<cd>& ab, int64_t base, int64_t ldab, int64_t kv, int64_t i, int64_t c) {
  return ab[base + (kv + i - c) + c * ldab];
}

// Banded LU with partial pivoting of one plane (zgbtf2 algorithm); this takes
// the role of PARDISO phase 11+22 (FGMRES.cpp:274-284).
static void plane_factor(OrProblem& p, int64_t k) {
  int64_t n = p.n2, nx = p.nx, ny = p.ny;
  int64_t kl = p.kl, ku = p.ku, kv = kl + ku, ldab = p.ldab;
  int64_t base = (k - p.k_off) * ldab * n;
  int64_t* ipiv = &p.ipiv[(k - p.k_off) * n];
  for (int64_t c = 0; c < n; ++c)
    for (int64_t r = 0; r < ldab; ++r) p.ab[base + r + c * ldab] = 0.0;
  // scatter the matrix
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      int64_t l = i + nx * j;
      AB(p.ab, base, ldab, kv, l, l) = (p.cdg[0][i] + p.cdg[1][j]) + plane_kappa(p, k, l);
      if (j > 0) AB(p.ab, base, ldab, kv, l, l - nx) = p.cm[1][j];
      if (i > 0) AB(p.ab, base, ldab, kv, l, l - 1) = p.cm[0][i];
      if (i < nx - 1) AB(p.ab, base, ldab, kv, l, l + 1) = p.cp[0][i];
      if (j < ny - 1) AB(p.ab, base, ldab, kv, l, l + nx) = p.cp[1][j];
    }
  int64_t ju = 0;
  for (int64_t j = 0; j < n; ++j) {
    int64_t km = std::min(kl, n - 1 - j);
    int64_t jp = 0;
    double best = -1.0;
    for (int64_t q = 0; q <= km; ++q) {
      cd v = AB(p.ab, base, ldab, kv, j + q, j);
      double a = std::fabs(v.real()) + std::fabs(v.imag());
      if (a > best) { best = a; jp = q; }
    }
    ipiv[j] = j + jp;
    ju = std::max(ju, std::min(j + ku + jp, n - 1));
    cd piv = AB(p.ab, base, ldab, kv, j + jp, j);
    if (piv == cd(0, 0)) continue;  // singular column: left to the solve
    if (jp != 0)
      for (int64_t c = j; c <= ju; ++c)
        std::swap(AB(p.ab, base, ldab, kv, j, c), AB(p.ab, base, ldab, kv, j + jp, c));
    cd inv = 1.0 / AB(p.ab, base, ldab, kv, j, j);
    for (int64_t q = 1; q <= km; ++q) AB(p.ab, base, ldab, kv, j + q, j) *= inv;
    for (int64_t c = j + 1; c <= ju; ++c) {
      cd ujc = AB(p.ab, base, ldab, kv, j, c);
      if (ujc == cd(0, 0)) continue;
      for (int64_t q = 1; q <= km; ++q)
        AB(p.ab, base, ldab, kv, j + q, c) -= AB(p.ab, base, ldab, kv, j + q, j) * ujc;
    }
  }
}

// Plane solve with the factors (zgbtrs, TRANS='N' or 'C').  PARDISO phase 33,
// functions.cpp:5959.
static void plane_solve(OrProblem& p, int64_t k, cd* b, bool adjoint) {
  int64_t n = p.n2, kl = p.kl, ku = p.ku, kv = kl + ku, ldab = p.ldab;
  int64_t base = (k - p.k_off) * ldab * n;
  const int64_t* ipiv = &p.ipiv[(k - p.k_off) * n];
  if (!adjoint) {
    for (int64_t j = 0; j < n - 1; ++j) {
      int64_t km = std::min(kl, n - 1 - j);
      int64_t l = ipiv[j];
      if (l != j) std::swap(b[l], b[j]);
      cd bj = b[j];
      for (int64_t q = 1; q <= km; ++q) b[j + q] -= AB(p.ab, base, ldab, kv, j + q, j) * bj;
    }
    for (int64_t j = n - 1; j >= 0; --j) {
      b[j] /= AB(p.ab, base, ldab, kv, j, j);
      cd bj = b[j];
      for (int64_t i = std::max<int64_t>(0, j - kv); i < j; ++i) b[i] -= AB(p.ab, base, ldab, kv, i, j) * bj;
    }
  } else {
    for (int64_t j = 0; j < n; ++j) {
      cd s = b[j];
      for (int64_t i = std::max<int64_t>(0, j - kv); i < j; ++i)
        s -= std::conj(AB(p.ab, base, ldab, kv, i, j)) * b[i];
      b[j] = s / std::conj(AB(p.ab, base, ldab, kv, j, j));
    }
    for (int64_t j = n - 2; j >= 0; --j) {
      int64_t km = std::min(kl, n - 1 - j);
      cd s = b[j];
      for (int64_t q = 1; q <= km; ++q) s -= std::conj(AB(p.ab, base, ldab, kv, j + q, j)) * b[j + q];
      b[j] = s;
      int64_t l = ipiv[j];
      if (l != j) std::swap(b[l], b[j]);
    }
  }
}

static void ensure_factored(OrProblem& p) {
  if (p.factored) return;
  p.k_off = 0;
  p.kl = p.nx; p.ku = p.nx; p.ldab = 2 * p.kl + p.ku + 1;
  p.ab.assign((size_t)(p.nz * p.ldab * p.n2), cd(0, 0));
  p.ipiv.assign((size_t)(p.nz * p.n2), 0);
#pragma omp parallel for schedule(dynamic)
  for (int64_t k = 0; k < p.nz; ++k)
    if (p.solve[k]) plane_factor(p, k);
  p.factored = true;
}

// ------------------------------------------------------- 3D preconditioner
// L0^{-1} f = F^{-1} diag(M_k^{-1}) F f.  Solve3DSparseUsingFT_CSR,
// functions.cpp:5837-6061.  adjoint: (L0^{-1})^H = F^{-1} diag(M_k^{-H}) F.
static void precond_apply(OrProblem& p, const cd* f, cd* x, bool adjoint) {
  ensure_factored(p);
  std::vector<cd> t(p.N);
  fft_z(p, f, t.data(), -1);
#pragma omp parallel for schedule(dynamic)
  for (int64_t k = 0; k < p.nz; ++k) {
    if (p.solve[k]) plane_solve(p, k, &t[k * p.n2], adjoint);
    else std::fill(t.begin() + k * p.n2, t.begin() + (k + 1) * p.n2, cd(0.0, 0.0));
  }
  fft_z(p, t.data(), x, +1);
}

// L0 u = F^{-1} diag(M_k) F u.  Multiply3DSparseUsingFT, functions.cpp:6247-6318.
static void L0_apply(OrProblem& p, const cd* u, cd* y) {
  std::vector<cd> t(p.N), s(p.N);
  fft_z(p, u, t.data(), -1);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < p.nz; ++k)
    if (p.solve[k]) plane_matvec(p, k, &t[k * p.n2], &s[k * p.n2]);
  fft_z(p, s.data(), y, +1);
}

// 7-point PML stencil with mass term w^2/c^2 (beta3D, functions.cpp:2242-2303,
// Dirichlet closure as in GenSparseMatrixOnline3DwithPML :2961-3161).
// adjoint: y = A^H u.
static void stencil_apply(const OrProblem& p, const cd* u, cd* y, bool adjoint) {
  int64_t nx = p.nx, ny = p.ny, nz = p.nz, n2 = p.n2;
  double w2 = p.d.omega * p.d.omega;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        int64_t l = i + nx * j + n2 * k;
        double cc = p.c[l];
        cd diag = p.cdg[0][i] + p.cdg[1][j] + p.cdg[2][k] + w2 / (cc * cc);
        cd s;
        if (!adjoint) {
          s = diag * u[l];
          if (i > 0) s += p.cm[0][i] * u[l - 1];
          if (i < nx - 1) s += p.cp[0][i] * u[l + 1];
          if (j > 0) s += p.cm[1][j] * u[l - nx];
          if (j < ny - 1) s += p.cp[1][j] * u[l + nx];
          if (k > 0) s += p.cm[2][k] * u[l - n2];
          if (k < nz - 1) s += p.cp[2][k] * u[l + n2];
        } else {
          // (A^H)(l, l') = conj(A(l', l)): the neighbour's coefficient toward l.
          s = std::conj(diag) * u[l];
          if (i > 0) s += std::conj(p.cp[0][i - 1]) * u[l - 1];
          if (i < nx - 1) s += std::conj(p.cm[0][i + 1]) * u[l + 1];
          if (j > 0) s += std::conj(p.cp[1][j - 1]) * u[l - nx];
          if (j < ny - 1) s += std::conj(p.cm[1][j + 1]) * u[l + nx];
          if (k > 0) s += std::conj(p.cp[2][k - 1]) * u[l - n2];
          if (k < nz - 1) s += std::conj(p.cm[2][k + 1]) * u[l + n2];
        }
        y[l] = s;
      }
}

// True operator A: REF = L0 - diag(delta-L) (implicit operator of
// ApplyCoeffMatrixA_CSR, functions.cpp:5554-5575); S7 = 7-point stencil.
static void op_apply(OrProblem& p, const cd* u, cd* y, bool adjoint) {
  if (p.d.mode == 1) { stencil_apply(p, u, y, adjoint); return; }
  if (adjoint) {
    // (L0 - D)^H = L0^H - D^H; L0^H = F^{-1} diag(M_k^H) F.  M_k^H applied via
    // the transposed-conjugate 5-point stencil.
    std::vector<cd> t(p.N), s(p.N);
    fft_z(p, u, t.data(), -1);
    int64_t nx = p.nx, ny = p.ny;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < p.nz; ++k) {
      if (!p.solve[k]) continue;  // skipped planes are zero in L0 (and L0^H)
      const cd* x = &t[k * p.n2];
      cd* yy = &s[k * p.n2];
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
          int64_t l = i + nx * j;
          cd diag = (p.cdg[0][i] + p.cdg[1][j]) + plane_kappa(p, k, l);
          cd a = std::conj(diag) * x[l];
          if (j > 0) a += std::conj(p.cp[1][j - 1]) * x[l - nx];
          if (i > 0) a += std::conj(p.cp[0][i - 1]) * x[l - 1];
          if (i < nx - 1) a += std::conj(p.cm[0][i + 1]) * x[l + 1];
          if (j < ny - 1) a += std::conj(p.cm[1][j + 1]) * x[l + nx];
          yy[l] = a;
        }
    }
    fft_z(p, s.data(), y, +1);
    for (int64_t l = 0; l < p.N; ++l) y[l] -= std::conj(p.dL[l]) * u[l];
    return;
  }
  L0_apply(p, u, y);
  for (int64_t l = 0; l < p.N; ++l) y[l] -= p.dL[l] * u[l];
}

// Right-preconditioned operator in w-space.
// REF: g = w - deltaL * L0^{-1} w  (ApplyCoeffMatrixA_CSR, functions.cpp:5554-5575,
//      with OpTwoMatrices '*' and '-', functions.cpp:6852-6858, 6844-6850).
// S7:  g = A L0^{-1} w.
static void op_pre_apply(OrProblem& p, const cd* w, cd* g) {
  if (p.d.mode == 0) {
    precond_apply(p, w, g, false);
    for (int64_t l = 0; l < p.N; ++l) g[l] = w[l] - p.dL[l] * g[l];
  } else {
    std::vector<cd> t(p.N);
    precond_apply(p, w, t.data(), false);
    stencil_apply(p, t.data(), g, false);
  }
}

// ------------------------------------------------------------ vector helpers
static double nrm2(int64_t n, const cd* x) {
  // dznrm2-equivalent (scaled sum of squares).
  double scale = 0.0, ssq = 1.0;
  for (int64_t i = 0; i < n; ++i) {
    double v[2] = {x[i].real(), x[i].imag()};
    for (double a : v) {
      if (a != 0.0) {
        double t = std::fabs(a);
        if (scale < t) { ssq = 1.0 + ssq * (scale / t) * (scale / t); scale = t; }
        else ssq += (t / scale) * (t / scale);
      }
    }
  }
  return scale * std::sqrt(ssq);
}

// zdotc(x, y) = sum conj(x) y
static cd dotc(int64_t n, const cd* x, const cd* y) {
  cd s(0, 0);
  for (int64_t i = 0; i < n; ++i) s += std::conj(x[i]) * y[i];
  return s;
}

// Physical-box mask: reducePML3D (functions.cpp:238-263) keeps
// pml <= i < n - pml in x, y and spg <= k < nz - spg in z; here both layer
// counts are removed on every axis.
static bool in_phys(const OrProblem& p, int64_t l) {
  int64_t k = l / p.n2, r = l - k * p.n2, j = r / p.nx, i = r - j * p.nx;
  const or_desc& d = p.d;
  int64_t lo[3], hi[3], idx[3] = {i, j, k};
  for (int a = 0; a < 3; ++a) {
    lo[a] = d.pml[a] + d.spg[a];
    hi[a] = d.n[a] - d.pml[a] - d.spg[a];
    if (idx[a] < lo[a] || idx[a] >= hi[a]) return false;
  }
  return true;
}

static double masked_nrm2(const OrProblem& p, const cd* x, int mask) {
  if (mask == 0) return nrm2(p.N, x);
  std::vector<cd> t;
  t.reserve(p.N);
  for (int64_t l = 0; l < p.N; ++l) if (in_phys(p, l)) t.push_back(x[l]);
  return nrm2((int64_t)t.size(), t.data());
}

// Householder QR least squares min ||b - H y||, H m x n (m >= n), column-major
// with leading dimension ld; the zgels call at FGMRES.cpp:567.  y returned in
// b[0..n).  Overwrites H.
static void lsq_householder(int64_t m, int64_t n, cd* H, int64_t ld, cd* b) {
  for (int64_t j = 0; j < n; ++j) {
    double nrm = 0.0;
    for (int64_t i = j; i < m; ++i) nrm += std::norm(H[i + ld * j]);
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) continue;
    cd x0 = H[j + ld * j];
    double ax0 = std::abs(x0);
    cd phase = (ax0 == 0.0) ? cd(1, 0) : x0 / ax0;
    cd alpha = -phase * nrm;
    // v = x - alpha e1
    std::vector<cd> v(m - j);
    for (int64_t i = j; i < m; ++i) v[i - j] = H[i + ld * j];
    v[0] -= alpha;
    double vn = 0.0;
    for (auto& e : v) vn += std::norm(e);
    if (vn == 0.0) continue;
    // apply (I - 2 v v^H / v^H v) to columns j..n-1 and to b
    for (int64_t c = j; c < n; ++c) {
      cd s(0, 0);
      for (int64_t i = j; i < m; ++i) s += std::conj(v[i - j]) * H[i + ld * c];
      s *= 2.0 / vn;
      for (int64_t i = j; i < m; ++i) H[i + ld * c] -= s * v[i - j];
    }
    cd s(0, 0);
    for (int64_t i = j; i < m; ++i) s += std::conj(v[i - j]) * b[i];
    s *= 2.0 / vn;
    for (int64_t i = j; i < m; ++i) b[i] -= s * v[i - j];
  }
  for (int64_t j = n - 1; j >= 0; --j) {
    cd s = b[j];
    for (int64_t c = j + 1; c < n; ++c) s -= H[j + ld * c] * b[c];
    b[j] = s / H[j + ld * j];
  }
}

// ------------------------------------------------------------- Krylov drivers
struct Hist {
  double* buf; int cap; int n;
  void push(double v) { if (buf && n < cap) buf[n] = v; ++n; }
};

// Restarted right-preconditioned GMRES on (I - deltaL L0^{-1}) w = f exactly as
// FGMRES.cpp:458-661:
//  * Arnoldi coefficient zdotc(w, V_i) = w^H V_i (ortho 0, FGMRES.cpp:511) or
//    the textbook V_i^H w (ortho 1);
//  * zgels least squares on the (j+2)x(j+1) Hessenberg with rhs beta e1
//    (FGMRES.cpp:551-569) and x_k = x_0 + V y (FGMRES.cpp:575);
//  * stopping on the true residual ||A_pre x_k - f|| over the mask divided by
//    the full-domain ||f|| (FGMRES.cpp:579-592, 650);
//  * breakdown |h_{j+1,j}| < thresh (FGMRES.cpp:519-527);
//  * u = L0^{-1} x (FGMRES.cpp:601).
// In S7 mode the same driver runs on A L0^{-1}.
static void gmres_ref(OrProblem& p, const cd* f, cd* u, int m, int rest, double thresh, double tol,
                      int mask, int ortho, int monitor, Hist& hist, or_stats* st) {
  int64_t N = p.N;
  std::vector<cd> x0(N), x_init(N, cd(0, 0)), w(N), work(N), V((size_t)N * (m + 1));
  std::vector<cd> H((size_t)(m + 1) * m, cd(0, 0)), Hg((size_t)(m + 1) * m), eb(m + 1);
  std::vector<cd> xs_prev, xs;
  int ldh = m + 1;
  int finish = 0, total = 0;
  double Res = 0, RelRes = 0, beta = 0, diff = 0;
  memset(st, 0, sizeof(*st));
  double norm_f = nrm2(N, f);
  for (int r = 0; r < rest; ++r) {
    if (finish) continue;
    st->restarts++;
    x0 = x_init;
    op_pre_apply(p, x0.data(), w.data());
    for (int64_t l = 0; l < N; ++l) work[l] = f[l] - w[l];
    beta = nrm2(N, work.data());
    for (int64_t l = 0; l < N; ++l) V[l] = work[l] / beta;
    for (int j = 0; j < m; ++j) {
      op_pre_apply(p, &V[(size_t)N * j], w.data());
      total++;
      for (int i = 0; i <= j; ++i) {
        cd hij = (ortho == 0) ? dotc(N, w.data(), &V[(size_t)N * i]) : dotc(N, &V[(size_t)N * i], w.data());
        H[i + ldh * j] = hij;
        for (int64_t l = 0; l < N; ++l) w[l] -= hij * V[(size_t)N * i + l];
      }
      double hn = nrm2(N, w.data());
      H[j + 1 + ldh * j] = hn;
      if (std::fabs(hn) < thresh) { st->breakdown = 1; break; }
      double da = 1.0 / hn;
      for (int64_t l = 0; l < N; ++l) V[(size_t)N * (j + 1) + l] = w[l] * da;
      x0 = x_init;
      std::fill(eb.begin(), eb.end(), cd(0, 0));
      eb[0] = beta;
      Hg = H;
      lsq_householder(j + 2, j + 1, Hg.data(), ldh, eb.data());
      for (int i = 0; i <= j; ++i)
        for (int64_t l = 0; l < N; ++l) x0[l] += eb[i] * V[(size_t)N * i + l];
      op_pre_apply(p, x0.data(), w.data());
      for (int64_t l = 0; l < N; ++l) w[l] -= f[l];
      Res = masked_nrm2(p, w.data(), mask);
      RelRes = Res / norm_f;
      hist.push(RelRes);
      if (monitor) {
        xs.resize(N);
        if (p.d.mode == 0) precond_apply(p, x0.data(), xs.data(), false);
        else precond_apply(p, x0.data(), xs.data(), false);
        if (!xs_prev.empty()) {
          std::vector<cd> a, b;
          for (int64_t l = 0; l < N; ++l) if (in_phys(p, l)) { a.push_back(xs[l] - xs_prev[l]); b.push_back(xs_prev[l]); }
          diff = nrm2((int64_t)a.size(), a.data()) / nrm2((int64_t)b.size(), b.data());
        }
        xs_prev = xs;
      }
      if (RelRes < tol) { finish = 1; break; }
    }
    x_init = x0;
  }
  precond_apply(p, x0.data(), u, false);
  st->iters = total; st->converged = finish; st->rel_res = RelRes; st->res = Res;
  st->diff_sol = diff; st->n_hist = hist.n;
}

// Flexible GMRES (Saad) with modified Gram-Schmidt and complex Givens rotations,
// right preconditioner M^{-1} = L0^{-1} stored per step (Z), on the true operator
// A (or A^H with (L0^{-1})^H when adjoint).  Inner stop on the Givens estimate,
// restart stop on the true residual over the mask.
static void fgmres(OrProblem& p, const cd* f, cd* u, int m, int rest, double tol, int mask,
                   int adjoint, Hist& hist, or_stats* st) {
  int64_t N = p.N;
  std::vector<cd> x(N, cd(0, 0)), r(N), w(N), V((size_t)N * (m + 1)), Z((size_t)N * m);
  std::vector<cd> H((size_t)(m + 1) * m, cd(0, 0)), g(m + 1), cs(m), sn(m);
  int ldh = m + 1, total = 0, conv = 0;
  double rel = 0, res = 0;
  memset(st, 0, sizeof(*st));
  double norm_f = nrm2(N, f);
  if (norm_f == 0.0) { std::fill(u, u + N, cd(0, 0)); st->converged = 1; return; }
  for (int cyc = 0; cyc < rest; ++cyc) {
    op_apply(p, x.data(), r.data(), adjoint);
    for (int64_t l = 0; l < N; ++l) r[l] = f[l] - r[l];
    res = masked_nrm2(p, r.data(), mask);
    rel = res / norm_f;
    if (rel < tol) { conv = 1; break; }
    st->restarts++;
    double beta = nrm2(N, r.data());
    for (int64_t l = 0; l < N; ++l) V[l] = r[l] / beta;
    std::fill(g.begin(), g.end(), cd(0, 0));
    g[0] = beta;
    int jj = 0;
    for (int j = 0; j < m; ++j) {
      precond_apply(p, &V[(size_t)N * j], &Z[(size_t)N * j], adjoint != 0);
      op_apply(p, &Z[(size_t)N * j], w.data(), adjoint);
      total++;
      for (int i = 0; i <= j; ++i) {
        cd hij = dotc(N, &V[(size_t)N * i], w.data());
        H[i + ldh * j] = hij;
        for (int64_t l = 0; l < N; ++l) w[l] -= hij * V[(size_t)N * i + l];
      }
      double hn = nrm2(N, w.data());
      H[j + 1 + ldh * j] = hn;
      for (int i = 0; i < j; ++i) {
        cd a = H[i + ldh * j], b = H[i + 1 + ldh * j];
        H[i + ldh * j] = cs[i] * a + sn[i] * b;
        H[i + 1 + ldh * j] = -std::conj(sn[i]) * a + cs[i] * b;
      }
      cd a = H[j + ldh * j], b = H[j + 1 + ldh * j];
      double aa = std::abs(a), bb = std::abs(b);
      if (bb == 0.0) { cs[j] = 1.0; sn[j] = 0.0; }
      else if (aa == 0.0) { cs[j] = 0.0; sn[j] = std::conj(b) / bb; }
      else {
        double t = std::sqrt(aa * aa + bb * bb);
        cs[j] = aa / t;
        sn[j] = (a / aa) * std::conj(b) / t;
      }
      H[j + ldh * j] = cs[j] * a + sn[j] * b;
      H[j + 1 + ldh * j] = 0.0;
      g[j + 1] = -std::conj(sn[j]) * g[j];
      g[j] = cs[j] * g[j];
      jj = j + 1;
      double est = std::abs(g[j + 1]) / norm_f;
      hist.push(est);
      if (est < tol || hn == 0.0) break;
      for (int64_t l = 0; l < N; ++l) V[(size_t)N * (j + 1) + l] = w[l] / hn;
    }
    std::vector<cd> y(jj);
    for (int i = jj - 1; i >= 0; --i) {
      cd s = g[i];
      for (int c = i + 1; c < jj; ++c) s -= H[i + ldh * c] * y[c];
      y[i] = s / H[i + ldh * i];
    }
    for (int i = 0; i < jj; ++i)
      for (int64_t l = 0; l < N; ++l) x[l] += y[i] * Z[(size_t)N * i + l];
  }
  if (!conv) {
    op_apply(p, x.data(), r.data(), adjoint);
    for (int64_t l = 0; l < N; ++l) r[l] = f[l] - r[l];
    res = masked_nrm2(p, r.data(), mask);
    rel = res / norm_f;
    conv = rel < tol;
  }
  std::copy(x.begin(), x.end(), u);
  st->iters = total; st->converged = conv; st->rel_res = rel; st->res = res; st->n_hist = hist.n;
}

// Right-preconditioned BiCGStab, BcGSTAB.cpp:397-568: w-space recurrences on
// A_pre with shadow residual 2 r_0 (:430-432) and ratio_dot coefficients
// (functions.cpp:2803-2813); x updated with L0^{-1} p and L0^{-1} s; residual
// A_pre(L0 x) - f over the mask (:495-508); stop when the residual
// (absolute when tol_abs, else relative to ||f||) drops below tol (:564).
static void bicgstab(OrProblem& p, const cd* f, cd* u, int maxit, double tol, int tol_abs, int mask,
                     Hist& hist, or_stats* st) {
  int64_t N = p.N;
  std::vector<cd> x0(N, cd(0, 0)), p0(N), s0(N), yn(N), zn(N), Ayn(N), Azn(N), r0(N), rn(N), r0st(N), g(N), w(N);
  memset(st, 0, sizeof(*st));
  st->restarts = 1;
  double norm_f = nrm2(N, f);
  op_pre_apply(p, x0.data(), w.data());
  for (int64_t l = 0; l < N; ++l) r0[l] = f[l] - w[l];
  p0 = r0;
  for (int64_t l = 0; l < N; ++l) r0st[l] = 2.0 * r0[l];
  double Res = 0, RelRes = 0;
  int it = 0, conv = 0;
  for (int j = 0; j < maxit; ++j) {
    op_pre_apply(p, p0.data(), Ayn.data());
    precond_apply(p, p0.data(), yn.data(), false);
    cd alpha_n = -(dotc(N, r0st.data(), r0.data()) / dotc(N, r0st.data(), Ayn.data()));
    for (int64_t l = 0; l < N; ++l) s0[l] = r0[l] + alpha_n * Ayn[l];
    op_pre_apply(p, s0.data(), Azn.data());
    precond_apply(p, s0.data(), zn.data(), false);
    cd omega_n = -(dotc(N, Azn.data(), s0.data()) / dotc(N, Azn.data(), Azn.data()));
    for (int64_t l = 0; l < N; ++l) rn[l] = s0[l] + omega_n * Azn[l];
    alpha_n = -alpha_n;
    omega_n = -omega_n;
    for (int64_t l = 0; l < N; ++l) x0[l] += alpha_n * yn[l] + omega_n * zn[l];
    cd beta_n = (alpha_n / omega_n) * (dotc(N, r0st.data(), rn.data()) / dotc(N, r0st.data(), r0.data()));
    omega_n = -omega_n;
    for (int64_t l = 0; l < N; ++l) p0[l] = rn[l] + beta_n * (p0[l] + omega_n * Ayn[l]);
    r0 = rn;
    it++;
    if (p.d.mode == 0) {
      L0_apply(p, x0.data(), g.data());
      op_pre_apply(p, g.data(), w.data());
    } else {
      op_apply(p, x0.data(), w.data(), 0);
    }
    for (int64_t l = 0; l < N; ++l) w[l] -= f[l];
    Res = masked_nrm2(p, w.data(), mask);
    RelRes = Res / norm_f;
    hist.push(RelRes);
    if ((tol_abs ? Res : RelRes) < tol) { conv = 1; break; }
  }
  std::copy(x0.begin(), x0.end(), u);
  st->iters = it; st->converged = conv; st->res = Res; st->rel_res = RelRes; st->n_hist = hist.n;
}

// ------------------------------------------------------------------- C API
extern "C" {

void* or_create(const or_desc* d, const double* c) {
  OrProblem* p = new OrProblem();
  p->d = *d;
  p->nx = d->n[0]; p->ny = d->n[1]; p->nz = d->n[2];
  p->n2 = p->nx * p->ny; p->N = p->n2 * p->nz;
  p->c.assign(c, c + p->N);
  for (int a = 0; a < 3; ++a) axis_tables(*d, a, p->cm[a], p->cp[a], p->cdg[a]);
  // In REF mode z carries no PML (main.cpp:126): the z tables are unused there.
  compute_c0(*p);
  p->lamz.resize(p->nz);
  for (int64_t k = 0; k < p->nz; ++k) p->lamz[k] = z_symbol(*d, k);
  plane_flags(*p);
  p->kappa.resize(p->n2);
  for (int64_t ij = 0; ij < p->n2; ++ij) {
    cd kx = (d->c_bg > 0.0) ? cd(d->omega / d->c_bg, 0.0) : d->omega / p->c0[ij];
    p->kappa[ij] = kx * kx * cd(1.0, d->beta);
  }
  if (d->mode == 0) compute_deltaL(*p);
  return p;
}

void or_destroy(void* h) { delete (OrProblem*)h; }

void or_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void or_alpha(int64_t n, int pml, double omega, double pml_c, double s, double* out) {
  cd a = alpha_at(n, pml, omega, pml_c, s);
  out[0] = a.real(); out[1] = a.imag();
}

void or_tables(void* h, int ax, double* cm, double* cp, double* cdg) {
  OrProblem* p = (OrProblem*)h;
  for (int64_t i = 0; i < p->d.n[ax]; ++i) {
    cm[2 * i] = p->cm[ax][i].real(); cm[2 * i + 1] = p->cm[ax][i].imag();
    cp[2 * i] = p->cp[ax][i].real(); cp[2 * i + 1] = p->cp[ax][i].imag();
    cdg[2 * i] = p->cdg[ax][i].real(); cdg[2 * i + 1] = p->cdg[ax][i].imag();
  }
}

void or_c0(void* h, double* out) { OrProblem* p = (OrProblem*)h; memcpy(out, p->c0.data(), p->n2 * sizeof(cd)); }
void or_kappa(void* h, double* out) { OrProblem* p = (OrProblem*)h; memcpy(out, p->kappa.data(), p->n2 * sizeof(cd)); }
void or_plane_solved(void* h, int* out) {
  OrProblem* p = (OrProblem*)h;
  for (int64_t k = 0; k < p->nz; ++k) out[k] = p->solve[k];
}
void or_lamz(void* h, double* out) { OrProblem* p = (OrProblem*)h; memcpy(out, p->lamz.data(), p->nz * sizeof(double)); }
void or_deltaL(void* h, double* out) {
  OrProblem* p = (OrProblem*)h;
  if (p->dL.empty()) memset(out, 0, p->N * sizeof(cd));
  else memcpy(out, p->dL.data(), p->N * sizeof(cd));
}

void or_fft_line(int64_t n, const double* in, double* out, int sign) {
  std::vector<cd> scratch;
  fft_rec((const cd*)in, 1, n, (cd*)out, sign, scratch);
  if (sign > 0) for (int64_t k = 0; k < n; ++k) ((cd*)out)[k] /= (double)n;
}
void or_dft_brute(int64_t n, const double* in, double* out, int sign) { dft_brute((const cd*)in, n, (cd*)out, sign); }
void or_fft_z(void* h, const double* in, double* out, int sign) { fft_z(*(OrProblem*)h, (const cd*)in, (cd*)out, sign); }

void or_plane_matvec(void* h, int64_t k, const double* in, double* out) {
  plane_matvec(*(OrProblem*)h, k, (const cd*)in, (cd*)out);
}
void or_plane_solve(void* h, int64_t k, double* b, int adjoint) {
  OrProblem* p = (OrProblem*)h;
  ensure_factored(*p);
  plane_solve(*p, k, (cd*)b, adjoint != 0);
}
void or_factor(void* h) { ensure_factored(*(OrProblem*)h); }
// Factor only planes [k0, k1) (CPU baseline sampling): holds just those planes
// and leaves the problem unfactored for the full-path functions.
void or_factor_planes(void* h, int64_t k0, int64_t k1) {
  OrProblem* p = (OrProblem*)h;
  p->factored = false;
  p->k_off = k0;
  p->kl = p->nx; p->ku = p->nx; p->ldab = 2 * p->kl + p->ku + 1;
  p->ab.assign((size_t)((k1 - k0) * p->ldab * p->n2), cd(0, 0));
  p->ipiv.assign((size_t)((k1 - k0) * p->n2), 0);
#pragma omp parallel for schedule(dynamic)
  for (int64_t k = k0; k < k1; ++k) plane_factor(*p, k);
}
void or_solve_planes(void* h, int64_t k0, int64_t k1, double* b) {
  OrProblem* p = (OrProblem*)h;
#pragma omp parallel for schedule(dynamic)
  for (int64_t k = k0; k < k1; ++k) plane_solve(*p, k, (cd*)b + (k - k0) * p->n2, false);
}
void or_precond_apply(void* h, const double* in, double* out, int adjoint) {
  precond_apply(*(OrProblem*)h, (const cd*)in, (cd*)out, adjoint != 0);
}
void or_L0_apply(void* h, const double* in, double* out) { L0_apply(*(OrProblem*)h, (const cd*)in, (cd*)out); }
void or_op_apply(void* h, const double* in, double* out, int adjoint) {
  op_apply(*(OrProblem*)h, (const cd*)in, (cd*)out, adjoint);
}
void or_op_pre_apply(void* h, const double* in, double* out) {
  op_pre_apply(*(OrProblem*)h, (const cd*)in, (cd*)out);
}
double or_masked_nrm2(void* h, const double* x, int mask) { return masked_nrm2(*(OrProblem*)h, (const cd*)x, mask); }
void or_lsq(int64_t m, int64_t n, double* H, int64_t ld, double* b) { lsq_householder(m, n, (cd*)H, ld, (cd*)b); }

void or_gmres_ref(void* h, const double* f, double* u, int m, int rest, double thresh, double tol,
                  int mask, int ortho, int monitor, double* hist, int cap, or_stats* st) {
  Hist hs{hist, cap, 0};
  gmres_ref(*(OrProblem*)h, (const cd*)f, (cd*)u, m, rest, thresh, tol, mask, ortho, monitor, hs, st);
}
void or_fgmres(void* h, const double* f, double* u, int m, int rest, double tol, int mask, int adjoint,
               double* hist, int cap, or_stats* st) {
  Hist hs{hist, cap, 0};
  fgmres(*(OrProblem*)h, (const cd*)f, (cd*)u, m, rest, tol, mask, adjoint, hs, st);
}
void or_bicgstab(void* h, const double* f, double* u, int maxit, double tol, int tol_abs, int mask,
                 double* hist, int cap, or_stats* st) {
  Hist hs{hist, cap, 0};
  bicgstab(*(OrProblem*)h, (const cd*)f, (cd*)u, maxit, tol, tol_abs, mask, hs, st);
}

// Receiver misfit J = 1/2 sum_s ||P u_s - d_s||^2 and its gradient with respect
// to m = 1/c^2 (S7 operator A = S + w^2 diag(m)):
//   A u_s = f_s,  A^H lambda_s = P^T (P u_s - d_s),
//   dJ/dm = -w^2 sum_s Re(conj(u_s) lambda_s).
// Forward and adjoint solves use flexible GMRES with L0^{-1} / (L0^{-1})^H.
int or_misfit_grad(void* h, int nsrc, const double* f, int nrec, const int64_t* rec, const double* data,
                   int m, int rest, double tol, double* J, double* grad, double* u_out) {
  OrProblem& p = *(OrProblem*)h;
  int64_t N = p.N;
  double w2 = p.d.omega * p.d.omega;
  std::vector<cd> u(N), lam(N), rhs(N);
  std::fill(grad, grad + N, 0.0);
  double Jt = 0.0;
  int ok = 1;
  for (int s = 0; s < nsrc; ++s) {
    Hist hs{nullptr, 0, 0};
    or_stats st;
    fgmres(p, (const cd*)f + (size_t)N * s, u.data(), m, rest, tol, 0, 0, hs, &st);
    ok &= st.converged;
    if (u_out) memcpy(u_out + (size_t)2 * N * s, u.data(), N * sizeof(cd));
    std::fill(rhs.begin(), rhs.end(), cd(0, 0));
    const cd* d = (const cd*)data + (size_t)nrec * s;
    for (int q = 0; q < nrec; ++q) {
      cd r = u[rec[q]] - d[q];
      Jt += 0.5 * std::norm(r);
      rhs[rec[q]] += r;
    }
    fgmres(p, rhs.data(), lam.data(), m, rest, tol, 0, 1, hs, &st);
    ok &= st.converged;
    for (int64_t l = 0; l < N; ++l) grad[l] += -w2 * (std::conj(u[l]) * lam[l]).real();
  }
  *J = Jt;
  return ok;
}

}  // extern "C"
