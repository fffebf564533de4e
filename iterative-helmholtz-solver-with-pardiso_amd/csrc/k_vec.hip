// k_vec.hip -- FP64 complex vector kernels of the Krylov drivers.
//
// Dots and norms (zdotc / dznrm2 in FGMRES.cpp:479-519) reduce in two stages
// with a fixed block count, so results do not depend on scheduling: stage 1
// writes one partial per block, stage 2 sums them in index order.  The
// Gram-Schmidt coefficients stay in device memory between the dot and the
// axpy, so an Arnoldi step needs no host round trip.
#include "hz_internal.h"

namespace hz {
namespace {

constexpr int RB = 1024;  // stage-1 blocks per vector
constexpr int RBA = 512;  // blocks of the Gram-Schmidt update pass (measured best at 200^3: 1024 -> 512)
constexpr int RT = 256;   // threads per block

struct Box {
  int nx, ny, nz;
  int x0, x1, y0, y1, z0, z1;
  int on;
};

__device__ __forceinline__ bool in_box(const Box& b, int64_t l) {
  if (!b.on) return true;
  const int64_t n2 = (int64_t)b.nx * b.ny;
  const int k = (int)(l / n2);
  const int64_t r = l - (int64_t)k * n2;
  const int j = (int)(r / b.nx), i = (int)(r - (int64_t)j * b.nx);
  return i >= b.x0 && i < b.x1 && j >= b.y0 && j < b.y1 && k >= b.z0 && k < b.z1;
}

// Basis vectors are streamed once per pass and never reused from cache: the
// Gram-Schmidt passes load them nontemporally (the update pass runs ~10-13%
// faster at 200^3 than with ordinary loads).
__device__ __forceinline__ double2 ldv(const double2* p) {
  typedef double dv2 __attribute__((ext_vector_type(2)));
  const dv2 t = __builtin_nontemporal_load(reinterpret_cast<const dv2*>(p));
  return make_double2(t.x, t.y);
}

__device__ __forceinline__ double2 block_sum(double2 v) {
  __shared__ double2 sm[RT / 64];
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off);
    v.y += __shfl_xor(v.y, off);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = v;
  __syncthreads();
  double2 s = make_double2(0, 0);
  if (threadIdx.x == 0)
    for (int i = 0; i < RT / 64; ++i) { s.x += sm[i].x; s.y += sm[i].y; }
  return s;
}

// partial[v * RB + b] = sum conj(x_v) y_v  (y == nullptr: sum |x_v|^2)
__global__ __launch_bounds__(RT) void k_dot_partial(int64_t n, const double2* __restrict__ x, int64_t xs,
                                                    const double2* __restrict__ y, int64_t ys, Box box,
                                                    double2* __restrict__ partial) {
  const int v = blockIdx.y;
  const double2* xv = x + (int64_t)v * xs;
  const double2* yv = y ? y + (int64_t)v * ys : nullptr;
  double sx = 0.0, sy = 0.0;
  const int64_t stride = (int64_t)RB * RT;
  for (int64_t l = (int64_t)blockIdx.x * RT + threadIdx.x; l < n; l += stride) {
    if (!in_box(box, l)) continue;
    const double2 a = xv[l];
    if (yv) {
      const double2 b = yv[l];
      sx += a.x * b.x + a.y * b.y;
      sy += a.x * b.y - a.y * b.x;
    } else {
      sx += a.x * a.x + a.y * a.y;
    }
  }
  const double2 s = block_sum(make_double2(sx, sy));
  if (threadIdx.x == 0) partial[(int64_t)v * RB + blockIdx.x] = s;
}

// out[v] = sum of RB partials (mode 1: sqrt of the real part, imaginary 0)
// sums of nb partials per vector
__global__ __launch_bounds__(RT) void k_dot_final(const double2* __restrict__ partial, int mode,
                                                  double2* __restrict__ out, int nb, int64_t pz, int64_t oz,
                                                  uint64_t mask) {
  // blockIdx.z: a batch member (partials pz, outputs oz apart; members not in mask skip)
  if (!((mask >> blockIdx.z) & 1ull)) return;
  partial += (int64_t)blockIdx.z * pz;
  out += (int64_t)blockIdx.z * oz;
  const int v = blockIdx.x;
  double sx = 0.0, sy = 0.0;
  for (int b = threadIdx.x; b < nb; b += RT) {
    sx += partial[(int64_t)v * nb + b].x;
    sy += partial[(int64_t)v * nb + b].y;
  }
  const double2 s = block_sum(make_double2(sx, sy));
  if (threadIdx.x == 0) out[v] = mode == 1 ? make_double2(sqrt(s.x), 0.0) : s;
}

// w -= coef * x, then partial[b] = sum conj(y) w  (y == nullptr: sum |w|^2):
// one modified Gram-Schmidt update fused with the next coefficient's dot.
__global__ __launch_bounds__(RT) void k_axpy_dot_partial(int64_t n, const double2* __restrict__ coef,
                                                         const double2* __restrict__ x, double2* __restrict__ w,
                                                         const double2* __restrict__ y,
                                                         double2* __restrict__ partial) {
  const double2 a = coef[0];
  double sx = 0.0, sy = 0.0;
  const int64_t stride = (int64_t)RB * RT;
  for (int64_t l = (int64_t)blockIdx.x * RT + threadIdx.x; l < n; l += stride) {
    const double2 xv = x[l];
    double2 wv = w[l];
    wv.x -= a.x * xv.x - a.y * xv.y;
    wv.y -= a.x * xv.y + a.y * xv.x;
    w[l] = wv;
    if (y) {
      const double2 b = y[l];
      sx += b.x * wv.x + b.y * wv.y;
      sy += b.x * wv.y - b.y * wv.x;
    } else {
      sx += wv.x * wv.x + wv.y * wv.y;
    }
  }
  const double2 s = block_sum(make_double2(sx, sy));
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}


// ------------------------------------------------ one-pass Gram-Schmidt blocks
// Multi-dot over a group of MD vectors: r_q = V_q^H w and (vj != nullptr)
// l_q = V_q^H vj for q in the group; partial[(group * RB + block) * 2 MD + ...].
constexpr int MD = 12;

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ __launch_bounds__(RT) void k_multidot_partial(int64_t n, int nv, const double2* __restrict__ V,
                                                         int64_t vs, const double2* __restrict__ w,
                                                         const double2* __restrict__ vj,
                                                         double2* __restrict__ partial, int64_t ms, int64_t ps,
                                                         uint64_t mask) {
  // blockIdx.z: a batch member (vectors ms, partials ps apart; members not in mask skip)
  if (!((mask >> blockIdx.z) & 1ull)) return;
  V += (int64_t)blockIdx.z * ms;
  w += (int64_t)blockIdx.z * ms;
  if (vj) vj += (int64_t)blockIdx.z * ms;
  partial += (int64_t)blockIdx.z * ps;
  const int group = blockIdx.y;
  const int q0 = group * MD;
  const int nq = min(MD, nv - q0);
  double ar[MD][2], al[MD][2];
#pragma unroll
  for (int g = 0; g < MD; ++g) { ar[g][0] = ar[g][1] = al[g][0] = al[g][1] = 0.0; }
  const int64_t stride = (int64_t)RB * RT;
  for (int64_t l = (int64_t)blockIdx.x * RT + threadIdx.x; l < n; l += stride) {
    const double2 wv = w[l];
    const double2 jv = vj ? vj[l] : make_double2(0, 0);
#pragma unroll
    for (int g = 0; g < MD; ++g) {
      if (g < nq) {
        const double2 a = ldv(V + (int64_t)(q0 + g) * vs + l);
        ar[g][0] += a.x * wv.x + a.y * wv.y;
        ar[g][1] += a.x * wv.y - a.y * wv.x;
        al[g][0] += a.x * jv.x + a.y * jv.y;
        al[g][1] += a.x * jv.y - a.y * jv.x;
      }
    }
  }
  __shared__ double sm[RT / 64][MD * 4];
  const int lane = threadIdx.x & 63, wv_ = threadIdx.x >> 6;
#pragma unroll
  for (int g = 0; g < MD; ++g) {
    const double s0 = wave_sum(ar[g][0]), s1 = wave_sum(ar[g][1]);
    const double s2 = wave_sum(al[g][0]), s3 = wave_sum(al[g][1]);
    if (lane == 0) { sm[wv_][4 * g] = s0; sm[wv_][4 * g + 1] = s1; sm[wv_][4 * g + 2] = s2; sm[wv_][4 * g + 3] = s3; }
  }
  __syncthreads();
  if (threadIdx.x < MD * 4) {
    double t = 0.0;
    for (int i = 0; i < RT / 64; ++i) t += sm[i][threadIdx.x];
    // layout: [r_0..r_{MD-1}, l_0..l_{MD-1}] as double2
    const int g = threadIdx.x / 4, c = threadIdx.x % 4;
    double* out = reinterpret_cast<double*>(partial + ((int64_t)group * RB + blockIdx.x) * 2 * MD);
    const int slot = (c < 2) ? g : MD + g;
    out[2 * slot + (c & 1)] = t;
  }
}

__device__ __forceinline__ double basis_scale(const double2* S, int k) { return k == 0 ? 1.0 : S[k].x; }

// out_r[q] = s_q s_jv (sum of vector q's r partials), out_l likewise: the
// dots of the stored basis vectors U_q (v_q = s_q U_q, s_0 = 1, S[k].x = s_k)
__global__ __launch_bounds__(RT) void k_multidot_final(const double2* __restrict__ partial, int nv,
                                                       double2* __restrict__ out_r, double2* __restrict__ out_l,
                                                       int64_t ps, int64_t os, uint64_t mask,
                                                       const double2* __restrict__ S, int jv) {
  if (!((mask >> blockIdx.z) & 1ull)) return;
  partial += (int64_t)blockIdx.z * ps;
  out_r += (int64_t)blockIdx.z * os;
  if (out_l) out_l += (int64_t)blockIdx.z * os;
  S += (int64_t)blockIdx.z * os;
  const int q = blockIdx.x;
  const int group = q / MD, g = q % MD;
  double sx = 0, sy = 0, tx = 0, ty = 0;
  for (int b = threadIdx.x; b < RB; b += RT) {
    const double2* pb = partial + ((int64_t)group * RB + b) * 2 * MD;
    sx += pb[g].x; sy += pb[g].y;
    tx += pb[MD + g].x; ty += pb[MD + g].y;
  }
  const double2 s = block_sum(make_double2(sx, sy));
  __syncthreads();
  const double2 t = block_sum(make_double2(tx, ty));
  if (threadIdx.x == 0) {
    const double f = basis_scale(S, q) * basis_scale(S, jv);
    out_r[q] = make_double2(s.x * f, s.y * f);
    if (out_l) out_l[q] = make_double2(t.x * f, t.y * f);
  }
}

// y = s_jv w - sum_q coef_q s_q U_q (stored basis vectors U_q, v_q = s_q U_q),
// partial[b] = sum |y|^2
__global__ __launch_bounds__(RT) void k_multi_axpy_norm_partial(int64_t n, int nv, const double2* __restrict__ V,
                                                                int64_t vs, const double2* __restrict__ coef,
                                                                const double2* __restrict__ w,
                                                                double2* __restrict__ partial, int64_t ms,
                                                                int64_t os, uint64_t mask,
                                                                const double2* __restrict__ S, int jv,
                                                                double2* __restrict__ y) {
  // blockIdx.z: a batch member (vectors ms, coefficients os, partials gridDim.x apart)
  if (!((mask >> blockIdx.z) & 1ull)) return;
  V += (int64_t)blockIdx.z * ms;
  w += (int64_t)blockIdx.z * ms;
  y += (int64_t)blockIdx.z * ms;
  coef += (int64_t)blockIdx.z * os;
  S += (int64_t)blockIdx.z * os;
  partial += (int64_t)blockIdx.z * gridDim.x;
  extern __shared__ double2 cf[];
  for (int q = threadIdx.x; q < nv; q += RT) {
    const double2 c = coef[q];
    const double sq = basis_scale(S, q);
    cf[q] = make_double2(c.x * sq, c.y * sq);
  }
  const double sj = basis_scale(S, jv);
  __syncthreads();
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t l = (int64_t)blockIdx.x * RT + threadIdx.x; l < n; l += stride) {
    const double2 wl = w[l];
    double2 s = make_double2(wl.x * sj, wl.y * sj);
    int q = 0;
    for (; q + 3 < nv; q += 4) {
      const double2 a0 = ldv(V + (int64_t)q * vs + l), a1 = ldv(V + (int64_t)(q + 1) * vs + l);
      const double2 a2 = ldv(V + (int64_t)(q + 2) * vs + l), a3 = ldv(V + (int64_t)(q + 3) * vs + l);
      const double2 c0 = cf[q], c1 = cf[q + 1], c2 = cf[q + 2], c3 = cf[q + 3];
      s.x -= c0.x * a0.x - c0.y * a0.y; s.y -= c0.x * a0.y + c0.y * a0.x;
      s.x -= c1.x * a1.x - c1.y * a1.y; s.y -= c1.x * a1.y + c1.y * a1.x;
      s.x -= c2.x * a2.x - c2.y * a2.y; s.y -= c2.x * a2.y + c2.y * a2.x;
      s.x -= c3.x * a3.x - c3.y * a3.y; s.y -= c3.x * a3.y + c3.y * a3.x;
    }
    for (; q < nv; ++q) {
      const double2 a = ldv(V + (int64_t)q * vs + l), c = cf[q];
      s.x -= c.x * a.x - c.y * a.y;
      s.y -= c.x * a.y + c.y * a.x;
    }
    y[l] = s;
    acc += s.x * s.x + s.y * s.y;
  }
  const double2 t = block_sum(make_double2(acc, 0.0));
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void k_axpy_dev(int64_t n, const double2* __restrict__ coef, double sgn, int cj,
                           const double2* __restrict__ x, double2* __restrict__ y) {
  double2 a = coef[0];
  a.x *= sgn;
  a.y = (cj ? -a.y : a.y) * sgn;
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    const double2 xv = x[l];
    double2 yv = y[l];
    yv.x += a.x * xv.x - a.y * xv.y;
    yv.y += a.x * xv.y + a.y * xv.x;
    y[l] = yv;
  }
}

// y = x / coef.x; with GUARD, y = 0 unless coef.x > thr (a breakdown or a
// non-finite norm leaves a finite vector)
template <bool GUARD>
__global__ void k_scale_inv_dev(int64_t n, const double2* __restrict__ coef, double thr, const double2* __restrict__ x,
                                double2* __restrict__ y, int64_t xs, int64_t cs, uint64_t mask) {
  // blockIdx.z: a batch member (vectors xs, coefficients cs apart)
  if (!((mask >> blockIdx.z) & 1ull)) return;
  x += (int64_t)blockIdx.z * xs;
  y += (int64_t)blockIdx.z * xs;
  coef += (int64_t)blockIdx.z * cs;
  const double c = coef[0].x;
  const double s = GUARD ? (c > thr ? 1.0 / c : 0.0) : 1.0 / c;
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    const double2 xv = x[l];
    y[l] = make_double2(xv.x * s, xv.y * s);
  }
}

// Inverse compact-WY correction of Arnoldi step j: h = (I + L)^{-1} r, where
// L is the strict lower part of V^H V (row i holds V_i^H V_c, c < i).  Row j,
// conj(l_c) with l_c = V_c^H V_j from this step's dot pass, is stored into L
// (leading dimension ldl) for the later steps of the cycle.  One wave: the
// rows 0..j go to LDS, lane i holds h_i, and the forward substitution runs
// over the columns c = 0..j-1 (h_i -= L_ic h_c, c increasing), the finished
// h_c broadcast from lane c.  Needs j < 64 and (j + 1)^2 double2 of LDS.
__global__ __launch_bounds__(64) void k_icwy(int j, int ldl, const double2* __restrict__ r,
                                             const double2* __restrict__ l, double2* __restrict__ L,
                                             double2* __restrict__ h, int64_t os, uint64_t mask) {
  // blockIdx.z: a batch member (all four arrays os apart)
  if (!((mask >> blockIdx.z) & 1ull)) return;
  r += (int64_t)blockIdx.z * os;
  l += (int64_t)blockIdx.z * os;
  L += (int64_t)blockIdx.z * os;
  h += (int64_t)blockIdx.z * os;
  extern __shared__ double2 sL[];
  const int lane = threadIdx.x;
  const int n = j + 1;
  if (lane < j) L[(size_t)j * ldl + lane] = make_double2(l[lane].x, -l[lane].y);
  for (int e = lane; e < n * n; e += 64) {
    const int i = e / n, c = e - i * n;
    double2 v = make_double2(0.0, 0.0);
    if (c < i) v = i == j ? make_double2(l[c].x, -l[c].y) : L[(size_t)i * ldl + c];
    sL[e] = v;
  }
  __syncthreads();
  double2 hi = lane < n ? r[lane] : make_double2(0.0, 0.0);
  for (int c = 0; c < j; ++c) {
    const double hx = __shfl(hi.x, c), hy = __shfl(hi.y, c);
    if (lane > c && lane < n) {
      const double2 a = sL[lane * n + c];
      hi.x -= a.x * hx - a.y * hy;
      hi.y -= a.x * hy + a.y * hx;
    }
  }
  if (lane < n) h[lane] = hi;
}

__global__ void k_lincomb(int64_t n, int nv, const double2* __restrict__ c, const double2* __restrict__ X,
                          int64_t xs, double2* __restrict__ y, int acc) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    double2 s = acc ? y[l] : make_double2(0, 0);
    for (int v = 0; v < nv; ++v) {
      const double2 cv = c[v], xv = X[(int64_t)v * xs + l];
      s.x += cv.x * xv.x - cv.y * xv.y;
      s.y += cv.x * xv.y + cv.y * xv.x;
    }
    y[l] = s;
  }
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
// a / b by Smith's algorithm (scaled by the larger component of b)
__device__ __forceinline__ double2 zdiv(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.x * r + b.y;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

// Complex Givens least-squares update of Arnoldi step j of member q (the
// per-iteration rotation of FGMRES's Hessenberg column and of g = ||r|| e_1);
// H: column j of the Hessenberg matrix, cs / sn: the cosines and sines, g:
// the least-squares right-hand side (global memory or LDS copies).  The
// earlier rotations (c_i real, s_i complex) act on rows (i, i+1) as
// [c s; -conj(s) c]; the new one zeroes h_{j+1,j}.
__device__ void givens_member(int q, int j, int m, double2* H, double2* cs, double2* sn, double2* g,
                              const double2* __restrict__ beta, const double2* __restrict__ nf, double tol,
                              double brk, double2* __restrict__ est) {
  if (j == 0) {
    g[0] = make_double2(beta[q].x, 0.0);
    for (int i = 1; i <= m; ++i) g[i] = make_double2(0.0, 0.0);
  }
  const double hn = H[j + 1].x;
  for (int i = 0; i < j; ++i) {
    const double2 u = H[i], w2 = H[i + 1];
    const double c = cs[i].x;
    const double2 sv = sn[i];
    const double2 a = zmul(sv, w2), b = zmul(zconj(sv), u);
    H[i] = make_double2(c * u.x + a.x, c * u.y + a.y);
    H[i + 1] = make_double2(c * w2.x - b.x, c * w2.y - b.y);
  }
  const double2 u = H[j], w2 = H[j + 1];
  const double aa = hypot(u.x, u.y), bb = hypot(w2.x, w2.y);
  double c;
  double2 sv;
  if (bb == 0.0) {
    c = 1.0;
    sv = make_double2(0.0, 0.0);
  } else if (aa == 0.0) {
    c = 0.0;
    sv = make_double2(w2.x / bb, -w2.y / bb);
  } else {
    const double t = sqrt(aa * aa + bb * bb);
    c = aa / t;
    const double2 ph = make_double2(u.x / aa, u.y / aa);
    const double2 z = zmul(ph, zconj(w2));
    sv = make_double2(z.x / t, z.y / t);
  }
  cs[j] = make_double2(c, 0.0);
  sn[j] = sv;
  const double2 a = zmul(sv, w2);
  H[j] = make_double2(c * u.x + a.x, c * u.y + a.y);
  H[j + 1] = make_double2(0.0, 0.0);
  const double2 gj = g[j], b = zmul(zconj(sv), gj);
  g[j + 1] = make_double2(-b.x, -b.y);
  g[j] = make_double2(c * gj.x, c * gj.y);
  const double e = hypot(b.x, b.y) / nf[q].x;
  const double status = !isfinite(e) ? 3.0 : (e < tol ? 1.0 : (hn <= brk ? 2.0 : 0.0));
  est[(int64_t)q * m + j] = make_double2(e, status);
}

// The Givens update of step j, one thread per member q of `mask` (member q's
// scalars at base + q sb)
__global__ __launch_bounds__(64) void k_givens(int j, int m, int ldh, uint64_t mask, double2* __restrict__ base,
                                              int64_t sb, int offH, int offCS, int offSN, int offG,
                                              const double2* __restrict__ beta, const double2* __restrict__ nf,
                                              double tol, double brk, double2* __restrict__ est) {
  const int q = threadIdx.x;
  if (q >= 64 || !((mask >> q) & 1ull)) return;
  double2* mb = base + sb * q;
  givens_member(q, j, m, mb + offH + (int64_t)ldh * j, mb + offCS, mb + offSN, mb + offG, beta, nf, tol, brk, est);
}

// The end of Arnoldi step j in one block per member q = blockIdx.z of `mask`:
// h_{j+1,j} = sqrt(sum of the nb update-pass partials, partial + q nb), the
// next basis vector's scale S[j + 1].x = 1 / h_{j+1,j} (0 where h_{j+1,j} <=
// brk, a breakdown), then the member's Givens update on LDS copies of column j
// of H and of the rotations, so its serial loop waits on LDS (j < ICWY_MAX)
__global__ __launch_bounds__(RT) void k_norm_givens(const double2* __restrict__ partial, int nb, int j, int m,
                                                    int ldh, uint64_t mask, double2* __restrict__ base, int64_t sb,
                                                    int offH, int offS, int offCS, int offSN, int offG,
                                                    const double2* __restrict__ beta, const double2* __restrict__ nf,
                                                    double tol, double brk, double2* __restrict__ est) {
  __shared__ double2 sh[ICWY_MAX + 1], scs[ICWY_MAX], ssn[ICWY_MAX];
  const int q = blockIdx.z;
  if (!((mask >> q) & 1ull)) return;
  partial += (int64_t)q * nb;
  double2* mb = base + sb * q;
  double2* Hc = mb + offH + (int64_t)ldh * j;
  for (int i = threadIdx.x; i <= j; i += RT) sh[i] = Hc[i];
  for (int i = threadIdx.x; i < j; i += RT) {
    scs[i] = mb[offCS + i];
    ssn[i] = mb[offSN + i];
  }
  double sx = 0.0;
  for (int b = threadIdx.x; b < nb; b += RT) sx += partial[b].x;
  const double2 s = block_sum(make_double2(sx, 0.0));  // its barrier also publishes the copies
  if (threadIdx.x != 0) return;
  const double hn = sqrt(s.x);
  sh[j + 1] = make_double2(hn, 0.0);
  mb[offS + j + 1] = make_double2(hn > brk ? 1.0 / hn : 0.0, 0.0);
  givens_member(q, j, m, sh, scs, ssn, mb + offG, beta, nf, tol, brk, est);
  for (int i = 0; i <= j + 1; ++i) Hc[i] = sh[i];
  mb[offCS + j] = scs[j];
  mb[offSN + j] = ssn[j];
}

// y_i = (g_i - sum_{c > i} R_ic y_c) / R_ii for i = jj-1 .. 0 (R: the rotated
// Hessenberg matrix, leading dimension ldh); with S, each y_i is then
// multiplied by the basis scale s_i (the solution update combines the
// preconditioned stored vectors M U_i, v_i = s_i U_i)
__global__ void k_backsub(int jj, int ldh, const double2* __restrict__ H, const double2* __restrict__ g,
                          double2* __restrict__ y, const double2* __restrict__ S, int stage) {
  // stage: the block's threads copy R (jj x jj) into LDS first, and y is
  // formed there, so the serial loop waits on LDS rather than on memory
  extern __shared__ double2 sm[];
  const double2* R = H;
  double2* yy = y;
  int ld = ldh;
  if (stage) {
    for (int e = threadIdx.x; e < jj * jj; e += blockDim.x) {
      const int c = e / jj, i = e - c * jj;
      sm[e] = H[i + (int64_t)ldh * c];
    }
    __syncthreads();
    R = sm;
    yy = sm + jj * jj;
    ld = jj;
  }
  if (threadIdx.x != 0) return;
  for (int i = jj - 1; i >= 0; --i) {
    double2 s = g[i];
    for (int c = i + 1; c < jj; ++c) {
      const double2 t = zmul(R[i + (int64_t)ld * c], yy[c]);
      s.x -= t.x;
      s.y -= t.y;
    }
    yy[i] = zdiv(s, R[i + (int64_t)ld * i]);
  }
  for (int i = 0; i < jj; ++i) {
    const double f = S ? basis_scale(S, i) : 1.0;
    y[i] = S ? make_double2(yy[i].x * f, yy[i].y * f) : yy[i];
  }
}

__global__ void k_axpby(int64_t n, double2 a, const double2* __restrict__ x, double2 b, double2* __restrict__ y) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    const double2 xv = x[l], yv = y[l];
    y[l] = make_double2(a.x * xv.x - a.y * xv.y + b.x * yv.x - b.y * yv.y,
                        a.x * xv.y + a.y * xv.x + b.x * yv.y + b.y * yv.x);
  }
}

__global__ void k_axpbypcz(int64_t n, const double2* __restrict__ x, double2 a, const double2* __restrict__ y,
                           double2 b, const double2* __restrict__ w, double2* __restrict__ z) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    const double2 xv = x[l], yv = y[l], wv = w[l];
    z[l] = make_double2(xv.x + a.x * yv.x - a.y * yv.y + b.x * wv.x - b.y * wv.y,
                        xv.y + a.x * yv.y + a.y * yv.x + b.x * wv.y + b.y * wv.x);
  }
}

__global__ void k_sub(int64_t n, const double2* __restrict__ a, const double2* __restrict__ b, double2* __restrict__ o) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    const double2 av = a[l], bv = b[l];
    o[l] = make_double2(av.x - bv.x, av.y - bv.y);
  }
}

__global__ void k_w_minus_dl(int64_t n, int64_t N, const double2* __restrict__ w, const double2* __restrict__ dL,
                             const double2* __restrict__ x, double2* __restrict__ o) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    const double2 wv = w[l], d = dL[l % N], xv = x[l];
    o[l] = make_double2(wv.x - (d.x * xv.x - d.y * xv.y), wv.y - (d.x * xv.y + d.y * xv.x));
  }
}

__global__ void k_grad_acc(int64_t n, double scale, const double2* __restrict__ u, const double2* __restrict__ lam,
                           double* __restrict__ g) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = u[l], b = lam[l];
    g[l] += scale * (a.x * b.x + a.y * b.y);
  }
}

// grad[l] += scale * sum_s Re(conj(u_s[l]) lam_s[l]) over ns fields (one pass)
__global__ void k_grad_acc_multi(int64_t n, int ns, double scale, const double2* __restrict__ u, int64_t us,
                                 const double2* __restrict__ lam, int64_t ls, double* __restrict__ g) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int s = 0; s < ns; ++s) {
      const double2 a = u[l + us * s], b = lam[l + ls * s];
      acc += a.x * b.x + a.y * b.y;
    }
    g[l] += scale * acc;
  }
}

// res[s][i] = U_s[idx[i]] - data[s][i]  (receiver residuals of ns fields)
__global__ void k_rec_residual(int64_t nrec, int ns, const int64_t* __restrict__ idx, const double2* __restrict__ U,
                               int64_t ustride, const double2* __restrict__ data, double2* __restrict__ res) {
  const int64_t tot = nrec * ns;
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < tot; l += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = l / nrec, i = l - s * nrec;
    const double2 u = U[ustride * s + idx[i]], d = data[l];
    res[l] = make_double2(u.x - d.x, u.y - d.y);
  }
}

// dst_s[idx[i]] = vals[s][i]  (receivers distinct; dst cleared beforehand)
__global__ void k_rec_scatter(int64_t nrec, int ns, const int64_t* __restrict__ idx, const double2* __restrict__ vals,
                              double2* __restrict__ dst, int64_t dstride) {
  const int64_t tot = nrec * ns;
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < tot; l += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = l / nrec, i = l - s * nrec;
    dst[dstride * s + idx[i]] = vals[l];
  }
}

__global__ void k_fill(int64_t n, double2 v, double2* __restrict__ y) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) y[l] = v;
}
__global__ void k_fill_real(int64_t n, double v, double* __restrict__ y) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) y[l] = v;
}

inline unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

Box make_box(hz_problem* p, int mask) {
  Box b;
  b.nx = (int)p->nx; b.ny = (int)p->ny; b.nz = (int)p->nz;
  const hz_desc& d = p->d;
  b.x0 = d.pml[0] + d.spg[0]; b.x1 = (int)p->nx - b.x0;
  b.y0 = d.pml[1] + d.spg[1]; b.y1 = (int)p->ny - b.y0;
  b.z0 = d.pml[2] + d.spg[2]; b.z1 = (int)p->nz - b.z0;
  b.on = mask == HZ_MASK_PHYS;
  return b;
}

void ensure_red(hz_problem* p, int nv) {
  const size_t need = (size_t)RB * nv * sizeof(double2);
  if (p->red.bytes < need) {
    p->red.release();
    p->red.alloc(need);
  }
}

}  // namespace

int reduce_blocks(int64_t) { return RB; }

void launch_dot_multi(hz_problem* p, int64_t n, int nv, const double2* x, int64_t xstride, const double2* y,
                      int64_t ystride, int mask, double2* out, hipStream_t s) {
  ensure_red(p, nv);
  Box box = make_box(p, mask);
  hipLaunchKernelGGL(k_dot_partial, dim3(RB, nv), dim3(RT), 0, s, n, x, xstride, y, ystride, box,
                     p->red.as<double2>());
  hipLaunchKernelGGL(k_dot_final, dim3(nv), dim3(RT), 0, s, p->red.as<double2>(), y ? 0 : 1, out, RB, 0, 0, ~0ull);
  HZ_HIP(hipGetLastError());
}

void launch_dot(hz_problem* p, int64_t n, const double2* x, const double2* y, int mask, double2* out,
                hipStream_t s) {
  launch_dot_multi(p, n, 1, x, 0, y, 0, mask, out, s);
}

void launch_axpy_dot(hz_problem* p, int64_t n, const double2* coef, const double2* x, double2* w,
                     const double2* y, double2* out, hipStream_t s) {
  ensure_red(p, 1);
  hipLaunchKernelGGL(k_axpy_dot_partial, dim3(RB), dim3(RT), 0, s, n, coef, x, w, y, p->red.as<double2>());
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RT), 0, s, p->red.as<double2>(), y ? 0 : 1, out, RB, 0, 0, ~0ull);
  HZ_HIP(hipGetLastError());
}

void launch_axpy_dev(int64_t n, const double2* coef, bool neg, bool conj_coef, const double2* x, double2* y,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_axpy_dev, dim3(grid_for(n)), dim3(256), 0, s, n, coef, neg ? -1.0 : 1.0, conj_coef ? 1 : 0, x, y);
  HZ_HIP(hipGetLastError());
}

void launch_scale_inv_dev(int64_t n, const double2* coef, const double2* x, double2* y, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_inv_dev<false>, dim3(grid_for(n)), dim3(256), 0, s, n, coef, 0.0, x, y, 0, 0, ~0ull);
  HZ_HIP(hipGetLastError());
}

void launch_scale_inv_guard(int64_t n, const double2* coef, double thr, const double2* x, double2* y, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_inv_dev<true>, dim3(grid_for(n)), dim3(256), 0, s, n, coef, thr, x, y, 0, 0, ~0ull);
  HZ_HIP(hipGetLastError());
}

void launch_gs_members(hz_problem* p, int64_t n, int nr, uint64_t mask, int j, int m, int ldh, const double2* V,
                       int64_t vs, const double2* w, double2* mem, int64_t msz, int offR, int offL, int offLg,
                       int offH, int offS, int offCS, int offSN, int offG, const double2* beta, const double2* nf,
                       double tol, double brk, double2* est, double2* vnext, hipStream_t s) {
  HZ_CHECK(j >= 0 && j < ICWY_MAX && j < m, HZ_ERR_ARG, "device ICWY step %d (at most %d, m %d)", j,
           ICWY_MAX - 1, m);
  HZ_CHECK(nr >= 1 && nr <= 64, HZ_ERR_ARG, "%d batch members (at most 64)", nr);
  const int nv = j + 1;
  const int ng = (nv + MD - 1) / MD;
  const int64_t ps = (int64_t)ng * RB * 2 * MD;  // multi-dot partials of one member
  HZ_CHECK(p->red.bytes >= (size_t)nr * (size_t)std::max<int64_t>(ps, RBA) * sizeof(double2), HZ_ERR_STATE,
           "reduction workspace not sized for %d members", nr);
  double2* red = p->red.as<double2>();
  double2* h = mem + offH + (int64_t)ldh * j;
  hipLaunchKernelGGL(k_multidot_partial, dim3(RB, ng, nr), dim3(RT), 0, s, n, nv, V, vs, w, V + (int64_t)j * vs, red,
                     n, ps, mask);
  hipLaunchKernelGGL(k_multidot_final, dim3(nv, 1, nr), dim3(RT), 0, s, red, nv, mem + offR, mem + offL, ps, msz,
                     mask, mem + offS, j);
  hipLaunchKernelGGL(k_icwy, dim3(1, 1, nr), dim3(64), (size_t)nv * nv * sizeof(double2), s, j, ldh, mem + offR,
                     mem + offL, mem + offLg, h, msz, mask);
  hipLaunchKernelGGL(k_multi_axpy_norm_partial, dim3(RBA, 1, nr), dim3(RT), nv * sizeof(double2), s, n, nv, V, vs, h,
                     w, red, n, msz, mask, mem + offS, j, vnext);
  hipLaunchKernelGGL(k_norm_givens, dim3(1, 1, nr), dim3(RT), 0, s, red, RBA, j, m, ldh, mask, mem, msz, offH, offS,
                     offCS, offSN, offG, beta, nf, tol, brk, est);
  HZ_HIP(hipGetLastError());
}

void launch_givens(int j, int m, int ldh, uint64_t mask, double2* base, int64_t sb, int offH, int offCS, int offSN,
                   int offG, const double2* beta, const double2* nf, double tol, double brk, double2* est,
                   hipStream_t s) {
  HZ_CHECK(j >= 0 && j < m, HZ_ERR_ARG, "Givens step %d of %d", j, m);
  hipLaunchKernelGGL(k_givens, dim3(1), dim3(64), 0, s, j, m, ldh, mask, base, sb, offH, offCS, offSN, offG, beta, nf,
                     tol, brk, est);
  HZ_HIP(hipGetLastError());
}

void launch_backsub(int jj, int ldh, const double2* H, const double2* g, double2* y, const double2* S,
                    hipStream_t s) {
  HZ_CHECK(jj >= 1 && jj < ldh, HZ_ERR_ARG, "back substitution over %d columns (ldh %d)", jj, ldh);
  // R and y staged in LDS up to 64 KB (jj <= 63)
  const size_t lds = (size_t)(jj * jj + jj) * sizeof(double2);
  const int stage = lds <= 64 * 1024 ? 1 : 0;
  hipLaunchKernelGGL(k_backsub, dim3(1), dim3(256), stage ? lds : 0, s, jj, ldh, H, g, y, S, stage);
  HZ_HIP(hipGetLastError());
}

void launch_lincomb(int64_t n, int nv, const double2* c, const double2* X, int64_t xstride, double2* y,
                    bool accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_lincomb, dim3(grid_for(n)), dim3(256), 0, s, n, nv, c, X, xstride, y, accumulate ? 1 : 0);
  HZ_HIP(hipGetLastError());
}

void launch_axpby(int64_t n, double2 a, const double2* x, double2 b, double2* y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, s, n, a, x, b, y);
  HZ_HIP(hipGetLastError());
}

void launch_axpbypcz(int64_t n, const double2* x, double2 a, const double2* y, double2 b, const double2* w,
                     double2* z, hipStream_t s) {
  hipLaunchKernelGGL(k_axpbypcz, dim3(grid_for(n)), dim3(256), 0, s, n, x, a, y, b, w, z);
  HZ_HIP(hipGetLastError());
}

void launch_sub(int64_t n, const double2* a, const double2* b, double2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(256), 0, s, n, a, b, out);
  HZ_HIP(hipGetLastError());
}

void launch_w_minus_dl(int64_t N, int nrhs, const double2* w, const double2* dL, const double2* x, double2* out,
                       hipStream_t s) {
  // nrhs vectors of N points back to back; dL (N entries) repeats per vector
  HZ_CHECK(N >= 1 && nrhs >= 1, HZ_ERR_ARG, "w - dL x over %d vectors of %lld points", nrhs, (long long)N);
  const int64_t n = N * nrhs;
  hipLaunchKernelGGL(k_w_minus_dl, dim3(grid_for(n)), dim3(256), 0, s, n, N, w, dL, x, out);
  HZ_HIP(hipGetLastError());
}

void launch_grad_acc(int64_t n, double scale, const double2* u, const double2* lam, double* grad, hipStream_t s) {
  hipLaunchKernelGGL(k_grad_acc, dim3(grid_for(n)), dim3(256), 0, s, n, scale, u, lam, grad);
  HZ_HIP(hipGetLastError());
}

void launch_grad_acc_multi(int64_t n, int ns, double scale, const double2* u, int64_t us, const double2* lam,
                           int64_t ls, double* grad, hipStream_t s) {
  hipLaunchKernelGGL(k_grad_acc_multi, dim3(grid_for(n)), dim3(256), 0, s, n, ns, scale, u, us, lam, ls, grad);
  HZ_HIP(hipGetLastError());
}

void launch_rec_residual(int64_t nrec, int ns, const int64_t* idx, const double2* U, int64_t ustride,
                         const double2* data, double2* res, hipStream_t s) {
  hipLaunchKernelGGL(k_rec_residual, dim3(grid_for(nrec * ns)), dim3(256), 0, s, nrec, ns, idx, U, ustride, data, res);
  HZ_HIP(hipGetLastError());
}

void launch_rec_scatter(int64_t nrec, int ns, const int64_t* idx, const double2* vals, double2* dst, int64_t dstride,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_rec_scatter, dim3(grid_for(nrec * ns)), dim3(256), 0, s, nrec, ns, idx, vals, dst, dstride);
  HZ_HIP(hipGetLastError());
}

void launch_fill(int64_t n, double2 v, double2* y, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, n, v, y);
  HZ_HIP(hipGetLastError());
}

void launch_fill_real(int64_t n, double v, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_real, dim3(grid_for(n)), dim3(256), 0, s, n, v, y);
  HZ_HIP(hipGetLastError());
}

}  // namespace hz
