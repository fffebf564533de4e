// k_fft.hip -- batched z-DFT of every (x,y) line (length nz, stride nx*ny).
//
// One workgroup owns LINES consecutive lines: it stages the LINES x nz tile in
// LDS (coalesced LINES*sizeof(complex) segments per z), runs a mixed-radix
// Stockham transform (radices 4, 2, 3, 5 and any small prime) ping-ponging
// between two LDS buffers, and writes the tile back, optionally converting
// between complex128 and complex64 and fusing the reference's
// g = w - deltaL * (L0^{-1} w) epilogue into the backward pass
// (functions.cpp:5571-5574).  Forward is unscaled with e^{-2 pi i jk/n},
// backward scales by 1/n (DFTI settings of functions.cpp:5864-5869).
#include <cstdlib>

#include "hz_internal.h"

namespace hz {
namespace {

template <typename T> struct Cx;
template <> struct Cx<float> { typedef float2 t; };
template <> struct Cx<double> { typedef double2 t; };

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
// multiply by sign * i
template <typename C> __device__ __forceinline__ C cmul_si(C a, int sign) {
  C r; r.x = -sign * a.y; r.y = sign * a.x; return r;
}
template <typename C, typename T> __device__ __forceinline__ C cscale(C a, T s) { C r; r.x = a.x * s; r.y = a.y * s; return r; }

template <typename C> __device__ __forceinline__ C load_c(const double2* p) { C r; r.x = p->x; r.y = p->y; return r; }
template <typename C> __device__ __forceinline__ C load_c(const float2* p) { C r; r.x = p->x; r.y = p->y; return r; }
// Nontemporal stores of the register-resident transform's output: at 200^3 the
// forward transform runs 55 -> 42 us and the kernels that follow it (plane
// solve, 7-point operator, Gram-Schmidt dot pass) 1-10% faster than with
// ordinary stores.
__device__ __forceinline__ void store_nt(double2* p, double2 v) {
  typedef double dv2 __attribute__((ext_vector_type(2)));
  dv2 t;
  t.x = v.x;
  t.y = v.y;
  __builtin_nontemporal_store(t, reinterpret_cast<dv2*>(p));
}
__device__ __forceinline__ void store_nt(float2* p, float2 v) {
  typedef float fv2 __attribute__((ext_vector_type(2)));
  fv2 t;
  t.x = v.x;
  t.y = v.y;
  __builtin_nontemporal_store(t, reinterpret_cast<fv2*>(p));
}

struct FftArgs {
  int n;
  int nstages;
  int radix[24];
};

// LDS row of one z index: LINES lines plus padding, so that the strided rows a
// Stockham stage writes (stride p rows in the first stage) spread over the banks.
template <typename T, int LINES>
struct Lds {
  static constexpr int LD = LINES + (sizeof(T) == 4 ? 2 : 1);
};

template <typename T, int LINES>
__device__ void stockham_lds(typename Cx<T>::t* x, typename Cx<T>::t* y, const FftArgs& a,
                             const typename Cx<T>::t* __restrict__ tw, int sign,
                             typename Cx<T>::t** result) {
  typedef typename Cx<T>::t C;
  constexpr int LD = Lds<T, LINES>::LD;
  const int n = a.n;
  int len = n, s = 1;
  for (int st = 0; st < a.nstages; ++st) {
    const int p = a.radix[st];
    const int m = len / p;
    const int nbf = LINES * (n / p);
    const float inv_s = 1.0f / (float)s;
    for (int b = threadIdx.x; b < nbf; b += blockDim.x) {
      const int l = b % LINES;
      const int qt = b / LINES;
      // q = qt / s exactly: (qt + 1/2) / s stays below the next integer for qt, s < 2^12
      const int q = (int)(((float)qt + 0.5f) * inv_s);
      const int t = qt - q * s;
      const int in0 = t + s * q;
      const int out0 = t + s * p * q;
      C v[5];
      if (p == 4) {
        C a0 = x[(in0) * LD + l], a1 = x[(in0 + s * m) * LD + l];
        C a2 = x[(in0 + 2 * s * m) * LD + l], a3 = x[(in0 + 3 * s * m) * LD + l];
        C t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = cmul_si(csub(a1, a3), sign);
        v[0] = cadd(t0, t2); v[1] = cadd(t1, t3); v[2] = csub(t0, t2); v[3] = csub(t1, t3);
      } else if (p == 2) {
        C a0 = x[in0 * LD + l], a1 = x[(in0 + s * m) * LD + l];
        v[0] = cadd(a0, a1); v[1] = csub(a0, a1);
      } else if (p == 3) {
        const T h = (T)0.86602540378443864676372317075294;
        C a0 = x[in0 * LD + l], a1 = x[(in0 + s * m) * LD + l], a2 = x[(in0 + 2 * s * m) * LD + l];
        C sm = cadd(a1, a2), df = cmul_si(cscale(csub(a1, a2), h), sign);
        C base = csub(a0, cscale(sm, (T)0.5));
        v[0] = cadd(a0, sm); v[1] = cadd(base, df); v[2] = csub(base, df);
      } else if (p == 5) {
        const T c1 = (T)0.30901699437494742410229341718282, c2 = (T)-0.80901699437494742410229341718282;
        const T s1 = (T)0.95105651629515357211643933337938, s2 = (T)0.58778525229247312916870595463907;
        C a0 = x[in0 * LD + l], a1 = x[(in0 + s * m) * LD + l], a2 = x[(in0 + 2 * s * m) * LD + l];
        C a3 = x[(in0 + 3 * s * m) * LD + l], a4 = x[(in0 + 4 * s * m) * LD + l];
        C b1 = cadd(a1, a4), b2 = cadd(a2, a3), d1 = csub(a1, a4), d2 = csub(a2, a3);
        C r1 = cadd(a0, cadd(cscale(b1, c1), cscale(b2, c2)));
        C r2 = cadd(a0, cadd(cscale(b1, c2), cscale(b2, c1)));
        C i1 = cmul_si(cadd(cscale(d1, s1), cscale(d2, s2)), sign);
        C i2 = cmul_si(csub(cscale(d1, s2), cscale(d2, s1)), sign);
        v[0] = cadd(a0, cadd(b1, b2)); v[1] = cadd(r1, i1); v[4] = csub(r1, i1); v[2] = cadd(r2, i2); v[3] = csub(r2, i2);
      }
      if (p <= 5) {
        y[out0 * LD + l] = v[0];
        for (int k = 1; k < p; ++k) {
          C w = tw[q * k * s];  // q s k < n: no wrap
          if (sign > 0) w.y = -w.y;
          y[(out0 + s * k) * LD + l] = cmul(v[k], w);
        }
      } else {
        // generic prime radix: direct sum with table twiddles
        const int np = n / p;
        for (int k = 0; k < p; ++k) {
          C acc; acc.x = 0; acc.y = 0;
          for (int r = 0; r < p; ++r) {
            C w = tw[(r * k * np) % n];
            if (sign > 0) w.y = -w.y;
            acc = cadd(acc, cmul(x[(in0 + r * s * m) * LD + l], w));
          }
          C w = tw[q * k * s];
          if (sign > 0) w.y = -w.y;
          y[(out0 + s * k) * LD + l] = cmul(acc, w);
        }
      }
    }
    __syncthreads();
    C* tmp = x; x = y; y = tmp;
    len = m;
    s *= p;
  }
  *result = x;
}

template <typename T, typename TI, typename TO, int EPI, int LINES>
__global__ __launch_bounds__(256) void k_fft_z(const TI* __restrict__ in, TO* __restrict__ out, int64_t n2,
                                               int64_t N, FftArgs a, const typename Cx<T>::t* __restrict__ twg,
                                               int sign, T scale, const double2* __restrict__ w,
                                               const double2* __restrict__ dL) {
  typedef typename Cx<T>::t C;
  constexpr int LD = Lds<T, LINES>::LD;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = a.n;
  C* buf0 = reinterpret_cast<C*>(smem_raw);
  C* buf1 = buf0 + LD * n;
  C* tw = buf1 + LD * n;  // twiddle table staged in LDS
  const int64_t rhs = blockIdx.y;
  const int64_t w0 = (int64_t)blockIdx.x * LINES;
  const TI* src = in + rhs * N;
  TO* dst = out + rhs * N;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
  for (int idx = threadIdx.x; idx < LINES * n; idx += blockDim.x) {
    const int l = idx % LINES, z = idx / LINES;
    const int64_t g = w0 + l;
    C v;
    v.x = 0; v.y = 0;
    if (g < n2) v = load_c<C>(src + g + n2 * z);
    buf0[z * LD + l] = v;
  }
  __syncthreads();
  C* res;
  stockham_lds<T, LINES>(buf0, buf1, a, tw, sign, &res);
  for (int idx = threadIdx.x; idx < LINES * n; idx += blockDim.x) {
    const int l = idx % LINES, z = idx / LINES;
    const int64_t g = w0 + l;
    if (g >= n2) continue;
    C v = res[z * LD + l];
    const int64_t gi = g + n2 * z;
    if (EPI == 1) {
      const double2 ww = w[rhs * N + gi], dd = dL[gi];
      const double vx = (double)v.x * (double)scale, vy = (double)v.y * (double)scale;
      double2 o;
      o.x = ww.x - (dd.x * vx - dd.y * vy);
      o.y = ww.y - (dd.x * vy + dd.y * vx);
      TO r; r.x = o.x; r.y = o.y;
      dst[gi] = r;
    } else {
      TO r;
      r.x = v.x * scale;
      r.y = v.y * scale;
      dst[gi] = r;
    }
  }
}

// ------------------------------------------- register-resident z-FFT
// For n = N1 * N2 with both factors small: z = z1 + N1 z2, k = k2 + N2 k1,
//   X[k2 + N2 k1] = sum_z1 W_n^{z1 k2} W_N1^{z1 k1} sum_z2 x[z1 + N1 z2] W_N2^{z2 k2}.
// A workgroup owns LINES consecutive (x,y) lines; thread (l, z1) loads the N2
// samples z1 + N1 z2 of line l (each load instruction reads LINES consecutive
// lines), transforms them in registers, applies W_n^{z1 k2} and stores them to
// LDS; after one barrier thread (l, k2) reads the N1 values of its column,
// transforms them in registers and writes X[k2 + N2 k1].  The in-register
// DFTs are compile-time mixed-radix Cooley-Tukey with constant twiddles.

// out[k] = sum_j in[j] e^{sign 2 pi i j k / N}  (in place, compile-time N)
template <typename T, int N, int SIGN>
struct RegDft {
  typedef typename Cx<T>::t C;
  static constexpr int P = (N % 4 == 0 && N != 4) ? 4 : (N % 2 == 0 && N != 2) ? 2 : (N % 3 == 0 && N != 3) ? 3
                           : (N % 5 == 0 && N != 5) ? 5 : N;
  static constexpr int M = N / P;
  __device__ __forceinline__ static void run(C* v) {
    if constexpr (P == N) {
      RegDft<T, N, SIGN>::prime(v);
    } else {
      // decimation in time: sub-transforms of the P residues, twiddles, P-point butterflies
      C t[N];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        C sub[M];
#pragma unroll
        for (int m = 0; m < M; ++m) sub[m] = v[p + P * m];
        RegDft<T, M, SIGN>::run(sub);
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const double ang = (double)SIGN * 2.0 * 3.14159265358979323846 * (double)(p * k) / (double)N;
          C w;
          w.x = (T)cos(ang);
          w.y = (T)sin(ang);
          t[p * M + k] = (p * k == 0) ? sub[k] : cmul(sub[k], w);
        }
      }
#pragma unroll
      for (int k = 0; k < M; ++k) {
        C col[P];
#pragma unroll
        for (int p = 0; p < P; ++p) col[p] = t[p * M + k];
        RegDft<T, P, SIGN>::run(col);
#pragma unroll
        for (int q = 0; q < P; ++q) v[k + M * q] = col[q];
      }
    }
  }
  __device__ __forceinline__ static void prime(C* v) {
    if constexpr (N == 1) {
    } else if constexpr (N == 2) {
      const C a = v[0], b = v[1];
      v[0] = cadd(a, b);
      v[1] = csub(a, b);
    } else if constexpr (N == 4) {
      const C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]), t2 = cadd(v[1], v[3]);
      const C t3 = cmul_si(csub(v[1], v[3]), SIGN);
      v[0] = cadd(t0, t2); v[1] = cadd(t1, t3); v[2] = csub(t0, t2); v[3] = csub(t1, t3);
    } else if constexpr (N == 3) {
      const T h = (T)0.86602540378443864676372317075294;
      const C sm = cadd(v[1], v[2]), df = cmul_si(cscale(csub(v[1], v[2]), h), SIGN);
      const C base = csub(v[0], cscale(sm, (T)0.5));
      v[0] = cadd(v[0], sm); v[1] = cadd(base, df); v[2] = csub(base, df);
    } else if constexpr (N == 5) {
      const T c1 = (T)0.30901699437494742410229341718282, c2 = (T)-0.80901699437494742410229341718282;
      const T s1 = (T)0.95105651629515357211643933337938, s2 = (T)0.58778525229247312916870595463907;
      const C a0 = v[0];
      const C b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]), d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
      const C r1 = cadd(a0, cadd(cscale(b1, c1), cscale(b2, c2)));
      const C r2 = cadd(a0, cadd(cscale(b1, c2), cscale(b2, c1)));
      const C i1 = cmul_si(cadd(cscale(d1, s1), cscale(d2, s2)), SIGN);
      const C i2 = cmul_si(csub(cscale(d1, s2), cscale(d2, s1)), SIGN);
      v[0] = cadd(a0, cadd(b1, b2)); v[1] = cadd(r1, i1); v[4] = csub(r1, i1); v[2] = cadd(r2, i2); v[3] = csub(r2, i2);
    } else {
      static_assert(N <= 5, "in-register DFT: unsupported prime factor");
    }
  }
};

template <typename T, typename TI, typename TO, int EPI, int N1, int N2, int LINES, int SIGN>
__global__ __launch_bounds__(LINES * N1) void k_fft_reg(const TI* __restrict__ in, TO* __restrict__ out, int64_t n2,
                                                        int64_t N, const typename Cx<T>::t* __restrict__ twg,
                                                        T scale, const double2* __restrict__ w,
                                                        const double2* __restrict__ dL) {
  typedef typename Cx<T>::t C;
  constexpr int n = N1 * N2;
  __shared__ C A[n * LINES];  // A[(z1 N2 + k2) LINES + l]
  __shared__ C tw[n];         // exp(-2 pi i t / n)
  const int tid = threadIdx.x;
  const int l = tid % LINES, z1 = tid / LINES;
  const int64_t rhs = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * LINES + l;
  const bool live = g < n2;
  const TI* src = in + rhs * N;
  TO* dst = out + rhs * N;
  for (int i = tid; i < n; i += LINES * N1) tw[i] = twg[i];
  // step A: N2-point transforms of the samples z1 + N1 z2
  C v[N2];
#pragma unroll
  for (int z2 = 0; z2 < N2; ++z2) {
    C x;
    x.x = 0; x.y = 0;
    if (live) x = load_c<C>(src + g + n2 * (int64_t)(z1 + N1 * z2));
    v[z2] = x;
  }
  RegDft<T, N2, SIGN>::run(v);
  __syncthreads();  // twiddle table staged
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) {
    C wt = tw[z1 * k2];  // z1 k2 < n: no wrap
    if (SIGN > 0) wt.y = -wt.y;
    A[(z1 * N2 + k2) * LINES + l] = (k2 == 0) ? v[k2] : cmul(v[k2], wt);
  }
  __syncthreads();
  // step C: N1-point transforms of the columns k2
  for (int item = tid; item < LINES * N2; item += LINES * N1) {
    const int lc = item % LINES, k2 = item / LINES;
    const int64_t gc = (int64_t)blockIdx.x * LINES + lc;
    C u[N1];
#pragma unroll
    for (int a = 0; a < N1; ++a) u[a] = A[(a * N2 + k2) * LINES + lc];
    RegDft<T, N1, SIGN>::run(u);
    if (gc >= n2) continue;
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int64_t gi = gc + n2 * (int64_t)(k2 + N2 * k1);
      const C vv = u[k1];
      if (EPI == 1) {
        const double2 ww = w[rhs * N + gi], dd = dL[gi];
        const double vx = (double)vv.x * (double)scale, vy = (double)vv.y * (double)scale;
        TO r;
        r.x = ww.x - (dd.x * vx - dd.y * vy);
        r.y = ww.y - (dd.x * vy + dd.y * vx);
        store_nt(dst + gi, r);
      } else {
        TO r;
        r.x = vv.x * scale;
        r.y = vv.y * scale;
        store_nt(dst + gi, r);
      }
    }
  }
}

template <typename T, typename TI, typename TO, int EPI, int N1, int N2>
bool try_fft_reg(hz_problem* p, int sign, int nrhs, const void* in, void* out, const double2* w, const double2* dL,
                 hipStream_t s) {
  if (p->fft.n != N1 * N2) return false;
  constexpr int LINES = sizeof(T) == 4 ? 64 : 32;
  const typename Cx<T>::t* tw =
      sizeof(T) == 8 ? (const typename Cx<T>::t*)p->fft.tw64.ptr : (const typename Cx<T>::t*)p->fft.tw32.ptr;
  dim3 grid((unsigned)((p->n2 + LINES - 1) / LINES), (unsigned)nrhs);
  const T scale = sign > 0 ? (T)(1.0 / (double)(N1 * N2)) : (T)1;
  if (sign < 0)
    hipLaunchKernelGGL((k_fft_reg<T, TI, TO, EPI, N1, N2, LINES, -1>), grid, dim3(LINES * N1), 0, s, (const TI*)in,
                       (TO*)out, p->n2, p->N, tw, scale, w, dL);
  else
    hipLaunchKernelGGL((k_fft_reg<T, TI, TO, EPI, N1, N2, LINES, 1>), grid, dim3(LINES * N1), 0, s, (const TI*)in,
                       (TO*)out, p->n2, p->N, tw, scale, w, dL);
  HZ_HIP(hipGetLastError());
  return true;
}

// lengths of the configurations: 64 = 8 x 8, 96 = 8 x 12, 128 = 8 x 16,
// 160 = 10 x 16, 192 = 12 x 16, 200 = 8 x 25
template <typename T, typename TI, typename TO, int EPI>
bool launch_reg(hz_problem* p, int sign, int nrhs, const void* in, void* out, const double2* w, const double2* dL,
                hipStream_t s) {
  return try_fft_reg<T, TI, TO, EPI, 8, 25>(p, sign, nrhs, in, out, w, dL, s) ||
         try_fft_reg<T, TI, TO, EPI, 12, 16>(p, sign, nrhs, in, out, w, dL, s) ||
         try_fft_reg<T, TI, TO, EPI, 10, 16>(p, sign, nrhs, in, out, w, dL, s) ||
         try_fft_reg<T, TI, TO, EPI, 8, 16>(p, sign, nrhs, in, out, w, dL, s) ||
         try_fft_reg<T, TI, TO, EPI, 8, 12>(p, sign, nrhs, in, out, w, dL, s) ||
         try_fft_reg<T, TI, TO, EPI, 8, 8>(p, sign, nrhs, in, out, w, dL, s);
}

template <typename T, typename TI, typename TO, int EPI, int LINES>
void launch_lines(hz_problem* p, int sign, int nrhs, const void* in, void* out, const double2* w,
                  const double2* dL, hipStream_t s) {
  FftArgs a;
  a.n = p->fft.n;
  a.nstages = p->fft.nstages;
  for (int i = 0; i < 24; ++i) a.radix[i] = i < a.nstages ? p->fft.radix[i] : 1;
  const typename Cx<T>::t* tw =
      sizeof(T) == 8 ? (const typename Cx<T>::t*)p->fft.tw64.ptr : (const typename Cx<T>::t*)p->fft.tw32.ptr;
  size_t lds = (2 * (size_t)Lds<T, LINES>::LD + 1) * a.n * sizeof(typename Cx<T>::t);
  HZ_CHECK(lds <= 160 * 1024, HZ_ERR_ARG, "nz=%d too long for the LDS-resident FFT", a.n);
  dim3 grid((unsigned)((p->n2 + LINES - 1) / LINES), (unsigned)nrhs);
  if (lds > 64 * 1024)
    set_max_lds((const void*)&k_fft_z<T, TI, TO, EPI, LINES>, lds);
  T scale = sign > 0 ? (T)(1.0 / (double)a.n) : (T)1;
  hipLaunchKernelGGL((k_fft_z<T, TI, TO, EPI, LINES>), grid, dim3(256), lds, s, (const TI*)in, (TO*)out, p->n2,
                     p->N, a, tw, sign, scale, w, dL);
  HZ_HIP(hipGetLastError());
}

template <typename T, typename TI, typename TO, int EPI>
void launch_typed(hz_problem* p, int lines, int sign, int nrhs, const void* in, void* out, const double2* w,
                  const double2* dL, hipStream_t s) {
  if (lines == 0 && launch_reg<T, TI, TO, EPI>(p, sign, nrhs, in, out, w, dL, s)) return;
  if (lines == 32 && fft_lines_fit(sizeof(T) == 8 ? HZ_PREC_FP64 : HZ_PREC_C64, p->fft.n, 32))
    launch_lines<T, TI, TO, EPI, 32>(p, sign, nrhs, in, out, w, dL, s);
  else if (lines == 8)
    launch_lines<T, TI, TO, EPI, 8>(p, sign, nrhs, in, out, w, dL, s);
  else
    launch_lines<T, TI, TO, EPI, 16>(p, sign, nrhs, in, out, w, dL, s);
}

}  // namespace

bool fft_reg_available(int nz) {
  return nz == 200 || nz == 192 || nz == 160 || nz == 128 || nz == 96 || nz == 64;
}

bool fft_lines_fit(int prec, int nz, int lines) {
  if (lines == 0) return fft_reg_available(nz);
  const size_t c = prec == HZ_PREC_FP64 ? 16 : 8;
  const size_t ld = lines + (prec == HZ_PREC_FP64 ? 1 : 2);
  return (2 * ld + 1) * nz * c <= 160 * 1024;
}

void launch_fft_z(hz_problem* p, int prec, int sign, int nrhs, const void* in, int in_prec, void* out,
                  int out_prec, int epi, const double2* w, const double2* dL, hipStream_t s) {
  // HZ_FFT_LINES=0|8|16|32 forces the lines per workgroup, 0 being the
  // register-resident kernel (timing and tests)
  int lines = p->tuning(prec, nrhs).fft_lines;
  if (lines < 0) lines = fft_reg_available((int)p->nz) ? 0 : 16;
  if (const char* e = getenv("HZ_FFT_LINES")) {
    const int v = atoi(e);
    if ((v == 0 || v == 8 || v == 16 || v == 32) && fft_lines_fit(prec, (int)p->nz, v)) lines = v;
  }
  launch_fft_z_lines(p, lines, prec, sign, nrhs, in, in_prec, out, out_prec, epi, w, dL, s);
}

void launch_fft_z_lines(hz_problem* p, int lines, int prec, int sign, int nrhs, const void* in, int in_prec,
                        void* out, int out_prec, int epi, const double2* w, const double2* dL, hipStream_t s) {
  HZ_CHECK(!(epi && out_prec != 0), HZ_ERR_ARG, "fft epilogue needs a complex128 output");
  if (prec == HZ_PREC_FP64) {
    HZ_CHECK(in_prec == 0 && out_prec == 0, HZ_ERR_ARG, "FP64 FFT takes complex128 in and out");
    if (epi) launch_typed<double, double2, double2, 1>(p, lines, sign, nrhs, in, out, w, dL, s);
    else launch_typed<double, double2, double2, 0>(p, lines, sign, nrhs, in, out, w, dL, s);
    return;
  }
  if (in_prec == 0 && out_prec == 1) launch_typed<float, double2, float2, 0>(p, lines, sign, nrhs, in, out, w, dL, s);
  else if (in_prec == 1 && out_prec == 0) {
    if (epi) launch_typed<float, float2, double2, 1>(p, lines, sign, nrhs, in, out, w, dL, s);
    else launch_typed<float, float2, double2, 0>(p, lines, sign, nrhs, in, out, w, dL, s);
  } else if (in_prec == 1 && out_prec == 1) launch_typed<float, float2, float2, 0>(p, lines, sign, nrhs, in, out, w, dL, s);
  else launch_typed<float, double2, double2, 0>(p, lines, sign, nrhs, in, out, w, dL, s);
}

}  // namespace hz
