// k_blocktri.hip -- factor and solve of the per-kz plane systems
//   M_k = tridiag_y( cm_y(j) I,  D_kj,  cp_y(j) I ),
//   D_kj = tridiag_x(cm_x, cdg_x + cdg_y(j) + kappa(.,j) - lambda_k, cp_x),
// i.e. the 5-point stretched-coordinate operator of one z-wavenumber ordered by
// y-lines (GenSparseMatrixOnline2DwithPMLFast, functions.cpp:3270-3370, with the
// per-plane diagonal shift of FGMRES.cpp:233-267).
//
// Factor (block Schur recursion over y-lines; the algorithm of
// DirFactFastDiagStructOnline, HODLR/BinaryTreesHODLR.cpp:2870-2969, with dense
// blocks and distinct lower/upper couplings):
//   S_0 = D_0,  S_j = D_j - cm_y(j) cp_y(j-1) G_{j-1},  G_j = S_j^{-1}.
// One workgroup per plane walks j; each inverse is an in-place blocked
// Gauss-Jordan with partial pivoting: a panel of NB columns is eliminated in
// LDS (pivot rows swapped eagerly across the whole matrix), then the other
// columns take the panel transform as one GEMM, A(:,c) <- P A(K,c) + [r not in
// K] A(r,c); the column interchanges are undone in reverse order at the end.
//
// Solve (DirSolveFastDiagStruct, HODLR/BinaryTreesHODLR.cpp:2971-3018):
//   forward  t_0 = b_0,  t_j = b_j - cm_y(j) G_{j-1} t_{j-1}
//   backward x_{ny-1} = G_{ny-1} t_{ny-1},  x_j = G_j (t_j - cp_y(j) x_{j+1})
// and for M_k^H:
//   forward  y_0 = G_0^H b_0,  y_j = G_j^H (b_j - conj(cp_y(j-1)) y_{j-1})
//   backward x_{ny-1} = y_{ny-1},  x_j = y_j - G_j^H conj(cm_y(j+1)) x_{j+1}.
// One workgroup per (plane, right-hand side); each step is a GEMV that streams
// the nx x nx block G_j once from HBM.
#include "hz_internal.h"

namespace hz {
namespace {

template <typename T> struct Cx;
template <> struct Cx<float> { typedef float2 t; };
template <> struct Cx<double> { typedef double2 t; };

template <typename C> __device__ __forceinline__ C mk(double x, double y) { C r; r.x = x; r.y = y; return r; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
template <typename C> __device__ __forceinline__ C cfma(C a, C b, C c) {  // a*b + c
  C r; r.x = a.x * b.x - a.y * b.y + c.x; r.y = a.x * b.y + a.y * b.x + c.y; return r;
}
template <typename C> __device__ __forceinline__ C cfma_conj(C a, C b, C c) {  // conj(a)*b + c
  C r; r.x = a.x * b.x + a.y * b.y + c.x; r.y = a.x * b.y - a.y * b.x + c.y; return r;
}
template <typename C> __device__ __forceinline__ C cinv(C a) {
  // 1/a with scaling to avoid overflow in |a|^2
  double ax = a.x, ay = a.y;
  double s = fmax(fabs(ax), fabs(ay));
  C r;
  if (s == 0.0) { r.x = ax / 0.0; r.y = 0; return r; }
  double bx = ax / s, by = ay / s;
  double d = s * (bx * bx + by * by);
  r.x = bx / d; r.y = -by / d;
  return r;
}

struct PlaneTabs {
  const double2* cmx;
  const double2* cpx;
  const double2* cdx;
  const double2* cdy;
  const double2* kappa;
  const double* lamz;
  const double2* sig;
};

constexpr int FACT_THREADS = 512;
constexpr int TC = 32;   // columns per trailing-update tile
constexpr int CB = 4;    // columns per thread in the trailing update

template <typename T>
__global__ __launch_bounds__(FACT_THREADS) void k_factor(typename Cx<T>::t* __restrict__ G, PlaneTabs tb, int n,
                                                         int ny, int nb, int* __restrict__ status) {
  typedef typename Cx<T>::t C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* P = reinterpret_cast<C*>(smem_raw);     // nb * n, column-major panel
  C* AK = P + (size_t)nb * n;                // TC * nb (AK[cl * nb + kk])
  C* prow = AK + (size_t)TC * nb;            // nb
  C* pcol = prow + nb;                       // n
  int* piv = reinterpret_cast<int*>(pcol + n);  // n
  int* sh = piv + n;                         // scratch ints

  const int k = blockIdx.x;
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6;
  const size_t nn = (size_t)n * n;
  const double lam = tb.lamz[k];
  int singular = 0;
  if (tid == 0) {
    sh[0] = 0;
    sh[1] = 0;
  }
  __syncthreads();

  for (int j = 0; j < ny; ++j) {
    C* Gj = G + ((size_t)k * ny + j) * nn;
    const C* Gp = Gj - nn;
    // ---- S_j = D_j - sig_j G_{j-1}
    const double2 sg = tb.sig[j];
    const double2 cdy = tb.cdy[j];
    for (size_t idx = tid; idx < nn; idx += nt) {
      const int r = (int)(idx % n), c = (int)(idx / n);
      double vx = 0.0, vy = 0.0;
      if (j > 0) {
        const C g = Gp[idx];
        vx = -(sg.x * (double)g.x - sg.y * (double)g.y);
        vy = -(sg.x * (double)g.y + sg.y * (double)g.x);
      }
      if (r == c) {
        const double2 kp = tb.kappa[r + (size_t)n * j];
        vx += tb.cdx[r].x + cdy.x + kp.x - lam;
        vy += tb.cdx[r].y + cdy.y + kp.y;
      } else if (c == r - 1) {
        vx += tb.cmx[r].x; vy += tb.cmx[r].y;
      } else if (c == r + 1) {
        vx += tb.cpx[r].x; vy += tb.cpx[r].y;
      }
      Gj[idx] = mk<C>(vx, vy);
    }
    __syncthreads();

    // ---- in-place blocked Gauss-Jordan inversion of Gj
    for (int p0 = 0; p0 < n; p0 += nb) {
      const int kb = min(nb, n - p0);
      for (int idx = tid; idx < n * kb; idx += nt) {
        const int r = idx % n, c = idx / n;
        P[c * n + r] = Gj[r + (size_t)n * (p0 + c)];
      }
      __syncthreads();
      for (int kk = 0; kk < kb; ++kk) {
        const int col = p0 + kk;
        if (wave == 0) {
          double best = -1.0;
          int bi = col;
          for (int r = col + lane; r < n; r += 64) {
            const C v = P[kk * n + r];
            const double a = fabs((double)v.x) + fabs((double)v.y);
            if (a > best) { best = a; bi = r; }
          }
          for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off);
            const int oi = __shfl_xor(bi, off);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
          }
          if (lane == 0) {
            sh[0] = bi;
            piv[col] = bi;
            if (best == 0.0) sh[1] = 1;
          }
        }
        __syncthreads();
        const int pr = sh[0];
        if (pr != col) {
          for (int c = tid; c < kb; c += nt) {
            const C t = P[c * n + col];
            P[c * n + col] = P[c * n + pr];
            P[c * n + pr] = t;
          }
          for (int c = tid; c < n - kb; c += nt) {
            const int cc = c < p0 ? c : c + kb;
            const C t = Gj[col + (size_t)n * cc];
            Gj[col + (size_t)n * cc] = Gj[pr + (size_t)n * cc];
            Gj[pr + (size_t)n * cc] = t;
          }
          __syncthreads();
        }
        for (int c = tid; c < kb; c += nt) prow[c] = P[c * n + col];
        for (int r = tid; r < n; r += nt) pcol[r] = P[kk * n + r];
        __syncthreads();
        const C pinv = cinv(pcol[col]);
        for (int idx = tid; idx < n * kb; idx += nt) {
          const int r = idx % n, c = idx / n;
          C v;
          if (r == col) {
            v = (c == kk) ? pinv : cmul(prow[c], pinv);
          } else {
            const C f = pcol[r];
            const C fp = cmul(f, pinv);
            if (c == kk) {
              v.x = -fp.x; v.y = -fp.y;
            } else {
              const C a = P[c * n + r];
              const C pc = prow[c];
              v.x = a.x - (fp.x * pc.x - fp.y * pc.y);
              v.y = a.y - (fp.x * pc.y + fp.y * pc.x);
            }
          }
          P[c * n + r] = v;
        }
        __syncthreads();
      }
      // panel back to global
      for (int idx = tid; idx < n * kb; idx += nt) {
        const int r = idx % n, c = idx / n;
        Gj[r + (size_t)n * (p0 + c)] = P[c * n + r];
      }
      // trailing update of the columns outside the panel
      const int nrest = n - kb;
      for (int c0 = 0; c0 < nrest; c0 += TC) {
        const int tc = min(TC, nrest - c0);
        __syncthreads();
        for (int idx = tid; idx < kb * tc; idx += nt) {
          const int kk = idx % kb, cl = idx / kb;
          const int cg = c0 + cl;
          const int cc = cg < p0 ? cg : cg + kb;
          AK[cl * nb + kk] = Gj[(p0 + kk) + (size_t)n * cc];
        }
        __syncthreads();
        const int nq = (tc + CB - 1) / CB;
        for (int idx = tid; idx < n * nq; idx += nt) {
          const int r = idx % n, q = idx / n;
          const bool inK = (r >= p0) && (r < p0 + kb);
          C acc[CB];
          int ccs[CB];
#pragma unroll
          for (int b = 0; b < CB; ++b) {
            const int cl = q * CB + b;
            const int cg = c0 + cl;
            ccs[b] = cg < p0 ? cg : cg + kb;
            if (cl < tc && !inK) acc[b] = Gj[r + (size_t)n * ccs[b]];
            else { acc[b].x = 0; acc[b].y = 0; }
          }
          for (int kk = 0; kk < kb; ++kk) {
            const C pv = P[kk * n + r];
#pragma unroll
            for (int b = 0; b < CB; ++b) acc[b] = cfma(pv, AK[(q * CB + b) * nb + kk], acc[b]);
          }
#pragma unroll
          for (int b = 0; b < CB; ++b)
            if (q * CB + b < tc) Gj[r + (size_t)n * ccs[b]] = acc[b];
        }
      }
      __syncthreads();
    }
    if (tid == 0 && sh[1]) singular = 1;
    // ---- undo the column interchanges in reverse order
    for (int col = n - 1; col >= 0; --col) {
      const int pr = piv[col];
      if (pr != col) {
        for (int r = tid; r < n; r += nt) {
          const C t = Gj[r + (size_t)n * col];
          Gj[r + (size_t)n * col] = Gj[r + (size_t)n * pr];
          Gj[r + (size_t)n * pr] = t;
        }
        __syncthreads();
      }
    }
    __syncthreads();
  }
  if (tid == 0 && singular) atomicOr(status, 1);
}

// ------------------------------------------------------------------ solve
constexpr int APPLY_THREADS = 512;

// y = G v (G column-major n x n), v in LDS; result to LDS `out`.
// rpt rows per thread, groups of threads split the columns.
template <typename C>
__device__ __forceinline__ void gemv_n(const C* __restrict__ G, const C* v, C* out, C* part, int n) {
  const int tid = threadIdx.x;
  const int tpg = n;                      // one row per thread
  const int ng = max(1, (int)blockDim.x / tpg);
  const int g = tid / tpg, r = tid - g * tpg;
  const int cper = (n + ng - 1) / ng;
  if (g < ng && r < n) {
    const int c0 = g * cper, c1 = min(n, c0 + cper);
    C acc0; acc0.x = 0; acc0.y = 0;
    C acc1 = acc0;
    int c = c0;
    for (; c + 1 < c1; c += 2) {
      const C a0 = G[r + (size_t)n * c];
      const C a1 = G[r + (size_t)n * (c + 1)];
      acc0 = cfma(a0, v[c], acc0);
      acc1 = cfma(a1, v[c + 1], acc1);
    }
    if (c < c1) acc0 = cfma(G[r + (size_t)n * c], v[c], acc0);
    acc0.x += acc1.x; acc0.y += acc1.y;
    part[g * n + r] = acc0;
  }
  __syncthreads();
  for (int rr = tid; rr < n; rr += blockDim.x) {
    C s = part[rr];
    for (int q = 1; q < ng; ++q) { s.x += part[q * n + rr].x; s.y += part[q * n + rr].y; }
    out[rr] = s;
  }
  __syncthreads();
}

// y = G^H v: output r is column r of G dotted (conjugated) with v.
template <typename C>
__device__ __forceinline__ void gemv_h(const C* __restrict__ G, const C* v, C* out, int n) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int r = wave; r < n; r += nw) {
    const C* col = G + (size_t)n * r;
    double sx = 0.0, sy = 0.0;
    for (int c = lane; c < n; c += 64) {
      const C a = col[c], b = v[c];
      sx += (double)a.x * b.x + (double)a.y * b.y;
      sy += (double)a.x * b.y - (double)a.y * b.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
      sx += __shfl_xor(sx, off);
      sy += __shfl_xor(sy, off);
    }
    if (lane == 0) { out[r].x = sx; out[r].y = sy; }
  }
  __syncthreads();
}

template <typename T, bool ADJ>
__global__ __launch_bounds__(APPLY_THREADS) void k_apply(const typename Cx<T>::t* __restrict__ G,
                                                         typename Cx<T>::t* __restrict__ X, PlaneTabs tb,
                                                         const double2* __restrict__ cmy,
                                                         const double2* __restrict__ cpy, int n, int ny,
                                                         int64_t n2, int64_t N) {
  typedef typename Cx<T>::t C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* v = reinterpret_cast<C*>(smem_raw);  // n
  C* cur = v + n;                          // n
  C* part = cur + n;                       // ngroups * n
  const int k = blockIdx.x;
  const int64_t rhs = blockIdx.y;
  C* x = X + rhs * N + (int64_t)k * n2;
  const C* Gk = G + (size_t)k * ny * n * n;
  const size_t nn = (size_t)n * n;
  const int tid = threadIdx.x;
  (void)tb;
  if (!ADJ) {
    // forward: cur = t_{j-1}
    for (int i = tid; i < n; i += blockDim.x) cur[i] = x[i];
    __syncthreads();
    for (int j = 1; j < ny; ++j) {
      gemv_n<C>(Gk + (size_t)(j - 1) * nn, cur, v, part, n);  // v = G_{j-1} t_{j-1}
      const double2 cm = cmy[j];
      for (int i = tid; i < n; i += blockDim.x) {
        const C b = x[(size_t)n * j + i];
        const C y = v[i];
        C t;
        t.x = b.x - (cm.x * y.x - cm.y * y.y);
        t.y = b.y - (cm.x * y.y + cm.y * y.x);
        x[(size_t)n * j + i] = t;
        cur[i] = t;
      }
      __syncthreads();
    }
    // backward: cur holds t_{ny-1}
    gemv_n<C>(Gk + (size_t)(ny - 1) * nn, cur, v, part, n);
    for (int i = tid; i < n; i += blockDim.x) {
      x[(size_t)n * (ny - 1) + i] = v[i];
      cur[i] = v[i];
    }
    __syncthreads();
    for (int j = ny - 2; j >= 0; --j) {
      const double2 cp = cpy[j];
      for (int i = tid; i < n; i += blockDim.x) {
        const C t = x[(size_t)n * j + i];
        const C xn = cur[i];
        C w;
        w.x = t.x - (cp.x * xn.x - cp.y * xn.y);
        w.y = t.y - (cp.x * xn.y + cp.y * xn.x);
        v[i] = w;
      }
      __syncthreads();
      gemv_n<C>(Gk + (size_t)j * nn, v, cur, part, n);
      for (int i = tid; i < n; i += blockDim.x) x[(size_t)n * j + i] = cur[i];
      __syncthreads();
    }
  } else {
    // forward: y_j = G_j^H (b_j - conj(cp_y(j-1)) y_{j-1})
    for (int i = tid; i < n; i += blockDim.x) v[i] = x[i];
    __syncthreads();
    gemv_h<C>(Gk, v, cur, n);
    for (int i = tid; i < n; i += blockDim.x) x[i] = cur[i];
    __syncthreads();
    for (int j = 1; j < ny; ++j) {
      const double2 cp = cpy[j - 1];
      for (int i = tid; i < n; i += blockDim.x) {
        const C b = x[(size_t)n * j + i];
        const C y = cur[i];
        C w;  // b - conj(cp) * y
        w.x = b.x - (cp.x * y.x + cp.y * y.y);
        w.y = b.y - (cp.x * y.y - cp.y * y.x);
        v[i] = w;
      }
      __syncthreads();
      gemv_h<C>(Gk + (size_t)j * nn, v, cur, n);
      for (int i = tid; i < n; i += blockDim.x) x[(size_t)n * j + i] = cur[i];
      __syncthreads();
    }
    // backward: x_j = y_j - G_j^H conj(cm_y(j+1)) x_{j+1}; cur holds x_{ny-1}
    for (int j = ny - 2; j >= 0; --j) {
      const double2 cm = cmy[j + 1];
      for (int i = tid; i < n; i += blockDim.x) {
        const C xn = cur[i];
        C w;  // conj(cm) * x_{j+1}
        w.x = cm.x * xn.x + cm.y * xn.y;
        w.y = cm.x * xn.y - cm.y * xn.x;
        v[i] = w;
      }
      __syncthreads();
      gemv_h<C>(Gk + (size_t)j * nn, v, part, n);
      for (int i = tid; i < n; i += blockDim.x) {
        const C y = x[(size_t)n * j + i];
        C o;
        o.x = y.x - part[i].x;
        o.y = y.y - part[i].y;
        x[(size_t)n * j + i] = o;
        cur[i] = o;
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------ fast solve
// M_k x = b with the factors streamed in a register pipeline.  One workgroup of
// 1024 threads per (plane, rhs).  Thread (g, ru) owns load unit ru (16 bytes:
// RPL = 2 rows of complex64 or 1 row of complex128) and the g-th contiguous
// column range of every block; it walks its columns in chunks of CH loads and
// issues the first chunk of the NEXT block before the cross-group reduction
// and the barrier of the current step, so the serial part of each step runs
// under that chunk's memory latency.  The block sequence of one solve is
// G_0 .. G_{ny-2} (forward) then G_{ny-1} .. G_0 (backward): 2 ny - 1 GEMVs.
template <typename T> struct Unit;
template <> struct Unit<float> { typedef float4 t; };
template <> struct Unit<double> { typedef double2 t; };

template <typename T, int RPL>
__device__ __forceinline__ void unit_fma(const typename Unit<T>::t& a, typename Cx<T>::t vc, T* acc) {
  if (RPL == 2) {
    const float4& q = reinterpret_cast<const float4&>(a);
    acc[0] += q.x * vc.x - q.y * vc.y;
    acc[1] += q.x * vc.y + q.y * vc.x;
    acc[2] += q.z * vc.x - q.w * vc.y;
    acc[3] += q.z * vc.y + q.w * vc.x;
  } else {
    acc[0] += (T)a.x * vc.x - (T)a.y * vc.y;
    acc[1] += (T)a.x * vc.y + (T)a.y * vc.x;
  }
}

template <typename T, int RPL, int CH>
__global__ __launch_bounds__(1024) void k_solve_fast(const typename Cx<T>::t* __restrict__ G,
                                                     typename Cx<T>::t* __restrict__ X,
                                                     const double2* __restrict__ cmy,
                                                     const double2* __restrict__ cpy, int n, int ny, int64_t n2,
                                                     int64_t N, int nru, int ncg, int cpg) {
  typedef typename Cx<T>::t C;
  typedef typename Unit<T>::t U;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* v = reinterpret_cast<C*>(smem_raw);  // n
  C* part = v + n;                        // ncg * n
  const int k = blockIdx.x;
  const int64_t rhs = blockIdx.y;
  C* x = X + rhs * N + (int64_t)k * n2;
  const size_t nn = (size_t)n * n;
  const C* Gk = G + (size_t)k * ny * nn;
  const int tid = threadIdx.x;
  const int g = tid / nru, ru = tid - g * nru;
  const bool worker = g < ncg;
  const int c0 = worker ? g * cpg : 0;
  const int c1 = worker ? min(n, c0 + cpg) : 0;
  const int ncol = max(0, c1 - c0);
  const int nq = (ncol + CH - 1) / CH;
  const int nsteps = 2 * ny - 1;
  // unit base offset (in C elements) of this thread within a block
  const size_t ubase = (size_t)ru * RPL + (size_t)n * c0;

  auto blk_of = [&](int s) { return s <= ny - 1 ? s : 2 * ny - 2 - s; };
  auto load_chunk = [&](const C* B, int q, U* buf) {
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int cc = q * CH + i;
      if (cc < ncol) buf[i] = *reinterpret_cast<const U*>(B + ubase + (size_t)n * cc);
    }
  };

  // v = t_0 = b_0
  for (int i = tid; i < n; i += blockDim.x) v[i] = x[i];
  U bufA[CH], bufB[CH];
  if (worker && nq > 0) load_chunk(Gk, 0, bufA);
  __syncthreads();

  for (int s = 0; s < nsteps; ++s) {
    const C* B = Gk + (size_t)blk_of(s) * nn;
    const C* Bn = (s + 1 < nsteps) ? Gk + (size_t)blk_of(s + 1) * nn : nullptr;
    // prefetch of the epilogue operand (b_{s+1} forward, t_{j-1} backward)
    C pre;
    pre.x = 0; pre.y = 0;
    int slot_next = -1;
    if (s < ny - 1) slot_next = s + 1;                       // forward: read b_{s+1}
    else if (2 * ny - 2 - s > 0) slot_next = 2 * ny - 3 - s;  // backward: read t_{j-1}
    if (tid < n && slot_next >= 0) pre = x[(size_t)n * slot_next + tid];

    T acc[2 * RPL];
#pragma unroll
    for (int i = 0; i < 2 * RPL; ++i) acc[i] = 0;
    if (worker) {
      for (int q = 0; q < nq; ++q) {
        // the next chunk: the rest of this block, else the head of the next one
        if (q + 1 < nq) load_chunk(B, q + 1, bufB);
        else if (Bn) load_chunk(Bn, 0, bufB);
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          const int cc = q * CH + i;
          if (cc < ncol) unit_fma<T, RPL>(bufA[i], v[c0 + cc], acc);
        }
#pragma unroll
        for (int i = 0; i < CH; ++i) bufA[i] = bufB[i];
      }
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        C o;
        o.x = acc[2 * r];
        o.y = acc[2 * r + 1];
        part[(size_t)g * n + ru * RPL + r] = o;
      }
    }
    __syncthreads();
    if (tid < n) {
      C y;
      y.x = 0; y.y = 0;
      for (int q = 0; q < ncg; ++q) {
        const C t = part[(size_t)q * n + tid];
        y.x += t.x;
        y.y += t.y;
      }
      C nv;
      if (s < ny - 1) {
        // t_{s+1} = b_{s+1} - cm_y(s+1) y
        const double2 cm = cmy[s + 1];
        nv.x = pre.x - (T)(cm.x * y.x - cm.y * y.y);
        nv.y = pre.y - (T)(cm.x * y.y + cm.y * y.x);
        x[(size_t)n * (s + 1) + tid] = nv;
      } else {
        const int j = 2 * ny - 2 - s;
        x[(size_t)n * j + tid] = y;  // x_j
        if (j > 0) {
          // v for x_{j-1}: t_{j-1} - cp_y(j-1) x_j
          const double2 cp = cpy[j - 1];
          nv.x = pre.x - (T)(cp.x * y.x - cp.y * y.y);
          nv.y = pre.y - (T)(cp.x * y.y + cp.y * y.x);
        } else {
          nv = y;
        }
      }
      v[tid] = nv;
    }
    __syncthreads();
  }
}

// Prefetching variant: each worker thread loads up to PF units of a block with
// unconditional (index-clamped) loads and masks the tail in the FMA, so the
// compiler keeps the loads in flight; the first PF units of the NEXT block are
// issued right after the current block's FMAs, before the partial-sum write,
// the barriers and the serial epilogue of the step.  With PF = 20 and 1024
// threads one 320 KB complex64 block per CU is in flight across the step
// boundary.
template <typename T, int RPL, int PF>
__global__ __launch_bounds__(1024) void k_solve_pf(const typename Cx<T>::t* __restrict__ G,
                                                   typename Cx<T>::t* __restrict__ X,
                                                   const double2* __restrict__ cmy, const double2* __restrict__ cpy,
                                                   int n, int ny, int64_t n2, int64_t N, int nru, int ncg, int cpg) {
  typedef typename Cx<T>::t C;
  typedef typename Unit<T>::t U;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* v = reinterpret_cast<C*>(smem_raw);  // n
  C* part = v + n;                        // ncg * n
  const int k = blockIdx.x;
  const int64_t rhs = blockIdx.y;
  C* x = X + rhs * N + (int64_t)k * n2;
  const size_t nn = (size_t)n * n;
  const C* Gk = G + (size_t)k * ny * nn;
  const int tid = threadIdx.x;
  const int g = tid / nru, ru = tid - g * nru;
  const int c0 = g * cpg;
  const bool worker = (g < ncg) && (c0 < n);
  const int ncol = worker ? min(n, c0 + cpg) - c0 : 0;
  const int nchunk = (ncol + PF - 1) / PF;
  const int nsteps = 2 * ny - 1;
  const size_t ubase = (size_t)ru * RPL + (size_t)n * c0;

  U buf[PF];
  auto issue = [&](const C* B, int chunk) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int cc = min(chunk * PF + i, ncol - 1);
      buf[i] = *reinterpret_cast<const U*>(B + ubase + (size_t)n * cc);
    }
  };
  T acc[2 * RPL];
  auto consume = [&](int chunk) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int cc = chunk * PF + i;
      C vc = v[c0 + min(cc, ncol - 1)];
      const bool ok = cc < ncol;
      vc.x = ok ? vc.x : (T)0;
      vc.y = ok ? vc.y : (T)0;
      unit_fma<T, RPL>(buf[i], vc, acc);
    }
  };
  auto blk_of = [&](int s) { return s <= ny - 1 ? s : 2 * ny - 2 - s; };

  for (int i = tid; i < n; i += blockDim.x) v[i] = x[i];
  if (worker) issue(Gk, 0);
  __syncthreads();

  for (int s = 0; s < nsteps; ++s) {
    const C* B = Gk + (size_t)blk_of(s) * nn;
    C pre;
    pre.x = 0; pre.y = 0;
    int slot_next = -1;
    if (s < ny - 1) slot_next = s + 1;
    else if (2 * ny - 2 - s > 0) slot_next = 2 * ny - 3 - s;
    if (tid < n && slot_next >= 0) pre = x[(size_t)n * slot_next + tid];
#pragma unroll
    for (int i = 0; i < 2 * RPL; ++i) acc[i] = 0;
    if (worker) {
      consume(0);
      for (int ch = 1; ch < nchunk; ++ch) {
        issue(B, ch);
        consume(ch);
      }
      if (s + 1 < nsteps) issue(Gk + (size_t)blk_of(s + 1) * nn, 0);
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        C o;
        o.x = acc[2 * r];
        o.y = acc[2 * r + 1];
        part[(size_t)g * n + ru * RPL + r] = o;
      }
    }
    __syncthreads();
    if (tid < n) {
      C y;
      y.x = 0; y.y = 0;
      for (int q = 0; q < ncg; ++q) {
        const C t = part[(size_t)q * n + tid];
        y.x += t.x;
        y.y += t.y;
      }
      C nv;
      if (s < ny - 1) {
        const double2 cm = cmy[s + 1];
        nv.x = pre.x - (T)(cm.x * y.x - cm.y * y.y);
        nv.y = pre.y - (T)(cm.x * y.y + cm.y * y.x);
        x[(size_t)n * (s + 1) + tid] = nv;
      } else {
        const int j = 2 * ny - 2 - s;
        x[(size_t)n * j + tid] = y;
        if (j > 0) {
          const double2 cp = cpy[j - 1];
          nv.x = pre.x - (T)(cp.x * y.x - cp.y * y.y);
          nv.y = pre.y - (T)(cp.x * y.y + cp.y * y.x);
        } else {
          nv = y;
        }
      }
      v[tid] = nv;
    }
    __syncthreads();
  }
}

PlaneTabs plane_tabs(hz_problem* p) {
  PlaneTabs t;
  t.cmx = p->tab(0, 0);
  t.cpx = p->tab(0, 1);
  t.cdx = p->tab(0, 2);
  t.cdy = p->tab(1, 2);
  t.kappa = p->d_kappa.as<double2>();
  t.lamz = p->d_lamz.as<double>();
  t.sig = p->d_sig.as<double2>();
  return t;
}

template <typename T>
size_t factor_lds(int n, int nb) {
  typedef typename Cx<T>::t C;
  return sizeof(C) * ((size_t)nb * n + (size_t)TC * nb + nb + n) + sizeof(int) * (n + 4);
}

}  // namespace

void launch_factor(hz_problem* p, int prec, hipStream_t s) {
  const int n = (int)p->nx, ny = (int)p->ny;
  Factors& f = p->fac[prec];
  f.elem = prec == HZ_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  const size_t bytes = (size_t)p->nz * ny * (size_t)n * n * f.elem;
  if (f.G.bytes != bytes) {
    f.G.release();
    f.G.alloc(bytes);
  }
  DevBuf st;
  st.alloc(sizeof(int));
  HZ_HIP(hipMemsetAsync(st.ptr, 0, sizeof(int), s));
  PlaneTabs tb = plane_tabs(p);
  int nb = 32;
  if (prec == HZ_PREC_FP64) {
    while (nb > 4 && factor_lds<double>(n, nb) > 160 * 1024) nb /= 2;
    const size_t lds = factor_lds<double>(n, nb);
    HZ_CHECK(lds <= 160 * 1024, HZ_ERR_ARG, "nx=%d too large for the LDS panel", n);
    HZ_HIP(hipFuncSetAttribute((const void*)&k_factor<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_factor<double>, dim3((unsigned)p->nz), dim3(FACT_THREADS), lds, s, f.G.as<double2>(), tb,
                       n, ny, nb, st.as<int>());
  } else {
    while (nb > 4 && factor_lds<float>(n, nb) > 160 * 1024) nb /= 2;
    const size_t lds = factor_lds<float>(n, nb);
    HZ_CHECK(lds <= 160 * 1024, HZ_ERR_ARG, "nx=%d too large for the LDS panel", n);
    HZ_HIP(hipFuncSetAttribute((const void*)&k_factor<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_factor<float>, dim3((unsigned)p->nz), dim3(FACT_THREADS), lds, s, f.G.as<float2>(), tb, n,
                       ny, nb, st.as<int>());
  }
  HZ_HIP(hipGetLastError());
  int hs = 0;
  HZ_HIP(hipMemcpyAsync(&hs, st.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  HZ_HIP(hipStreamSynchronize(s));
  HZ_CHECK(hs == 0, HZ_ERR_SINGULAR, "zero pivot while factoring a plane");
  f.ready = true;
}

void launch_plane_solve(hz_problem* p, int prec, int nrhs, void* x, bool adjoint, hipStream_t s) {
  Factors& f = p->fac[prec];
  HZ_CHECK(f.ready, HZ_ERR_STATE, "preconditioner of precision %d not factored", prec);
  const int n = (int)p->nx, ny = (int)p->ny;
  const int ng = std::max(1, APPLY_THREADS / n);
  PlaneTabs tb = plane_tabs(p);
  const double2* cmy = p->tab(1, 0);
  const double2* cpy = p->tab(1, 1);
  dim3 grid((unsigned)p->nz, (unsigned)nrhs);
  if (!adjoint && ny >= 1) {
    constexpr int PF = 20;
    if (prec == HZ_PREC_C64 && n % 2 == 0) {
      const int nru = n / 2, ncg = std::max(1, 1024 / nru), cpg = (n + ncg - 1) / ncg;
      const size_t lds = sizeof(float2) * (size_t)n * (1 + ncg);
      HZ_HIP(hipFuncSetAttribute((const void*)&k_solve_pf<float, 2, PF>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL((k_solve_pf<float, 2, PF>), grid, dim3(1024), lds, s, f.G.as<float2>(), (float2*)x, cmy, cpy,
                         n, ny, p->n2, p->N, nru, ncg, cpg);
      HZ_HIP(hipGetLastError());
      return;
    }
    if (prec == HZ_PREC_FP64) {
      const int nru = n, ncg = std::max(1, 1024 / nru), cpg = (n + ncg - 1) / ncg;
      const size_t lds = sizeof(double2) * (size_t)n * (1 + ncg);
      HZ_HIP(hipFuncSetAttribute((const void*)&k_solve_pf<double, 1, PF>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL((k_solve_pf<double, 1, PF>), grid, dim3(1024), lds, s, f.G.as<double2>(), (double2*)x, cmy,
                         cpy, n, ny, p->n2, p->N, nru, ncg, cpg);
      HZ_HIP(hipGetLastError());
      return;
    }
  }
  if (prec == HZ_PREC_FP64) {
    const size_t lds = sizeof(double2) * (size_t)n * (2 + ng);
    HZ_CHECK(lds <= 160 * 1024, HZ_ERR_ARG, "nx too large for the apply kernel");
    if (adjoint) {
      HZ_HIP(hipFuncSetAttribute((const void*)&k_apply<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL((k_apply<double, true>), grid, dim3(APPLY_THREADS), lds, s, f.G.as<double2>(), (double2*)x, tb,
                         cmy, cpy, n, ny, p->n2, p->N);
    } else {
      HZ_HIP(hipFuncSetAttribute((const void*)&k_apply<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL((k_apply<double, false>), grid, dim3(APPLY_THREADS), lds, s, f.G.as<double2>(), (double2*)x,
                         tb, cmy, cpy, n, ny, p->n2, p->N);
    }
  } else {
    const size_t lds = sizeof(float2) * (size_t)n * (2 + ng);
    if (adjoint) {
      HZ_HIP(hipFuncSetAttribute((const void*)&k_apply<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL((k_apply<float, true>), grid, dim3(APPLY_THREADS), lds, s, f.G.as<float2>(), (float2*)x, tb,
                         cmy, cpy, n, ny, p->n2, p->N);
    } else {
      HZ_HIP(hipFuncSetAttribute((const void*)&k_apply<float, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL((k_apply<float, false>), grid, dim3(APPLY_THREADS), lds, s, f.G.as<float2>(), (float2*)x, tb,
                         cmy, cpy, n, ny, p->n2, p->N);
    }
  }
  HZ_HIP(hipGetLastError());
}

}  // namespace hz
