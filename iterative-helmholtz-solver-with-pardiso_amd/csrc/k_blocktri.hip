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
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <type_traits>

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

typedef float hz_v4f __attribute__((ext_vector_type(4)));
constexpr int FACT_THREADS = 512;
constexpr int TCM = 64;  // columns per trailing-update chunk on the matrix cores (complex64)
constexpr int TC = 32;   // columns per trailing-update tile
constexpr int CB = 4;    // columns per thread in the trailing update

// mode 0: the top-down chain G_j = (D_j - sig_j G_{j-1})^{-1}, j = jbeg .. jend-1
//         (from G_{jbeg-1} when jbeg > 0), into G ([slot][j]), one workgroup per slot.
// mode 1: the same in workgroups 0 .. nslot-1 and, in workgroups nslot ..
//         2 nslot-1, the bottom-up chain of the twisted factors
//         B_j = (D_j - sig_{j+1} B_{j+1})^{-1}, j = ny-1 .. m+1, into Gb
//         ([slot][j - m], ny - m blocks a slot).
// mode 2: the middle block T_m = (D_m - sig_m G_{m-1} - sig_{m+1} B_{m+1})^{-1}
//         into Gb ([slot][0]) from both chains.
template <typename T>
__global__ __launch_bounds__(FACT_THREADS) void k_factor(typename Cx<T>::t* __restrict__ G, PlaneTabs tb, int n,
                                                         int ny, int nb, int* __restrict__ status,
                                                         const int* __restrict__ splane, int mode, int m,
                                                         typename Cx<T>::t* __restrict__ Gb, int jbeg, int jend) {
  typedef typename Cx<T>::t C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* P = reinterpret_cast<C*>(smem_raw);     // nb * n, column-major panel
  C* AK = P + (size_t)nb * n;                // (TCM or TC) * nb (AK[cl * nb + kk])
  C* prow = AK + (size_t)(sizeof(T) == 4 ? TCM : TC) * nb;  // nb
  C* pcol = prow + nb;                       // n
  int* piv = reinterpret_cast<int*>(pcol + n);  // n
  int* sh = piv + n;                         // scratch ints

  const int nsl = mode == 1 ? (int)gridDim.x / 2 : (int)gridDim.x;
  const bool up = mode == 1 && (int)blockIdx.x >= nsl;
  const int slot = up ? (int)blockIdx.x - nsl : (int)blockIdx.x;
  const bool rev = up || mode == 2;
  const int k = splane[slot];  // the plane whose z symbol this factor slot carries
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

  const int nblk = mode == 2 ? 1 : (up ? ny - 1 - m : jend - jbeg);
  C* const Gs = rev ? Gb + (size_t)slot * (ny - m) * nn : G + (size_t)slot * ny * nn;
  for (int jj = 0; jj < nblk; ++jj) {
    const int j = mode == 2 ? m : (up ? ny - 1 - jj : jbeg + jj);
    C* Gj = Gs + (size_t)(rev ? j - m : j) * nn;
    // ---- S_j = D_j - sig_j G_{j-1} (bottom-up: - sig_{j+1} B_{j+1}; the
    //      middle block both)
    const bool hasp = rev ? j + 1 < ny : j > 0;
    const C* Gp = rev ? Gj + nn : Gj - nn;
    const double2 sg = hasp ? tb.sig[rev ? j + 1 : j] : make_double2(0, 0);
    const bool mid = mode == 2 && m > 0;
    const C* Gm = mid ? G + ((size_t)slot * ny + (m - 1)) * nn : nullptr;
    const double2 sgm = mid ? tb.sig[m] : make_double2(0, 0);
    const double2 cdy = tb.cdy[j];
    for (size_t idx = tid; idx < nn; idx += nt) {
      const int r = (int)(idx % n), c = (int)(idx / n);
      double vx = 0.0, vy = 0.0;
      if (hasp) {
        const C g = Gp[idx];
        vx = -(sg.x * (double)g.x - sg.y * (double)g.y);
        vy = -(sg.x * (double)g.y + sg.y * (double)g.x);
      }
      if (mid) {
        const C g = Gm[idx];
        vx -= sgm.x * (double)g.x - sgm.y * (double)g.y;
        vy -= sgm.x * (double)g.y + sgm.y * (double)g.x;
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
      // trailing update of the columns outside the panel: G[:, rest] = P AK
      // (+ the old value outside the panel rows), AK = the panel rows of those
      // columns; complex64 on the matrix cores, complex128 on the vector units
      const int nrest = n - kb;
      if constexpr (sizeof(T) == 4) {
        // 16-row x 8-column tasks, C = [Re Y | Im Y] from two FP32 MFMAs per
        // k-slice: Re(P) [Re AK | Im AK] + Im(P) [-Im AK | Re AK]
        const int li = lane & 15, lg = lane >> 4;
        const int RTn = (n + 15) / 16;
        for (int c0 = 0; c0 < nrest; c0 += TCM) {
          const int tc = min(TCM, nrest - c0);
          __syncthreads();
          for (int idx = tid; idx < kb * tc; idx += nt) {
            const int kk = idx % kb, cl = idx / kb;
            const int cg = c0 + cl;
            const int cc = cg < p0 ? cg : cg + kb;
            AK[cl * nb + kk] = Gj[(p0 + kk) + (size_t)n * cc];
          }
          __syncthreads();
          const int CGn = (tc + 7) / 8;
          for (int task = wave; task < RTn * CGn; task += nt / 64) {
            const int rt = task % RTn, cgp = task / RTn;
            hz_v4f acc = {0.f, 0.f, 0.f, 0.f};
            for (int ks = 0; 4 * ks < kb; ++ks) {
              const int kq = 4 * ks + lg;
              C ga, bv;
              ga.x = ga.y = bv.x = bv.y = 0;
              if (kq < kb) {
                ga = P[kq * n + 16 * rt + li];  // rows past n read the next column (discarded)
                bv = AK[(cgp * 8 + (li & 7)) * nb + kq];
              }
              const float b1 = li < 8 ? bv.x : bv.y;
              const float b2 = li < 8 ? -bv.y : bv.x;
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ga.x, b1, acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ga.y, b2, acc, 0, 0, 0);
            }
            float im[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) im[r] = __shfl_xor(acc[r], 8);
            const int col = cgp * 8 + li;
            if (li < 8 && col < tc) {
              const int cg = c0 + col;
              const int cc = cg < p0 ? cg : cg + kb;
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int row = 16 * rt + 4 * lg + r;
                if (row < n) {
                  C* dst = &Gj[row + (size_t)n * cc];
                  C y;
                  y.x = acc[r];
                  y.y = im[r];
                  if (row < p0 || row >= p0 + kb) {
                    const C old = *dst;
                    y.x += old.x;
                    y.y += old.y;
                  }
                  *dst = y;
                }
              }
            }
          }
        }
      } else {
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
                                                         int64_t n2, int64_t N,
    const int* __restrict__ fmap) {
  typedef typename Cx<T>::t C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* v = reinterpret_cast<C*>(smem_raw);  // n
  C* cur = v + n;                          // n
  C* part = cur + n;                       // ngroups * n
  const int k = blockIdx.x;
  const int kf = fmap[k];  // factor slot; -1: a skipped plane (zeroed after the solve)
  if (kf < 0) return;
  const int64_t rhs = blockIdx.y;
  C* x = X + rhs * N + (int64_t)k * n2;
  const C* Gk = G + (size_t)kf * ny * n * n;
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

// acc (RPL rows, re/im interleaved) += unit a (RPL complex entries) * vc, as
// chains of fused multiply-adds into the accumulators (4 FMAs per complex
// product instead of a product followed by an add)
template <typename T, int RPL>
__device__ __forceinline__ void unit_fma_fused(const typename Unit<T>::t& a, typename Cx<T>::t vc, T* acc) {
  typedef typename Cx<T>::t C;
  const C* ag = reinterpret_cast<const C*>(&a);
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    acc[2 * r] = fma(ag[r].x, vc.x, acc[2 * r]);
    acc[2 * r] = fma(-ag[r].y, vc.y, acc[2 * r]);
    acc[2 * r + 1] = fma(ag[r].x, vc.y, acc[2 * r + 1]);
    acc[2 * r + 1] = fma(ag[r].y, vc.x, acc[2 * r + 1]);
  }
}

template <typename T, int RPL, int CH>
__global__ __launch_bounds__(1024) void k_solve_fast(const typename Cx<T>::t* __restrict__ G,
                                                     typename Cx<T>::t* __restrict__ X,
                                                     const double2* __restrict__ cmy,
                                                     const double2* __restrict__ cpy, int n, int ny, int64_t n2,
                                                     int64_t N, int nru, int ncg, int cpg,
    const int* __restrict__ fmap) {
  typedef typename Cx<T>::t C;
  typedef typename Unit<T>::t U;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* v = reinterpret_cast<C*>(smem_raw);  // n
  C* part = v + n;                        // ncg * n
  const int k = blockIdx.x;
  const int kf = fmap[k];  // factor slot; -1: a skipped plane (zeroed after the solve)
  if (kf < 0) return;
  const int64_t rhs = blockIdx.y;
  C* x = X + rhs * N + (int64_t)k * n2;
  const size_t nn = (size_t)n * n;
  const C* Gk = G + (size_t)kf * ny * nn;
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

// ------------------------------------------------------------ LDS-DMA solve
// The same forward/backward sweep with the factor blocks streamed through LDS
// by the gfx950 direct-to-LDS loads (global_load_lds_dwordx4).  A block is one
// contiguous column-major array; it is cut into tiles of TCOL whole columns
// (R rounds of 512 x 16 bytes) that flow through a ring of four LDS slots with
// two tiles in flight at all times -- also across the reduction, barriers and
// serial epilogue that end each step.  The ring slots are distinct LDS objects
// and the ring is unrolled four ways, so each DMA and each read names its slot
// statically; waits are explicit (s_waitcnt vmcnt(k R)) and barriers are raw
// s_barrier, which keeps in-flight DMA out of every wait except the one that
// needs it.  512 threads: thread (cg, ru) owns load unit ru (16 bytes = RPL
// rows) of TCOL/NCG columns of each tile.
namespace dma {
#define HZ_LDS __attribute__((address_space(3)))
constexpr int NT = 512;

__device__ __forceinline__ constexpr int wait_vm(int vm) {
  // s_waitcnt: vmcnt = vm, expcnt and lgkmcnt at their maxima (no wait)
  return (vm & 0xF) | (0x7 << 4) | (0xF << 8) | (((vm >> 4) & 3) << 14);
}
constexpr int WAIT_LGKM0 = 0xF | (0x7 << 4) | (0x0 << 8) | (0x3 << 14);  // lgkmcnt(0) only

__device__ __forceinline__ void raw_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
}  // namespace dma

template <typename T, int RPL, int NCG, int TCOL, int NX>
__global__ __launch_bounds__(dma::NT) void k_solve_dma(const typename Cx<T>::t* __restrict__ G,
                                                       typename Cx<T>::t* __restrict__ X,
                                                       const double2* __restrict__ cmy,
                                                       const double2* __restrict__ cpy, int ny, int64_t n2,
                                                       int64_t N,
    const int* __restrict__ fmap) {
  typedef typename Cx<T>::t C;
  typedef typename Unit<T>::t U;
  constexpr int n = NX;
  constexpr int nru = n / RPL;
  constexpr int units = TCOL * nru;                   // 16-byte units per tile
  constexpr int R = (units + dma::NT - 1) / dma::NT;  // DMA rounds per tile
  constexpr int PU = n * (int)sizeof(C) / 16;         // 16-byte units of one y-line
  static_assert(n % RPL == 0 && n % TCOL == 0 && TCOL % NCG == 0 && NCG * nru <= dma::NT, "tile split");
  static_assert(PU <= dma::NT && (n * (int)sizeof(C)) % 16 == 0, "line DMA");
  static_assert(2 * R <= 63, "vmcnt range");
  __shared__ U s0[R * dma::NT], s1[R * dma::NT], s2[R * dma::NT], s3[R * dma::NT];
  __shared__ U pbuf[dma::NT];   // epilogue operand of the current step (first PU units used)
  const C* pline = reinterpret_cast<const C*>(pbuf);
  __shared__ C v[n];
  __shared__ C part[NCG * n];

  const int k = blockIdx.x;
  const int kf = fmap[k];  // factor slot; -1: a skipped plane (zeroed after the solve)
  if (kf < 0) return;
  const int64_t rhs = blockIdx.y;
  C* x = X + rhs * N + (int64_t)k * n2;
  const size_t nn = (size_t)n * n;
  const C* Gk = G + (size_t)kf * ny * nn;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int cg = tid / nru, ru = tid - cg * nru;
  const bool worker = cg < NCG;
  constexpr int CPT = TCOL / NCG;                     // columns of a tile per worker
  constexpr int tps = n / TCOL;                       // tiles per block
  const int nsteps = 2 * ny - 1;
  const int TT = nsteps * tps;                        // tiles of the whole solve
  auto blk_of = [&](int s) { return s <= ny - 1 ? s : 2 * ny - 2 - s; };

  // DMA of tile Tc into slot SLOT: a flat copy of TCOL*n complex (contiguous).
  // Tiles past the end re-read tile 0 so that every wave always has exactly
  // two tiles of DMA behind the one it waits for.  (Macros, so that every DMA
  // and every read names its LDS slot object directly.)
#define HZ_DMA_TILE(TC_, SLOT)                                                                          \
  {                                                                                                     \
    int tc_ = (TC_) >= TT ? 0 : (TC_);  /* past the end: re-read tile 0 */                             \
    const int s_ = tc_ / tps, t_ = tc_ - s_ * tps;                                                      \
    const unsigned char* base_ = reinterpret_cast<const unsigned char*>(Gk + (size_t)blk_of(s_) * nn) + \
                                 (size_t)t_ * TCOL * n * sizeof(C);                                     \
    _Pragma("unroll") for (int r_ = 0; r_ < R; ++r_) {                                                  \
      const int pos_ = r_ * dma::NT + wv * 64 + lane;                                                   \
      const int u_ = pos_ < units ? pos_ : units - 1;                                                   \
      __builtin_amdgcn_global_load_lds((const void*)(base_ + (size_t)u_ * 16),                          \
                                       (HZ_LDS void*)(&SLOT[r_ * dma::NT + wv * 64]), 16, 0, 0);        \
    }                                                                                                   \
  }
  // DMA of one y-line of the plane vector (slot J_) into pbuf
#define HZ_DMA_LINE(J_)                                                                                   \
  {                                                                                                       \
    const unsigned char* base_ = reinterpret_cast<const unsigned char*>(x + (size_t)n * (J_));            \
    const int u_ = min(wv * 64 + lane, PU - 1);                                                           \
    __builtin_amdgcn_global_load_lds((const void*)(base_ + (size_t)u_ * 16), (HZ_LDS void*)(&pbuf[wv * 64]), \
                                     16, 0, 0);                                                           \
  }
  T acc[2 * RPL];
#define HZ_PROCESS(TC_, SLOT, NEXT)                                                   \
  {                                                                                   \
    const int Tc = (TC_);                                                             \
    const int s = Tc / tps, t = Tc - s * tps;                                         \
    __builtin_amdgcn_s_waitcnt(dma::wait_vm(2 * R));                                  \
    /* reads of the slot that the DMA below refills are done in every wave */       \
    __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);                                      \
    dma::raw_barrier();                                                               \
    if (Tc < TT && t == 0) {                                                          \
      _Pragma("unroll") for (int i = 0; i < 2 * RPL; ++i) acc[i] = 0;                 \
      int slot_line = 0;                                                              \
      if (s < ny - 1) slot_line = s + 1;                                              \
      else if (2 * ny - 2 - s > 0) slot_line = 2 * ny - 3 - s;                        \
      HZ_DMA_LINE(slot_line)                                                          \
    }                                                                                 \
    if (Tc < TT && worker) {                                                          \
      _Pragma("unroll") for (int j = 0; j < CPT; ++j) {                               \
        const int cl = cg * CPT + j;                                                  \
        unit_fma<T, RPL>(SLOT[cl * nru + ru], v[t * TCOL + cl], acc);                  \
      }                                                                               \
    }                                                                                 \
    HZ_DMA_TILE(Tc + 3, NEXT)                                                         \
    if (Tc < TT && t == tps - 1) HZ_EPILOGUE(s)                                      \
  }

  // end of a step: cross-group reduction, t_{s+1} / x_j, next vector
#define HZ_EPILOGUE(S_)                                                         \
  {                                                                             \
    const int se_ = (S_);                                                       \
    if (worker) {                                                               \
      _Pragma("unroll") for (int r = 0; r < RPL; ++r) {                         \
        C o;                                                                    \
        o.x = acc[2 * r];                                                       \
        o.y = acc[2 * r + 1];                                                   \
        part[cg * n + ru * RPL + r] = o;                                        \
      }                                                                         \
    }                                                                           \
    /* the line DMA of this step is older than the last 3 tiles of DMA */     \
    __builtin_amdgcn_s_waitcnt(dma::wait_vm(3 * R));                            \
    __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);                                \
    dma::raw_barrier();                                                         \
    if (tid < n) {                                                              \
      C y;                                                                      \
      y.x = 0; y.y = 0;                                                         \
      _Pragma("unroll") for (int q = 0; q < NCG; ++q) {                         \
        const C u = part[q * n + tid];                                          \
        y.x += u.x;                                                             \
        y.y += u.y;                                                             \
      }                                                                         \
      const C pre = pline[tid];                                                 \
      C nv;                                                                     \
      if (se_ < ny - 1) {                                                       \
        const double2 cm = cmy[se_ + 1];                                        \
        nv.x = pre.x - (T)(cm.x * y.x - cm.y * y.y);                            \
        nv.y = pre.y - (T)(cm.x * y.y + cm.y * y.x);                            \
        x[(size_t)n * (se_ + 1) + tid] = nv;                                    \
      } else {                                                                  \
        const int j = 2 * ny - 2 - se_;                                         \
        x[(size_t)n * j + tid] = y;                                             \
        if (j > 0) {                                                            \
          const double2 cp = cpy[j - 1];                                        \
          nv.x = pre.x - (T)(cp.x * y.x - cp.y * y.y);                          \
          nv.y = pre.y - (T)(cp.x * y.y + cp.y * y.x);                          \
        } else {                                                                \
          nv = y;                                                               \
        }                                                                       \
      }                                                                         \
      v[tid] = nv;                                                              \
    }                                                                           \
    __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);                                \
  }

  for (int i = tid; i < n; i += dma::NT) v[i] = x[i];
  __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
  HZ_DMA_TILE(0, s0)
  HZ_DMA_TILE(1, s1)
  HZ_DMA_TILE(2, s2)
  // the tile stream is padded to a multiple of four stages; stages past the
  // end only keep the DMA accounting uniform
  for (int T0 = 0; T0 < TT; T0 += 4) {
    HZ_PROCESS(T0, s0, s3)
    HZ_PROCESS(T0 + 1, s1, s0)
    HZ_PROCESS(T0 + 2, s2, s1)
    HZ_PROCESS(T0 + 3, s3, s2)
  }
#undef HZ_EPILOGUE
#undef HZ_PROCESS
#undef HZ_DMA_LINE
#undef HZ_DMA_TILE
  // drain the trailing DMA before the workgroup retires
  __builtin_amdgcn_s_waitcnt(dma::wait_vm(0));
}

template <typename T, int RPL, int NCG, int TCOL, int NX>
bool try_launch_dma(int n, int ny, dim3 grid, hipStream_t s, const void* G, void* x, const double2* cmy,
                    const double2* cpy, int64_t n2, int64_t N, const int* fmap) {
  if (n != NX) return false;
  typedef typename Cx<T>::t C;
  hipLaunchKernelGGL((k_solve_dma<T, RPL, NCG, TCOL, NX>), grid, dim3(dma::NT), 0, s, (const C*)G, (C*)x, cmy, cpy,
                     ny, n2, N, fmap);
  HZ_HIP(hipGetLastError());
  return true;
}

// ------------------------------------------------ multi-RHS LDS-DMA solve
// NR right-hand sides of one plane share every factor tile: the tile is
// streamed once through the LDS ring and each thread applies its 16-byte unit
// (RPL rows of one column) to all NR vectors.  Thread tid = ru * NCG + cg owns
// row unit ru and column group cg; the NCG partial sums of a row unit sit in
// adjacent lanes and are combined with xor shuffles, so no LDS reduction
// buffer is needed.  The tile DMA scatters units so that the LDS slot is read
// as SLOT[j * NW + tid] (contiguous per wave): LDS position p = j * NW + tid
// holds unit (cg * CPT + j) * nru + ru of the column-major tile.  The vectors
// sit in LDS as v[col][q] (right-hand sides fastest).  In the epilogue lane cg
// finishes right-hand sides cg, cg + NCG, ...; the operand it needs (one
// 16-byte unit of the next y-line per right-hand side) was fetched by the
// lane's own DMA at the start of the step, so only the wave's own wait count
// guards it.
template <typename T, int RPL, int NCG, int TCOL, int NX, int NR>
struct MrShape {
  typedef typename Cx<T>::t C;
  static constexpr int n = NX;
  static constexpr int nru = n / RPL;
  static constexpr int NW = NCG * nru;                  // worker threads
  static constexpr int CPT = TCOL / NCG;                // columns of a tile per thread
  static constexpr int tps = n / TCOL;                  // tiles per block
  static constexpr int units = TCOL * nru;              // 16-byte units per tile
  static constexpr int R = (units + dma::NT - 1) / dma::NT;
  static constexpr int KQ = (NR + NCG - 1) / NCG;       // right-hand sides per lane in the epilogue
  static constexpr int WE = (tps < 3 ? tps : 3) * R;    // epilogue wait: tiles issued after the line DMA
  static constexpr size_t lds = 4 * (size_t)R * dma::NT * 16 + (size_t)NR * n * sizeof(C) +
                                (size_t)(dma::NT / 64) * KQ * 64 * 16;
  static_assert(n % RPL == 0 && n % TCOL == 0 && TCOL % NCG == 0, "tile split");
  static_assert(NW <= dma::NT && (NCG & (NCG - 1)) == 0 && NCG <= 64, "thread split");
  static_assert(2 * R <= 63 && WE <= 63, "vmcnt range");
  static_assert(lds <= 160 * 1024, "LDS budget");
};

template <typename T, int RPL, int NCG, int TCOL, int NX, int NR>
__global__ __launch_bounds__(dma::NT) void k_solve_mr(const typename Cx<T>::t* __restrict__ G,
                                                      typename Cx<T>::t* __restrict__ X,
                                                      const double2* __restrict__ cmy,
                                                      const double2* __restrict__ cpy, int ny, int64_t n2, int64_t N,
    const int* __restrict__ fmap) {
  typedef MrShape<T, RPL, NCG, TCOL, NX, NR> S;
  typedef typename Cx<T>::t C;
  typedef typename Unit<T>::t U;
  constexpr int n = S::n, nru = S::nru, NW = S::NW, CPT = S::CPT, tps = S::tps, units = S::units, R = S::R;
  constexpr int KQ = S::KQ;
  __shared__ U s0[R * dma::NT], s1[R * dma::NT], s2[R * dma::NT], s3[R * dma::NT];
  __shared__ U pbuf[(dma::NT / 64) * KQ * 64];
  __shared__ C v[n * NR];

  const int k = blockIdx.x;
  const int kf = fmap[k];  // factor slot; -1: a skipped plane (zeroed after the solve)
  if (kf < 0) return;
  const int64_t rhs0 = (int64_t)blockIdx.y * NR;
  C* xk = X + rhs0 * N + (int64_t)k * n2;   // right-hand side q of this plane: xk + q * N
  const size_t nn = (size_t)n * n;
  const C* Gk = G + (size_t)kf * ny * nn;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int cg = tid & (NCG - 1), ru = tid / NCG;
  const bool worker = tid < NW;
  const int nsteps = 2 * ny - 1;
  const int TT = nsteps * tps;
  auto blk_of = [&](int s) { return s <= ny - 1 ? s : 2 * ny - 2 - s; };

  // tile-relative source unit of each DMA round of this lane (same for every tile)
  int uoff[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int p = r * dma::NT + tid;
    int u = units - 1;
    if (p < units) {
      const int j = p / NW, rem = p - j * NW;
      const int rr = rem / NCG, cc = rem - rr * NCG;
      u = (cc * CPT + j) * nru + rr;
    }
    uoff[r] = u;
  }
  // the epilogue right-hand sides of this lane and its 16-byte unit of a line
  const int ru_c = worker ? ru : 0;

#define HZ_MR_TILE(TC_, SLOT)                                                                           \
  {                                                                                                     \
    const int tc_ = (TC_) >= TT ? 0 : (TC_); /* past the end: re-read tile 0 */                        \
    const int s_ = tc_ / tps, t_ = tc_ - s_ * tps;                                                      \
    const unsigned char* base_ = reinterpret_cast<const unsigned char*>(Gk + (size_t)blk_of(s_) * nn) + \
                                 (size_t)t_ * TCOL * n * sizeof(C);                                     \
    _Pragma("unroll") for (int r_ = 0; r_ < R; ++r_) {                                                  \
      __builtin_amdgcn_global_load_lds((const void*)(base_ + (size_t)uoff[r_] * 16),                    \
                                       (HZ_LDS void*)(&SLOT[r_ * dma::NT + wv * 64]), 16, 0, 0);        \
    }                                                                                                   \
  }
#define HZ_MR_LINE(J_)                                                                                  \
  {                                                                                                     \
    _Pragma("unroll") for (int kq_ = 0; kq_ < KQ; ++kq_) {                                              \
      const int q_ = cg + NCG * kq_;                                                                    \
      const C* src_ = xk + (int64_t)(q_ < NR ? q_ : 0) * N + (size_t)n * (J_) + (size_t)ru_c * RPL;     \
      __builtin_amdgcn_global_load_lds((const void*)src_, (HZ_LDS void*)(&pbuf[(wv * KQ + kq_) * 64]),  \
                                       16, 0, 0);                                                       \
    }                                                                                                   \
  }
  T acc[NR * 2 * RPL];
#define HZ_MR_EPILOGUE(S_)                                                                      \
  {                                                                                             \
    const int se_ = (S_);                                                                       \
    _Pragma("unroll") for (int off_ = 1; off_ < NCG; off_ <<= 1) {                              \
      _Pragma("unroll") for (int i_ = 0; i_ < NR * 2 * RPL; ++i_) acc[i_] += __shfl_xor(acc[i_], off_); \
    }                                                                                           \
    __builtin_amdgcn_s_waitcnt(dma::wait_vm(S::WE));                                            \
    __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);                                                \
    dma::raw_barrier(); /* every wave is done reading v of this step */                        \
    if (worker) {                                                                               \
      _Pragma("unroll") for (int kq_ = 0; kq_ < KQ; ++kq_) {                                    \
        const int q_ = cg + NCG * kq_;                                                          \
        if (q_ < NR) {                                                                          \
          const U pu_ = pbuf[(wv * KQ + kq_) * 64 + lane];                                      \
          const C* pre_ = reinterpret_cast<const C*>(&pu_);                                     \
          C* xq_ = xk + (int64_t)q_ * N;                                                        \
          U out_;                                                                               \
          C* o_ = reinterpret_cast<C*>(&out_);                                                  \
          _Pragma("unroll") for (int r_ = 0; r_ < RPL; ++r_) {                                  \
            C y_;                                                                               \
            y_.x = acc[(q_ * RPL + r_) * 2];                                                    \
            y_.y = acc[(q_ * RPL + r_) * 2 + 1];                                                \
            C nv_;                                                                              \
            if (se_ < ny - 1) {                                                                 \
              const double2 cm = cmy[se_ + 1];                                                  \
              nv_.x = pre_[r_].x - (T)(cm.x * y_.x - cm.y * y_.y);                              \
              nv_.y = pre_[r_].y - (T)(cm.x * y_.y + cm.y * y_.x);                              \
              o_[r_] = nv_;                                                                     \
            } else {                                                                            \
              const int j = 2 * ny - 2 - se_;                                                   \
              o_[r_] = y_;                                                                      \
              if (j > 0) {                                                                      \
                const double2 cp = cpy[j - 1];                                                  \
                nv_.x = pre_[r_].x - (T)(cp.x * y_.x - cp.y * y_.y);                            \
                nv_.y = pre_[r_].y - (T)(cp.x * y_.y + cp.y * y_.x);                            \
              } else {                                                                          \
                nv_ = y_;                                                                       \
              }                                                                                 \
            }                                                                                   \
            v[(ru * RPL + r_) * NR + q_] = nv_;                                                 \
          }                                                                                     \
          const int line_ = se_ < ny - 1 ? se_ + 1 : 2 * ny - 2 - se_;                          \
          *reinterpret_cast<U*>(xq_ + (size_t)n * line_ + (size_t)ru * RPL) = out_;             \
        }                                                                                       \
      }                                                                                         \
    }                                                                                           \
    __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);                                                \
  }
#define HZ_MR_STAGE(TC_, SLOT, NEXT)                                                    \
  {                                                                                     \
    const int Tc = (TC_);                                                               \
    const int s = Tc / tps, t = Tc - s * tps;                                           \
    __builtin_amdgcn_s_waitcnt(dma::wait_vm(2 * R));                                    \
    __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);                                        \
    dma::raw_barrier();                                                                 \
    if (Tc < TT && t == 0) {                                                            \
      _Pragma("unroll") for (int i = 0; i < NR * 2 * RPL; ++i) acc[i] = 0;              \
      int slot_line = 0;                                                                \
      if (s < ny - 1) slot_line = s + 1;                                                \
      else if (2 * ny - 2 - s > 0) slot_line = 2 * ny - 3 - s;                          \
      HZ_MR_LINE(slot_line)                                                             \
    }                                                                                   \
    if (Tc < TT && worker) {                                                            \
      _Pragma("unroll") for (int j = 0; j < CPT; ++j) {                                 \
        const U a = SLOT[j * NW + tid];                                                 \
        const C* vc = &v[(t * TCOL + cg * CPT + j) * NR];                               \
        _Pragma("unroll") for (int q = 0; q < NR; ++q)                                  \
          unit_fma<T, RPL>(a, vc[q], &acc[q * 2 * RPL]);                                \
      }                                                                                 \
    }                                                                                   \
    HZ_MR_TILE(Tc + 3, NEXT)                                                            \
    if (Tc < TT && t == tps - 1) HZ_MR_EPILOGUE(s)                                      \
  }

  for (int i = tid; i < n * NR; i += dma::NT) {
    const int q = i % NR, row = i / NR;
    v[i] = xk[(int64_t)q * N + row];
  }
  __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
  HZ_MR_TILE(0, s0)
  HZ_MR_TILE(1, s1)
  HZ_MR_TILE(2, s2)
  for (int T0 = 0; T0 < TT; T0 += 4) {
    HZ_MR_STAGE(T0, s0, s3)
    HZ_MR_STAGE(T0 + 1, s1, s0)
    HZ_MR_STAGE(T0 + 2, s2, s1)
    HZ_MR_STAGE(T0 + 3, s3, s2)
  }
#undef HZ_MR_STAGE
#undef HZ_MR_EPILOGUE
#undef HZ_MR_LINE
#undef HZ_MR_TILE
  __builtin_amdgcn_s_waitcnt(dma::wait_vm(0));
}

// ------------------------------------------- 8-RHS streaming solve on MFMA
// Same LDS ring as k_solve_mr, for groups of 8 complex64 right-hand sides, with
// the block product Y = G V (n x 8 complex) on the matrix cores: per 16-row x
// 4-column tile of G two v_mfma_f32_16x16x4f32 accumulate
//   C += Re(G) [Re V | Im V]  +  Im(G) [-Im V | Re V]  =  [Re Y | Im Y]
// into one 16 x 16 FP32 tile (columns 0-7 real parts of the 8 vectors, 8-15
// imaginary parts).  Operand layout of 16x16x4 f32 (wave64): A lane l holds
// A[l % 16][l / 16], B lane l holds B[l / 16][l % 16], C lane l holds rows
// 4 (l / 16) + 0..3 of column l % 16.  The B operands of a k-slice come from
// the vectors in LDS once per wave and serve every row tile of the wave.
// Wave w owns row tiles w and w + 8.

template <int NX, int TCOL>
struct MfShape {
  static constexpr int n = NX;
  static constexpr int RT = (n + 15) / 16;                    // row tiles
  static constexpr int KS = TCOL / 4;                         // k-slices per tile
  static constexpr int tps = n / TCOL;
  static constexpr int units = TCOL * n / 2;                  // 16-byte units per tile
  static constexpr int R = (units + dma::NT - 1) / dma::NT;
  static constexpr int WE = (tps < 3 ? tps : 3) * R;
  static constexpr int TPW = (RT + 7) / 8;                    // row tiles per wave (max)
  static constexpr size_t lds = 4 * (size_t)R * dma::NT * 16 + (size_t)n * 8 * sizeof(float2) +
                                (size_t)8 * TPW * 64 * 16;
  static_assert(n % TCOL == 0 && TCOL % 4 == 0 && n % 2 == 0, "tile split");
  static_assert(TPW <= 2, "row tiles per wave");
  // the last row tile may read up to 15 rows past the block inside the slot
  static_assert((size_t)(TCOL - 1) * n + 16 * RT <= (size_t)R * dma::NT * 2, "slot slack for the last row tile");
  static_assert(2 * R <= 63 && WE <= 63, "vmcnt range");
  static_assert(lds <= 160 * 1024, "LDS budget");
};

// ADJ: the adjoint sweep in conjugated variables with G_j^H = conj(D_j^-1 G_j D_j)
// (G_j D_j symmetric, D_j = diag(dx(i) dy(j)) from symd = dx | dy | 1/dx | 1/dy):
// the vector v holds D_j u_j, the product z = G_j v gives Y_j = D_j^-1 z; forward
// u_{j+1} = conj(b_{j+1}) - cp_j Y_j (Y_j stored), backward X_j = Y_j - D_j^-1 z,
// x_j = conj(X_j), u = cm_j X_j.  Same block sequence as the forward solve.
template <int NX, int TCOL, bool ADJ>
__global__ __launch_bounds__(dma::NT) void k_solve_mfma8(const float2* __restrict__ G, float2* __restrict__ X,
                                                         const double2* __restrict__ cmy,
                                                         const double2* __restrict__ cpy,
                                                         const double2* __restrict__ symd, int ny, int64_t n2,
                                                         int64_t N,
    const int* __restrict__ fmap) {
  typedef MfShape<NX, TCOL> S;
  typedef float4 U;
  constexpr int n = S::n, RT = S::RT, KS = S::KS, tps = S::tps, units = S::units, R = S::R, TPW = S::TPW;
  constexpr int NR = 8;
  __shared__ U s0[R * dma::NT], s1[R * dma::NT], s2[R * dma::NT], s3[R * dma::NT];
  __shared__ U pbuf[8 * TPW * 64];
  __shared__ float2 v[n * NR];
  __shared__ float2 sdx[ADJ ? n : 1], sdxi[ADJ ? n : 1];  // dx, 1/dx (adjoint only)

  const int k = blockIdx.x;
  const int kf = fmap[k];  // factor slot; -1: a skipped plane (zeroed after the solve)
  if (kf < 0) return;
  const int64_t rhs0 = (int64_t)blockIdx.y * NR;
  float2* xk = X + rhs0 * N + (int64_t)k * n2;
  const size_t nn = (size_t)n * n;
  const float2* Gk = G + (size_t)kf * ny * nn;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;  // operand row/column index, k / row-group index
  const int nsteps = 2 * ny - 1;
  const int TT = nsteps * tps;
  auto blk_of = [&](int s) { return s <= ny - 1 ? s : 2 * ny - 2 - s; };

#define HZ_MF_TILE(TC_, SLOT)                                                                           \
  {                                                                                                     \
    const int tc_ = (TC_) >= TT ? 0 : (TC_); /* past the end: re-read tile 0 */                        \
    const int s_ = tc_ / tps, t_ = tc_ - s_ * tps;                                                      \
    const unsigned char* base_ = reinterpret_cast<const unsigned char*>(Gk + (size_t)blk_of(s_) * nn) + \
                                 (size_t)t_ * TCOL * n * sizeof(float2);                                \
    _Pragma("unroll") for (int r_ = 0; r_ < R; ++r_) {                                                  \
      const int p_ = r_ * dma::NT + tid;                                                                \
      const int u_ = p_ < units ? p_ : units - 1;                                                       \
      __builtin_amdgcn_global_load_lds((const void*)(base_ + (size_t)u_ * 16),                          \
                                       (HZ_LDS void*)(&SLOT[r_ * dma::NT + wv * 64]), 16, 0, 0);        \
    }                                                                                                   \
  }
  // the next line's operand for this wave's epilogue: lane l loads unit
  // 8 rt + 2 g + (l & 1) of right-hand side q, consumer (g, q) = (l >> 1) / 8, % 8
#define HZ_MF_LINE(J_)                                                                                  \
  {                                                                                                     \
    _Pragma("unroll") for (int a_ = 0; a_ < TPW; ++a_) {                                                \
      const int rt_ = wv + 8 * a_;                                                                      \
      if (rt_ < RT) {                                                                                   \
        const int c_ = lane >> 1, g_ = c_ >> 3, q_ = c_ & 7;                                            \
        int u_ = 8 * rt_ + 2 * g_ + (lane & 1);                                                         \
        u_ = u_ < n / 2 ? u_ : n / 2 - 1;                                                               \
        const float2* src_ = xk + (int64_t)q_ * N + (size_t)n * (J_) + 2 * (size_t)u_;                  \
        __builtin_amdgcn_global_load_lds((const void*)src_, (HZ_LDS void*)(&pbuf[(wv * TPW + a_) * 64]), \
                                         16, 0, 0);                                                     \
      }                                                                                                 \
    }                                                                                                   \
  }
  hz_v4f acc[TPW];
#define HZ_MF_EPILOGUE(S_)                                                                                \
  {                                                                                                       \
    const int se_ = (S_);                                                                                 \
    __builtin_amdgcn_s_waitcnt(dma::wait_vm(S::WE));                                                      \
    __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);                                                          \
    dma::raw_barrier(); /* every wave is done reading v of this step */                                  \
    _Pragma("unroll") for (int a_ = 0; a_ < TPW; ++a_) {                                                  \
      const int rt_ = wv + 8 * a_;                                                                        \
      float im_[4];                                                                                       \
      _Pragma("unroll") for (int r_ = 0; r_ < 4; ++r_) im_[r_] = __shfl_xor(acc[a_][r_], 8);             \
      if (rt_ < RT && li < 8) {                                                                           \
        const int q_ = li;                                                                                \
        const U p0_ = pbuf[(wv * TPW + a_) * 64 + 2 * (8 * lg + q_)];                                     \
        const U p1_ = pbuf[(wv * TPW + a_) * 64 + 2 * (8 * lg + q_) + 1];                                 \
        const float2 pre_[4] = {make_float2(p0_.x, p0_.y), make_float2(p0_.z, p0_.w),                   \
                                make_float2(p1_.x, p1_.y), make_float2(p1_.z, p1_.w)};                  \
        float2 out_[4];                                                                                   \
        _Pragma("unroll") for (int r_ = 0; r_ < 4; ++r_) {                                                \
          const int row_ = 16 * rt_ + 4 * lg + r_;                                                        \
          const float2 y_ = make_float2(acc[a_][r_], im_[r_]);                                            \
          float2 nv_;                                                                                     \
          if (ADJ) {                                                                                      \
            const int jc_ = blk_of(se_);                                                                  \
            const double2 dyi_ = symd[2 * n + ny + jc_];                                                  \
            const float2 sxi_ = sdxi[row_ < n ? row_ : 0];                                                \
            const double sx_ = (double)sxi_.x * dyi_.x - (double)sxi_.y * dyi_.y;                         \
            const double sy_ = (double)sxi_.x * dyi_.y + (double)sxi_.y * dyi_.x;                         \
            const double zx_ = y_.x * sx_ - y_.y * sy_, zy_ = y_.x * sy_ + y_.y * sx_;                    \
            double ux_, uy_;                                                                              \
            int jn_;                                                                                      \
            if (se_ < ny - 1) {                                                                           \
              const double2 cp = cpy[se_];                                                                \
              out_[r_] = make_float2((float)zx_, (float)zy_);                                             \
              ux_ = (double)pre_[r_].x - (cp.x * zx_ - cp.y * zy_);                                       \
              uy_ = -(double)pre_[r_].y - (cp.x * zy_ + cp.y * zx_);                                      \
              jn_ = se_ + 1;                                                                              \
            } else {                                                                                      \
              const double Xx_ = se_ >= ny ? (double)pre_[r_].x - zx_ : zx_;                              \
              const double Xy_ = se_ >= ny ? (double)pre_[r_].y - zy_ : zy_;                              \
              out_[r_] = make_float2((float)Xx_, (float)(-Xy_));                                          \
              const double2 cm = cmy[jc_];                                                                \
              ux_ = cm.x * Xx_ - cm.y * Xy_;                                                              \
              uy_ = cm.x * Xy_ + cm.y * Xx_;                                                              \
              jn_ = jc_ > 0 ? jc_ - 1 : 0;                                                                \
            }                                                                                             \
            const double2 dyn_ = symd[n + jn_];                                                           \
            const float2 sxn_ = sdx[row_ < n ? row_ : 0];                                                 \
            const double tx_ = (double)sxn_.x * dyn_.x - (double)sxn_.y * dyn_.y;                         \
            const double ty_ = (double)sxn_.x * dyn_.y + (double)sxn_.y * dyn_.x;                         \
            nv_ = make_float2((float)(ux_ * tx_ - uy_ * ty_), (float)(ux_ * ty_ + uy_ * tx_));            \
          } else if (se_ < ny - 1) {                                                                      \
            const double2 cm = cmy[se_ + 1];                                                              \
            nv_.x = pre_[r_].x - (float)(cm.x * y_.x - cm.y * y_.y);                                      \
            nv_.y = pre_[r_].y - (float)(cm.x * y_.y + cm.y * y_.x);                                      \
            out_[r_] = nv_;                                                                               \
          } else {                                                                                        \
            const int j = 2 * ny - 2 - se_;                                                               \
            out_[r_] = y_;                                                                                \
            if (j > 0) {                                                                                  \
              const double2 cp = cpy[j - 1];                                                              \
              nv_.x = pre_[r_].x - (float)(cp.x * y_.x - cp.y * y_.y);                                    \
              nv_.y = pre_[r_].y - (float)(cp.x * y_.y + cp.y * y_.x);                                    \
            } else {                                                                                      \
              nv_ = y_;                                                                                   \
            }                                                                                             \
          }                                                                                               \
          if (row_ < n) v[row_ * NR + q_] = nv_;                                                          \
        }                                                                                                 \
        const int line_ = ADJ ? blk_of(se_) : (se_ < ny - 1 ? se_ + 1 : 2 * ny - 2 - se_);                \
        float2* xq_ = xk + (int64_t)q_ * N + (size_t)n * line_ + 16 * rt_ + 4 * lg;                       \
        if (16 * rt_ + 4 * lg + 1 < n)                                                                    \
          *reinterpret_cast<U*>(xq_) = make_float4(out_[0].x, out_[0].y, out_[1].x, out_[1].y);           \
        if (16 * rt_ + 4 * lg + 3 < n)                                                                    \
          *reinterpret_cast<U*>(xq_ + 2) = make_float4(out_[2].x, out_[2].y, out_[3].x, out_[3].y);       \
      }                                                                                                   \
    }                                                                                                     \
    __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);                                                          \
  }
#define HZ_MF_STAGE(TC_, SLOT, NEXT)                                                          \
  {                                                                                           \
    const int Tc = (TC_);                                                                     \
    const int s = Tc / tps, t = Tc - s * tps;                                                 \
    __builtin_amdgcn_s_waitcnt(dma::wait_vm(2 * R));                                          \
    __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);                                              \
    dma::raw_barrier();                                                                       \
    if (Tc < TT && t == 0) {                                                                  \
      _Pragma("unroll") for (int a = 0; a < TPW; ++a) acc[a] = hz_v4f{0.f, 0.f, 0.f, 0.f};    \
      int slot_line = 0;                                                                      \
      if (s < ny - 1) slot_line = s + 1;                                                      \
      else if (ADJ && s >= ny) slot_line = 2 * ny - 2 - s;                                    \
      else if (!ADJ && 2 * ny - 2 - s > 0) slot_line = 2 * ny - 3 - s;                        \
      HZ_MF_LINE(slot_line)                                                                   \
    }                                                                                         \
    if (Tc < TT) {                                                                            \
      const float2* sl = reinterpret_cast<const float2*>(SLOT);                               \
      _Pragma("unroll") for (int ks = 0; ks < KS; ++ks) {                                     \
        const float2 vv = v[(t * TCOL + 4 * ks + lg) * NR + (li & 7)];                        \
        const float b1 = li < 8 ? vv.x : vv.y;                                                \
        const float b2 = li < 8 ? -vv.y : vv.x;                                               \
        _Pragma("unroll") for (int a = 0; a < TPW; ++a) {                                     \
          const int rt = wv + 8 * a;                                                          \
          if (rt < RT) {                                                                      \
            const float2 ga = sl[(4 * ks + lg) * n + 16 * rt + li];                           \
            acc[a] = __builtin_amdgcn_mfma_f32_16x16x4f32(ga.x, b1, acc[a], 0, 0, 0);         \
            acc[a] = __builtin_amdgcn_mfma_f32_16x16x4f32(ga.y, b2, acc[a], 0, 0, 0);         \
          }                                                                                   \
        }                                                                                     \
      }                                                                                       \
    }                                                                                         \
    HZ_MF_TILE(Tc + 3, NEXT)                                                                  \
    if (Tc < TT && t == tps - 1) HZ_MF_EPILOGUE(s)                                            \
  }

  if (ADJ) {
    for (int i = tid; i < n; i += dma::NT) {
      const double2 a = symd[i], b = symd[n + ny + i];
      sdx[i] = make_float2((float)a.x, (float)a.y);
      sdxi[i] = make_float2((float)b.x, (float)b.y);
    }
    __syncthreads();
  }
  for (int i = tid; i < n * NR; i += dma::NT) {
    const int q = i % NR, row = i / NR;
    const float2 b = xk[(int64_t)q * N + row];
    if (ADJ) {
      // v = D_0 conj(b_0)
      const float2 a = sdx[row];
      const double2 d0 = symd[n];
      const double tx = (double)a.x * d0.x - (double)a.y * d0.y, ty = (double)a.x * d0.y + (double)a.y * d0.x;
      v[i] = make_float2((float)(b.x * tx + b.y * ty), (float)(b.x * ty - b.y * tx));
    } else {
      v[i] = b;
    }
  }
  __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
  HZ_MF_TILE(0, s0)
  HZ_MF_TILE(1, s1)
  HZ_MF_TILE(2, s2)
  for (int T0 = 0; T0 < TT; T0 += 4) {
    HZ_MF_STAGE(T0, s0, s3)
    HZ_MF_STAGE(T0 + 1, s1, s0)
    HZ_MF_STAGE(T0 + 2, s2, s1)
    HZ_MF_STAGE(T0 + 3, s3, s2)
  }
#undef HZ_MF_STAGE
#undef HZ_MF_EPILOGUE
#undef HZ_MF_LINE
#undef HZ_MF_TILE
  __builtin_amdgcn_s_waitcnt(dma::wait_vm(0));
}

template <int NX, int TCOL>
bool try_launch_mfma8(int n, int ny, int nz, int groups, hipStream_t s, const void* G, void* x, const double2* cmy,
                      const double2* cpy, const double2* symd, bool adjoint, int64_t n2, int64_t N, const int* fmap) {
  if (n != NX) return false;
  const dim3 grid((unsigned)nz, (unsigned)groups);
  if (adjoint)
    hipLaunchKernelGGL((k_solve_mfma8<NX, TCOL, true>), grid, dim3(dma::NT), 0, s, (const float2*)G, (float2*)x, cmy,
                       cpy, symd, ny, n2, N, fmap);
  else
    hipLaunchKernelGGL((k_solve_mfma8<NX, TCOL, false>), grid, dim3(dma::NT), 0, s, (const float2*)G, (float2*)x,
                       cmy, cpy, symd, ny, n2, N, fmap);
  HZ_HIP(hipGetLastError());
  return true;
}

bool launch_mfma8_any(int n, int ny, int nz, int groups, hipStream_t s, const void* G, void* x, const double2* cmy,
                      const double2* cpy, const double2* symd, bool adjoint, int64_t n2, int64_t N, const int* fmap) {
  return try_launch_mfma8<200, 20>(n, ny, nz, groups, s, G, x, cmy, cpy, symd, adjoint, n2, N, fmap) ||
         try_launch_mfma8<192, 16>(n, ny, nz, groups, s, G, x, cmy, cpy, symd, adjoint, n2, N, fmap) ||
         try_launch_mfma8<160, 20>(n, ny, nz, groups, s, G, x, cmy, cpy, symd, adjoint, n2, N, fmap) ||
         try_launch_mfma8<128, 32>(n, ny, nz, groups, s, G, x, cmy, cpy, symd, adjoint, n2, N, fmap) ||
         try_launch_mfma8<96, 32>(n, ny, nz, groups, s, G, x, cmy, cpy, symd, adjoint, n2, N, fmap) ||
         try_launch_mfma8<64, 32>(n, ny, nz, groups, s, G, x, cmy, cpy, symd, adjoint, n2, N, fmap);
}

// ------------------------------------------ symmetric packed plane solve
// D_j^{-1} M_k is complex symmetric for D_j = diag(dx(i) dy(j)), hence so is
// K_j = G_j D_j, and G_j v = K_j (D_j^{-1} v).  The packed form keeps the upper
// block triangle of K_j in TB x TB tiles, block column J holding rows
// [0, (J+1) TB) column-major, with block columns J and T-1-J stored next to
// each other so that every pair (one DMA tile) has the same size (T+1) TB^2:
// 52% of the full block at T = 25.  A stored element K(r, c) feeds y_r += K v'_c
// and, below the diagonal tile, also y_c += K v'_r; the second product is a
// dot over the rows of column c, reduced across the wave with xor shuffles and
// accumulated per wave in LDS.  v' = D_j^{-1} v is formed when the next vector
// is written.
template <int NX, int TB>
struct SymLayout {
  static constexpr int T = NX / TB;                     // block columns
  static constexpr int NPAIR = T / 2;                   // pair tiles
  static constexpr int TPS = (T + 1) / 2;               // tiles per block (with the middle one when T is odd)
  static constexpr int PAIR = (T + 1) * TB * TB;        // complex per pair tile
  static constexpr int BS = TB * TB * T * (T + 1) / 2;  // complex per packed block
  static_assert(NX % TB == 0, "tile split");
};

// P[(s ny + j) BS + ...] = packed X_j D_j of factor slot s, X_j = G_j; with
// G2 (the twisted factors) X_j = G2[s][j - m] for j >= m
template <typename C, int NX, int TB>
__global__ __launch_bounds__(256) void k_pack_sym(const C* __restrict__ G, C* __restrict__ P,
                                                  const double2* __restrict__ dx, const double2* __restrict__ dy,
                                                  int ny, const C* __restrict__ G2, int m) {
  typedef SymLayout<NX, TB> L;
  const int blk = blockIdx.y;
  const int j = blk % ny;
  const int t = blockIdx.x;
  const int J1 = t, J2 = L::T - 1 - t;
  const bool two = J2 != J1;
  const int H1 = (J1 + 1) * TB, H2 = (J2 + 1) * TB;
  const int sz = H1 * TB + (two ? H2 * TB : 0);
  const C* Gb = (G2 && j >= m) ? G2 + ((size_t)(blk / ny) * (ny - m) + (j - m)) * NX * NX : G + (size_t)blk * NX * NX;
  C* Pb = P + (size_t)blk * L::BS + (size_t)t * L::PAIR;
  const double2 dyj = dy[j];
  for (int e = threadIdx.x; e < sz; e += blockDim.x) {
    int J, cl, r;
    if (e < H1 * TB) {
      J = J1; cl = e / H1; r = e - cl * H1;
    } else {
      const int e2 = e - H1 * TB;
      J = J2; cl = e2 / H2; r = e2 - cl * H2;
    }
    const int c = J * TB + cl;
    const C g = Gb[(size_t)c * NX + r];
    const double2 dc = dx[c];
    const double sx = dc.x * dyj.x - dc.y * dyj.y, sy = dc.x * dyj.y + dc.y * dyj.x;
    C o;
    o.x = (double)g.x * sx - (double)g.y * sy;
    o.y = (double)g.x * sy + (double)g.y * sx;
    Pb[e] = o;
  }
}

// Sum of a float over each 16-lane row of a wave (every lane of the row gets
// it): DPP adds with quad xor 1, quad xor 2, half-row mirror and row mirror.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
}
// one transpose-reduce step: CTRL pairs each lane with a lane whose bit
// (hi) differs; the lane with the bit set keeps b and adds the partner's b,
// the other keeps a and adds the partner's a, so two values per lane become one
template <int CTRL>
__device__ __forceinline__ float pair_step(float a, float b, bool hi) {
  const float keep = hi ? b : a, send = hi ? a : b;
  return keep + dpp_f<CTRL>(send);
}
// 16-lane row sums of 8 values per lane in 22 operations: afterwards lane l
// holds the row sum of x[l & 7] (lanes l and l + 8 alike)
__device__ __forceinline__ float row16_sum8(const float* x, int lane) {
  const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4;
  const float w0 = pair_step<0xB1>(x[0], x[1], b0), w1 = pair_step<0xB1>(x[2], x[3], b0);
  const float w2 = pair_step<0xB1>(x[4], x[5], b0), w3 = pair_step<0xB1>(x[6], x[7], b0);
  const float y0 = pair_step<0x4E>(w0, w1, b1), y1 = pair_step<0x4E>(w2, w3, b1);
  float z = pair_step<0x124>(y0, y1, b2);  // row_ror:4
  z += dpp_f<0x128>(z);                    // row_ror:8
  return z;
}
// the same for 4 values: lane l holds the row sum of x[l & 3]
__device__ __forceinline__ float row16_sum4(const float* x, int lane) {
  const bool b0 = lane & 1, b1 = lane & 2;
  const float w0 = pair_step<0xB1>(x[0], x[1], b0), w1 = pair_step<0xB1>(x[2], x[3], b0);
  float z = pair_step<0x4E>(w0, w1, b1);
  z += dpp_f<0x124>(z);  // row_ror:4
  z += dpp_f<0x128>(z);  // row_ror:8
  return z;
}

template <typename T, int RPL, int NCG, int TB, int NX>
struct SymShape {
  typedef SymLayout<NX, TB> L;
  typedef typename Cx<T>::t C;
  static constexpr int S = 8;                               // ring slots
  static constexpr int n = NX, nru = n / RPL;
  static constexpr int RUP = dma::NT / NCG;                 // thread slots per column group
  static constexpr int CPB = TB / NCG;                      // columns per thread per block column
  static constexpr int units = L::PAIR / RPL;               // 16-byte units per pair tile
  static constexpr int R = (units + dma::NT - 1) / dma::NT;
  static constexpr int PU = n * (int)sizeof(C) / 16;        // units of one y-line
  static constexpr int NQ = 4 * (RUP / 64);                 // partial sums of a transposed product
  static constexpr size_t lds = (size_t)S * R * dma::NT * 16 + (size_t)NCG * n * sizeof(C) +
                                (2 + NQ) * (size_t)n * sizeof(C) + 3 * 64 * 16 + S * 16;
  static_assert(nru <= RUP && RUP % 64 == 0 && TB % NCG == 0 && TB % RPL == 0, "thread split");
  static_assert(CPB == 2 && RPL == 2 && sizeof(T) == 4, "transposed values: 2 block columns x 2 columns x re/im");
  static_assert(RUP % 64 == 0 && (dma::NT / 64) % NCG == 0, "row units: whole waves per column group");
  static_assert(PU <= 128, "line DMA of two waves");
  static_assert((S - 1) * R <= 63, "vmcnt range");
  static_assert(L::TPS >= S - 1, "the epilogue wait assumes a step of at least S-1 tiles");
  static_assert(lds <= 160 * 1024, "LDS budget");
};

// OUT_[0..1] (re, im) += sum over the unit's RPL rows of A_(row) * W_(row):
// one column's transposed contribution from a 16-byte unit
#define HZ_SY_UNIT_DOT(A_, W_, OUT_)                                                               \
  {                                                                                                \
    const C* ag_ = reinterpret_cast<const C*>(&(A_));                                              \
    _Pragma("unroll") for (int i_ = 0; i_ < RPL; ++i_) {                                           \
      (OUT_)[0] = fma(ag_[i_].x, (W_)[i_].x, (OUT_)[0]);                                           \
      (OUT_)[0] = fma(-ag_[i_].y, (W_)[i_].y, (OUT_)[0]);                                          \
      (OUT_)[1] = fma(ag_[i_].x, (W_)[i_].y, (OUT_)[1]);                                           \
      (OUT_)[1] = fma(ag_[i_].y, (W_)[i_].x, (OUT_)[1]);                                           \
    }                                                                                              \
  }

// LDS DMA of 16 bytes per lane issued by inline assembly (M0 = the wave's
// LDS destination, lanes at 16-byte steps).  The compiler does not track it:
// reads of the ring below address it through a runtime slot offset and get no
// automatic waits; every wait on these transfers is an explicit s_waitcnt.
// (M0 is named as clobbered; the compiler sets M0 itself before each of its
// own uses, and this code has no dynamically indexed register arrays.)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void dma16_untracked(const void* g, unsigned lds_wave) {
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "s"(__builtin_amdgcn_readfirstlane(lds_wave))
               : "memory", "m0");
}
#pragma clang diagnostic pop

// Symmetric packed plane solve with the step unrolled: the TPS tiles of a
// block are a compile-time loop, so every tile parameter (block columns,
// stored and off-diagonal rows, offsets, tile size) is a constant and the
// per-tile scalar work is the ring slot, one DMA address and the waits.  The
// 8-slot ring is one LDS array addressed by a runtime slot; its transfers are
// untracked DMAs with explicit waits.  Products as in k_solve_sym: each unit
// serves its rows (direct product, registers) and its columns against the
// lane's own v' rows (transposed product, transpose-reduce row sums into NQ
// partials).
template <int NX, bool ADJ>
__global__ __launch_bounds__(dma::NT) void k_solve_sym2(const float2* __restrict__ P, float2* __restrict__ X,
                                                        const double2* __restrict__ cmy,
                                                        const double2* __restrict__ cpy,
                                                        const double2* __restrict__ symd, int ny, int64_t n2,
                                                        int64_t N,
    const int* __restrict__ fmap) {
  typedef float T;
  typedef float2 C;
  typedef float4 U;
  constexpr int RPL = 2, NCG = 4, TB = 8;
  typedef SymShape<T, RPL, NCG, TB, NX> S_;
  typedef SymLayout<NX, TB> L;
  constexpr int n = S_::n, nru = S_::nru, CPB = S_::CPB, R = S_::R, PU = S_::PU, NQ = S_::NQ;
  constexpr int TPS = L::TPS, NPAIR = L::NPAIR, PAIR = L::PAIR, BS = L::BS, T_ = L::T;
  constexpr int NS = S_::S;           // ring slots (a power of two)
  constexpr int ZU = R * dma::NT;     // zero unit of a slot
  constexpr int SU = ZU + 1;          // units per slot
  static_assert((NS & (NS - 1)) == 0, "ring slots: power of two");
  __shared__ U ring[NS * SU];
  __shared__ C part[NCG * n];
  __shared__ __align__(16) T tdot[NQ * n * 2];
  __shared__ C v[n];
  __shared__ U pbuf[3 * 64];
  __shared__ C sdxi[n];
  const C* pline = reinterpret_cast<const C*>(pbuf);
  const double2* pscal = reinterpret_cast<const double2*>(&pbuf[128]);

  const int k = blockIdx.x;
  const int kf = fmap[k];  // factor slot; -1: a skipped plane (zeroed after the solve)
  if (kf < 0) return;
  const int64_t rhs = blockIdx.y;
  C* x = X + rhs * N + (int64_t)k * n2;
  const C* Pk = P + (size_t)kf * ny * BS;
  const double2* dxi = symd + n + ny;
  const double2* dyi = symd + 2 * n + ny;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int cg = wv % NCG, rb = (wv / NCG) * 64, ru = rb + lane;
  const bool worker = ru < nru;
  const int nsteps = 2 * ny - 1;
  auto blk_of = [&](int st) { return st <= ny - 1 ? st : 2 * ny - 2 - st; };
  const unsigned ring_lds = (unsigned)(size_t)(&ring[0]);
  const unsigned pbuf_lds = (unsigned)(size_t)(&pbuf[0]);

  for (int i = tid; i < NQ * n * 2; i += dma::NT) tdot[i] = 0;
  if (tid < NS) ring[tid * SU + ZU] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int i = tid; i < n; i += dma::NT) {
    const double2 a = dxi[i];
    sdxi[i] = make_float2((T)a.x, (T)a.y);
  }
  __syncthreads();
  for (int i = tid; i < n; i += dma::NT) {
    const C x0 = x[i];
    if (ADJ) {
      v[i] = make_float2(x0.x, -x0.y);  // u_0 = conj(b_0)
    } else {
      const C a = sdxi[i];
      const double2 b = dyi[blk_of(0)];
      const double2 di = make_double2((double)a.x * b.x - (double)a.y * b.y, (double)a.x * b.y + (double)a.y * b.x);
      v[i] = make_float2((T)((double)x0.x * di.x - (double)x0.y * di.y),
                         (T)((double)x0.x * di.y + (double)x0.y * di.x));
    }
  }
  __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);

  // tile tt of the block at byte address blk into ring slot sl (tiles past the
  // end of the sweep re-read tile 0 so that every stage issues R transfers)
  auto tile_dma = [&](const char* blk, int tt, int sl) {
    const int tu = (tt < NPAIR ? PAIR : (NPAIR + 1) * TB * TB) / RPL;
    const char* base = blk + (size_t)tt * PAIR * sizeof(C);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int u = min(r * dma::NT + tid, tu - 1);
      dma16_untracked(base + (size_t)u * 16, ring_lds + (unsigned)((sl * SU + r * dma::NT + wv * 64) * 16));
    }
  };
  auto blk_addr = [&](int st) {
    return reinterpret_cast<const char*>(Pk + (size_t)blk_of(st < nsteps ? st : 0) * BS);
  };
#pragma unroll
  for (int g = 0; g < NS - 1; ++g) tile_dma(blk_addr(0), g, g);

  T acc[2 * RPL];
  U vr;
  int slot = 0;
  for (int st = 0; st < nsteps; ++st) {
    const char* b0 = blk_addr(st);
    const char* b1 = blk_addr(st + 1);
#pragma unroll
    for (int t = 0; t < TPS; ++t) {
      __builtin_amdgcn_s_waitcnt(dma::wait_vm((NS - 2) * R));
      __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
      dma::raw_barrier();
      const U* SL = ring + slot * SU;
      if (t == 0) {
#pragma unroll
        for (int i = 0; i < 2 * RPL; ++i) acc[i] = 0;
        vr = reinterpret_cast<const U*>(v)[worker ? ru : 0];
        if (!worker) vr = make_float4(0.f, 0.f, 0.f, 0.f);
        // next y-line (waves 0-1) and the step's coefficients (wave 2)
        int line = 0;
        if (st < ny - 1) line = st + 1;
        else if (ADJ && st >= ny) line = 2 * ny - 2 - st;  // Y_j of the backward block
        else if (!ADJ && 2 * ny - 2 - st > 0) line = 2 * ny - 3 - st;
        if (wv < 2) {
          const char* lb = reinterpret_cast<const char*>(x + (size_t)n * line);
          dma16_untracked(lb + (size_t)min(tid, PU - 1) * 16, pbuf_lds + (unsigned)(wv * 64 * 16));
        } else if (wv == 2) {
          const double2* src;
          if (ADJ) {
            // cm_y and cp_y of the current block, 1/d_y of the current block
            const int jc = blk_of(st);
            src = lane == 1 ? cpy + min(st, ny - 1) : (lane == 2 ? dyi + jc : cmy + jc);
          } else {
            const int sn = min(st + 1, nsteps - 1);
            const int jb = min(max(2 * ny - 3 - st, 0), ny - 1);
            src = lane == 1 ? cpy + jb : (lane == 2 ? dyi + blk_of(sn) : cmy + min(st + 1, ny - 1));
          }
          dma16_untracked(src, pbuf_lds + 128 * 16);
        }
      }
      {
        constexpr int dummy = 0;
        (void)dummy;
        const int J1 = t, J2 = T_ - 1 - t;
        const int Hu1 = (J1 + 1) * TB / RPL, Hu2 = (J2 + 1) * TB / RPL;
        const int nu1 = J1 * TB / RPL, nu2 = J2 * TB / RPL;
        const int ub2 = (J1 + 1) * TB * TB / RPL;
        const bool two = J2 != J1;
        const C* vrow = reinterpret_cast<const C*>(&vr);
        const C zc = make_float2(0.f, 0.f);
        if (rb < Hu1) {
          const bool ok1 = worker && ru < Hu1;
          const bool ok2 = worker && two && ru < Hu2;
          const int g1 = ok1 ? cg * CPB * Hu1 + ru : ZU;
          const int g2 = ok2 ? ub2 + cg * CPB * Hu2 + ru : ZU;
          const int st1 = ok1 ? Hu1 : 0, st2 = ok2 ? Hu2 : 0;
          U a1[CPB], a2[CPB];
#pragma unroll
          for (int cc = 0; cc < CPB; ++cc) {
            a1[cc] = SL[g1 + cc * st1];
            a2[cc] = SL[g2 + cc * st2];
          }
          const C* v1 = &v[J1 * TB + cg * CPB];
          const C* v2 = &v[J2 * TB + cg * CPB];
          C w1[RPL], w2[RPL];
#pragma unroll
          for (int i = 0; i < RPL; ++i) {
            w1[i] = ru < nu1 ? vrow[i] : zc;
            w2[i] = ru < nu2 ? vrow[i] : zc;
          }
          T tv[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) tv[i] = 0;
#pragma unroll
          for (int cc = 0; cc < CPB; ++cc) {
            unit_fma_fused<T, RPL>(a1[cc], v1[cc], acc);
            unit_fma_fused<T, RPL>(a2[cc], v2[cc], acc);
            HZ_SY_UNIT_DOT(a1[cc], w1, &tv[2 * cc])
            HZ_SY_UNIT_DOT(a2[cc], w2, &tv[4 + 2 * cc])
          }
          const T sm = row16_sum8(tv, lane);
          const int kk = lane & 7;
          if ((lane & 8) == 0 && (two || kk < 4)) {
            const int col = ((kk & 4) ? J2 : J1) * TB + cg * CPB + ((kk >> 1) & 1);
            const int q = (rb >> 6) * 4 + (lane >> 4);
            tdot[(q * n + col) * 2 + (kk & 1)] = sm;
          }
        } else if (rb < Hu2 && two) {
          const bool ok2 = worker && ru < Hu2;
          const int g2 = ok2 ? ub2 + cg * CPB * Hu2 + ru : ZU;
          const int st2 = ok2 ? Hu2 : 0;
          U a2[CPB];
#pragma unroll
          for (int cc = 0; cc < CPB; ++cc) a2[cc] = SL[g2 + cc * st2];
          const C* v2 = &v[J2 * TB + cg * CPB];
          C w2[RPL];
#pragma unroll
          for (int i = 0; i < RPL; ++i) w2[i] = ru < nu2 ? vrow[i] : zc;
          T tv[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) tv[i] = 0;
#pragma unroll
          for (int cc = 0; cc < CPB; ++cc) {
            unit_fma_fused<T, RPL>(a2[cc], v2[cc], acc);
            HZ_SY_UNIT_DOT(a2[cc], w2, &tv[2 * cc])
          }
          const T sm = row16_sum4(tv, lane);
          const int kk = lane & 3;
          if ((lane & 12) == 0) {
            const int col = J2 * TB + cg * CPB + (kk >> 1);
            const int q = (rb >> 6) * 4 + (lane >> 4);
            tdot[(q * n + col) * 2 + (kk & 1)] = sm;
          }
        }
      }
      {
        // the tile NS - 1 stages ahead into the slot the previous stage read
        constexpr int dummy2 = 0;
        (void)dummy2;
        const int ta = t + NS - 1;
        const int tt = ta % TPS;
        const char* blk = (ta / TPS) == 0 ? b0 : b1;
        const bool past = (ta / TPS) != 0 && st + 1 >= nsteps;
        tile_dma(past ? reinterpret_cast<const char*>(Pk) : blk, past ? 0 : tt, (slot + NS - 1) & (NS - 1));
      }
      if (t == TPS - 1) {
        if (worker) {
#pragma unroll
          for (int r = 0; r < RPL; ++r) part[cg * n + ru * RPL + r] = make_float2(acc[2 * r], acc[2 * r + 1]);
        }
        __builtin_amdgcn_s_waitcnt(dma::wait_vm((NS - 1) * R));
        __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
        dma::raw_barrier();
        if (tid < n) {
          C y = make_float2(0.f, 0.f);
#pragma unroll
          for (int q = 0; q < NCG; ++q) {
            const C u = part[q * n + tid];
            y.x += u.x; y.y += u.y;
          }
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const C u = reinterpret_cast<const C*>(tdot)[q * n + tid];
            y.x += u.x; y.y += u.y;
          }
          const C pre = pline[tid];
          if (ADJ) {
            // conjugated space: z = P_j u, Z = D_j^-1 z; forward Y_j = Z (stored),
            // u = conj(b_{j+1}) - cp_j Y_j; backward X_j = Y_j - Z, x_j = conj(X_j),
            // u = cm_j X_j
            const C a = sdxi[tid];
            const double2 b = pscal[2];
            const double2 di =
                make_double2((double)a.x * b.x - (double)a.y * b.y, (double)a.x * b.y + (double)a.y * b.x);
            const double zx = (double)y.x * di.x - (double)y.y * di.y, zy = (double)y.x * di.y + (double)y.y * di.x;
            C u;
            if (st < ny - 1) {
              const double2 cp = pscal[1];
              x[(size_t)n * st + tid] = make_float2((T)zx, (T)zy);
              u.x = (T)((double)pre.x - (cp.x * zx - cp.y * zy));
              u.y = (T)(-(double)pre.y - (cp.x * zy + cp.y * zx));
            } else {
              double X_x = zx, X_y = zy;
              int j = ny - 1;
              if (st >= ny) {
                j = 2 * ny - 2 - st;
                X_x = (double)pre.x - zx;
                X_y = (double)pre.y - zy;
              }
              x[(size_t)n * j + tid] = make_float2((T)X_x, (T)(-X_y));
              const double2 cm = pscal[0];
              u.x = (T)(cm.x * X_x - cm.y * X_y);
              u.y = (T)(cm.x * X_y + cm.y * X_x);
            }
            if (st + 1 < nsteps) v[tid] = u;
          } else {
          C nv;
          if (st < ny - 1) {
            const double2 cm = pscal[0];
            nv.x = pre.x - (T)(cm.x * y.x - cm.y * y.y);
            nv.y = pre.y - (T)(cm.x * y.y + cm.y * y.x);
            x[(size_t)n * (st + 1) + tid] = nv;
          } else {
            const int j = 2 * ny - 2 - st;
            x[(size_t)n * j + tid] = y;
            if (j > 0) {
              const double2 cp = pscal[1];
              nv.x = pre.x - (T)(cp.x * y.x - cp.y * y.y);
              nv.y = pre.y - (T)(cp.x * y.y + cp.y * y.x);
            } else {
              nv = y;
            }
          }
          if (st + 1 < nsteps) {
            const C a = sdxi[tid];
            const double2 b = pscal[2];
            const double2 di =
                make_double2((double)a.x * b.x - (double)a.y * b.y, (double)a.x * b.y + (double)a.y * b.x);
            v[tid] = make_float2((T)((double)nv.x * di.x - (double)nv.y * di.y),
                                 (T)((double)nv.x * di.y + (double)nv.y * di.x));
          }
          }
        }
        __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
      }
      slot = (slot + 1) & (NS - 1);
    }
  }
  __builtin_amdgcn_s_waitcnt(dma::wait_vm(0));
}

// lane l ^ 16 and lane l ^ 32 of a wave (gfx950 permlane swaps)
__device__ __forceinline__ float xor16_f(float x, int lane) {
  const unsigned u = __builtin_bit_cast(unsigned, x);
  const auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false);
  return __builtin_bit_cast(float, (lane & 16) ? r[0] : r[1]);
}
__device__ __forceinline__ float xor32_f(float x, int lane) {
  const unsigned u = __builtin_bit_cast(unsigned, x);
  const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
  return __builtin_bit_cast(float, (lane & 32) ? r[0] : r[1]);
}
// ---- twisted packed solve of both planes of a factor slot
// Planes k and nz - k share a factor slot.  The twisted factors split every
// slot's chain at m: blocks j < m hold the top-down inverses G_j, block m
// T_m = (D_m - sig_m G_{m-1} - sig_{m+1} B_{m+1})^{-1} and blocks j > m the
// bottom-up inverses B_j = (D_j - sig_{j+1} B_{j+1})^{-1} (sig_j = cm_y(j)
// cp_y(j-1)), all packed like the forward factors.  Phase 0 (forward):
//   top:    t_0 = f_0, t_j = f_j - cm_y(j) G_{j-1} t_{j-1} (0 < j < m), y_top = G_{m-1} t_{m-1}
//   bottom: b_{ny-1} = f_{ny-1}, b_j = f_j - cp_y(j) B_{j+1} b_{j+1} (m < j < ny-1), y_bot = B_{m+1} b_{m+1}
// Phase 1 (backward):
//   x_m = T_m (f_m - cm_y(m) y_top - cp_y(m) y_bot),
//   top x_j = G_j (t_j - cp_y(j) x_{j+1}) down to j = 0, bottom x_j = B_j (b_j - cm_y(j) x_{j-1}) up to ny-1.
// One workgroup per (slot, half) carries both planes of the slot: every packed
// tile is read once for two right-hand sides and 2 nslot workgroups walk
// chains of about ny/2 blocks (the one-plane kernel: nz workgroups, 2 ny - 1
// blocks).  y_top, y_bot and a copy of f_m pass between the phases in scr
// (phase 1 writes x_m into line m, which the other half reads as f_m).
//
// Waves: wave (cg, rb), cg = wave % 4, rb = 64 (wave / 4), takes columns
// 2cg, 2cg + 1 of both block columns of a tile pair (J1 = t, J2 = T - 1 - t)
// and row units rb .. rb + 63 (a unit is two rows of one column).  Both
// products run on the matrix cores (v_mfma_f32_4x4x1f32: 16 blocks of 4x4x1;
// D(lane 4b+j, reg i) += A(lane 4b+i) B(lane 4b+j)).  The four real columns
// of every product are the two planes' real and imaginary parts, so no block
// column idles:
//   direct, y_r += K(r, c) v'_c: lane l holds row unit rb + l (rows 2(rb+l)
//     and 2(rb+l)+1); A = Re or Im K of one of those rows, B = the components
//     of v'_c, so D(lane 4b+j, reg i) is row 2(rb+4b+i) (+1), component j,
//     accumulated over the step's tiles;
//   transposed, y_c += K(r, c) v'_r (rows above the diagonal tile): lane 4b+j
//     holds (row unit rb+16k+b, column c_j) of the wave's four columns;
//     A = the components of v'_r, B = Re or Im K, so D(lane 4b+j, reg i) sums
//     block b's rows for column c_j, component i, and is reduced over the 16
//     blocks at the end of the tile.
// Components: 0, 1 = Re, Im of plane A; 2, 3 = Re, Im of plane B.  Im K
// multiplies the rotated components (-Im A, Re A, -Im B, Re B).
//
// Tile rings: each wave moves exactly the part of every tile pair that it
// reads into a ring of its own, so tiles need no barrier: a wave waits for its
// own transfers only, and the workgroup meets twice per block step (partial
// sums, then the new right-hand side).  Slot layouts (16-byte units; segments
// 68 units apart, so the transposed reads of four columns fall in distinct
// 4-bank groups):
//   rb = 0:  [0, Hu1), [68, 68 + Hu1)   columns 2cg, 2cg + 1 of J1
//            [136, 200), [204, 268)    columns 2cg, 2cg + 1 of J2, row units
//                                      0 .. min(64, Hu2); unit 272 holds zeros
//   rb = 64: [0, 64), [68, 132)        columns 2cg, 2cg + 1 of J2, row units
//                                      64 .. Hu2; unit 136 holds zeros
// (Hu = 4 (J + 1) stored row units of block column J.)  Every tile costs a
// fixed number of transfers per wave (4, or 2 for rb = 64; lanes past a
// segment repeat its last unit, tiles without rows for the wave repeat the
// tile's first unit), so the vmcnt waits count tiles exactly.  The waves of
// rb = 64 hold no row of J1 (Hu1 <= 64) and skip the tiles whose J2 rows end
// before unit 64.
template <int NX, int PH>
__global__ __launch_bounds__(dma::NT) void k_solve_twist_pw(const float2* __restrict__ P, float2* __restrict__ X,
                                                            const double2* __restrict__ cmy,
                                                            const double2* __restrict__ cpy,
                                                            const double2* __restrict__ symd, int ny, int m, int nz,
                                                            int64_t n2, int64_t N, const int2* __restrict__ slotp,
                                                            float2* __restrict__ scr) {
  typedef float T;
  typedef float2 C;
  typedef float4 U;
  constexpr int RPL = 2, NCG = 4, TB = 8;
  typedef SymLayout<NX, TB> L;
  constexpr int n = NX, T_ = L::T, TPS = L::TPS, PAIR = L::PAIR, BS = L::BS;
  constexpr int PU = n * (int)sizeof(C) / 16;  // units of one y-line
  constexpr int NW = 5;                        // ring slots per wave
  // column segments of a slot start at units 0, 68, 136, 204 (distinct
  // 4-bank groups: the transposed reads of four columns are conflict-free);
  // the zero unit of a slot is its last unit
  constexpr int SEG = 68;
  constexpr int SLO = 273, SHI = 137;          // units per slot (rb = 0, rb = 64)
  constexpr int LLO = 4, LHI = 2;              // transfers per tile
  constexpr bool HAS_HI = n / RPL > 64;
  // transposed blocks of the second row half that can hold a stored row
  constexpr int KT64 = 4 * (T_ - 1) <= 64 ? 0 : (4 * (T_ - 1) - 64 + 15) / 16;
  static_assert(n / RPL <= 128 && KT64 <= 4, "two row halves of 64 units");
  static_assert(TPS * TB / RPL <= 64, "block column J1 lies in the first row half");
  static_assert(NW <= TPS && TPS * LLO <= 63, "vmcnt range");
  static_assert(PU <= 128, "line DMA of two waves per plane");
  static_assert(2 * n <= dma::NT, "one thread per plane and row in the step epilogue");
  static_assert((size_t)(4 * NW * (SLO + SHI) + (NCG + 2 + 2) * n + 5 * 64) * 16 + n * 8 <= 160 * 1024,
                "LDS budget");
  __shared__ U ring[4 * NW * SLO + 4 * NW * SHI];
  __shared__ U part[NCG * n];   // direct products per column group
  __shared__ U tdot[2 * n];     // transposed products per row half
  __shared__ U vq[n], vqr[n];   // v' components and rotated components
  __shared__ U pbuf[5 * 64];    // lines of planes A (units 0..127) and B (128..255); coefficients at 256..319
  __shared__ C sdxi[n];
  const double2* pscal = reinterpret_cast<const double2*>(&pbuf[256]);
  const T* vqf = reinterpret_cast<const T*>(vq);
  const T* vqrf = reinterpret_cast<const T*>(vqr);
  T* tdf = reinterpret_cast<T*>(tdot);
  T* pf = reinterpret_cast<T*>(part);

  const int slot = blockIdx.x >> 1, half = blockIdx.x & 1;
  const int2 sp = slotp[slot];
  const bool twoB = sp.y >= 0;
  const int pl0 = __builtin_amdgcn_readfirstlane(sp.x);
  const int pl1 = __builtin_amdgcn_readfirstlane(twoB ? sp.y : sp.x);
  const int64_t rhs = blockIdx.y;
  const int nrhs = gridDim.y;
  C* const xs0 = X + rhs * N + (int64_t)pl0 * n2;
  C* const xs1 = X + rhs * N + (int64_t)pl1 * n2;
  auto xs = [&](int s) { return s ? xs1 : xs0; };
  // scr: [what][rhs][plane][n], what = 0 y_top, 1 y_bot, 2 f_m
  auto scr_at = [&](int what, int s) { return scr + (((int64_t)what * nrhs + rhs) * nz + (s ? pl1 : pl0)) * n; };
  const C* Pk = P + (size_t)slot * ny * BS;
  const double2* dxi = symd + n + ny;
  const double2* dyi = symd + 2 * n + ny;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int cg = wv % NCG;
  const int bq = lane >> 2, jq = lane & 3;  // the lane's MFMA block and index in the block
  const int nsteps = PH == 0 ? (half == 0 ? m : ny - 1 - m) : (half == 0 ? m + 1 : ny - m);
  auto blk_of = [&](int st) {
    if (PH == 0) return half == 0 ? st : ny - 1 - st;
    return half == 0 ? m - st : m + st;
  };
  // block whose vector the update after step st forms (-1: none; -2: y goes to scr)
  auto next_of = [&](int st) {
    const int jb = blk_of(st);
    if (PH == 0) {
      if (half == 0) return jb + 1 < m ? jb + 1 : -2;
      return jb - 1 > m ? jb - 1 : -2;
    }
    if (half == 0) return jb > 0 ? jb - 1 : -1;
    return jb < ny - 1 ? jb + 1 : -1;
  };
  const unsigned ring_lds = (unsigned)(size_t)(&ring[0]);
  const unsigned pbuf_lds = (unsigned)(size_t)(&pbuf[0]);

  if (PH == 0 && half == 0) {
    // f_m for phase 1 (phase 1 overwrites line m with x_m)
    for (int i = tid; i < 2 * n; i += dma::NT) {
      const int s = i / n, c = i - s * n;
      if (s == 0 || twoB) scr_at(2, s)[c] = xs(s)[(size_t)m * n + c];
    }
  }
  if (nsteps <= 0) return;
  for (int i = tid; i < 2 * n; i += dma::NT) tdot[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (tid < 8 * NW) {
    const int w = tid / NW, s = tid - w * NW;
    const int z = w < 4 ? (w * NW + s) * SLO + SLO - 1 : 4 * NW * SLO + ((w - 4) * NW + s) * SHI + SHI - 1;
    ring[z] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  for (int i = tid; i < n; i += dma::NT) {
    const double2 a = dxi[i];
    sdxi[i] = make_float2((T)a.x, (T)a.y);
  }
  __syncthreads();
  {
    const int j0 = blk_of(0);
    const double2 dy0 = dyi[j0];
    for (int c = tid; c < n; c += dma::NT) {
      T w[4];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        double fx, fy;
        if (PH == 0) {
          const C f = xs(s)[(size_t)j0 * n + c];
          fx = f.x;
          fy = f.y;
        } else {
          // f_m - cm_y(m) y_top - cp_y(m) y_bot
          const C f = scr_at(2, s)[c];
          fx = f.x;
          fy = f.y;
          if (m > 0) {
            const C yt = scr_at(0, s)[c];
            const double2 cm = cmy[m];
            fx -= cm.x * yt.x - cm.y * yt.y;
            fy -= cm.x * yt.y + cm.y * yt.x;
          }
          if (m < ny - 1) {
            const C yb = scr_at(1, s)[c];
            const double2 cp = cpy[m];
            fx -= cp.x * yb.x - cp.y * yb.y;
            fy -= cp.x * yb.y + cp.y * yb.x;
          }
        }
        const C a = sdxi[c];
        const double dx = (double)a.x * dy0.x - (double)a.y * dy0.y, dy = (double)a.x * dy0.y + (double)a.y * dy0.x;
        w[2 * s] = (T)(fx * dx - fy * dy);
        w[2 * s + 1] = (T)(fx * dy + fy * dx);
      }
      vq[c] = make_float4(w[0], w[1], w[2], w[3]);
      vqr[c] = make_float4(-w[1], w[0], -w[3], w[2]);
    }
  }
  __syncthreads();

  auto blk_addr = [&](int st) {
    return reinterpret_cast<const char*>(Pk + (size_t)blk_of(st < nsteps ? st : 0) * BS);
  };
  const hz_v4f z4 = {0.f, 0.f, 0.f, 0.f};

  auto run = [&](auto hi_tag) {
    constexpr bool HI = decltype(hi_tag)::value;
    constexpr int SZ = HI ? SHI : SLO, ZU = SZ - 1, LD = HI ? LHI : LLO;
    constexpr int rb = HI ? 64 : 0, q2 = HI ? 1 : 0;
    const int wq = HI ? wv - 4 : wv;
    const int wbase = HI ? 4 * NW * SLO + wq * NW * SHI : wq * NW * SLO;
    const U* WR = ring + wbase;
    const unsigned wr_lds = ring_lds + (unsigned)wbase * 16u;
    const int ca = cg * 2;
    // the wave's transfers of tile tt of a block into ring slot sl
    auto tile_dma = [&](const char* blk, int tt, int sl) {
      const int J1 = tt, J2 = T_ - 1 - tt;
      const bool two = J2 != J1;
      const int Hu1 = (J1 + 1) * TB / RPL, Hu2 = (J2 + 1) * TB / RPL;
      const int ub2 = (J1 + 1) * TB * TB / RPL;
      const char* base = blk + (size_t)tt * PAIR * sizeof(C);
      const unsigned dst = wr_lds + (unsigned)(sl * SZ) * 16u;
      int u[LD];
      if (!HI) {
        const int L2 = Hu2 < 64 ? Hu2 : 64;
        u[0] = ca * Hu1 + min(lane, Hu1 - 1);
        u[1] = (ca + 1) * Hu1 + min(lane, Hu1 - 1);
        u[2] = two ? ub2 + ca * Hu2 + min(lane, L2 - 1) : 0;
        u[3] = two ? ub2 + (ca + 1) * Hu2 + min(lane, L2 - 1) : 0;
      } else {
        const bool has = two && Hu2 > 64;
        u[0] = has ? ub2 + ca * Hu2 + 64 + min(lane, Hu2 - 65) : 0;
        u[1] = has ? ub2 + (ca + 1) * Hu2 + 64 + min(lane, Hu2 - 65) : 0;
      }
#pragma unroll
      for (int i = 0; i < LD; ++i) dma16_untracked(base + (size_t)u[i] * 16, dst + (unsigned)(i * SEG * 16));
    };
#pragma unroll
    for (int g = 0; g < NW; ++g) tile_dma(blk_addr(0), g, g);

    hz_v4f dE1 = z4, dO1 = z4, dE2 = z4, dO2 = z4;  // direct: even / odd rows, block columns J1 / J2
    T va[4][2], vr[4][2];                          // transposed A operands of the step
    int rslot = 0;
    for (int st = 0; st < nsteps; ++st) {
      const char* b0 = blk_addr(st);
      const char* b1 = blk_addr(st + 1);
      const int jb = blk_of(st), jn = next_of(st);
#pragma unroll
      for (int t = 0; t < TPS; ++t) {
        // this tile's transfers are complete once at most the NW - 1 younger
        // tiles' transfers are outstanding
        __builtin_amdgcn_s_waitcnt(dma::wait_vm((NW - 1) * LD));
        asm volatile("" ::: "memory");
        const U* SL = WR + rslot * SZ;
        if (t == 0) {
          dE1 = z4;
          dO1 = z4;
          dE2 = z4;
          dO2 = z4;
          // rows 2 (rb + 16 k + b) + e of the transposed products, component i
#pragma unroll
          for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
              const int row = 2 * (rb + 16 * k + bq) + e;
              const int ix = (row < n ? row : 0) * 4 + jq;
              va[k][e] = row < n ? vqf[ix] : 0.f;
              vr[k][e] = row < n ? vqrf[ix] : 0.f;
            }
          // the lines of this step's update (waves 0-1 plane A, 2-3 plane B)
          // and its coefficients (wave 4: coupling, 1/dy of the next block)
          if (jn >= 0) {
            if (!HI) {
              const int s = wv >> 1;
              const char* lb = reinterpret_cast<const char*>(xs(s) + (size_t)jn * n);
              dma16_untracked(lb + (size_t)min((wv & 1) * 64 + lane, PU - 1) * 16,
                              pbuf_lds + (unsigned)(wv * 64 * 16));
            } else if (wv == 4) {
              const double2* src;
              if (lane == 1) src = dyi + jn;
              else src = (PH == 0 ? (half == 0) : (half != 0)) ? cmy + jn : cpy + jn;
              dma16_untracked(src, pbuf_lds + 256 * 16);
            }
          }
        }
        {
          const int J1 = t, J2 = T_ - 1 - t;
          const int Hu1 = (J1 + 1) * TB / RPL, Hu2 = (J2 + 1) * TB / RPL;
          const int nu1 = J1 * TB / RPL, nu2 = J2 * TB / RPL;
          const bool two = J2 != J1;
          if (!HI || (HAS_HI && two && Hu2 > 64)) {
            constexpr int KT = HI ? KT64 : 4;
            // ---- the tile's operands
            // direct: lane = row unit rb + lane; transposed: lane (b, j) <->
            // (row unit rb + 16 k + b, column c_j)
            U a1[2], a2[2], uu[4];
            T bv1[2], br1[2], bv2[2], br2[2];
            const bool lo = jq < 2;
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              if (!HI) {
                a1[cc] = SL[lane < Hu1 ? cc * SEG + lane : ZU];
                a2[cc] = SL[two && lane < Hu2 ? (2 + cc) * SEG + lane : ZU];
                const int c1 = (J1 * TB + ca + cc) * 4 + jq;
                bv1[cc] = vqf[c1];
                br1[cc] = vqrf[c1];
              } else {
                a2[cc] = SL[lane < Hu2 - 64 ? cc * SEG + lane : ZU];
              }
              if (two) {
                const int c2 = (J2 * TB + ca + cc) * 4 + jq;
                bv2[cc] = vqf[c2];
                br2[cc] = vqrf[c2];
              }
            }
#pragma unroll
            for (int k = 0; k < KT; ++k) {
              const int rk = rb + 16 * k + bq;
              int ix;
              if (!HI) ix = lo ? (rk < nu1 ? jq * SEG + rk : ZU) : (two && rk < nu2 ? jq * SEG + rk : ZU);
              else ix = !lo && rk < nu2 ? (jq - 2) * SEG + rk - 64 : ZU;
              uu[k] = SL[ix];
            }
            // ---- direct products
            if (!HI) {
#pragma unroll
              for (int cc = 0; cc < 2; ++cc) {
                dE1 = __builtin_amdgcn_mfma_f32_4x4x1f32(a1[cc].x, bv1[cc], dE1, 0, 0, 0);
                dO1 = __builtin_amdgcn_mfma_f32_4x4x1f32(a1[cc].z, bv1[cc], dO1, 0, 0, 0);
                dE1 = __builtin_amdgcn_mfma_f32_4x4x1f32(a1[cc].y, br1[cc], dE1, 0, 0, 0);
                dO1 = __builtin_amdgcn_mfma_f32_4x4x1f32(a1[cc].w, br1[cc], dO1, 0, 0, 0);
              }
            }
            if (two) {
#pragma unroll
              for (int cc = 0; cc < 2; ++cc) {
                dE2 = __builtin_amdgcn_mfma_f32_4x4x1f32(a2[cc].x, bv2[cc], dE2, 0, 0, 0);
                dO2 = __builtin_amdgcn_mfma_f32_4x4x1f32(a2[cc].z, bv2[cc], dO2, 0, 0, 0);
                dE2 = __builtin_amdgcn_mfma_f32_4x4x1f32(a2[cc].y, br2[cc], dE2, 0, 0, 0);
                dO2 = __builtin_amdgcn_mfma_f32_4x4x1f32(a2[cc].w, br2[cc], dO2, 0, 0, 0);
              }
            }
            if (KT > 0) {
              // ---- transposed products
              hz_v4f tA = z4, tB = z4;
#pragma unroll
              for (int k = 0; k < KT; ++k) {
                tA = __builtin_amdgcn_mfma_f32_4x4x1f32(va[k][0], uu[k].x, tA, 0, 0, 0);
                tB = __builtin_amdgcn_mfma_f32_4x4x1f32(vr[k][0], uu[k].y, tB, 0, 0, 0);
                tA = __builtin_amdgcn_mfma_f32_4x4x1f32(va[k][1], uu[k].z, tA, 0, 0, 0);
                tB = __builtin_amdgcn_mfma_f32_4x4x1f32(vr[k][1], uu[k].w, tB, 0, 0, 0);
              }
              // sum over the 16 blocks (lane bits 2..5): afterwards lane l holds
              // component (l >> 2) & 3 of column c_(l & 3)
              const T w0 = pair_step<0x124>(tA[0] + tB[0], tA[1] + tB[1], lane & 4);
              const T w1 = pair_step<0x124>(tA[2] + tB[2], tA[3] + tB[3], lane & 4);
              T sm = pair_step<0x128>(w0, w1, lane & 8);
              sm += xor16_f(sm, lane);
              sm += xor32_f(sm, lane);
              if (lane < 16 && (jq < 2 || two)) {
                const int col = (jq < 2 ? J1 : J2) * TB + ca + (jq & 1);
                tdf[(q2 * n + col) * 4 + ((lane >> 2) & 3)] = sm;
              }
            }
          }
        }
        {
          // the tile NW stages ahead into this slot once its reads are done
          __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
          asm volatile("" ::: "memory");
          const int ta = t + NW;
          const int tt = ta % TPS;
          const char* blk = (ta / TPS) == 0 ? b0 : b1;
          const bool past = (ta / TPS) != 0 && st + 1 >= nsteps;
          tile_dma(past ? reinterpret_cast<const char*>(Pk) : blk, tt, rslot);
        }
        rslot = rslot + 1 == NW ? 0 : rslot + 1;
        if (t == TPS - 1) {
          {
            // direct partials: lane 4b+j, reg i = rows 2(rb+4b+i), +1, component j
            const hz_v4f dE = dE1 + dE2, dO = dO1 + dO2;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int row = 2 * (rb + 4 * bq + i);
              if (row < n) {
                pf[(cg * n + row) * 4 + jq] = dE[i];
                pf[(cg * n + row + 1) * 4 + jq] = dO[i];
              }
            }
          }
          // the step's line transfers precede its TPS tiles' transfers
          __builtin_amdgcn_s_waitcnt(dma::wait_vm(TPS * LD));
          __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
          dma::raw_barrier();
          // one thread per (plane, row): the plane's partial sums, the
          // update of the next right-hand side and the stores
          if (tid < 2 * n) {
            const int s = tid >= n ? 1 : 0, r = tid - s * n;
            const C* part2 = reinterpret_cast<const C*>(part);
            const C* tdot2 = reinterpret_cast<const C*>(tdot);
            C ys = make_float2(0.f, 0.f);
#pragma unroll
            for (int q = 0; q < NCG; ++q) {
              const C u = part2[(q * n + r) * 2 + s];
              ys.x += u.x;
              ys.y += u.y;
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
              const C u = tdot2[(q * n + r) * 2 + s];
              ys.x += u.x;
              ys.y += u.y;
            }
            const C a = sdxi[r];
            const double2 cf = pscal[0], b = pscal[1];
            const double2 di =
                make_double2((double)a.x * b.x - (double)a.y * b.y, (double)a.x * b.y + (double)a.y * b.x);
            const bool wr = s == 0 || twoB;
            C* x = xs(s);
            if (wr && PH == 1 && (half == 0 || st > 0)) x[(size_t)jb * n + r] = ys;  // x_jb
            if (wr && jn == -2) scr_at(half, s)[r] = ys;                             // y_top / y_bot
            if (jn >= 0) {
              const C pre = reinterpret_cast<const C*>(pbuf + 128 * s)[r];
              C nv;
              nv.x = pre.x - (T)(cf.x * ys.x - cf.y * ys.y);
              nv.y = pre.y - (T)(cf.x * ys.y + cf.y * ys.x);
              if (wr && PH == 0) x[(size_t)jn * n + r] = nv;  // t_jn / b_jn
              const T w0 = (T)((double)nv.x * di.x - (double)nv.y * di.y);
              const T w1 = (T)((double)nv.x * di.y + (double)nv.y * di.x);
              reinterpret_cast<C*>(vq)[r * 2 + s] = make_float2(w0, w1);
              reinterpret_cast<C*>(vqr)[r * 2 + s] = make_float2(-w1, w0);
            }
          }
          __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
          dma::raw_barrier();
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(dma::wait_vm(0));
  };
  if (wv < 4) run(std::false_type{});
  else run(std::true_type{});
}

// ---- complex128 form of the step-unrolled packed solve
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ double pair_step_d(double a, double b, bool hi) {
  const double keep = hi ? b : a, send = hi ? a : b;
  return keep + dpp_d<CTRL>(send);
}
// wave sums of NV (8 or 16) values per lane: afterwards lane l holds the sum
// over all 64 lanes of x[l & (NV - 1)] (transpose-reduce over each 16-lane
// row, then the four rows by lane exchanges)
template <int NV>
__device__ __forceinline__ double wave_sum_values_d(const double* x, int lane) {
  static_assert(NV == 8 || NV == 16, "values per lane");
  const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4, b3 = lane & 8;
  double w[NV / 2];
#pragma unroll
  for (int i = 0; i < NV / 2; ++i) w[i] = pair_step_d<0xB1>(x[2 * i], x[2 * i + 1], b0);  // xor 1
  double y[NV / 4];
#pragma unroll
  for (int i = 0; i < NV / 4; ++i) y[i] = pair_step_d<0x4E>(w[2 * i], w[2 * i + 1], b1);  // xor 2
  double u;
  if constexpr (NV == 16) {
    const double z0 = pair_step_d<0x124>(y[0], y[1], b2);  // row_ror:4
    const double z1 = pair_step_d<0x124>(y[2], y[3], b2);
    u = pair_step_d<0x128>(z0, z1, b3);                   // row_ror:8
  } else {
    u = pair_step_d<0x124>(y[0], y[1], b2);
    u += dpp_d<0x128>(u);
  }
  u += __shfl_xor(u, 16);
  u += __shfl_xor(u, 32);
  return u;
}

// Complex128 packed plane solve (forward), same structure as k_solve_sym2: a
// 16-byte unit is one complex128 row entry; two column groups of 4 columns
// per block column; lane (cg, ru) takes row ru = rb + lane; the transposed
// values (2 block columns x 4 columns x re/im) are summed over the wave and
// stored as one partial per row wave (4 per column).
template <int NX>
__global__ __launch_bounds__(dma::NT) void k_solve_sym2d(const double2* __restrict__ P, double2* __restrict__ X,
                                                         const double2* __restrict__ cmy,
                                                         const double2* __restrict__ cpy,
                                                         const double2* __restrict__ symd, int ny, int64_t n2,
                                                         int64_t N,
    const int* __restrict__ fmap) {
  typedef double T;
  typedef double2 C;
  typedef double2 U;
  constexpr int NCG = 2, TB = 8;  // one complex128 row per 16-byte unit
  typedef SymLayout<NX, TB> L;
  constexpr int n = NX, nru = n, CPB = TB / NCG, RUP = dma::NT / NCG, NQ = RUP / 64;
  constexpr int TPS = L::TPS, NPAIR = L::NPAIR, PAIR = L::PAIR, BS = L::BS, T_ = L::T;
  constexpr int R = (PAIR + dma::NT - 1) / dma::NT;  // DMA rounds per tile (PAIR units)
  constexpr int ZU = R * dma::NT, SU = ZU + 1;
  constexpr int PU = n;                              // units of one y-line
  constexpr int LW = (PU + 63) / 64;                 // waves fetching the next line
  constexpr int OTHER = NCG * n * 16 + NQ * n * 16 + 2 * n * 16 + (LW + 1) * 64 * 16;
  constexpr int NS_LDS = (160 * 1024 - OTHER) / (SU * 16);
  constexpr int NS = NS_LDS < TPS + 1 ? NS_LDS : TPS + 1;
  static_assert(NS >= 3 && (NS - 1) * R <= 63 && nru <= RUP && LW < 8, "ring and thread split");
  static_assert(CPB == 4, "transposed values: 2 block columns x 4 columns x re/im");
  __shared__ U ring[NS * SU];
  __shared__ C part[NCG * n];
  __shared__ __align__(16) T tdot[NQ * n * 2];
  __shared__ C v[n];
  __shared__ U pbuf[(LW + 1) * 64];
  __shared__ C sdxi[n];
  const C* pline = pbuf;
  const double2* pscal = &pbuf[LW * 64];

  const int k = blockIdx.x;
  const int kf = fmap[k];  // factor slot; -1: a skipped plane (zeroed after the solve)
  if (kf < 0) return;
  const int64_t rhs = blockIdx.y;
  C* x = X + rhs * N + (int64_t)k * n2;
  const C* Pk = P + (size_t)kf * ny * BS;
  const double2* dxi = symd + n + ny;
  const double2* dyi = symd + 2 * n + ny;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int cg = wv % NCG, rb = (wv / NCG) * 64, ru = rb + lane;
  const bool worker = ru < nru;
  const int nsteps = 2 * ny - 1;
  auto blk_of = [&](int st) { return st <= ny - 1 ? st : 2 * ny - 2 - st; };
  const unsigned ring_lds = (unsigned)(size_t)(&ring[0]);
  const unsigned pbuf_lds = (unsigned)(size_t)(&pbuf[0]);
  auto cmulc = [](C a, C b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };

  for (int i = tid; i < NQ * n * 2; i += dma::NT) tdot[i] = 0;
  if (tid < NS) ring[tid * SU + ZU] = make_double2(0.0, 0.0);
  for (int i = tid; i < n; i += dma::NT) sdxi[i] = dxi[i];
  __syncthreads();
  for (int i = tid; i < n; i += dma::NT) v[i] = cmulc(x[i], cmulc(sdxi[i], dyi[blk_of(0)]));
  __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);

  auto tile_dma = [&](const char* blk, int tt, int sl) {
    const int tu = tt < NPAIR ? PAIR : (NPAIR + 1) * TB * TB;
    const char* base = blk + (size_t)tt * PAIR * sizeof(C);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int u = min(r * dma::NT + tid, tu - 1);
      dma16_untracked(base + (size_t)u * 16, ring_lds + (unsigned)((sl * SU + r * dma::NT + wv * 64) * 16));
    }
  };
  auto blk_addr = [&](int st) {
    return reinterpret_cast<const char*>(Pk + (size_t)blk_of(st < nsteps ? st : 0) * BS);
  };
#pragma unroll
  for (int g = 0; g < NS - 1; ++g) tile_dma(blk_addr(0), g, g);

  C acc = make_double2(0.0, 0.0);
  C vr = make_double2(0.0, 0.0);
  int slot = 0;
  for (int st = 0; st < nsteps; ++st) {
    const char* b0 = blk_addr(st);
    const char* b1 = blk_addr(st + 1);
#pragma unroll
    for (int t = 0; t < TPS; ++t) {
      __builtin_amdgcn_s_waitcnt(dma::wait_vm((NS - 2) * R));
      __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
      dma::raw_barrier();
      const U* SL = ring + slot * SU;
      if (t == 0) {
        acc = make_double2(0.0, 0.0);
        vr = v[worker ? ru : 0];
        if (!worker) vr = make_double2(0.0, 0.0);
        int line = 0;
        if (st < ny - 1) line = st + 1;
        else if (2 * ny - 2 - st > 0) line = 2 * ny - 3 - st;
        if (wv < LW) {
          const char* lb = reinterpret_cast<const char*>(x + (size_t)n * line);
          dma16_untracked(lb + (size_t)min(tid, PU - 1) * 16, pbuf_lds + (unsigned)(wv * 64 * 16));
        } else if (wv == LW) {
          const int sn = min(st + 1, nsteps - 1);
          const int jb = min(max(2 * ny - 3 - st, 0), ny - 1);
          const double2* src = lane == 1 ? cpy + jb : (lane == 2 ? dyi + blk_of(sn) : cmy + min(st + 1, ny - 1));
          dma16_untracked(src, pbuf_lds + (unsigned)(LW * 64 * 16));
        }
      }
      {
        const int J1 = t, J2 = T_ - 1 - t;
        const int Hu1 = (J1 + 1) * TB, Hu2 = (J2 + 1) * TB;
        const int nu1 = J1 * TB, nu2 = J2 * TB;
        const int ub2 = (J1 + 1) * TB * TB;
        const bool two = J2 != J1;
        const C zc = make_double2(0.0, 0.0);
        if (rb < Hu1) {
          const bool ok1 = worker && ru < Hu1;
          const bool ok2 = worker && two && ru < Hu2;
          const int g1 = ok1 ? cg * CPB * Hu1 + ru : ZU;
          const int g2 = ok2 ? ub2 + cg * CPB * Hu2 + ru : ZU;
          const int st1 = ok1 ? Hu1 : 0, st2 = ok2 ? Hu2 : 0;
          C a1[CPB], a2[CPB];
#pragma unroll
          for (int cc = 0; cc < CPB; ++cc) {
            a1[cc] = SL[g1 + cc * st1];
            a2[cc] = SL[g2 + cc * st2];
          }
          const C w1 = ru < nu1 ? vr : zc, w2 = ru < nu2 ? vr : zc;
          T tv[16];
#pragma unroll
          for (int cc = 0; cc < CPB; ++cc) {
            const C v1 = v[J1 * TB + cg * CPB + cc], v2 = v[J2 * TB + cg * CPB + cc];
            acc.x = fma(a1[cc].x, v1.x, acc.x);
            acc.x = fma(-a1[cc].y, v1.y, acc.x);
            acc.y = fma(a1[cc].x, v1.y, acc.y);
            acc.y = fma(a1[cc].y, v1.x, acc.y);
            acc.x = fma(a2[cc].x, v2.x, acc.x);
            acc.x = fma(-a2[cc].y, v2.y, acc.x);
            acc.y = fma(a2[cc].x, v2.y, acc.y);
            acc.y = fma(a2[cc].y, v2.x, acc.y);
            tv[2 * cc] = fma(-a1[cc].y, w1.y, a1[cc].x * w1.x);
            tv[2 * cc + 1] = fma(a1[cc].y, w1.x, a1[cc].x * w1.y);
            tv[8 + 2 * cc] = fma(-a2[cc].y, w2.y, a2[cc].x * w2.x);
            tv[8 + 2 * cc + 1] = fma(a2[cc].y, w2.x, a2[cc].x * w2.y);
          }
          const T sm = wave_sum_values_d<16>(tv, lane);
          const int kk = lane & 15;
          if (lane < 16 && (two || kk < 8)) {
            const int col = ((kk & 8) ? J2 : J1) * TB + cg * CPB + ((kk >> 1) & 3);
            tdot[((rb >> 6) * n + col) * 2 + (kk & 1)] = sm;
          }
        } else if (rb < Hu2 && two) {
          const bool ok2 = worker && ru < Hu2;
          const int g2 = ok2 ? ub2 + cg * CPB * Hu2 + ru : ZU;
          const int st2 = ok2 ? Hu2 : 0;
          const C w2 = ru < nu2 ? vr : zc;
          T tv[8];
#pragma unroll
          for (int cc = 0; cc < CPB; ++cc) {
            const C a2 = SL[g2 + cc * st2];
            const C v2 = v[J2 * TB + cg * CPB + cc];
            acc.x = fma(a2.x, v2.x, acc.x);
            acc.x = fma(-a2.y, v2.y, acc.x);
            acc.y = fma(a2.x, v2.y, acc.y);
            acc.y = fma(a2.y, v2.x, acc.y);
            tv[2 * cc] = fma(-a2.y, w2.y, a2.x * w2.x);
            tv[2 * cc + 1] = fma(a2.y, w2.x, a2.x * w2.y);
          }
          const T sm = wave_sum_values_d<8>(tv, lane);
          const int kk = lane & 7;
          if (lane < 8) {
            const int col = J2 * TB + cg * CPB + (kk >> 1);
            tdot[((rb >> 6) * n + col) * 2 + (kk & 1)] = sm;
          }
        }
      }
      {
        const int ta = t + NS - 1;
        const int tt = ta % TPS;
        const char* blk = (ta / TPS) == 0 ? b0 : b1;
        const bool past = (ta / TPS) != 0 && st + 1 >= nsteps;
        tile_dma(past ? reinterpret_cast<const char*>(Pk) : blk, past ? 0 : tt, slot == 0 ? NS - 1 : slot - 1);
      }
      if (t == TPS - 1) {
        if (worker) part[cg * n + ru] = acc;
        __builtin_amdgcn_s_waitcnt(dma::wait_vm((NS - 1) * R));
        __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
        dma::raw_barrier();
        if (tid < n) {
          C y = make_double2(0.0, 0.0);
#pragma unroll
          for (int q = 0; q < NCG; ++q) {
            const C u = part[q * n + tid];
            y.x += u.x; y.y += u.y;
          }
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            y.x += tdot[(q * n + tid) * 2];
            y.y += tdot[(q * n + tid) * 2 + 1];
          }
          const C pre = pline[tid];
          C nv;
          if (st < ny - 1) {
            const double2 cm = pscal[0];
            nv = make_double2(pre.x - (cm.x * y.x - cm.y * y.y), pre.y - (cm.x * y.y + cm.y * y.x));
            x[(size_t)n * (st + 1) + tid] = nv;
          } else {
            const int j = 2 * ny - 2 - st;
            x[(size_t)n * j + tid] = y;
            if (j > 0) {
              const double2 cp = pscal[1];
              nv = make_double2(pre.x - (cp.x * y.x - cp.y * y.y), pre.y - (cp.x * y.y + cp.y * y.x));
            } else {
              nv = y;
            }
          }
          if (st + 1 < nsteps) v[tid] = cmulc(nv, cmulc(sdxi[tid], pscal[2]));
        }
        __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
      }
      slot = slot + 1 == NS ? 0 : slot + 1;
    }
  }
  __builtin_amdgcn_s_waitcnt(dma::wait_vm(0));
}

template <int NX>
bool try_solve_sym2d(hz_problem* p, int prec, int nrhs, void* x, hipStream_t s) {
  if (p->nx != NX || prec != HZ_PREC_FP64) return false;
  hipLaunchKernelGGL((k_solve_sym2d<NX>), dim3((unsigned)p->nz, (unsigned)nrhs), dim3(dma::NT), 0, s,
                     p->fac[prec].Gsym.as<double2>(), (double2*)x, p->tab(1, 0), p->tab(1, 1),
                     p->d_symd.as<double2>(), (int)p->ny, p->n2, p->N, p->d_fmap.as<int>());
  HZ_HIP(hipGetLastError());
  return true;
}

// ---- 8 complex64 right-hand sides on the packed factors, FP32 MFMA
// Same packed tiles, untracked DMA ring and step-unrolled tile loop as
// k_solve_sym2.  V' = D_j^-1 V (n x 8 complex) sits in LDS.  Direct product of
// block column J: 16-row tiles of its stored rows times the 8 columns (two
// k-slices of 4), C = [Re Y | Im Y] from Re(P) [Re V | Im V] + Im(P) [-Im V | Re V],
// accumulated in registers over the step by the wave owning the row tile.
// Transposed product of the pair: the 16 columns of both block columns as the
// MFMA rows, their off-diagonal rows as k (rows past a column's off-diagonal
// part read the slot's zero unit), k-steps spread over the 8 waves and summed
// into yT by LDS float atomics.
template <int NX>
__global__ __launch_bounds__(dma::NT) void k_solve_sym_mfma8(const float2* __restrict__ P, float2* __restrict__ X,
                                                             const double2* __restrict__ cmy,
                                                             const double2* __restrict__ cpy,
                                                             const double2* __restrict__ symd, int ny, int64_t n2,
                                                             int64_t N,
    const int* __restrict__ fmap) {
  typedef float4 U;
  constexpr int TB = 8, NR = 8;
  typedef SymLayout<NX, TB> L;
  constexpr int n = NX;
  constexpr int TPS = L::TPS, NPAIR = L::NPAIR, PAIR = L::PAIR, BS = L::BS, T_ = L::T;
  constexpr int R = (PAIR / 2 + dma::NT - 1) / dma::NT;  // DMA rounds per tile (PAIR / 2 units)
  constexpr int ZU = R * dma::NT, SU = ZU + 1;
  constexpr int ZC = 2 * ZU;                             // complex index of the slot's zero unit
  constexpr int RT = (n + 15) / 16, TPW = (RT + 7) / 8;  // row tiles, per wave
  constexpr int NS = 6;                                  // ring slots
  static_assert(TPW <= 2 && (NS - 1) * R <= 63 && TPS >= NS - 1, "shape");
  __shared__ U ring[NS * SU];
  __shared__ float2 v[n * NR];            // V' (row-major, 8 right-hand sides)
  __shared__ __align__(16) float yT[n * NR * 2];  // transposed products of the step
  // each wave's partial transposed tile (16 columns x [Re | Im] of 8 RHS),
  // double-buffered by tile parity and folded into yT in the next stage
  __shared__ float tscr[2][8][256];
  __shared__ U pbuf[8 * TPW * 64];
  __shared__ float2 sdxi[n];

  const int k = blockIdx.x;
  const int kf = fmap[k];  // factor slot; -1: a skipped plane (zeroed after the solve)
  if (kf < 0) return;
  const int64_t rhs0 = (int64_t)blockIdx.y * NR;
  float2* xk = X + rhs0 * N + (int64_t)k * n2;
  const float2* Pk = P + (size_t)kf * ny * BS;
  const double2* dxi = symd + n + ny;
  const double2* dyi = symd + 2 * n + ny;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int nsteps = 2 * ny - 1;
  auto blk_of = [&](int st) { return st <= ny - 1 ? st : 2 * ny - 2 - st; };
  const unsigned ring_lds = (unsigned)(size_t)(&ring[0]);
  const unsigned pbuf_lds = (unsigned)(size_t)(&pbuf[0]);

  for (int i = tid; i < n * NR * 2; i += dma::NT) yT[i] = 0.f;
  if (tid < NS) ring[tid * SU + ZU] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int i = tid; i < n; i += dma::NT) {
    const double2 a = dxi[i];
    sdxi[i] = make_float2((float)a.x, (float)a.y);
  }
  __syncthreads();
  {
    const double2 b = dyi[blk_of(0)];
    for (int i = tid; i < n * NR; i += dma::NT) {
      const int q = i % NR, row = i / NR;
      const float2 x0 = xk[(int64_t)q * N + row];
      const float2 a = sdxi[row];
      const double dx = (double)a.x * b.x - (double)a.y * b.y, dy = (double)a.x * b.y + (double)a.y * b.x;
      v[i] = make_float2((float)(x0.x * dx - x0.y * dy), (float)(x0.x * dy + x0.y * dx));
    }
  }
  __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);

  auto tile_dma = [&](const char* blk, int tt, int sl) {
    const int tu = (tt < NPAIR ? PAIR : (NPAIR + 1) * TB * TB) / 2;
    const char* base = blk + (size_t)tt * PAIR * sizeof(float2);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int u = min(r * dma::NT + tid, tu - 1);
      dma16_untracked(base + (size_t)u * 16, ring_lds + (unsigned)((sl * SU + r * dma::NT + wv * 64) * 16));
    }
  };
  auto blk_addr = [&](int st) {
    return reinterpret_cast<const char*>(Pk + (size_t)blk_of(st < nsteps ? st : 0) * BS);
  };
#pragma unroll
  for (int g = 0; g < NS - 1; ++g) tile_dma(blk_addr(0), g, g);

  hz_v4f acc[TPW];
  int slot = 0;
  for (int st = 0; st < nsteps; ++st) {
    const char* b0 = blk_addr(st);
    const char* b1 = blk_addr(st + 1);
#pragma unroll
    for (int t = 0; t < TPS; ++t) {
      __builtin_amdgcn_s_waitcnt(dma::wait_vm((NS - 2) * R));
      __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
      dma::raw_barrier();
      const float2* SLc = reinterpret_cast<const float2*>(ring + slot * SU);
      // fold the previous tile's partial transposed products into yT (the last
      // tile of a step is folded by the step's epilogue)
      auto fold = [&](int tp) {
        const int Jp1 = tp, Jp2 = T_ - 1 - tp;
        if (tid < 256) {
          const int m = tid >> 4, nn = tid & 15;
          if (m < 8 || Jp2 != Jp1) {
            float sum = 0.f;
#pragma unroll
            for (int w = 0; w < 8; ++w) sum += tscr[tp & 1][w][tid];
            const int col = (m < 8 ? Jp1 : Jp2) * TB + (m & 7);
            yT[(col * NR + (nn & 7)) * 2 + (nn >> 3)] += sum;
          }
        }
      };
      if (t > 0) fold(t - 1);
      if (t == 0) {
#pragma unroll
        for (int a = 0; a < TPW; ++a) acc[a] = hz_v4f{0.f, 0.f, 0.f, 0.f};
        int line = 0;
        if (st < ny - 1) line = st + 1;
        else if (2 * ny - 2 - st > 0) line = 2 * ny - 3 - st;
        // the epilogue operand of this wave's row tiles: lane l loads unit
        // 8 rt + 2 g + (l & 1) of right-hand side q, (g, q) = ((l >> 1) / 8, (l >> 1) % 8)
#pragma unroll
        for (int a = 0; a < TPW; ++a) {
          const int rt = wv + 8 * a;
          if (rt < RT) {
            const int c = lane >> 1, g = c >> 3, q = c & 7;
            int u = 8 * rt + 2 * g + (lane & 1);
            u = u < n / 2 ? u : n / 2 - 1;
            const float2* src = xk + (int64_t)q * N + (size_t)n * line + 2 * (size_t)u;
            dma16_untracked(src, pbuf_lds + (unsigned)((wv * TPW + a) * 64 * 16));
          }
        }
      }
      {
        const int J1 = t, J2 = T_ - 1 - t;
        const bool two = J2 != J1;
        const int Hc1 = (J1 + 1) * TB, Hc2 = (J2 + 1) * TB;  // stored rows of a column
        const int ub2 = (J1 + 1) * TB * TB;                   // complex offset of block column J2
        // direct products of this wave's row tiles
#pragma unroll
        for (int a = 0; a < TPW; ++a) {
          const int rt = wv + 8 * a;
          if (rt < RT) {
#pragma unroll
            for (int part = 0; part < 2; ++part) {
              const int J = part ? J2 : J1, Hc = part ? Hc2 : Hc1, ub = part ? ub2 : 0;
              if ((part == 0 || two) && 16 * rt < Hc) {
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                  const int row = 16 * rt + li, cc = 4 * ks + lg;
                  const float2 ga = SLc[row < Hc ? ub + cc * Hc + row : ZC];
                  const float2 vv = v[(J * TB + cc) * NR + (li & 7)];
                  const float b1 = li < 8 ? vv.x : vv.y;
                  const float b2 = li < 8 ? -vv.y : vv.x;
                  acc[a] = __builtin_amdgcn_mfma_f32_16x16x4f32(ga.x, b1, acc[a], 0, 0, 0);
                  acc[a] = __builtin_amdgcn_mfma_f32_16x16x4f32(ga.y, b2, acc[a], 0, 0, 0);
                }
              }
            }
          }
        }
        // transposed products: C[m][n], m = column of the pair (0-7 J1, 8-15 J2)
        const int nk = 2 * (two ? J2 : J1);  // k-steps of 4 off-diagonal rows
        {
          hz_v4f ct = hz_v4f{0.f, 0.f, 0.f, 0.f};
          for (int kk = wv; kk < nk; kk += 8) {
            const int r = 4 * kk + lg;
            int idx = ZC;
            if (li < 8) {
              if (r < TB * J1) idx = li * Hc1 + r;
            } else if (two) {
              idx = ub2 + (li - 8) * Hc2 + r;  // r < TB J2 here
            }
            const float2 ga = SLc[idx];
            const float2 vv = v[r * NR + (li & 7)];
            const float b1 = li < 8 ? vv.x : vv.y;
            const float b2 = li < 8 ? -vv.y : vv.x;
            ct = __builtin_amdgcn_mfma_f32_16x16x4f32(ga.x, b1, ct, 0, 0, 0);
            ct = __builtin_amdgcn_mfma_f32_16x16x4f32(ga.y, b2, ct, 0, 0, 0);
          }
#pragma unroll
          for (int i = 0; i < 4; ++i) tscr[t & 1][wv][(4 * lg + i) * 16 + li] = ct[i];
        }
      }
      {
        const int ta = t + NS - 1;
        const int tt = ta % TPS;
        const char* blk = (ta / TPS) == 0 ? b0 : b1;
        const bool past = (ta / TPS) != 0 && st + 1 >= nsteps;
        tile_dma(past ? reinterpret_cast<const char*>(Pk) : blk, past ? 0 : tt, slot == 0 ? NS - 1 : slot - 1);
      }
      if (t == TPS - 1) {
        __builtin_amdgcn_s_waitcnt(dma::wait_vm((NS - 1) * R));
        __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
        dma::raw_barrier();  // every wave's products of this step are done
        fold(t);
        __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
        dma::raw_barrier();
        const int jn = blk_of(min(st + 1, nsteps - 1));
        const double2 dyn = dyi[jn];
#pragma unroll
        for (int a = 0; a < TPW; ++a) {
          const int rt = wv + 8 * a;
          float im[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) im[i] = __shfl_xor(acc[a][i], 8);
          if (rt < RT && li < 8) {
            const int q = li;
            const U p0 = pbuf[(wv * TPW + a) * 64 + 2 * (8 * lg + q)];
            const U p1 = pbuf[(wv * TPW + a) * 64 + 2 * (8 * lg + q) + 1];
            const float2 pre[4] = {make_float2(p0.x, p0.y), make_float2(p0.z, p0.w), make_float2(p1.x, p1.y),
                                   make_float2(p1.z, p1.w)};
            float2 out[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int row = 16 * rt + 4 * lg + i;
              const int rr = row < n ? row : 0;
              float* yt = &yT[(rr * NR + q) * 2];
              float2 y = make_float2(acc[a][i] + yt[0], im[i] + yt[1]);
              if (row < n) {
                yt[0] = 0.f;
                yt[1] = 0.f;
              }
              float2 nv;
              if (st < ny - 1) {
                const double2 cm = cmy[st + 1];
                nv.x = pre[i].x - (float)(cm.x * y.x - cm.y * y.y);
                nv.y = pre[i].y - (float)(cm.x * y.y + cm.y * y.x);
                out[i] = nv;
              } else {
                const int j = 2 * ny - 2 - st;
                out[i] = y;
                if (j > 0) {
                  const double2 cp = cpy[j - 1];
                  nv.x = pre[i].x - (float)(cp.x * y.x - cp.y * y.y);
                  nv.y = pre[i].y - (float)(cp.x * y.y + cp.y * y.x);
                } else {
                  nv = y;
                }
              }
              if (row < n && st + 1 < nsteps) {
                const float2 sa = sdxi[row];
                const double dx = (double)sa.x * dyn.x - (double)sa.y * dyn.y;
                const double dy = (double)sa.x * dyn.y + (double)sa.y * dyn.x;
                v[row * NR + q] = make_float2((float)(nv.x * dx - nv.y * dy), (float)(nv.x * dy + nv.y * dx));
              }
            }
            const int line = st < ny - 1 ? st + 1 : 2 * ny - 2 - st;
            float2* xq = xk + (int64_t)q * N + (size_t)n * line + 16 * rt + 4 * lg;
            if (16 * rt + 4 * lg + 1 < n)
              *reinterpret_cast<U*>(xq) = make_float4(out[0].x, out[0].y, out[1].x, out[1].y);
            if (16 * rt + 4 * lg + 3 < n)
              *reinterpret_cast<U*>(xq + 2) = make_float4(out[2].x, out[2].y, out[3].x, out[3].y);
          }
        }
        __builtin_amdgcn_s_waitcnt(dma::WAIT_LGKM0);
      }
      slot = slot + 1 == NS ? 0 : slot + 1;
    }
  }
  __builtin_amdgcn_s_waitcnt(dma::wait_vm(0));
}

template <int NX>
bool try_solve_sym_mfma8(hz_problem* p, int groups, void* x, hipStream_t s) {
  if (p->nx != NX) return false;
  hipLaunchKernelGGL((k_solve_sym_mfma8<NX>), dim3((unsigned)p->nz, (unsigned)groups), dim3(dma::NT), 0, s,
                     p->fac[HZ_PREC_C64].Gsym.as<float2>(), (float2*)x, p->tab(1, 0), p->tab(1, 1),
                     p->d_symd.as<double2>(), (int)p->ny, p->n2, p->N, p->d_fmap.as<int>());
  HZ_HIP(hipGetLastError());
  return true;
}

bool solve_sym_mfma8_any(hz_problem* p, int groups, void* x, hipStream_t s) {
  return try_solve_sym_mfma8<200>(p, groups, x, s) || try_solve_sym_mfma8<192>(p, groups, x, s) ||
         try_solve_sym_mfma8<160>(p, groups, x, s) || try_solve_sym_mfma8<128>(p, groups, x, s);
}

template <typename T, int RPL, int NCG, int TB, int NX>
bool try_pack_sym(hz_problem* p, int prec, hipStream_t s, bool alloc_only) {
  if (p->nx != NX) return false;
  typedef typename Cx<T>::t C;
  typedef SymLayout<NX, TB> L;
  Factors& f = p->fac[prec];
  const size_t bytes = (size_t)p->nslot * p->ny * L::BS * sizeof(C);
  if (f.Gsym.bytes != bytes) {
    f.Gsym.release();
    f.Gsym.alloc(bytes);
  }
  const double2* dx = p->d_symd.as<double2>();
  const double2* dy = dx + NX;
  if (p->nslot == 0 || alloc_only) return true;
  hipLaunchKernelGGL((k_pack_sym<C, NX, TB>), dim3(L::TPS, (unsigned)(p->nslot * p->ny)), dim3(256), 0, s, f.G.as<C>(),
                     f.Gsym.as<C>(), dx, dy, (int)p->ny, (const C*)nullptr, 0);
  HZ_HIP(hipGetLastError());
  return true;
}

// the twisted factors (G for j < m, Gb from m on) packed like Gsym into Gtw
template <int NX>
bool try_pack_twist(hz_problem* p, hipStream_t s) {
  if (p->nx != NX) return false;
  typedef SymLayout<NX, 8> L;
  Factors& f = p->fac[HZ_PREC_C64];
  const size_t bytes = (size_t)p->nslot * p->ny * L::BS * sizeof(float2);
  if (f.Gtw.bytes != bytes) {
    f.Gtw.release();
    f.Gtw.alloc(bytes);
  }
  const double2* dx = p->d_symd.as<double2>();
  hipLaunchKernelGGL((k_pack_sym<float2, NX, 8>), dim3(L::TPS, (unsigned)(p->nslot * p->ny)), dim3(256), 0, s,
                     f.G.as<float2>(), f.Gtw.as<float2>(), dx, dx + NX, (int)p->ny, f.Gb.as<float2>(), f.tw_m);
  HZ_HIP(hipGetLastError());
  return true;
}

bool pack_twist_any(hz_problem* p, hipStream_t s) {
  return try_pack_twist<200>(p, s) || try_pack_twist<192>(p, s) || try_pack_twist<160>(p, s) ||
         try_pack_twist<128>(p, s);
}

// both phases of the twisted solve, TW_RHS right-hand sides per launch pair
template <int NX>
bool try_solve_twist(hz_problem* p, int nrhs, void* x, hipStream_t s) {
  if (p->nx != NX) return false;
  const Factors& f = p->fac[HZ_PREC_C64];
  const float2* P = f.Gtw.as<float2>();
  const int2* sp = p->d_slotp.as<int2>();
  float2* scr = f.tw_scr.as<float2>();
  for (int r0 = 0; r0 < nrhs; r0 += TW_RHS) {
    const int nr = std::min(TW_RHS, nrhs - r0);
    float2* xr = reinterpret_cast<float2*>(x) + (size_t)r0 * p->N;
    const dim3 grid((unsigned)(2 * p->nslot), (unsigned)nr);
    hipLaunchKernelGGL((k_solve_twist_pw<NX, 0>), grid, dim3(dma::NT), 0, s, P, xr, p->tab(1, 0), p->tab(1, 1),
                       p->d_symd.as<double2>(), (int)p->ny, f.tw_m, (int)p->nz, p->n2, p->N, sp, scr);
    hipLaunchKernelGGL((k_solve_twist_pw<NX, 1>), grid, dim3(dma::NT), 0, s, P, xr, p->tab(1, 0), p->tab(1, 1),
                       p->d_symd.as<double2>(), (int)p->ny, f.tw_m, (int)p->nz, p->n2, p->N, sp, scr);
    HZ_HIP(hipGetLastError());
  }
  return true;
}

bool solve_twist_any(hz_problem* p, int nrhs, void* x, hipStream_t s) {
  return try_solve_twist<200>(p, nrhs, x, s) || try_solve_twist<192>(p, nrhs, x, s) ||
         try_solve_twist<160>(p, nrhs, x, s) || try_solve_twist<128>(p, nrhs, x, s);
}

template <int NX>
bool try_solve_sym2(hz_problem* p, int prec, int nrhs, void* x, bool adjoint, hipStream_t s) {
  if (p->nx != NX || prec != HZ_PREC_C64) return false;
  const dim3 grid((unsigned)p->nz, (unsigned)nrhs);
  const float2* P = p->fac[prec].Gsym.as<float2>();
  if (adjoint)
    hipLaunchKernelGGL((k_solve_sym2<NX, true>), grid, dim3(dma::NT), 0, s, P, (float2*)x, p->tab(1, 0), p->tab(1, 1),
                       p->d_symd.as<double2>(), (int)p->ny, p->n2, p->N, p->d_fmap.as<int>());
  else
    hipLaunchKernelGGL((k_solve_sym2<NX, false>), grid, dim3(dma::NT), 0, s, P, (float2*)x, p->tab(1, 0),
                       p->tab(1, 1), p->d_symd.as<double2>(), (int)p->ny, p->n2, p->N, p->d_fmap.as<int>());
  HZ_HIP(hipGetLastError());
  return true;
}

#define HZ_SYM_WIDTHS(X)                 \
  X(HZ_PREC_C64, float, 2, 4, 8, 200)    \
  X(HZ_PREC_C64, float, 2, 4, 8, 192)    \
  X(HZ_PREC_C64, float, 2, 4, 8, 160)    \
  X(HZ_PREC_C64, float, 2, 4, 8, 128)

bool pack_sym_any(hz_problem* p, int prec, hipStream_t s, bool alloc_only = false) {
#define HZ_SYM_PACK(P, T, RPL, NCG, TB, NX) \
  if (prec == P && try_pack_sym<T, RPL, NCG, TB, NX>(p, prec, s, alloc_only)) return true;
  HZ_SYM_WIDTHS(HZ_SYM_PACK)
#undef HZ_SYM_PACK
  if (prec == HZ_PREC_FP64)
    return try_pack_sym<double, 1, 2, 8, 200>(p, prec, s, alloc_only) ||
           try_pack_sym<double, 1, 2, 8, 192>(p, prec, s, alloc_only) ||
           try_pack_sym<double, 1, 2, 8, 160>(p, prec, s, alloc_only) ||
           try_pack_sym<double, 1, 2, 8, 128>(p, prec, s, alloc_only);
  return false;
}

bool solve_sym_any(hz_problem* p, int prec, int nrhs, void* x, bool adjoint, hipStream_t s) {
#define HZ_SYM_SOLVE(P, T, RPL, NCG, TB, NX) \
  if (prec == P && try_solve_sym2<NX>(p, prec, nrhs, x, adjoint, s)) return true;
  HZ_SYM_WIDTHS(HZ_SYM_SOLVE)
#undef HZ_SYM_SOLVE
  if (prec == HZ_PREC_FP64 && !adjoint)
    return try_solve_sym2d<200>(p, prec, nrhs, x, s) || try_solve_sym2d<192>(p, prec, nrhs, x, s) ||
           try_solve_sym2d<160>(p, prec, nrhs, x, s) || try_solve_sym2d<128>(p, prec, nrhs, x, s);
  return false;
}

// Launch of one right-hand-side group size NR over nrhs_g * NR vectors.
template <typename T, int RPL, int NCG, int TCOL, int NX, int NR>
void launch_mr(int ny, int nz, int groups, hipStream_t s, const void* G, void* x, const double2* cmy,
               const double2* cpy, int64_t n2, int64_t N, const int* fmap) {
  typedef typename Cx<T>::t C;
  hipLaunchKernelGGL((k_solve_mr<T, RPL, NCG, TCOL, NX, NR>), dim3((unsigned)nz, (unsigned)groups), dim3(dma::NT), 0,
                     s, (const C*)G, (C*)x, cmy, cpy, ny, n2, N, fmap);
  HZ_HIP(hipGetLastError());
}

// All nrhs vectors: as many full groups of the cap (<= NRMAX) as fit, then
// the remainder in groups of 4, 2 and 1; returns false when NX has no
// instantiation.
template <typename T, int RPL, int NCG, int TCOL, int NX, int NRMAX>
bool try_launch_mr(int n, int ny, int nz, int nrhs, int cap, hipStream_t s, const void* G, void* x,
                   const double2* cmy, const double2* cpy, int64_t n2, int64_t N, const int* fmap) {
  if (n != NX) return false;
  typedef typename Cx<T>::t C;
  C* xp = (C*)x;
  int done = 0;
  auto groups = [&](int g, int count) {
    switch (g) {
      case 8: launch_mr<T, RPL, NCG, TCOL, NX, (NRMAX >= 8 ? 8 : 1)>(ny, nz, count, s, G, xp + done * N, cmy, cpy, n2, N, fmap); break;
      case 4: launch_mr<T, RPL, NCG, TCOL, NX, (NRMAX >= 4 ? 4 : 1)>(ny, nz, count, s, G, xp + done * N, cmy, cpy, n2, N, fmap); break;
      case 2: launch_mr<T, RPL, NCG, TCOL, NX, 2>(ny, nz, count, s, G, xp + done * N, cmy, cpy, n2, N, fmap); break;
      default: launch_mr<T, RPL, NCG, TCOL, NX, 1>(ny, nz, count, s, G, xp + done * N, cmy, cpy, n2, N, fmap); break;
    }
    done += g * count;
  };
  int g = 1;
  while (g * 2 <= std::min(cap, NRMAX)) g *= 2;  // largest power of two <= cap
  if (nrhs / g > 0) groups(g, nrhs / g);
  for (int r = g / 2; r >= 1; r /= 2)
    if (nrhs - done >= r) groups(r, 1);
  HZ_CHECK(done == nrhs, HZ_ERR_STATE, "plane solve covered %d of %d right-hand sides", done, nrhs);
  return true;
}

// Plane-solve kernel family: HZ_SOLVE_KERNEL=mr|dma|fast|mfma|mfma_sym|sym|twist forces one for
// timing; otherwise the tuned (or default) choice applies.
enum SolveKernel { SK_DEFAULT, SK_MR, SK_DMA, SK_FAST, SK_MFMA, SK_SYM, SK_MFMA_SYM, SK_TWIST };
SolveKernel solve_kernel_choice() {
  const char* v = getenv("HZ_SOLVE_KERNEL");
  if (!v || !v[0]) return SK_DEFAULT;
  if (!strcmp(v, "mr")) return SK_MR;
  if (!strcmp(v, "dma")) return SK_DMA;
  if (!strcmp(v, "fast")) return SK_FAST;
  if (!strcmp(v, "mfma")) return SK_MFMA;
  if (!strcmp(v, "mfma_sym")) return SK_MFMA_SYM;
  if (!strcmp(v, "sym")) return SK_SYM;
  if (!strcmp(v, "twist")) return SK_TWIST;
  return SK_DEFAULT;
}

// The multi-RHS kernel for the grid widths it is instantiated for.
#define HZ_MR_WIDTHS(X)                                                                      \
  X(HZ_PREC_C64, float, 2, 4, 20, 200, 8)                                                     \
  X(HZ_PREC_C64, float, 2, 4, 16, 192, 8)                                                     \
  X(HZ_PREC_C64, float, 2, 4, 20, 160, 8)                                                     \
  X(HZ_PREC_C64, float, 2, 8, 32, 128, 8)                                                     \
  X(HZ_PREC_C64, float, 2, 8, 32, 96, 8)                                                      \
  X(HZ_PREC_C64, float, 2, 16, 32, 64, 8)                                                     \
  X(HZ_PREC_FP64, double, 1, 2, 10, 200, 4)                                                   \
  X(HZ_PREC_FP64, double, 1, 2, 8, 192, 4)                                                    \
  X(HZ_PREC_FP64, double, 1, 2, 10, 160, 4)                                                   \
  X(HZ_PREC_FP64, double, 1, 4, 16, 128, 4)                                                   \
  X(HZ_PREC_FP64, double, 1, 4, 16, 96, 4)                                                    \
  X(HZ_PREC_FP64, double, 1, 8, 32, 64, 4)

bool launch_mr_any(int prec, int n, int ny, int nz, int nrhs, int cap, hipStream_t s, const void* G, void* x,
                   const double2* cmy, const double2* cpy, int64_t n2, int64_t N, const int* fmap) {
  if (reinterpret_cast<uintptr_t>(x) % 16 != 0) return false;
#define HZ_MR_TRY(P, T, RPL, NCG, TCOL, NX, NRMAX)                                                        \
  if (prec == P && try_launch_mr<T, RPL, NCG, TCOL, NX, NRMAX>(n, ny, nz, nrhs, cap, s, G, x, cmy, cpy, n2, N, fmap)) \
    return true;
  HZ_MR_WIDTHS(HZ_MR_TRY)
#undef HZ_MR_TRY
  return false;
}

}  // namespace

bool sym_solve_available(int prec, int nx) {
  return (prec == HZ_PREC_C64 || prec == HZ_PREC_FP64) && (nx == 200 || nx == 192 || nx == 160 || nx == 128);
}

bool twist_solve_available(const hz_problem* p, int prec) {
  return prec == HZ_PREC_C64 && p->fac[prec].tw_ready;
}

namespace {
// whether plane_solve_kernels takes the twisted solve for this call (the only
// kernel that runs on an incomplete top-down chain)
bool uses_twist(const hz_problem* p, int prec, int nrhs, bool adjoint, const Tuning& t, const void* x) {
  const Factors& f = p->fac[prec];
  if (adjoint || prec != HZ_PREC_C64 || !f.tw_ready || reinterpret_cast<uintptr_t>(x) % 16 != 0) return false;
  const SolveKernel sk = solve_kernel_choice();
  if (sk == SK_TWIST) return true;
  if (sk != SK_DEFAULT) return false;
  return t.solve_family == HZ_SOLVE_TWIST || (t.solve_family == HZ_SOLVE_DEFAULT && nrhs == 1);
}
}  // namespace

bool mfma_solve_available(int prec, int nx) {
  return prec == HZ_PREC_C64 && (nx == 200 || nx == 192 || nx == 160 || nx == 128 || nx == 96 || nx == 64);
}

bool single_ring_available(int prec, int nx) {
  (void)prec;
  return nx == 200;
}

bool stream_solve_available(int prec, int nx) {
#define HZ_MR_HAS(P, T, RPL, NCG, TCOL, NX, NRMAX) \
  if (prec == P && nx == NX) return true;
  HZ_MR_WIDTHS(HZ_MR_HAS)
#undef HZ_MR_HAS
  return false;
}

namespace {

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
  const size_t tc = sizeof(T) == 4 ? TCM : TC;  // trailing-update chunk (matrix cores for complex64)
  return sizeof(C) * ((size_t)nb * n + tc * nb + nb + n) + sizeof(int) * (n + 4);
}

}  // namespace

void factor_begin(hz_problem* p, int prec, hipStream_t s) {
  const int n = (int)p->nx, ny = (int)p->ny;
  Factors& f = p->fac[prec];
  f.ready = false;
  f.sym_ready = false;
  f.tw_ready = false;
  f.hodlr.release();
  // twisted factors (complex64, the packed widths, at most two planes a slot):
  // the bottom-up chain runs beside the top-down one in the same launch
  const bool twist = prec == HZ_PREC_C64 && ny >= 2 && p->slot_pairs && sym_solve_available(prec, n) &&
                     !getenv("HZ_NO_TWIST") && !getenv("HZ_NO_SYM_PACK");
  f.tw_m = twist ? ny / 2 : 0;
  f.top_full = !twist || getenv("HZ_FACTOR_FULL") != nullptr;
  if (twist) {
    const size_t gb = (size_t)p->nslot * (ny - f.tw_m) * (size_t)n * n * sizeof(float2);
    if (f.Gb.bytes != gb) {
      f.Gb.release();
      f.Gb.alloc(gb);
    }
  } else {
    f.Gb.release();
    f.Gtw.release();
    f.tw_scr.release();
  }
  f.elem = prec == HZ_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  const size_t bytes = (size_t)p->nslot * ny * (size_t)n * n * f.elem;
  if (f.G.bytes != bytes) {
    f.G.release();
    f.G.alloc(bytes);
  }
  // status word: device int at the end of the scalar block's misc area, copied
  // to pinned host memory
  grow(p->scal, 4096 * sizeof(double2));
  int* st = reinterpret_cast<int*>(dscal(p, 63));
  p->pin_misc.reserve(64);
  *p->pin_misc.as<int>() = -1;
  launch_fill(1, make_double2(0, 0), dscal(p, 63), s);
  PlaneTabs tb = plane_tabs(p);
  int nb = 32;
  if (prec == HZ_PREC_FP64) {
    while (nb > 4 && factor_lds<double>(n, nb) > 160 * 1024) nb /= 2;
    const size_t lds = factor_lds<double>(n, nb);
    HZ_CHECK(lds <= 160 * 1024, HZ_ERR_ARG, "nx=%d too large for the LDS panel", n);
    set_max_lds((const void*)&k_factor<double>, lds);
    if (p->nslot > 0)
      hipLaunchKernelGGL(k_factor<double>, dim3((unsigned)p->nslot), dim3(FACT_THREADS), lds, s, f.G.as<double2>(),
                         tb, n, ny, nb, st, p->d_splane.as<int>(), 0, 0, (double2*)nullptr, 0, ny);
  } else {
    while (nb > 4 && factor_lds<float>(n, nb) > 160 * 1024) nb /= 2;
    const size_t lds = factor_lds<float>(n, nb);
    HZ_CHECK(lds <= 160 * 1024, HZ_ERR_ARG, "nx=%d too large for the LDS panel", n);
    set_max_lds((const void*)&k_factor<float>, lds);
    if (p->nslot > 0) {
      // with the twisted factors the top-down chain stops at tw_m - 1 (the
      // rest is completed on the first use of a one-directional kernel)
      hipLaunchKernelGGL(k_factor<float>, dim3((unsigned)(twist ? 2 * p->nslot : p->nslot)), dim3(FACT_THREADS), lds,
                         s, f.G.as<float2>(), tb, n, ny, nb, st, p->d_splane.as<int>(), twist ? 1 : 0, f.tw_m,
                         f.Gb.as<float2>(), 0, f.top_full ? ny : f.tw_m);
      if (twist)
        hipLaunchKernelGGL(k_factor<float>, dim3((unsigned)p->nslot), dim3(FACT_THREADS), lds, s, f.G.as<float2>(), tb,
                           n, ny, nb, st, p->d_splane.as<int>(), 2, f.tw_m, f.Gb.as<float2>(), 0, 0);
    }
  }
  HZ_HIP(hipGetLastError());
  HZ_HIP(hipMemcpyAsync(p->pin_misc.ptr, st, sizeof(int), hipMemcpyDeviceToHost, s));
}

void factor_end(hz_problem* p, int prec, hipStream_t s) {
  Factors& f = p->fac[prec];
  HZ_HIP(hipStreamSynchronize(s));
  HZ_CHECK(*p->pin_misc.as<int>() == 0, HZ_ERR_SINGULAR, "zero pivot while factoring a plane");
  // symmetric packed copy for the widths the packed solve is instantiated for
  f.sym_ready = false;
  if (sym_solve_available(prec, (int)p->nx) && !getenv("HZ_NO_SYM_PACK")) {
    // with an incomplete top-down chain only the storage is set up here
    const bool ok = pack_sym_any(p, prec, s, !f.top_full);
    f.sym_ready = ok && f.top_full;
    HZ_HIP(hipStreamSynchronize(s));
  }
  // twisted packed copy; the unpacked bottom-up blocks are not kept
  f.tw_ready = false;
  if (f.Gb.ptr && p->nslot > 0) {
    f.tw_ready = pack_twist_any(p, s);
    const size_t sb = (size_t)3 * TW_RHS * p->nz * p->nx * sizeof(float2);
    if (f.tw_scr.bytes != sb) {
      f.tw_scr.release();
      f.tw_scr.alloc(sb);
    }
    HZ_HIP(hipStreamSynchronize(s));
    f.Gb.release();
  }
  f.ready = true;
}

void launch_factor(hz_problem* p, int prec, hipStream_t s) {
  factor_begin(p, prec, s);
  factor_end(p, prec, s);
}

void ensure_dense(hz_problem* p, int prec, hipStream_t s) {
  Factors& f = p->fac[prec];
  if (!f.ready || f.top_full) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HZ_HIP(hipStreamIsCapturing(s, &cs));
  HZ_CHECK(cs == hipStreamCaptureStatusNone, HZ_ERR_STATE,
           "the one-directional plane factors are completed on their first use, which cannot run under stream "
           "capture: run one solve of this kind before capturing");
  const int n = (int)p->nx, ny = (int)p->ny;
  int* st = reinterpret_cast<int*>(dscal(p, 63));
  *p->pin_misc.as<int>() = -1;
  launch_fill(1, make_double2(0, 0), dscal(p, 63), s);
  int nb = 32;
  while (nb > 4 && factor_lds<float>(n, nb) > 160 * 1024) nb /= 2;
  const size_t lds = factor_lds<float>(n, nb);
  if (p->nslot > 0)
    hipLaunchKernelGGL(k_factor<float>, dim3((unsigned)p->nslot), dim3(FACT_THREADS), lds, s, f.G.as<float2>(),
                       plane_tabs(p), n, ny, nb, st, p->d_splane.as<int>(), 0, f.tw_m, (float2*)nullptr, f.tw_m, ny);
  HZ_HIP(hipGetLastError());
  HZ_HIP(hipMemcpyAsync(p->pin_misc.ptr, st, sizeof(int), hipMemcpyDeviceToHost, s));
  HZ_HIP(hipStreamSynchronize(s));
  HZ_CHECK(*p->pin_misc.as<int>() == 0, HZ_ERR_SINGULAR, "zero pivot while factoring a plane");
  f.top_full = true;
  if (sym_solve_available(prec, n) && !getenv("HZ_NO_SYM_PACK")) {
    f.sym_ready = pack_sym_any(p, prec, s);
    HZ_HIP(hipStreamSynchronize(s));
  }
}

void launch_plane_solve(hz_problem* p, int prec, int nrhs, void* x, bool adjoint, hipStream_t s) {
  launch_plane_solve_with(p, prec, nrhs, x, adjoint, p->tuning(prec, nrhs), s);
}

namespace {
__global__ void k_zero_skipped(unsigned char* __restrict__ x, const int* __restrict__ fmap, int64_t n2, int64_t N,
                               int elem) {
  const int k = blockIdx.x;
  if (fmap[k] >= 0) return;
  const int64_t q = blockIdx.y;
  // elem is 8 or 16 bytes: clear in 8-byte words
  unsigned long long* w = reinterpret_cast<unsigned long long*>(x + ((size_t)q * N + (size_t)k * n2) * elem);
  const int64_t nw = n2 * (elem / 8);
  for (int64_t i = threadIdx.x; i < nw; i += blockDim.x) w[i] = 0ull;
}

// the tuned kernel family for every plane with a factor slot
void plane_solve_kernels(hz_problem* p, int prec, int nrhs, void* x, bool adjoint, const Tuning& t, hipStream_t s);
}  // namespace

void launch_zero_skipped(hz_problem* p, int nrhs, void* x, size_t elem, hipStream_t s) {
  if (p->nskip == 0) return;
  HZ_CHECK(elem == 8 || elem == 16, HZ_ERR_ARG, "element of %zu bytes", elem);
  hipLaunchKernelGGL(k_zero_skipped, dim3((unsigned)p->nz, (unsigned)nrhs), dim3(256), 0, s, (unsigned char*)x,
                     p->d_fmap.as<int>(), p->n2, p->N, (int)elem);
  HZ_HIP(hipGetLastError());
}

void launch_plane_solve_with(hz_problem* p, int prec, int nrhs, void* x, bool adjoint, const Tuning& t,
                             hipStream_t s) {
  HZ_CHECK(p->fac[prec].ready, HZ_ERR_STATE, "preconditioner of precision %d not factored", prec);
  if (p->nslot > 0) plane_solve_kernels(p, prec, nrhs, x, adjoint, t, s);
  launch_zero_skipped(p, nrhs, x, prec == HZ_PREC_FP64 ? sizeof(double2) : sizeof(float2), s);
}

namespace {
void plane_solve_kernels(hz_problem* p, int prec, int nrhs, void* x, bool adjoint, const Tuning& t, hipStream_t s) {
  Factors& f = p->fac[prec];
  HZ_CHECK(f.ready, HZ_ERR_STATE, "preconditioner of precision %d not factored", prec);
  if (f.hodlr.active) {
    // compressed factors: forward sweeps on the HODLR records; adjoint sweeps
    // on the dense factors when they were kept
    if (!adjoint) {
      launch_hodlr_solve(p, prec, nrhs, x, s);
      return;
    }
    HZ_CHECK(f.G.ptr != nullptr, HZ_ERR_STATE,
             "adjoint plane solve: the dense factors were released by HZ_HODLR_FREE_DENSE");
  }
  if (!f.top_full && !uses_twist(p, prec, nrhs, adjoint, t, x)) ensure_dense(p, prec, s);
  const int n = (int)p->nx, ny = (int)p->ny;
  const int ng = std::max(1, APPLY_THREADS / n);
  PlaneTabs tb = plane_tabs(p);
  const double2* cmy = p->tab(1, 0);
  const double2* cpy = p->tab(1, 1);
  dim3 grid((unsigned)p->nz, (unsigned)nrhs);
  SolveKernel sk = solve_kernel_choice();
  // adjoint sweeps (the tuned families are forward kernels): complex64 groups
  // of 8 on the matrix cores with G_j^H = conj(D_j^-1 G_j D_j), the rest on the
  // packed symmetric factors with G_j^H = conj(D_j^-1 P_j conj(.)), unless the
  // environment asks for another kernel family
  if (adjoint && ny >= 1 && prec == HZ_PREC_C64 && reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
      (sk == SK_DEFAULT || sk == SK_SYM || sk == SK_MFMA || sk == SK_TWIST)) {
    if (sk != SK_SYM && nrhs >= 8 && mfma_solve_available(prec, n)) {
      const int g = nrhs / 8;
      launch_mfma8_any(n, ny, (int)p->nz, g, s, f.G.ptr, x, cmy, cpy, p->d_symd.as<double2>(), true, p->n2, p->N,
                       p->d_fmap.as<int>());
      nrhs -= 8 * g;
      if (nrhs == 0) return;
      x = reinterpret_cast<float2*>(x) + (size_t)8 * g * p->N;
      grid = dim3((unsigned)p->nz, (unsigned)nrhs);
    }
    if (sk != SK_MFMA && f.sym_ready && solve_sym_any(p, prec, nrhs, x, true, s)) return;
  }
  if (sk == SK_DEFAULT && t.solve_family == HZ_SOLVE_STREAM) sk = SK_MR;
  if (sk == SK_DEFAULT && t.solve_family == HZ_SOLVE_CHUNKED) sk = SK_FAST;
  if (sk == SK_DEFAULT && t.solve_family == HZ_SOLVE_STREAM_1) sk = SK_DMA;
  if (sk == SK_DEFAULT && t.solve_family == HZ_SOLVE_STREAM_MFMA) sk = SK_MFMA;
  if (sk == SK_DEFAULT && t.solve_family == HZ_SOLVE_SYM) sk = SK_SYM;
  if (sk == SK_DEFAULT && t.solve_family == HZ_SOLVE_TWIST) sk = SK_TWIST;
  // untuned default for one right-hand side: the twisted two-plane packed
  // solve, else the symmetric packed stream, where they are built
  if (sk == SK_DEFAULT && t.solve_family == HZ_SOLVE_DEFAULT && nrhs == 1 && f.tw_ready) sk = SK_TWIST;
  if (sk == SK_DEFAULT && t.solve_family == HZ_SOLVE_DEFAULT && nrhs == 1 && f.sym_ready) sk = SK_SYM;
  if (!adjoint && sk == SK_TWIST && f.tw_ready && reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
      solve_twist_any(p, nrhs, x, s))
    return;
  if (sk == SK_TWIST) sk = SK_SYM;
  if (!adjoint && ny >= 1 && sk == SK_SYM && f.sym_ready && reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
      solve_sym_any(p, prec, nrhs, x, false, s))
    return;
  if (sk == SK_SYM) sk = SK_DEFAULT;

  // untuned defaults: complex64 groups of 8 on the matrix cores, one
  // right-hand side on the single-RHS ring where they are instantiated
  if (sk == SK_DEFAULT && t.solve_family == HZ_SOLVE_DEFAULT && prec == HZ_PREC_C64 && nrhs >= 8 &&
      mfma_solve_available(prec, n))
    sk = SK_MFMA;
  if (sk == SK_DEFAULT && nrhs == 1 && single_ring_available(prec, n)) sk = SK_DMA;
  if (!adjoint && ny >= 1 && (sk == SK_MFMA || sk == SK_MFMA_SYM) && prec == HZ_PREC_C64 && nrhs >= 8 &&
      mfma_solve_available(prec, n) && reinterpret_cast<uintptr_t>(x) % 16 == 0) {
    // groups of 8 on the matrix cores, the remainder on the VALU streaming
    // kernel; HZ_SOLVE_KERNEL=mfma_sym takes the packed symmetric factors
    // instead of the full blocks (half the bytes, the same MFMA work plus the
    // transposed tiles' idle rows: slower at 8 right-hand sides)
    const int g = nrhs / 8;
    if (!(sk == SK_MFMA_SYM && f.sym_ready && solve_sym_mfma8_any(p, g, x, s)))
      launch_mfma8_any(n, ny, (int)p->nz, g, s, f.G.ptr, x, cmy, cpy, p->d_symd.as<double2>(), false, p->n2, p->N,
                       p->d_fmap.as<int>());
    const int rest = nrhs - 8 * g;
    if (rest == 0) return;
    x = reinterpret_cast<float2*>(x) + (size_t)8 * g * p->N;
    nrhs = rest;
    grid = dim3((unsigned)p->nz, (unsigned)nrhs);
    sk = SK_MR;
  }
  if (!adjoint && ny >= 1 && (sk == SK_DEFAULT || sk == SK_MR || sk == SK_MFMA) &&
      launch_mr_any(prec, n, ny, (int)p->nz, nrhs, (sk == SK_MFMA || sk == SK_MFMA_SYM) ? 4 : t.group_cap, s, f.G.ptr, x, cmy, cpy, p->n2,
                    p->N, p->d_fmap.as<int>()))
    return;
  if (!adjoint && ny >= 1 && sk == SK_DMA && reinterpret_cast<uintptr_t>(x) % 16 == 0) {
    // single-RHS LDS-DMA ring for the configurations it is instantiated for
    const void* Gp = f.G.ptr;
    if (prec == HZ_PREC_C64 &&
        (try_launch_dma<float, 2, 5, 20, 200>(n, ny, grid, s, Gp, x, cmy, cpy, p->n2, p->N, p->d_fmap.as<int>())))
      return;
    if (prec == HZ_PREC_FP64 &&
        (try_launch_dma<double, 1, 2, 10, 200>(n, ny, grid, s, Gp, x, cmy, cpy, p->n2, p->N, p->d_fmap.as<int>())))
      return;
  }
  if (!adjoint && ny >= 1) {
    // generic chunked kernel for every other split
    constexpr int CH = 8;
    if (prec == HZ_PREC_C64 && n % 2 == 0) {
      const int nru = n / 2, ncg = std::max(1, 1024 / nru), cpg = (n + ncg - 1) / ncg;
      const size_t lds = sizeof(float2) * (size_t)n * (1 + ncg);
      set_max_lds((const void*)&k_solve_fast<float, 2, CH>, lds);
      hipLaunchKernelGGL((k_solve_fast<float, 2, CH>), grid, dim3(1024), lds, s, f.G.as<float2>(), (float2*)x, cmy, cpy,
                         n, ny, p->n2, p->N, nru, ncg, cpg, p->d_fmap.as<int>());
      HZ_HIP(hipGetLastError());
      return;
    }
    if (prec == HZ_PREC_FP64) {
      const int nru = n, ncg = std::max(1, 1024 / nru), cpg = (n + ncg - 1) / ncg;
      const size_t lds = sizeof(double2) * (size_t)n * (1 + ncg);
      set_max_lds((const void*)&k_solve_fast<double, 1, CH>, lds);
      hipLaunchKernelGGL((k_solve_fast<double, 1, CH>), grid, dim3(1024), lds, s, f.G.as<double2>(), (double2*)x, cmy,
                         cpy, n, ny, p->n2, p->N, nru, ncg, cpg, p->d_fmap.as<int>());
      HZ_HIP(hipGetLastError());
      return;
    }
  }
  if (prec == HZ_PREC_FP64) {
    const size_t lds = sizeof(double2) * (size_t)n * (2 + ng);
    HZ_CHECK(lds <= 160 * 1024, HZ_ERR_ARG, "nx too large for the apply kernel");
    if (adjoint) {
      set_max_lds((const void*)&k_apply<double, true>, lds);
      hipLaunchKernelGGL((k_apply<double, true>), grid, dim3(APPLY_THREADS), lds, s, f.G.as<double2>(), (double2*)x, tb,
                         cmy, cpy, n, ny, p->n2, p->N, p->d_fmap.as<int>());
    } else {
      set_max_lds((const void*)&k_apply<double, false>, lds);
      hipLaunchKernelGGL((k_apply<double, false>), grid, dim3(APPLY_THREADS), lds, s, f.G.as<double2>(), (double2*)x,
                         tb, cmy, cpy, n, ny, p->n2, p->N, p->d_fmap.as<int>());
    }
  } else {
    const size_t lds = sizeof(float2) * (size_t)n * (2 + ng);
    if (adjoint) {
      set_max_lds((const void*)&k_apply<float, true>, lds);
      hipLaunchKernelGGL((k_apply<float, true>), grid, dim3(APPLY_THREADS), lds, s, f.G.as<float2>(), (float2*)x, tb,
                         cmy, cpy, n, ny, p->n2, p->N, p->d_fmap.as<int>());
    } else {
      set_max_lds((const void*)&k_apply<float, false>, lds);
      hipLaunchKernelGGL((k_apply<float, false>), grid, dim3(APPLY_THREADS), lds, s, f.G.as<float2>(), (float2*)x, tb,
                         cmy, cpy, n, ny, p->n2, p->N, p->d_fmap.as<int>());
    }
  }
  HZ_HIP(hipGetLastError());
}
}  // namespace

}  // namespace hz
