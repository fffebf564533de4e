// k_hodlr.hip -- HODLR-compressed plane factors and the plane solve on them.
//
// The reference's HODLR build (HODLR/BinaryTreesHODLR.cpp) stores every block
// inverse of the block-tridiagonal plane solve as a recursively compressed
// matrix: a node of order m > smallsize splits into n1 = m - ceil(m/2) leading
// and n2 = ceil(m/2) trailing rows, its lower off-diagonal block A21 is kept as
// U (n2 x p) VT (p x n1) from a truncated SVD (singular values S[j] with
// S[j] / S[0] >= eps, S folded into U; LowRankApproxStruct, :298-384), the
// diagonal blocks recurse and nodes of order <= smallsize stay dense
// (SymRecCompressStruct, :386-416).  In the symmetric form the upper block is
// A21^T (RecMultLStructWork2, :1095-1130).
//
// Here the compressed matrices are K_j = G_j D_j (complex symmetric, D_j the
// diagonal scaling of the packed solve), compressed from the dense factors:
//
// k_hodlr_compress: one workgroup per block inverse (persistent loop).  Each
//   off-diagonal block is copied to an FP64 workspace, reduced by column-pivoted
//   QR (modified Gram-Schmidt on the deflated block, stopped when the largest
//   remaining column is below 0.05 eps of the first pivot or at the input
//   precision's floor), and the small factor R is diagonalised by a one-sided
//   Jacobi SVD of R^H in LDS.  The singular values then give the rank by the
//   reference's rule; VT = Z^H and U = A Z (= U S) are computed from the
//   original block.
// k_hodlr_compact: packs the fixed-capacity staging records into one pool.
// k_solve_hodlr: the block-tridiagonal sweeps (DirSolveFastDiagStruct,
//   :2971-3018) with one workgroup per plane.  256 thread slots form a fixed
//   binary tree (a node of level l owns an aligned group of 256 >> l slots, its
//   halves the two aligned groups of 256 >> (l+1)), so that every low-rank
//   product is a sum over an aligned lane group followed by an exchange with
//   the partner group: in-wave DPP / shuffles for groups up to 32 lanes, LDS
//   for 64 and 128.  Thread slot t owns one plane row i and keeps, per block
//   inverse, its leaf row and its p_l values of every level in registers: the
//   record is stored [value][row] so that each value is one coalesced line.
//   The next block's record is loaded while the current one is applied.
#include "hz_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace hz {
namespace {

constexpr int HT = 256;     // threads per workgroup = slots of the row tree (n <= 256)
constexpr int MB = 128;     // largest side of an off-diagonal block
constexpr int RQMAX = 32;   // pivoted QR steps per block (rank cap before truncation)
constexpr int PCAP = 16;    // staging rows per level (largest stored rank)
constexpr int NSWEEP = 30;  // one-sided Jacobi sweeps at most

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) b
__device__ __forceinline__ double2 zcmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double wsum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
template <typename C>
__device__ __forceinline__ C from_z(double2 v) {
  C o;
  o.x = (decltype(o.x))v.x;
  o.y = (decltype(o.y))v.y;
  return o;
}
template <typename C>
__device__ __forceinline__ double2 to_z(C v) {
  return make_double2((double)v.x, (double)v.y);
}

// Circle-method round robin over rr (even) columns: pair s of round t.
__device__ __forceinline__ int rr_pos(int x, int t, int rr) { return x == 0 ? 0 : ((x - 1 + t) % (rr - 1)) + 1; }

// K(r, c) of block b of the packed symmetric layout (the upper block triangle of
// K = X D in 8 x 8 tiles, block columns J and T-1-J side by side, each holding
// rows [0, 8 (J+1)) column-major; below that, K(r, c) = K(c, r))
__device__ __forceinline__ size_t packed_at(int r, int c, int T) {
  if (r / 8 > c / 8) {
    const int t = r;
    r = c;
    c = t;
  }
  const int J = c / 8, H = 8 * (J + 1);
  const size_t base = J <= T - 1 - J ? (size_t)J * (T + 1) * 64 : (size_t)(T - 1 - J) * (T + 1) * 64 + (size_t)(T - J) * 64;
  return base + (size_t)(c - 8 * J) * H + r;
}

// G: dense blocks ([rec] n x n, column-major; K = G D); with packed != 0 the
// blocks are read from the packed symmetric form of K itself (n = 8 T)
template <typename CIN>
__global__ __launch_bounds__(HT) void k_hodlr_compress(const CIN* __restrict__ G, int packed, const double2* __restrict__ symd,
                                                       int n, int ny, int nrec, const int4* __restrict__ nodes, int nn,
                                                       const int2* __restrict__ leaves, int nl, int mleaf, int lev,
                                                       double eps, double floor_rel, CIN* __restrict__ stage,
                                                       unsigned char* __restrict__ ranks, int* __restrict__ over,
                                                       double2* __restrict__ wsA, double2* __restrict__ wsR) {
  __shared__ double2 M[RQMAX * MB];  // R^H, column i at i * MB
  __shared__ double2 qv[MB];
  __shared__ double cn[MB];
  __shared__ double redv[HT / 64];
  __shared__ int redi[HT / 64];
  __shared__ double sig[RQMAX];
  __shared__ int perm[RQMAX];
  __shared__ int flag[NSWEEP];
  __shared__ int lstart[HT], lsize[HT];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double2* A = wsA + (size_t)blockIdx.x * MB * MB;
  double2* R = wsR + (size_t)blockIdx.x * RQMAX * MB;
  for (int q = tid; q < nl; q += HT) {
    const int2 L = leaves[q];
    for (int r = 0; r < L.y; ++r) {
      lstart[L.x + r] = L.x;
      lsize[L.x + r] = L.y;
    }
  }
  __syncthreads();
  const double2* dx = symd;
  const double2* dy = symd + n;
  const size_t WS = (size_t)mleaf + (size_t)lev * PCAP;

  for (int rec = blockIdx.x; rec < nrec; rec += gridDim.x) {
    const int j = rec % ny;
    const int T = n / 8;
    const CIN* Gb = G + (size_t)rec * (packed ? (size_t)T * (T + 1) * 32 : (size_t)n * n);
    CIN* Sb = stage + (size_t)rec * n * WS;
    const double2 dyj = dy[j];
    auto K = [&](int r, int c) {
      if (packed) return to_z(Gb[packed_at(r, c, T)]);
      return zmul(to_z(Gb[(size_t)c * n + r]), zmul(dx[c], dyj));
    };
    // leaves: [c][i] = K(i, leaf start + c), zero past the leaf
    for (int e = tid; e < mleaf * n; e += HT) {
      const int c = e / n, i = e - c * n;
      const int a = lstart[i], m = lsize[i];
      Sb[e] = from_z<CIN>(c < m ? K(i, a + c) : make_double2(0, 0));
    }
    // level rows: zero where no node of the level covers a row (leaves of
    // smaller depth); the nodes overwrite their own rows
    for (size_t e = tid; e < (size_t)lev * PCAP * n; e += HT) Sb[(size_t)mleaf * n + e] = from_z<CIN>(make_double2(0, 0));
    for (int q = 0; q < nn; ++q) {
      const int4 nd = nodes[q];
      const int a = nd.x, m1 = nd.y, m2 = nd.z, lv = nd.w;
      // block A = K[a+m1 .., a ..] (m2 x m1), column-major, and its column norms
      for (int c = wv; c < m1; c += HT / 64) {
        double s = 0;
        for (int r = lane; r < m2; r += 64) {
          const double2 v = K(a + m1 + r, a + c);
          A[(size_t)c * m2 + r] = v;
          s += v.x * v.x + v.y * v.y;
        }
        s = wsum(s);
        if (lane == 0) cn[c] = s;
      }
      for (int i = tid; i < NSWEEP; i += HT) flag[i] = 0;
      __syncthreads();
      // ---- column-pivoted QR: R rows r_k = q_k^H A
      const int kmax = min(min(m1, m2), RQMAX);
      int rq = 0;
      double thr2 = 0.0, nrm02 = 0.0;
      const double eps2 = eps * eps;
      bool capped = false;
      for (;;) {
        double bv = tid < m1 ? cn[tid] : -1.0;
        int bi = tid < m1 ? tid : 0;
        for (int off = 32; off > 0; off >>= 1) {
          const double ov = __shfl_xor(bv, off);
          const int oi = __shfl_xor(bi, off);
          if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
          }
        }
        if (lane == 0) {
          redv[wv] = bv;
          redi[wv] = bi;
        }
        __syncthreads();
        bv = redv[0];
        bi = redi[0];
        for (int w = 1; w < HT / 64; ++w)
          if (redv[w] > bv || (redv[w] == bv && redi[w] < bi)) {
            bv = redv[w];
            bi = redi[w];
          }
        if (rq == 0) {
          nrm02 = bv;
          const double nrm0 = sqrt(bv);
          const double thr = fmax(0.05 * eps * nrm0, floor_rel * nrm0);
          thr2 = thr * thr;
        }
        if (!(bv > thr2) || bv == 0.0) break;
        if (rq == kmax) {
          // the block's rank at this eps exceeds the cap when the residual is
          // still above the eps level
          capped = rq < min(m1, m2) && bv > eps2 * nrm02;
          break;
        }
        const double inv = 1.0 / sqrt(bv);
        __syncthreads();  // everyone has read redv/redi
        for (int r = tid; r < m2; r += HT) {
          const double2 v = A[(size_t)bi * m2 + r];
          qv[r] = make_double2(v.x * inv, v.y * inv);
        }
        __syncthreads();
        for (int c = wv; c < m1; c += HT / 64) {
          double2 d = make_double2(0, 0);
          for (int r = lane; r < m2; r += 64) {
            const double2 t = zcmul(qv[r], A[(size_t)c * m2 + r]);
            d.x += t.x;
            d.y += t.y;
          }
          d.x = wsum(d.x);
          d.y = wsum(d.y);
          double s = 0;
          for (int r = lane; r < m2; r += 64) {
            const double2 qd = zmul(qv[r], d);
            double2 v = A[(size_t)c * m2 + r];
            v.x -= qd.x;
            v.y -= qd.y;
            A[(size_t)c * m2 + r] = v;
            s += v.x * v.x + v.y * v.y;
          }
          s = wsum(s);
          if (lane == 0) {
            cn[c] = s;
            R[(size_t)rq * MB + c] = d;
          }
        }
        __syncthreads();
        ++rq;
      }
      // ---- one-sided Jacobi SVD of R^H (m1 x rq): columns orthogonalised in LDS
      const int rr = rq + (rq & 1);
      for (int e = tid; e < rr * m1; e += HT) {
        const int i = e / m1, c = e - i * m1;
        const double2 v = i < rq ? R[(size_t)i * MB + c] : make_double2(0, 0);
        M[i * MB + c] = make_double2(v.x, -v.y);
      }
      __syncthreads();
      for (int sw = 0; sw < NSWEEP && rr >= 2; ++sw) {
        for (int t = 0; t < rr - 1; ++t) {
          for (int s = wv; s < rr / 2; s += HT / 64) {
            const int ci = rr_pos(s, t, rr), cj = rr_pos(rr - 1 - s, t, rr);
            double al = 0, be = 0;
            double2 ga = make_double2(0, 0);
            for (int c = lane; c < m1; c += 64) {
              const double2 u = M[ci * MB + c], w = M[cj * MB + c];
              al += u.x * u.x + u.y * u.y;
              be += w.x * w.x + w.y * w.y;
              const double2 g = zcmul(u, w);
              ga.x += g.x;
              ga.y += g.y;
            }
            al = wsum(al);
            be = wsum(be);
            ga.x = wsum(ga.x);
            ga.y = wsum(ga.y);
            const double g = sqrt(ga.x * ga.x + ga.y * ga.y);
            if (g > 0.0 && g > 1e-13 * sqrt(al) * sqrt(be)) {
              const double zeta = (be - al) / (2.0 * g);
              const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
              const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
              const double2 ph = make_double2(ga.x / g, -ga.y / g);  // exp(-i arg gamma)
              for (int c = lane; c < m1; c += 64) {
                const double2 u = M[ci * MB + c];
                const double2 w = zmul(M[cj * MB + c], ph);
                M[ci * MB + c] = make_double2(cs * u.x - sn * w.x, cs * u.y - sn * w.y);
                M[cj * MB + c] = make_double2(sn * u.x + cs * w.x, sn * u.y + cs * w.y);
              }
              if (lane == 0) flag[sw] = 1;
            }
          }
          __syncthreads();
        }
        if (flag[sw] == 0) break;
      }
      // ---- singular values, order, rank by S[t] / S[0] >= eps
      if (tid < rq) {
        double s = 0;
        for (int c = 0; c < m1; ++c) {
          const double2 u = M[tid * MB + c];
          s += u.x * u.x + u.y * u.y;
        }
        sig[tid] = sqrt(s);
      }
      __syncthreads();
      if (tid < rq) {
        int pos = 0;
        for (int i = 0; i < rq; ++i) pos += (sig[i] > sig[tid] || (sig[i] == sig[tid] && i < tid)) ? 1 : 0;
        perm[pos] = tid;
      }
      __syncthreads();
      int p = 0;
      if (rq > 0) {
        const double s0 = sig[perm[0]];
        if (s0 > 0.0)
          for (int t = 0; t < rq; ++t) {
            if (!(sig[perm[t]] / s0 >= eps)) break;
            p = t + 1;
          }
      }
      const bool too_many = p > PCAP;
      if (too_many) p = PCAP;
      if (tid == 0) {
        ranks[(size_t)rec * nn + q] = (unsigned char)p;
        if (capped || too_many) atomicAdd(over, 1);
      }
      // ---- VT rows (left half of the node's rows) and U columns (right half)
      CIN* X = Sb + (size_t)n * (mleaf + (size_t)lv * PCAP);
      for (int e = tid; e < PCAP * m1; e += HT) {
        const int t = e / m1, c = e - t * m1;
        double2 v = make_double2(0, 0);
        if (t < p) {
          const int ci = perm[t];
          const double is = 1.0 / sig[ci];
          const double2 u = M[ci * MB + c];
          v = make_double2(u.x * is, -u.y * is);
        }
        X[(size_t)t * n + a + c] = from_z<CIN>(v);
      }
      {
        // U[r][t] = sum_c A(r, c) z_t(c), z_t = column perm[t] / sigma; thread
        // (r, h) takes t = h, h + 2, ...
        constexpr int TH = PCAP / 2;
        const int r = tid % MB, h = tid / MB;
        if (r < m2) {
          double2 acc[TH];
#pragma unroll
          for (int u = 0; u < TH; ++u) acc[u] = make_double2(0, 0);
          for (int c = 0; c < m1; ++c) {
            const double2 av = K(a + m1 + r, a + c);
#pragma unroll
            for (int u = 0; u < TH; ++u) {
              const int t = h + 2 * u;
              if (t < p) {
                const double2 z = M[perm[t] * MB + c];
                const double2 pr = zmul(av, z);
                acc[u].x += pr.x;
                acc[u].y += pr.y;
              }
            }
          }
#pragma unroll
          for (int u = 0; u < TH; ++u) {
            const int t = h + 2 * u;
            double2 v = make_double2(0, 0);
            if (t < p) {
              const double is = 1.0 / sig[perm[t]];
              v = make_double2(acc[u].x * is, acc[u].y * is);
            }
            X[(size_t)t * n + a + m1 + r] = from_z<CIN>(v);
          }
        }
      }
      __syncthreads();
    }
  }
}

// Compressed record header: first complex of the record in the pool and the
// stored rank of every level.
struct __align__(16) HRec {
  long long off;
  unsigned char p[8];
};

// Record layout in the pool: values in pairs, [pair][row][2]: the leaf
// columns (ceil(mleaf/2) pairs), then per level its ceil(p_l/2) rank pairs;
// the odd value of an odd count is zero.
template <typename C>
__global__ __launch_bounds__(HT) void k_hodlr_compact(const C* __restrict__ stage, const HRec* __restrict__ recs,
                                                      C* __restrict__ pool, int n, int mleaf, int lev) {
  const int rec = blockIdx.x;
  const HRec h = recs[rec];
  const size_t WS = (size_t)mleaf + (size_t)lev * PCAP;
  const C* S = stage + (size_t)rec * n * WS;
  C* D = pool + h.off;
  const C zero = from_z<C>(make_double2(0, 0));
  const int mlp = (mleaf + 1) / 2;
  for (int e = threadIdx.x; e < 2 * mlp * n; e += HT) {
    const int c = 2 * (e / (2 * n)) + (e & 1), i = (e / 2) % n;
    D[e] = c < mleaf ? S[(size_t)c * n + i] : zero;
  }
  size_t o = (size_t)2 * mlp * n;
  for (int l = 0; l < lev; ++l) {
    const int pl = h.p[l], pp = (pl + 1) / 2;
    const C* src = S + (size_t)n * (mleaf + (size_t)l * PCAP);
    for (int e = threadIdx.x; e < 2 * pp * n; e += HT) {
      const int t = 2 * (e / (2 * n)) + (e & 1), i = (e / 2) % n;
      D[o + e] = t < pl ? src[(size_t)t * n + i] : zero;
    }
    o += (size_t)2 * pp * n;
  }
}

// ---------------------------------------------------------------- solve
template <int CTRL>
__device__ __forceinline__ float dpp_f1(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d1(double x) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
template <typename T, int CTRL>
__device__ __forceinline__ T dpp1(T x) {
  if constexpr (sizeof(T) == 4) return dpp_f1<CTRL>(x);
  else return dpp_d1<CTRL>(x);
}

// Value of lane ^ 16 (W = 16) or lane ^ 32 (W = 32) with the gfx950 row swaps
// (VALU, no LDS): permlane16_swap of a register with itself returns rows
// (0,0,2,2) and (1,1,3,3), permlane32_swap the halves (lo,lo) and (hi,hi).
template <int W>
__device__ __forceinline__ unsigned xorlane_u(unsigned u, int lane) {
  if constexpr (W == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    return (lane & 16) ? r[0] : r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    return (lane & 32) ? r[0] : r[1];
  }
}
template <int W, typename T>
__device__ __forceinline__ T xorlane(T x, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, xorlane_u<W>(__builtin_bit_cast(unsigned, x), lane));
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = xorlane_u<W>((unsigned)b, lane), hi = xorlane_u<W>((unsigned)(b >> 32), lane);
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  }
}

// Sum over the aligned group of G lanes (G a power of two, 2..64): every lane
// of the group gets it.  After the quad sums, the half-row and row mirrors
// reach the other quad and the other half row, the row swaps the other row
// and the other half of the wave.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T x, int lane) {
  if constexpr (G >= 2) x += dpp1<T, 0xB1>(x);   // quad_perm [1,0,3,2]
  if constexpr (G >= 4) x += dpp1<T, 0x4E>(x);   // quad_perm [2,3,0,1]
  if constexpr (G >= 8) x += dpp1<T, 0x141>(x);  // row_half_mirror
  if constexpr (G >= 16) x += dpp1<T, 0x140>(x); // row_mirror
  if constexpr (G >= 32) x += xorlane<16>(x, lane);
  if constexpr (G >= 64) x += xorlane<32>(x, lane);
  return x;
}
// the value held by the partner group (lanes xor G), G = 2..32, when the
// value is uniform over each group
template <int G, typename T>
__device__ __forceinline__ T group_partner(T x, int lane) {
  if constexpr (G == 2) return dpp1<T, 0x4E>(x);        // xor 2
  else if constexpr (G == 4) return dpp1<T, 0x141>(x);  // the other quad of the half row
  else if constexpr (G == 8) return dpp1<T, 0x140>(x);  // the other half of the row
  else if constexpr (G == 16) return xorlane<16>(x, lane);
  else return xorlane<32>(x, lane);
}

template <typename C>
struct CT;
template <>
struct CT<float2> {
  typedef float T;
};
template <>
struct CT<double2> {
  typedef double T;
};

template <typename C>
__device__ __forceinline__ C cfma(C a, C b, C c) {
  c.x = fma(a.x, b.x, fma(-a.y, b.y, c.x));
  c.y = fma(a.x, b.y, fma(a.y, b.x, c.y));
  return c;
}

// Two consecutive complex values of one row: the records keep values in pairs
// ([pair][row][2]) so that each lane loads 16 bytes per instruction (complex64).
template <typename C>
struct __align__(2 * sizeof(C)) CPair {
  C a, b;
};

template <typename C, int LEV, int PMAX>
struct HBuf {
  static constexpr int ML = HT >> LEV;
  C L[ML];
  C X[LEV][PMAX];
  C pre;
  int p[LEV];     // ranks of the record's levels
  double2 cf;     // coupling of the step's update: cm_y(j+1) forward, cp_y(j-1) backward
  double2 dyn;    // 1/dy of the next step's block
};

// The sweep of one plane: steps st = 0 .. 2 ny - 2 apply block jb(st) (the
// forward blocks 0 .. ny-1, then ny-2 .. 0 backward).  The record, the update
// coefficients and the line of step st + NB - 1 are loaded while step st runs.
template <typename C, int LEV, int PMAX, int NB>
__global__ __launch_bounds__(HT) void k_solve_hodlr(const C* __restrict__ pool, const HRec* __restrict__ recs,
                                                    C* __restrict__ X, const double2* __restrict__ cmy,
                                                    const double2* __restrict__ cpy,
                                                    const double2* __restrict__ symd, int n, int ny, int small,
                                                    int mleaf, int64_t n2, int64_t N, const int* __restrict__ fmap) {
  typedef typename CT<C>::T T;
  typedef HBuf<C, LEV, PMAX> B;
  typedef CPair<C> P2;
  constexpr int ML = B::ML;
  static_assert(ML % 2 == 0 && PMAX % 2 == 0, "values in pairs");
  __shared__ C v[HT];
  __shared__ C ex[2][PMAX][HT / 64];
  const int k = blockIdx.x;
  const int kf = __builtin_amdgcn_readfirstlane(fmap[k]);
  if (kf < 0) return;
  C* x = X + (int64_t)blockIdx.y * N + (int64_t)k * n2;
  const HRec* rk = recs + (size_t)kf * ny;
  const double2* dxi = symd + n + ny;
  const double2* dyi = symd + 2 * n + ny;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // this slot's leaf: walk the tree
  int la = 0, lm = n, sa = 0, S = HT;
  for (int l = 0; l < LEV; ++l) {
    if (lm <= small) break;
    const int m2 = (lm + 1) / 2, m1 = lm - m2;
    S >>= 1;
    if (tid < sa + S) {
      lm = m1;
    } else {
      la += m1;
      lm = m2;
      sa += S;
    }
  }
  const int row = tid - sa < lm ? la + (tid - sa) : -1;
  const int r0 = row >= 0 ? row : 0;
  // 1/dx of this row in a register (no per-lane load inside the sweep)
  const double2 dxr = row >= 0 ? dxi[row] : make_double2(0, 0);
  const int nsteps = 2 * ny - 1;
  const int mlp = (mleaf + 1) / 2;  // leaf pairs
  auto blk_of = [&](int st) { return st <= ny - 1 ? st : 2 * ny - 2 - st; };
  // block whose vector the update of step st forms (-1: none)
  auto next_blk = [&](int st) {
    if (st < ny - 1) return st + 1;
    const int j = blk_of(st);
    return j > 0 ? j - 1 : -1;
  };
  const C zc = from_z<C>(make_double2(0, 0));

  // Branch-free loads: every pair slot loads (past the record's sizes from a
  // clamped address in the same lines), so that the wait before a step
  // covers exactly what was issued one step earlier; the step masks the
  // slots past the record's sizes.
  auto load = [&](B& b, int st) {
    const int sc = min(st, nsteps - 1);
    const HRec h = rk[blk_of(sc)];
    const P2* base = reinterpret_cast<const P2*>(pool + h.off);
#pragma unroll
    for (int c = 0; c < ML / 2; ++c) {
      const P2 q = base[(size_t)(c < mlp ? c : 0) * n + r0];
      b.L[2 * c] = q.a;
      b.L[2 * c + 1] = q.b;
    }
    size_t o = (size_t)mlp * n;
#pragma unroll
    for (int l = 0; l < LEV; ++l) {
      const int pl = h.p[l], pp = (pl + 1) / 2;
      b.p[l] = pl;
#pragma unroll
      for (int t = 0; t < PMAX / 2; ++t) {
        const P2 q = base[o + (size_t)(t < pp ? t : 0) * n + r0];
        b.X[l][2 * t] = q.a;
        b.X[l][2 * t + 1] = q.b;
      }
      o += (size_t)pp * n;
    }
    const int jn = next_blk(sc);
    // the update reads f_{st+1} forward and tb_{j-1} backward: the line of jn
    b.pre = x[(size_t)max(jn, 0) * n + r0];
    b.cf = sc < ny - 1 ? cmy[sc + 1] : (jn >= 0 ? cpy[jn] : make_double2(0, 0));
    b.dyn = dyi[max(jn, 0)];
  };

  // rank slots in pairs (one uniform branch per pair) so that the two
  // reductions interleave
  auto level_pass = [&](const B& b, const C vr, C& y, auto lc) {
    constexpr int l = decltype(lc)::value;
    constexpr int G = HT >> (l + 1);
    const int pl = b.p[l];
#pragma unroll
    for (int t0 = 0; t0 < PMAX; t0 += 2) {
      if (t0 < pl) {
#pragma unroll
        for (int t = t0; t < t0 + 2; ++t) {
          const C xv = (row >= 0 && t < pl) ? b.X[l][t] : zc;
          C pr;
          pr.x = xv.x * vr.x - xv.y * vr.y;
          pr.y = xv.x * vr.y + xv.y * vr.x;
          if constexpr (G <= 32) {
            C s;
            s.x = group_sum<G>(pr.x, lane);
            s.y = group_sum<G>(pr.y, lane);
            C q;
            q.x = group_partner<G>(s.x, lane);
            q.y = group_partner<G>(s.y, lane);
            y = cfma(xv, q, y);
          } else {
            C s;
            s.x = group_sum<64>(pr.x, lane);
            s.y = group_sum<64>(pr.y, lane);
            if (lane == 0) ex[l][t][wv] = s;
          }
        }
      }
    }
  };
  auto level_finish = [&](const B& b, C& y, auto lc) {
    constexpr int l = decltype(lc)::value;
    const int pl = b.p[l];
#pragma unroll
    for (int t = 0; t < PMAX; ++t) {
      if (t < ((pl + 1) & ~1)) {
        const C xv = (row >= 0 && t < pl) ? b.X[l][t] : zc;
        C q;
        if constexpr (l == 1) {
          q = ex[1][t][wv ^ 1];
        } else {
          const C q0 = ex[0][t][wv ^ 2], q1 = ex[0][t][wv ^ 3];
          q.x = q0.x + q1.x;
          q.y = q0.y + q1.y;
        }
        y = cfma(xv, q, y);
      }
    }
  };

  // tb_{ny-2}, kept from the forward step that formed it: the turn step's
  // line may have been prefetched before that step stored it
  C tb_keep = zc;
  auto step = [&](int st, const B& b) {
    C y = zc;
    if (row >= 0) {
#pragma unroll
      for (int c = 0; c < ML; ++c)
        if (c < mleaf) y = cfma(b.L[c], v[min(la + c, n - 1)], y);
    }
    const C vr = row >= 0 ? v[row] : zc;
    if constexpr (LEV > 0) level_pass(b, vr, y, std::integral_constant<int, 0>{});
    if constexpr (LEV > 1) level_pass(b, vr, y, std::integral_constant<int, 1>{});
    if constexpr (LEV > 2) level_pass(b, vr, y, std::integral_constant<int, 2>{});
    if constexpr (LEV > 3) level_pass(b, vr, y, std::integral_constant<int, 3>{});
    if constexpr (LEV > 4) level_pass(b, vr, y, std::integral_constant<int, 4>{});
    if constexpr (LEV > 5) level_pass(b, vr, y, std::integral_constant<int, 5>{});
    __syncthreads();
    if constexpr (LEV >= 1) level_finish(b, y, std::integral_constant<int, 0>{});
    if constexpr (LEV >= 2) level_finish(b, y, std::integral_constant<int, 1>{});
    if (row >= 0) {
      const int i = row;
      const C pre = (st == ny - 1 && ny >= 3) ? tb_keep : b.pre;
      const int jn = next_blk(st);
      const double2 cf = b.cf;
      C nv = zc;
      if (st < ny - 1) {
        // forward: tb_{st+1} = f_{st+1} - cm_y(st+1) G_st tb_st
        nv.x = pre.x - (T)(cf.x * (double)y.x - cf.y * (double)y.y);
        nv.y = pre.y - (T)(cf.x * (double)y.y + cf.y * (double)y.x);
        x[(size_t)(st + 1) * n + i] = nv;
        if (st + 1 == ny - 2) tb_keep = nv;
      } else {
        // backward: x_j = G_j (..) stored; next vector tb_{j-1} - cp_y(j-1) x_j
        x[(size_t)blk_of(st) * n + i] = y;
        nv.x = pre.x - (T)(cf.x * (double)y.x - cf.y * (double)y.y);
        nv.y = pre.y - (T)(cf.x * (double)y.y + cf.y * (double)y.x);
      }
      if (jn >= 0) v[i] = from_z<C>(zmul(to_z(nv), zmul(dxr, b.dyn)));
    }
    __syncthreads();
  };

  // v' = D_0^{-1} f_0
  if (row >= 0) v[row] = from_z<C>(zmul(to_z(x[row]), zmul(dxr, dyi[0])));
  // named buffers, no copies between them
  B b0, b1, b2;
  if constexpr (NB == 1) {
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) {
      load(b0, st);
      step(st, b0);
    }
  } else if constexpr (NB == 2) {
    load(b0, 0);
    __syncthreads();
    for (int st = 0; st < nsteps; st += 2) {
      load(b1, st + 1);
      step(st, b0);
      if (st + 1 < nsteps) {
        load(b0, st + 2);
        step(st + 1, b1);
      }
    }
  } else {
    load(b0, 0);
    load(b1, 1);
    __syncthreads();
    for (int st = 0; st < nsteps; st += 3) {
      load(b2, st + 2);
      step(st, b0);
      if (st + 1 < nsteps) {
        load(b0, st + 3);
        step(st + 1, b1);
      }
      if (st + 2 < nsteps) {
        load(b1, st + 4);
        step(st + 2, b2);
      }
    }
  }
}

// The twisted sweep (records of the twisted factors: X_j = G_j below m, T_m,
// B_j above m), one half of one plane per workgroup.  Phase 0: the top half
// forms t_j downward (t_0 = f_0, t_j = f_j - cm_y(j) X_{j-1} t_{j-1}) and the
// bottom half b_j upward (b_j = f_j - cp_y(j) X_{j+1} b_{j+1}), leaving
// y_top = X_{m-1} t_{m-1}, y_bot = X_{m+1} b_{m+1} and a copy of f_m in scr.
// Phase 1: both halves form x_m = T_m (f_m - cm_y(m) y_top - cp_y(m) y_bot) and
// sweep outward (x_j = X_j (t_j - cp_y(j) x_{j+1}) above, x_j = X_j (b_j -
// cm_y(j) x_{j-1}) below).  Steps, tree products and prefetch as in
// k_solve_hodlr.
template <typename C, int LEV, int PMAX, int NB, int PH>
__global__ __launch_bounds__(HT) void k_solve_hodlr_tw(const C* __restrict__ pool, const HRec* __restrict__ recs,
                                                       C* __restrict__ X, const double2* __restrict__ cmy,
                                                       const double2* __restrict__ cpy,
                                                       const double2* __restrict__ symd, int n, int ny, int small,
                                                       int mleaf, int64_t n2, int64_t N, const int* __restrict__ fmap,
                                                       int m, int nz, C* __restrict__ scr) {
  typedef typename CT<C>::T T;
  typedef HBuf<C, LEV, PMAX> B;
  typedef CPair<C> P2;
  constexpr int ML = B::ML;
  static_assert(ML % 2 == 0 && PMAX % 2 == 0, "values in pairs");
  __shared__ C v[HT];
  __shared__ C ex[2][PMAX][HT / 64];
  const int k = blockIdx.x >> 1, half = blockIdx.x & 1;
  const int kf = __builtin_amdgcn_readfirstlane(fmap[k]);
  if (kf < 0) return;
  C* x = X + (int64_t)blockIdx.y * N + (int64_t)k * n2;
  // scr: [what][rhs][plane][n], what = 0 y_top, 1 y_bot, 2 f_m
  auto scr_at = [&](int what) { return scr + (((int64_t)what * gridDim.y + blockIdx.y) * nz + k) * n; };
  const HRec* rk = recs + (size_t)kf * ny;
  const double2* dxi = symd + n + ny;
  const double2* dyi = symd + 2 * n + ny;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // this slot's leaf: walk the tree
  int la = 0, lm = n, sa = 0, S = HT;
  for (int l = 0; l < LEV; ++l) {
    if (lm <= small) break;
    const int m2 = (lm + 1) / 2, m1 = lm - m2;
    S >>= 1;
    if (tid < sa + S) {
      lm = m1;
    } else {
      la += m1;
      lm = m2;
      sa += S;
    }
  }
  const int row = tid - sa < lm ? la + (tid - sa) : -1;
  const int r0 = row >= 0 ? row : 0;
  // 1/dx of this row in a register (no per-lane load inside the sweep)
  const double2 dxr = row >= 0 ? dxi[row] : make_double2(0, 0);
  const int nsteps = PH == 0 ? (half == 0 ? m : ny - 1 - m) : (half == 0 ? m + 1 : ny - m);
  const int mlp = (mleaf + 1) / 2;  // leaf pairs
  auto blk_of = [&](int st) {
    if (PH == 0) return half == 0 ? st : ny - 1 - st;
    return half == 0 ? m - st : m + st;
  };
  // block whose vector the update of step st forms (-1: none; -2: y goes to scr)
  auto next_blk = [&](int st) {
    const int jb = blk_of(st);
    if (PH == 0) {
      if (half == 0) return jb + 1 < m ? jb + 1 : -2;
      return jb - 1 > m ? jb - 1 : -2;
    }
    if (half == 0) return jb > 0 ? jb - 1 : -1;
    return jb < ny - 1 ? jb + 1 : -1;
  };
  // the update toward larger j couples through cm_y, toward smaller j through cp_y
  const bool usecm = (PH == 0) == (half == 0);
  if (PH == 0 && half == 0) {
    // f_m for phase 1 (phase 1 overwrites line m with x_m)
    for (int i = tid; i < n; i += HT) scr_at(2)[i] = x[(size_t)m * n + i];
  }
  if (nsteps <= 0) return;
  const C zc = from_z<C>(make_double2(0, 0));

  // Branch-free loads: every pair slot loads (past the record's sizes from a
  // clamped address in the same lines), so that the wait before a step
  // covers exactly what was issued one step earlier; the step masks the
  // slots past the record's sizes.
  auto load = [&](B& b, int st) {
    const int sc = min(st, nsteps - 1);
    const HRec h = rk[blk_of(sc)];
    const P2* base = reinterpret_cast<const P2*>(pool + h.off);
#pragma unroll
    for (int c = 0; c < ML / 2; ++c) {
      const P2 q = base[(size_t)(c < mlp ? c : 0) * n + r0];
      b.L[2 * c] = q.a;
      b.L[2 * c + 1] = q.b;
    }
    size_t o = (size_t)mlp * n;
#pragma unroll
    for (int l = 0; l < LEV; ++l) {
      const int pl = h.p[l], pp = (pl + 1) / 2;
      b.p[l] = pl;
#pragma unroll
      for (int t = 0; t < PMAX / 2; ++t) {
        const P2 q = base[o + (size_t)(t < pp ? t : 0) * n + r0];
        b.X[l][2 * t] = q.a;
        b.X[l][2 * t + 1] = q.b;
      }
      o += (size_t)pp * n;
    }
    const int jn = next_blk(sc);
    // the update reads the line of jn: f (phase 0) or t / b (phase 1)
    b.pre = x[(size_t)max(jn, 0) * n + r0];
    b.cf = jn >= 0 ? (usecm ? cmy[jn] : cpy[jn]) : make_double2(0, 0);
    b.dyn = dyi[max(jn, 0)];
  };

  // rank slots in pairs (one uniform branch per pair) so that the two
  // reductions interleave
  auto level_pass = [&](const B& b, const C vr, C& y, auto lc) {
    constexpr int l = decltype(lc)::value;
    constexpr int G = HT >> (l + 1);
    const int pl = b.p[l];
#pragma unroll
    for (int t0 = 0; t0 < PMAX; t0 += 2) {
      if (t0 < pl) {
#pragma unroll
        for (int t = t0; t < t0 + 2; ++t) {
          const C xv = (row >= 0 && t < pl) ? b.X[l][t] : zc;
          C pr;
          pr.x = xv.x * vr.x - xv.y * vr.y;
          pr.y = xv.x * vr.y + xv.y * vr.x;
          if constexpr (G <= 32) {
            C s;
            s.x = group_sum<G>(pr.x, lane);
            s.y = group_sum<G>(pr.y, lane);
            C q;
            q.x = group_partner<G>(s.x, lane);
            q.y = group_partner<G>(s.y, lane);
            y = cfma(xv, q, y);
          } else {
            C s;
            s.x = group_sum<64>(pr.x, lane);
            s.y = group_sum<64>(pr.y, lane);
            if (lane == 0) ex[l][t][wv] = s;
          }
        }
      }
    }
  };
  auto level_finish = [&](const B& b, C& y, auto lc) {
    constexpr int l = decltype(lc)::value;
    const int pl = b.p[l];
#pragma unroll
    for (int t = 0; t < PMAX; ++t) {
      if (t < ((pl + 1) & ~1)) {
        const C xv = (row >= 0 && t < pl) ? b.X[l][t] : zc;
        C q;
        if constexpr (l == 1) {
          q = ex[1][t][wv ^ 1];
        } else {
          const C q0 = ex[0][t][wv ^ 2], q1 = ex[0][t][wv ^ 3];
          q.x = q0.x + q1.x;
          q.y = q0.y + q1.y;
        }
        y = cfma(xv, q, y);
      }
    }
  };

  auto step = [&](int st, const B& b) {
    C y = zc;
    if (row >= 0) {
#pragma unroll
      for (int c = 0; c < ML; ++c)
        if (c < mleaf) y = cfma(b.L[c], v[min(la + c, n - 1)], y);
    }
    const C vr = row >= 0 ? v[row] : zc;
    if constexpr (LEV > 0) level_pass(b, vr, y, std::integral_constant<int, 0>{});
    if constexpr (LEV > 1) level_pass(b, vr, y, std::integral_constant<int, 1>{});
    if constexpr (LEV > 2) level_pass(b, vr, y, std::integral_constant<int, 2>{});
    if constexpr (LEV > 3) level_pass(b, vr, y, std::integral_constant<int, 3>{});
    if constexpr (LEV > 4) level_pass(b, vr, y, std::integral_constant<int, 4>{});
    if constexpr (LEV > 5) level_pass(b, vr, y, std::integral_constant<int, 5>{});
    __syncthreads();
    if constexpr (LEV >= 1) level_finish(b, y, std::integral_constant<int, 0>{});
    if constexpr (LEV >= 2) level_finish(b, y, std::integral_constant<int, 1>{});
    if (row >= 0) {
      const int i = row;
      const int jb = blk_of(st), jn = next_blk(st);
      if (PH == 1 && (half == 0 || st > 0)) x[(size_t)jb * n + i] = y;  // x_jb
      if (jn == -2) scr_at(half)[i] = y;                               // y_top / y_bot
      if (jn >= 0) {
        const double2 cf = b.cf;
        C nv;
        nv.x = b.pre.x - (T)(cf.x * (double)y.x - cf.y * (double)y.y);
        nv.y = b.pre.y - (T)(cf.x * (double)y.y + cf.y * (double)y.x);
        if (PH == 0) x[(size_t)jn * n + i] = nv;  // t_jn / b_jn
        v[i] = from_z<C>(zmul(to_z(nv), zmul(dxr, b.dyn)));
      }
    }
    __syncthreads();
  };

  // v' of the first block: D^{-1} f (phase 0), D_m^{-1} (f_m - cm_y(m) y_top -
  // cp_y(m) y_bot) (phase 1)
  if (row >= 0) {
    const int j0 = blk_of(0);
    double2 f;
    if (PH == 0) {
      f = to_z(x[(size_t)j0 * n + row]);
    } else {
      f = to_z(scr_at(2)[row]);
      if (m > 0) {
        const double2 yt = zmul(cmy[m], to_z(scr_at(0)[row]));
        f.x -= yt.x;
        f.y -= yt.y;
      }
      if (m < ny - 1) {
        const double2 yb = zmul(cpy[m], to_z(scr_at(1)[row]));
        f.x -= yb.x;
        f.y -= yb.y;
      }
    }
    v[row] = from_z<C>(zmul(f, zmul(dxr, dyi[j0])));
  }
  // named buffers, no copies between them
  B b0, b1, b2;
  if constexpr (NB == 1) {
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) {
      load(b0, st);
      step(st, b0);
    }
  } else if constexpr (NB == 2) {
    load(b0, 0);
    __syncthreads();
    for (int st = 0; st < nsteps; st += 2) {
      load(b1, st + 1);
      step(st, b0);
      if (st + 1 < nsteps) {
        load(b0, st + 2);
        step(st + 1, b1);
      }
    }
  } else {
    load(b0, 0);
    load(b1, 1);
    __syncthreads();
    for (int st = 0; st < nsteps; st += 3) {
      load(b2, st + 2);
      step(st, b0);
      if (st + 1 < nsteps) {
        load(b0, st + 3);
        step(st + 1, b1);
      }
      if (st + 2 < nsteps) {
        load(b1, st + 4);
        step(st + 2, b2);
      }
    }
  }
}

// ---------------------------------------------------------------- host
struct Tree {
  int lev = 0, mleaf = 0;
  std::vector<int4> nodes;   // a, n1, n2, level (breadth first)
  std::vector<int2> leaves;  // a, m
};

Tree make_tree(int n, int small) {
  Tree t;
  struct Nd {
    int a, m, l;
  };
  std::vector<Nd> todo{{0, n, 0}};
  while (!todo.empty()) {
    std::vector<Nd> nxt;
    for (const Nd& d : todo) {
      if (d.m <= small) {
        t.leaves.push_back(make_int2(d.a, d.m));
        t.mleaf = std::max(t.mleaf, d.m);
        t.lev = std::max(t.lev, d.l);
        continue;
      }
      const int n2 = (d.m + 1) / 2, n1 = d.m - n2;
      t.nodes.push_back(make_int4(d.a, n1, n2, d.l));
      nxt.push_back({d.a, n1, d.l + 1});
      nxt.push_back({d.a + n1, n2, d.l + 1});
    }
    todo.swap(nxt);
  }
  return t;
}

template <typename C, int LEV, int PMAX>
void launch_solve_lev(hz_problem* p, const Hodlr& h, int nrhs, void* x, hipStream_t s) {
  dim3 grid((unsigned)p->nz, (unsigned)nrhs);
  // registers: three records in flight for complex64 up to rank 4, two above
  constexpr int NB = sizeof(C) == 8 ? (PMAX <= 4 ? 3 : 2) : 1;
  hipLaunchKernelGGL((k_solve_hodlr<C, LEV, PMAX, NB>), grid, dim3(HT), 0, s, h.pool.as<C>(), h.recs.as<HRec>(),
                     (C*)x, p->tab(1, 0), p->tab(1, 1), p->d_symd.as<double2>(), (int)p->nx, (int)p->ny, h.small,
                     h.mleaf, p->n2, p->N, p->d_fmap.as<int>());
}

// the twisted sweep: phase 0 then phase 1, TW_RHS right-hand sides per pair
template <typename C, int LEV, int PMAX>
void launch_solve_lev_tw(hz_problem* p, const Hodlr& h, int nrhs, void* x, hipStream_t s) {
  constexpr int NB = sizeof(C) == 8 ? (PMAX <= 4 ? 3 : 2) : 1;
  const Factors& f = p->fac[HZ_PREC_C64];
  for (int r0 = 0; r0 < nrhs; r0 += TW_RHS) {
    const int nr = std::min(TW_RHS, nrhs - r0);
    C* xr = reinterpret_cast<C*>(x) + (size_t)r0 * p->N;
    const dim3 grid((unsigned)(2 * p->nz), (unsigned)nr);
    hipLaunchKernelGGL((k_solve_hodlr_tw<C, LEV, PMAX, NB, 0>), grid, dim3(HT), 0, s, h.pool.as<C>(),
                       h.recs.as<HRec>(), xr, p->tab(1, 0), p->tab(1, 1), p->d_symd.as<double2>(), (int)p->nx,
                       (int)p->ny, h.small, h.mleaf, p->n2, p->N, p->d_fmap.as<int>(), f.tw_m, (int)p->nz,
                       f.tw_scr.as<C>());
    hipLaunchKernelGGL((k_solve_hodlr_tw<C, LEV, PMAX, NB, 1>), grid, dim3(HT), 0, s, h.pool.as<C>(),
                       h.recs.as<HRec>(), xr, p->tab(1, 0), p->tab(1, 1), p->d_symd.as<double2>(), (int)p->nx,
                       (int)p->ny, h.small, h.mleaf, p->n2, p->N, p->d_fmap.as<int>(), f.tw_m, (int)p->nz,
                       f.tw_scr.as<C>());
  }
}

template <typename C, int PMAX>
void launch_solve_pmax(hz_problem* p, const Hodlr& h, int nrhs, void* x, hipStream_t s) {
  if constexpr (sizeof(C) == 8) {
    if (h.twisted) {
      switch (h.lev) {
        case 3: launch_solve_lev_tw<C, 3, PMAX>(p, h, nrhs, x, s); return;
        case 4: launch_solve_lev_tw<C, 4, PMAX>(p, h, nrhs, x, s); return;
        case 5: launch_solve_lev_tw<C, 5, PMAX>(p, h, nrhs, x, s); return;
        case 6: launch_solve_lev_tw<C, 6, PMAX>(p, h, nrhs, x, s); return;
        default: HZ_CHECK(false, HZ_ERR_ARG, "HODLR tree depth %d outside 3..6", h.lev);
      }
    }
  }
  switch (h.lev) {
    case 3: launch_solve_lev<C, 3, PMAX>(p, h, nrhs, x, s); break;
    case 4: launch_solve_lev<C, 4, PMAX>(p, h, nrhs, x, s); break;
    case 5: launch_solve_lev<C, 5, PMAX>(p, h, nrhs, x, s); break;
    case 6: launch_solve_lev<C, 6, PMAX>(p, h, nrhs, x, s); break;
    default: HZ_CHECK(false, HZ_ERR_ARG, "HODLR tree depth %d outside 3..6", h.lev);
  }
}

template <typename C>
void launch_solve_any(hz_problem* p, const Hodlr& h, int nrhs, void* x, hipStream_t s) {
  if (h.pmax <= 4) launch_solve_pmax<C, 4>(p, h, nrhs, x, s);
  else if (h.pmax <= 6) launch_solve_pmax<C, 6>(p, h, nrhs, x, s);
  else if (h.pmax <= 8) launch_solve_pmax<C, 8>(p, h, nrhs, x, s);
  else launch_solve_pmax<C, 16>(p, h, nrhs, x, s);
}

}  // namespace

void Hodlr::release() {
  pool.release();
  recs.release();
  ranks.clear();
  active = false;
}

void hodlr_compress(hz_problem* p, int prec, double eps, int small, int flags) {
  Factors& f = p->fac[prec];
  if (eps < 0.0) {
    f.hodlr.release();
    return;
  }
  const bool twisted = (flags & HZ_HODLR_TWISTED) != 0;
  if (twisted) {
    HZ_CHECK(prec == HZ_PREC_C64 && f.ready && f.tw_ready && f.Gtw.ptr, HZ_ERR_STATE,
             "twisted compression needs the twisted complex64 factors (widths 128, 160, 192, 200, ny >= 2)");
  } else {
    HZ_CHECK(f.ready && f.G.ptr, HZ_ERR_STATE, "compression needs the dense factors of precision %d", prec);
    ensure_dense(p, prec, p->stream);
  }
  HZ_CHECK(eps < 1.0, HZ_ERR_ARG, "eps=%g: the truncation rule needs 0 <= eps < 1", eps);
  const int n = (int)p->nx, ny = (int)p->ny;
  HZ_CHECK(n >= 2 && n <= HT, HZ_ERR_ARG, "HODLR plane solve: nx=%d outside 2..%d", n, HT);
  HZ_CHECK(small >= 1, HZ_ERR_ARG, "smallsize=%d must be positive", small);
  Tree t = make_tree(n, small);
  HZ_CHECK(!t.nodes.empty(), HZ_ERR_ARG, "smallsize=%d >= nx=%d: nothing to compress", small, n);
  HZ_CHECK(t.lev >= 3 && t.lev <= 6, HZ_ERR_ARG,
           "smallsize=%d gives a tree of depth %d at nx=%d (3..6 supported: leaves of nx/8 .. nx/64 rows)", small,
           t.lev, n);
  HZ_CHECK(t.mleaf <= (HT >> t.lev), HZ_ERR_ARG,
           "smallsize=%d: a leaf of %d rows above a depth-%d tree's %d slots (uneven leaf depths); use "
           "smallsize = ceil(nx / 2^L)",
           small, t.mleaf, t.lev, HT >> t.lev);
  f.hodlr.release();
  hipStream_t s = p->stream;
  const int nrec = p->nslot * ny, nn = (int)t.nodes.size();
  const size_t elem = f.elem;
  const size_t WS = (size_t)t.mleaf + (size_t)t.lev * PCAP;
  if (nrec == 0) {
    f.hodlr.active = true;
    f.hodlr.eps = eps;
    f.hodlr.small = small;
    f.hodlr.lev = t.lev;
    f.hodlr.mleaf = t.mleaf;
    f.hodlr.nn = nn;
    f.hodlr.pmax = 0;
    return;
  }
  DevBuf stage, dranks, dover, dnodes, dleaves, wsA, wsR;
  stage.alloc((size_t)nrec * n * WS * elem);
  dranks.alloc((size_t)nrec * nn);
  dover.alloc(sizeof(double));
  dnodes.alloc(t.nodes.size() * sizeof(int4));
  dleaves.alloc(t.leaves.size() * sizeof(int2));
  HZ_HIP(hipMemcpyAsync(dnodes.ptr, t.nodes.data(), t.nodes.size() * sizeof(int4), hipMemcpyHostToDevice, s));
  HZ_HIP(hipMemcpyAsync(dleaves.ptr, t.leaves.data(), t.leaves.size() * sizeof(int2), hipMemcpyHostToDevice, s));
  launch_fill_real(1, 0.0, reinterpret_cast<double*>(dover.ptr), s);  // zeroes the int counter
  const int gw = std::min(nrec, 1024);
  wsA.alloc((size_t)gw * MB * MB * sizeof(double2));
  wsR.alloc((size_t)gw * RQMAX * MB * sizeof(double2));
  // noise floor of the input factors relative to the largest column (the
  // pivoted QR stops there when 0.05 eps lies below it)
  const double floor_rel = prec == HZ_PREC_FP64 ? 1e-15 : 1e-7;
  if (prec == HZ_PREC_FP64)
    hipLaunchKernelGGL(k_hodlr_compress<double2>, dim3(gw), dim3(HT), 0, s, f.G.as<double2>(), 0,
                       p->d_symd.as<double2>(), n, ny, nrec, dnodes.as<int4>(), nn, dleaves.as<int2>(),
                       (int)t.leaves.size(), t.mleaf, t.lev, eps, floor_rel, stage.as<double2>(),
                       dranks.as<unsigned char>(), dover.as<int>(), wsA.as<double2>(), wsR.as<double2>());
  else
    hipLaunchKernelGGL(k_hodlr_compress<float2>, dim3(gw), dim3(HT), 0, s,
                       twisted ? f.Gtw.as<float2>() : f.G.as<float2>(), twisted ? 1 : 0,
                       p->d_symd.as<double2>(), n, ny, nrec, dnodes.as<int4>(), nn, dleaves.as<int2>(),
                       (int)t.leaves.size(), t.mleaf, t.lev, eps, floor_rel, stage.as<float2>(),
                       dranks.as<unsigned char>(), dover.as<int>(), wsA.as<double2>(), wsR.as<double2>());
  HZ_HIP(hipGetLastError());
  std::vector<unsigned char> rk((size_t)nrec * nn);
  int over = 0;
  HZ_HIP(hipMemcpyAsync(rk.data(), dranks.ptr, rk.size(), hipMemcpyDeviceToHost, s));
  HZ_HIP(hipMemcpyAsync(&over, dover.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  HZ_HIP(hipStreamSynchronize(s));
  wsA.release();
  wsR.release();
  HZ_CHECK(over == 0, HZ_ERR_ARG,
           "HODLR compression at eps=%g: %d blocks need more than rank %d; raise eps", eps, over, PCAP);
  // records: per level the largest rank of its nodes
  std::vector<HRec> recs(nrec);
  long long off = 0;
  int pmax = 0;
  double rsum = 0;
  for (int r = 0; r < nrec; ++r) {
    HRec& h = recs[r];
    memset(&h, 0, sizeof(h));
    h.off = off;
    for (int q = 0; q < nn; ++q) {
      const int l = t.nodes[q].w, pq = rk[(size_t)r * nn + q];
      h.p[l] = (unsigned char)std::max<int>(h.p[l], pq);
      pmax = std::max(pmax, pq);
      rsum += pq;
    }
    long long w = (t.mleaf + 1) / 2;
    for (int l = 0; l < t.lev; ++l) w += (h.p[l] + 1) / 2;
    off += 2 * w * n;  // pairs of complex values: records stay 16-byte aligned
  }
  Hodlr& H = f.hodlr;
  H.pool.alloc((size_t)off * elem);
  H.recs.alloc(recs.size() * sizeof(HRec));
  HZ_HIP(hipMemcpyAsync(H.recs.ptr, recs.data(), recs.size() * sizeof(HRec), hipMemcpyHostToDevice, s));
  if (prec == HZ_PREC_FP64)
    hipLaunchKernelGGL(k_hodlr_compact<double2>, dim3(nrec), dim3(HT), 0, s, stage.as<double2>(), H.recs.as<HRec>(),
                       H.pool.as<double2>(), n, t.mleaf, t.lev);
  else
    hipLaunchKernelGGL(k_hodlr_compact<float2>, dim3(nrec), dim3(HT), 0, s, stage.as<float2>(), H.recs.as<HRec>(),
                       H.pool.as<float2>(), n, t.mleaf, t.lev);
  HZ_HIP(hipGetLastError());
  HZ_HIP(hipStreamSynchronize(s));
  H.active = true;
  H.twisted = twisted;
  H.eps = eps;
  H.small = small;
  H.lev = t.lev;
  H.mleaf = t.mleaf;
  H.nn = nn;
  H.pmax = pmax;
  H.ranks.swap(rk);
  H.mean_rank = rsum / ((double)nrec * nn);
  H.bytes = (double)off * elem;
  if (flags & HZ_HODLR_FREE_DENSE) {
    f.G.release();
    f.Gsym.release();
    f.sym_ready = false;
    f.Gtw.release();
    if (!twisted) f.tw_scr.release();  // the twisted sweep keeps its workspace
    f.tw_ready = false;
  }
}

void launch_hodlr_solve(hz_problem* p, int prec, int nrhs, void* x, hipStream_t s) {
  const Hodlr& h = p->fac[prec].hodlr;
  if (p->nslot == 0) return;
  if (prec == HZ_PREC_FP64) launch_solve_any<double2>(p, h, nrhs, x, s);
  else launch_solve_any<float2>(p, h, nrhs, x, s);
  HZ_HIP(hipGetLastError());
}

}  // namespace hz
