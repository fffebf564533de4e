// k_stencil.hip -- FP64 stencil operators.
//
// k_stencil7: the 7-point stretched-coordinate PML operator with mass term
//   (A u)(i,j,k) = sum_axes [cm u(-1) + cp u(+1) + cdg u] + omega^2/c^2 u
// (beta3D, functions.cpp:2242-2303; Dirichlet closure: neighbours outside the
// grid are dropped).  The adjoint uses the tables acm(i) = conj(cp(i-1)),
// acp(i) = conj(cm(i+1)) and the conjugated diagonal.  Each thread owns one
// (x,y) column of a z-chunk and streams it upward, keeping u(k-1), u(k), u(k+1)
// in registers; x/y neighbours come through L1/L2.  A tile is a run of 256
// consecutive (x,y) columns (full waves for any nx), with blockIdx remapped so
// that each XCD walks a contiguous range of tiles.
//
// k_plane5: the per-kz 5-point operator M_k on z-transformed data
//   (M_k x)(i,j) = (cdg_x + cdg_y + kappa(i,j) - lambda_k) x + x/y couplings,
// the mkl_zcsrgemv of Multiply3DSparseUsingFT, functions.cpp:6284-6297.
#include "hz_internal.h"

namespace hz {
namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}

// Nontemporal store of the operator's output: in FGMRES the Gram-Schmidt dot
// pass that reads it next runs 459 -> 416 us at 200^3 (the store itself costs
// nothing extra).
__device__ __forceinline__ void store_nt(double2* p, double2 v) {
  typedef double dv2 __attribute__((ext_vector_type(2)));
  dv2 t;
  t.x = v.x;
  t.y = v.y;
  __builtin_nontemporal_store(t, reinterpret_cast<dv2*>(p));
}

struct AxisTabs {
  const double2* m;
  const double2* p;
  const double2* d;
};

constexpr int ST = 256, KZ = 16;  // threads per tile (a run of ST consecutive (x,y) columns), z chunk

template <bool ADJ, bool RESID>
__global__ __launch_bounds__(ST) void k_stencil7(const double2* __restrict__ u, double2* __restrict__ y,
                                                 const double2* __restrict__ f, const double* __restrict__ mass,
                                                 AxisTabs tx, AxisTabs ty, AxisTabs tz, int nx, int ny, int nz,
                                                 int64_t N, int gxy, int gz) {
  // blockIdx remap: consecutive linear ids go to different XCDs (round robin
  // over 8); give each XCD a contiguous range of tiles so that the y-neighbour
  // columns of a tile (nx columns away, in the neighbouring tiles) stay in the
  // same L2.
  const int nblk = gxy * gz;
  const int lin = blockIdx.x;
  const int per = (nblk + 7) / 8;
  const int xcd = lin % 8, slot = lin / 8;
  const int tile = xcd * per + slot;
  if (tile >= nblk) return;
  const int bxy = tile % gxy;
  const int bz_all = tile / gxy;
  const int zch = (nz + KZ - 1) / KZ;
  const int rhs = bz_all / zch, bz = bz_all % zch;

  const int64_t n2 = (int64_t)nx * ny;
  const int64_t ij = (int64_t)bxy * ST + threadIdx.x;  // one (x,y) column per thread, x fastest
  if (ij >= n2) return;
  const int j = (int)(ij / nx), i = (int)(ij - (int64_t)j * nx);
  const int k0 = bz * KZ;
  const int k1 = min(nz, k0 + KZ);
  u += (int64_t)rhs * N;
  y += (int64_t)rhs * N;
  if (RESID) f += (int64_t)rhs * N;

  const double2 cmx = tx.m[i], cpx = tx.p[i], cmy = ty.m[j], cpy = ty.p[j];
  double2 dxy = make_double2(tx.d[i].x + ty.d[j].x, tx.d[i].y + ty.d[j].y);
  const bool hxm = i > 0, hxp = i < nx - 1, hym = j > 0, hyp = j < ny - 1;
  int64_t l = ij + n2 * k0;
  double2 um = k0 > 0 ? u[l - n2] : make_double2(0, 0);
  double2 uc = u[l];
#pragma unroll 4  // more loads in flight per thread: 87 -> 82 us at 200^3
  for (int k = k0; k < k1; ++k, l += n2) {
    const double2 up = (k + 1 < nz) ? u[l + n2] : make_double2(0, 0);
    const double2 xm = hxm ? u[l - 1] : make_double2(0, 0);
    const double2 xp = hxp ? u[l + 1] : make_double2(0, 0);
    const double2 ym = hym ? u[l - nx] : make_double2(0, 0);
    const double2 yp = hyp ? u[l + nx] : make_double2(0, 0);
    double2 dg = make_double2(dxy.x + tz.d[k].x, dxy.y + tz.d[k].y);
    if (ADJ) dg.y = -dg.y;
    dg.x += mass[l];
    double2 s = cmul(dg, uc);
    s = cfma(cmx, xm, s);
    s = cfma(cpx, xp, s);
    s = cfma(cmy, ym, s);
    s = cfma(cpy, yp, s);
    s = cfma(tz.m[k], um, s);  // tz.m[0] multiplies a zero
    s = cfma(tz.p[k], up, s);
    if (RESID) {
      const double2 ff = f[l];
      s = make_double2(ff.x - s.x, ff.y - s.y);
    }
    store_nt(y + l, s);
    um = uc;
    uc = up;
  }
}

template <bool ADJ>
__global__ __launch_bounds__(256) void k_plane5(const double2* __restrict__ x, double2* __restrict__ y,
                                                AxisTabs tx, AxisTabs ty, const double2* __restrict__ kappa,
                                                const double* __restrict__ lamz, int nx, int ny, int nz,
                                                int64_t N, int64_t total) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int64_t n2 = (int64_t)nx * ny;
  const int64_t rhs = g / N;
  const int64_t l = g - rhs * N;
  const int64_t k = l / n2;
  const int64_t ij = l - k * n2;
  const int j = (int)(ij / nx), i = (int)(ij - (int64_t)j * nx);
  const double2* xs = x + rhs * N;
  double2 dg = make_double2(tx.d[i].x + ty.d[j].x + kappa[ij].x - lamz[k], tx.d[i].y + ty.d[j].y + kappa[ij].y);
  if (ADJ) dg.y = -dg.y;
  double2 s = cmul(dg, xs[l]);
  if (j > 0) s = cfma(ty.m[j], xs[l - nx], s);
  if (i > 0) s = cfma(tx.m[i], xs[l - 1], s);
  if (i < nx - 1) s = cfma(tx.p[i], xs[l + 1], s);
  if (j < ny - 1) s = cfma(ty.p[j], xs[l + nx], s);
  y[g] = s;
}

AxisTabs tabs(hz_problem* p, int ax, bool adj) {
  AxisTabs t;
  t.m = p->tab(ax, adj ? 3 : 0);
  t.p = p->tab(ax, adj ? 4 : 1);
  t.d = p->tab(ax, 2);
  return t;
}

}  // namespace

void launch_stencil(hz_problem* p, int nrhs, const double2* u, double2* y, bool adjoint, const double2* f,
                    hipStream_t s) {
  HZ_CHECK(p->d.mode == HZ_MODE_S7, HZ_ERR_STATE, "7-point stencil needs HZ_MODE_S7");
  const int nx = (int)p->nx, ny = (int)p->ny, nz = (int)p->nz;
  const int gxy = (int)((p->n2 + ST - 1) / ST), gz = ((nz + KZ - 1) / KZ) * nrhs;
  const int nblk = gxy * gz;
  const int launch = ((nblk + 7) / 8) * 8;
  AxisTabs tx = tabs(p, 0, adjoint), ty = tabs(p, 1, adjoint), tz = tabs(p, 2, adjoint);
  const double* mass = p->d_mass.as<double>();
  dim3 blk(ST);
  if (adjoint) {
    if (f) hipLaunchKernelGGL((k_stencil7<true, true>), dim3(launch), blk, 0, s, u, y, f, mass, tx, ty, tz, nx, ny, nz, p->N, gxy, gz);
    else hipLaunchKernelGGL((k_stencil7<true, false>), dim3(launch), blk, 0, s, u, y, f, mass, tx, ty, tz, nx, ny, nz, p->N, gxy, gz);
  } else {
    if (f) hipLaunchKernelGGL((k_stencil7<false, true>), dim3(launch), blk, 0, s, u, y, f, mass, tx, ty, tz, nx, ny, nz, p->N, gxy, gz);
    else hipLaunchKernelGGL((k_stencil7<false, false>), dim3(launch), blk, 0, s, u, y, f, mass, tx, ty, tz, nx, ny, nz, p->N, gxy, gz);
  }
  HZ_HIP(hipGetLastError());
}

void launch_plane_matvec(hz_problem* p, int nrhs, const double2* x, double2* y, bool adjoint, hipStream_t s) {
  const int64_t total = p->N * nrhs;
  const int64_t blocks = (total + 255) / 256;
  AxisTabs tx = tabs(p, 0, adjoint), ty = tabs(p, 1, adjoint);
  if (adjoint)
    hipLaunchKernelGGL((k_plane5<true>), dim3((unsigned)blocks), dim3(256), 0, s, x, y, tx, ty,
                       p->d_kappa.as<double2>(), p->d_lamz.as<double>(), (int)p->nx, (int)p->ny, (int)p->nz, p->N, total);
  else
    hipLaunchKernelGGL((k_plane5<false>), dim3((unsigned)blocks), dim3(256), 0, s, x, y, tx, ty,
                       p->d_kappa.as<double2>(), p->d_lamz.as<double>(), (int)p->nx, (int)p->ny, (int)p->nz, p->N, total);
  HZ_HIP(hipGetLastError());
  // planes that the SAVE_MEM rule skips are zero in L0 too (functions.cpp:6286-6295)
  launch_zero_skipped(p, nrhs, y, sizeof(double2), s);
}

}  // namespace hz
