// hz_internal.h -- internal state and kernel entry points of the hz library.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hz.h"

namespace hz {

// ------------------------------------------------------------ errors
void set_error(const char* fmt, ...);
struct Error {
  int code;
};

#define HZ_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ::hz::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      throw ::hz::Error{HZ_ERR_HIP};                                                   \
    }                                                                                  \
  } while (0)

#define HZ_CHECK(cond, code, ...)       \
  do {                                  \
    if (!(cond)) {                      \
      ::hz::set_error(__VA_ARGS__);     \
      throw ::hz::Error{code};          \
    }                                   \
  } while (0)

// ------------------------------------------------------------ device buffers
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void alloc(size_t b);
  void release();
  ~DevBuf() { release(); }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(ptr); }
};

// Pinned host memory (hipHostMalloc), grown outside the solve loops; the owner
// synchronises its stream before freeing it.
struct PinBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t b);
  void release();
  ~PinBuf() { release(); }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(ptr); }
};

// FFT plan for one length: radix sequence and twiddle tables.
struct FftPlan {
  int n = 0;
  int nstages = 0;
  int radix[24];
  DevBuf tw64;  // double2[n]: exp(-2 pi i t / n)
  DevBuf tw32;  // float2[n]
};

// Launch choices of one (precision, right-hand-side count), set by the
// autotuner; the defaults are used until it has run.
struct Tuning {
  int solve_family = 0;  // HZ_SOLVE_DEFAULT / _STREAM / _CHUNKED
  int group_cap = 8;     // largest right-hand-side group of the streaming solve (8, 4, 2, 1)
  int fft_lines = -1;    // z-FFT: 0 register-resident kernel, 8/16/32 lines per LDS workgroup, -1 automatic
};

// HODLR-compressed form of one precision's block inverses (k_hodlr.hip):
// pool of records [leaf rows][level 0 rows]...[level L-1 rows], each row n
// complex, one record per (factor slot, y-line) with its header in recs.
struct Hodlr {
  bool active = false;
  bool twisted = false;  // records of the twisted factors (applied by the two-phase sweep)
  double eps = 0.0;
  int small = 0, lev = 0, mleaf = 0, nn = 0, pmax = 0;
  double mean_rank = 0.0, bytes = 0.0;
  DevBuf pool, recs;
  std::vector<unsigned char> ranks;  // (slot, y-line, node) ranks, host copy
  void release();
};

// Plane factors of one precision.
struct Factors {
  bool ready = false;
  DevBuf G;        // nz * ny * nx * nx complex, [k][j] blocks, column-major nx x nx
  size_t elem = 0; // bytes per complex element
  // symmetric packed form of G_j D_j (upper block triangle of 8 x 8 tiles,
  // block columns paired J / T-1-J), for the widths it is built for
  bool sym_ready = false;
  DevBuf Gsym;
  // twisted factors (complex64): blocks j < tw_m the top-down inverses, block
  // tw_m the middle inverse, blocks j > tw_m the bottom-up inverses (Gb, dense,
  // [slot][j - tw_m]), packed like Gsym into Gtw; tw_scr passes y_top, y_bot and
  // f_m between the two phases (3 x TW_RHS x nz x nx complex64)
  bool tw_ready = false;
  int tw_m = 0;
  // the top-down chain G_j is complete for every j (with the twisted factors
  // it stops at tw_m - 1 until a one-directional kernel first needs it)
  bool top_full = true;
  DevBuf Gb, Gtw, tw_scr;
  Hodlr hodlr;
};
constexpr int TW_RHS = 64;  // right-hand sides per twisted solve launch

}  // namespace hz

struct hz_problem {
  hz_desc d;
  int64_t nx, ny, nz, n2, N;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;

  // host tables
  std::vector<double2> cm[3], cp[3], cdg[3];  // stencil coefficients per axis
  std::vector<double> lamz;                    // z symbol per plane
  // Plane map: fmap[k] = factor slot of plane k (planes with equal z symbols
  // share one, k and nz - k for the symmetric symbols), -1 for a plane skipped
  // by the SAVE_MEM rule (no factor, zero preconditioner output);
  // splane[s] = the plane whose symbol slot s factors.
  std::vector<int> fmap, splane;
  int nslot = 0, nskip = 0;
  hz::DevBuf d_fmap;     // int[nz]
  hz::DevBuf d_splane;   // int[nslot]
  // the planes of each slot (int2[nslot]: the first plane, the second or -1);
  // slot_pairs: no slot holds more than two planes
  hz::DevBuf d_slotp;
  bool slot_pairs = false;
  std::vector<double2> kappa;                  // (nx*ny) (w/cb)^2 (1 + i beta)
  std::vector<double2> c0;                     // (nx*ny) background speed

  // device tables: coefficient block [3 axes][cm, cp, cdg, acm, acp] (double2)
  hz::DevBuf d_tab;      // stencil tables (forward and adjoint)
  hz::DevBuf d_lamz;     // double[nz]
  hz::DevBuf d_kappa;    // double2[n2]
  hz::DevBuf d_mass;     // double[N] = omega^2 / c^2 (S7)
  hz::DevBuf d_dL;       // double2[N] delta-L (REF)
  hz::DevBuf d_dLc;      // double2[N] conj(delta-L) (REF adjoint)
  hz::DevBuf d_sig;      // double2[ny] Schur coupling cm_y(j) cp_y(j-1)
  std::vector<double2> symd[2];  // symmetrising scalings dx (nx), dy (ny)
  hz::DevBuf d_symd;     // double2: dx[nx] | dy[ny] | 1/dx[nx] | 1/dy[ny]

  hz::FftPlan fft;
  hz::Factors fac[2];

  // work buffers (grown on demand)
  hz::DevBuf w_a, w_b, w_c;    // c128 N * nrhs each
  hz::DevBuf red;              // reduction partials
  hz::DevBuf scal;             // device scalars (double2)
  hz::PinBuf pin;              // pinned host copies of Krylov scalars (statuses, norms)
  hz::PinBuf pin_misc;         // pinned host word of the factorisation status

  // autotuned launch choices per (precision, right-hand sides)
  std::map<std::pair<int, int>, hz::Tuning> tuned;
  hz::Tuning tuning(int prec, int nrhs) const {
    auto it = tuned.find({prec, nrhs});
    if (it != tuned.end()) return it->second;
    // a tuned count with the same group structure serves larger batches too
    hz::Tuning best;
    int best_n = 0;
    for (const auto& kv : tuned)
      if (kv.first.first == prec && kv.first.second <= nrhs && kv.first.second > best_n) {
        best = kv.second;
        best_n = kv.first.second;
      }
    return best;
  }

  // Krylov storage
  hz::DevBuf kV, kZ, kx, kr, kw, kaux;
  int k_m = 0, k_nrhs = 0;

  // misfit state
  std::vector<int64_t> rec;
  hz::DevBuf d_rec;     // receiver indices (int64)
  hz::DevBuf d_data;    // nsrc * nrec observed data
  hz::DevBuf d_res;     // batch * nrec receiver residuals
  int nsrc = 0;
  hz::DevBuf fields_u;  // nsrc * N (stored forward fields)
  hz::DevBuf src_f;     // nsrc * N copy of the sources (recompute policy)
  hz::DevBuf m_adj;     // adjoint right-hand side and multiplier (2 N)
  hz_krylov_opts mopts;
  bool fields_stored = false;
  double field_budget = 0.0;  // bytes for stored forward fields (HZ_FIELDS_AUTO); 0: half the free memory
  bool misfit_ready = false;

  // offsets into d_tab
  const double2* tab(int ax, int which) const {
    // which: 0 cm, 1 cp, 2 cdg, 3 adjoint cm (conj cp(i-1)), 4 adjoint cp (conj cm(i+1))
    const double2* base = d_tab.as<double2>();
    int64_t off = 0;
    for (int a = 0; a < ax; ++a) off += 5 * d.n[a];
    return base + off + which * d.n[ax];
  }
};

namespace hz {

// ------------------------------------------------------------ autotuner
void autotune(hz_problem* p, int prec, int nrhs, hz_tune_info* info);

// ------------------------------------------------------------ setup (host)
void setup_tables(hz_problem* p, const double* c_host);
// PML stretching factor 1 / (1 + i pml_c x^2 / omega) at (possibly half-integer)
// node s of an axis of n nodes with `pml` layer nodes per side (alpha,
// functions.cpp:1839-1878)
std::complex<double> stretch(int64_t n, int pml, double omega, double pml_c, double s);

// ------------------------------------------------------------ kernels
// z-FFT over every (x,y) line.  in/out: complex arrays of `in_prec`/`out_prec`
// (0: double2, 1: float2); compute precision `prec`.  sign -1 forward
// (unscaled), +1 backward (x 1/nz).  epi = 1 (backward only): out = w - dL*res
// with w (double2) and dL (double2) read at the same index.
void launch_fft_z(hz_problem* p, int prec, int sign, int nrhs, const void* in, int in_prec, void* out,
                  int out_prec, int epi, const double2* w, const double2* dL, hipStream_t s);
// the same with an explicit number of lines per workgroup (8, 16, 32)
void launch_fft_z_lines(hz_problem* p, int lines, int prec, int sign, int nrhs, const void* in, int in_prec,
                        void* out, int out_prec, int epi, const double2* w, const double2* dL, hipStream_t s);
// whether `lines` lines of length nz fit the LDS-resident FFT in precision prec
bool fft_lines_fit(int prec, int nz, int lines);

// S7 stencil y = A u (adjoint: A^H u); epi 1: y = f - A u.
void launch_stencil(hz_problem* p, int nrhs, const double2* u, double2* y, bool adjoint, const double2* f,
                    hipStream_t s);
// plane 5-point matvec on z-transformed data: y_k = M_k x_k (adjoint: M_k^H x_k).
void launch_plane_matvec(hz_problem* p, int nrhs, const double2* x, double2* y, bool adjoint, hipStream_t s);

// Raise a kernel's dynamic-LDS limit to at least `bytes` (hipFuncSetAttribute
// only when the limit grows, so that launches inside a stream capture make no
// attribute calls once the kernel has run eagerly).
void set_max_lds(const void* fn, size_t bytes);
// block-tridiagonal factorisation of every plane, precision prec.
void launch_factor(hz_problem* p, int prec, hipStream_t s);
// The same in two halves, so that one host thread can factor several problems
// at once: factor_begin enqueues the factorisation and the copy of its status
// word (pinned) on s; factor_end waits for it, checks the status and packs the
// symmetric factors.
void factor_begin(hz_problem* p, int prec, hipStream_t s);
void factor_end(hz_problem* p, int prec, hipStream_t s);
// zero the planes that the plane map skips, in nrhs vectors of complex of
// `elem` bytes (the preconditioner's output there, FGMRES.cpp:196-206 and
// functions.cpp:6023-6027)
void launch_zero_skipped(hz_problem* p, int nrhs, void* x, size_t elem, hipStream_t s);
// HODLR compression of precision prec's factors at eps with leaves of at most
// `small` rows (eps < 0: drop the compressed form); flags HZ_HODLR_FREE_DENSE
void hodlr_compress(hz_problem* p, int prec, double eps, int small, int flags);
// forward plane solves on the compressed factors
void launch_hodlr_solve(hz_problem* p, int prec, int nrhs, void* x, hipStream_t s);
// in-place plane solves on z-transformed data (prec 0: double2, 1: float2).
void launch_plane_solve(hz_problem* p, int prec, int nrhs, void* x, bool adjoint, hipStream_t s);
// the same with explicit launch choices (autotuner)
void launch_plane_solve_with(hz_problem* p, int prec, int nrhs, void* x, bool adjoint, const Tuning& t,
                             hipStream_t s);
// whether the streaming multi-RHS solve is instantiated for this width/precision
bool stream_solve_available(int prec, int nx);
// whether the symmetric packed streaming kernel is instantiated for this width/precision
bool sym_solve_available(int prec, int nx);
// whether the twisted two-plane packed solve is built for this problem and precision
bool twist_solve_available(const hz_problem* p, int prec);
// completes the top-down factor chain (and its packed copy) where the
// factorisation stopped it at the twisted factors' middle; synchronous, and an
// error under stream capture
void ensure_dense(hz_problem* p, int prec, hipStream_t s);
// whether the 8-RHS MFMA streaming kernel is instantiated for this width/precision
bool mfma_solve_available(int prec, int nx);
// whether the single-RHS ring kernel is instantiated for this width/precision
bool single_ring_available(int prec, int nx);

// ---- vector kernels (FP64 complex)
int reduce_blocks(int64_t n);
// partial sums of conj(x) . y  (or |x|^2 when y == nullptr) over a masked or full range,
// nv vectors laid out with stride N; result: out[v] (double2) on device.
void launch_dot(hz_problem* p, int64_t n, const double2* x, const double2* y, int mask, double2* out,
                hipStream_t s);
void launch_dot_multi(hz_problem* p, int64_t n, int nv, const double2* x, int64_t xstride, const double2* y,
                      int64_t ystride, int mask, double2* out, hipStream_t s);
// y = y + a * x where a = coef[0] * (neg ? -1 : 1) (device scalar)
void launch_axpy_dev(int64_t n, const double2* coef, bool neg, bool conj_coef, const double2* x, double2* y,
                     hipStream_t s);
// w -= coef[0] * x; out = y^H w (y == nullptr: out = ||w||), one pass
void launch_axpy_dot(hz_problem* p, int64_t n, const double2* coef, const double2* x, double2* w,
                     const double2* y, double2* out, hipStream_t s);
// y = x / |coef| (coef holds a norm in .x)
void launch_scale_inv_dev(int64_t n, const double2* coef, const double2* x, double2* y, hipStream_t s);
// y = x / coef.x if coef.x > thr, else y = 0
void launch_scale_inv_guard(int64_t n, const double2* coef, double thr, const double2* x, double2* y, hipStream_t s);
// Steps j of the device ICWY correction: j < ICWY_MAX (one 64-lane wave)
constexpr int ICWY_MAX = 64;
// Arnoldi step j < min(ICWY_MAX, m) after w = A M U_j, for every member of
// mask at once, on stored basis vectors U_k with v_k = s_k U_k (U_0 = v_0,
// s_0 = 1; s_k for k >= 1 in the step memory at offS).  Member q: basis
// vectors V + q n, vs apart; w at w + q n; step memory base mem + q msz holding
// r, l, the L rows, the Hessenberg columns, the scales, the Givens cosines and
// sines and g at offR, offL, offLg, offH, offS, offCS, offSN, offG.  Five
// launches: r = V^H w and l = V^H v_j in one pass over the basis; their sums
// (scaled by s_i s_j), one block per vector; h = (I + L)^{-1} r (row j of L
// stored from l), one wave per member; U_{j+1} = s_j w - sum h_i s_i U_i
// written to vnext + q n with the partial sums of ||U_{j+1}||^2 in a second
// pass; then, one block per member, h_{j+1,j} = ||U_{j+1}||, s_{j+1} =
// 1 / h_{j+1,j} (0 where h_{j+1,j} <= brk) and the Givens update of
// launch_givens (beta, nf, tol, est as there).  No pass normalises the new
// vector: its scale is folded into the next step's dots and coefficients and
// into the solution update (launch_backsub).  The reduction workspace must
// hold nr members' partials (ensure_krylov sizes it).
void launch_gs_members(hz_problem* p, int64_t n, int nr, uint64_t mask, int j, int m, int ldh, const double2* V,
                       int64_t vs, const double2* w, double2* mem, int64_t msz, int offR, int offL, int offLg,
                       int offH, int offS, int offCS, int offSN, int offG, const double2* beta, const double2* nf,
                       double tol, double brk, double2* est, double2* vnext, hipStream_t s);
// Complex Givens update of Arnoldi step j on the device, one thread per member
// q of `mask` (q < 64).  Member q's scalars start at base + q * sb: column j of
// H (leading dimension ldh) at offH + ldh * j, the cosines at offCS, sines at
// offSN, the least-squares right-hand side g at offG.  At j = 0, g = beta[q].x e_1.
// Writes est[q * m + j] = (|g_{j+1}| / nf[q].x, status): 0 go on, 1 below tol,
// 2 breakdown (h_{j+1,j} <= brk), 3 non-finite.
void launch_givens(int j, int m, int ldh, uint64_t mask, double2* base, int64_t sb, int offH, int offCS, int offSN,
                   int offG, const double2* beta, const double2* nf, double tol, double brk, double2* est,
                   hipStream_t s);
// y = R^{-1} g over the leading jj x jj triangle of the rotated H (one thread;
// R and y in LDS for jj <= 63)
// (S != nullptr: y_i *= s_i afterwards, S[i].x = s_i for i >= 1, s_0 = 1)
void launch_backsub(int jj, int ldh, const double2* H, const double2* g, double2* y, const double2* S,
                    hipStream_t s);
// y = sum_i c[i] * X_i  + (accumulate ? y : 0)
void launch_lincomb(int64_t n, int nv, const double2* c, const double2* X, int64_t xstride, double2* y,
                    bool accumulate, hipStream_t s);
// y = a*x + b*y with host scalars
void launch_axpby(int64_t n, double2 a, const double2* x, double2 b, double2* y, hipStream_t s);
// z = x + a*y + b*w (host scalars)
void launch_axpbypcz(int64_t n, const double2* x, double2 a, const double2* y, double2 b, const double2* w,
                     double2* z, hipStream_t s);
// out = a - b (elementwise)
void launch_sub(int64_t n, const double2* a, const double2* b, double2* out, hipStream_t s);
// out = w - dL .* x over nrhs vectors of N points (dL has N entries)
void launch_w_minus_dl(int64_t N, int nrhs, const double2* w, const double2* dL, const double2* x, double2* out,
                       hipStream_t s);
// grad[l] += scale * Re(conj(u[l]) * lam[l])
void launch_grad_acc_multi(int64_t n, int ns, double scale, const double2* u, int64_t us, const double2* lam,
                           int64_t ls, double* grad, hipStream_t s);
void launch_rec_residual(int64_t nrec, int ns, const int64_t* idx, const double2* U, int64_t ustride,
                         const double2* data, double2* res, hipStream_t s);
void launch_rec_scatter(int64_t nrec, int ns, const int64_t* idx, const double2* vals, double2* dst, int64_t dstride,
                        hipStream_t s);
void launch_grad_acc(int64_t n, double scale, const double2* u, const double2* lam, double* grad, hipStream_t s);
void launch_fill(int64_t n, double2 v, double2* y, hipStream_t s);
void launch_fill_real(int64_t n, double v, double* y, hipStream_t s);

// ------------------------------------------------------------ operators (hz_abi.cpp)
void grow(DevBuf& b, size_t bytes);
void ensure_work(hz_problem* p, int nrhs);
double2* dscal(hz_problem* p, int i);
// out = L0^{-1} in (adjoint: (L0^{-1})^H in); epi: REF out = in - deltaL L0^{-1} in
void precond_apply(hz_problem* p, int prec, int nrhs, const double2* in, double2* out, bool adjoint, bool epi);
// the same with the caller's work buffer (nrhs N complex128 for HZ_PREC_FP64,
// complex64 for HZ_PREC_C64); allocates nothing
void precond_apply_ws(hz_problem* p, int prec, int nrhs, const double2* in, double2* out, bool adjoint, bool epi,
                      void* work);
// out = A in (adjoint: A^H in), the true operator
void op_apply(hz_problem* p, int nrhs, const double2* in, double2* out, bool adjoint);
// r = f - A u
void residual(hz_problem* p, int nrhs, const double2* f, const double2* u, double2* r, bool adjoint);

// ------------------------------------------------------------ lockstep FGMRES (hz_krylov.cpp)
// Largest lockstep batch (right-hand sides sharing one batched apply).
constexpr int LOCKSTEP_MAX = 64;
// One problem's share of a lockstep solve: nrhs right-hand sides f -> u
// (device, stride N) in batches of `batch`, stats (host, nrhs).
struct LockJob {
  hz_problem* p = nullptr;
  int nrhs = 0;
  const double2* f = nullptr;
  double2* u = nullptr;
  hz_krylov_opts o{};
  hz_krylov_stats* st = nullptr;
  bool adjoint = false;
  int batch = 1;
  int rc = HZ_OK;   // outcome: HZ_OK or the code of the error that stopped it
  std::string msg;  // message of that error
};
// Runs every job from the calling thread: each problem's batches go one after
// the other on its own stream, the problems side by side.  Errors stop only
// their own job (rc, msg); the stream of a failed job is synchronised before
// its state is released.
void run_lockstep(std::vector<LockJob>& jobs);
// Krylov device storage of a lockstep batch of R members with restart m
// (basis, scalars, reduction partials), sized before the solve loop.
void ensure_krylov(hz_problem* p, int m, int R);

}  // namespace hz
