/* hz.h -- C ABI of the MI355X 3D Helmholtz solver hot path.
 *
 * The reference (dklyuchinskiy/Iterative-Helmholtz-solver-with-PARDISO, paths
 * below relative to "Iterative Helholtz solver with PARDISO/Source/") exposes
 * its hot path as free C++ functions with MKL PARDISO handles threaded through;
 * each entry point here names the reference function it replaces.
 *
 * Conventions
 *  - complex numbers are interleaved (re, im) doubles: hz_c128;
 *  - every vector argument is a DEVICE pointer holding nrhs consecutive
 *    vectors of N = nx*ny*nz points, layout l = i + nx*j + nx*ny*k
 *    (functions.cpp:225-230);
 *  - all calls are ordered on the problem's HIP stream; calls that return
 *    host-visible results (hz_fgmres, hz_bicgstab, norms, misfits) synchronise
 *    that stream before returning;
 *  - every function returns an HZ_* status; hz_last_error() gives the message
 *    of the last failure on the calling thread;
 *  - a problem handle is not thread-safe: one host thread per handle.
 */
#ifndef HZ_H
#define HZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double re, im;
} hz_c128;

enum {
  HZ_OK = 0,
  HZ_ERR_ARG = 1,           /* invalid argument or size */
  HZ_ERR_HIP = 2,           /* HIP runtime error */
  HZ_ERR_STATE = 3,         /* call out of order (e.g. apply before factor) */
  HZ_ERR_NOMEM = 4,         /* device allocation failed */
  HZ_ERR_NOT_CONVERGED = 5, /* Krylov driver hit its iteration limit */
  HZ_ERR_BREAKDOWN = 6,     /* Krylov breakdown before the tolerance */
  HZ_ERR_SINGULAR = 7,      /* zero pivot in a plane factorisation */
  HZ_ERR_NONFINITE = 8      /* NaN/Inf met in a residual */
};

/* operator modes */
enum {
  HZ_MODE_REF = 0, /* x/y PML, spectral periodic z, sponge delta-L: A = L0 - diag(deltaL) */
  HZ_MODE_S7 = 1   /* 7-point PML stencil on every face (beta3D, functions.cpp:2242-2303) */
};
/* z symbol of the preconditioner planes */
enum {
  HZ_KZ_REFERENCE = 0, /* FGMRES.cpp:182-190 map (k < nz/2-1 -> k, else nz-k) */
  HZ_KZ_SPECTRAL = 1,  /* min(k, nz-k) */
  HZ_KZ_FD = 2         /* (2 - 2cos(2 pi k/nz))/hz^2 */
};
/* background speed of the preconditioner */
enum {
  HZ_C0_REFERENCE = 0, /* SetSoundSpeed2D running sum, functions.cpp:1716-1728 */
  HZ_C0_HARMONIC = 1   /* per-column harmonic mean over the physical z range */
};
/* preconditioner precision */
enum {
  HZ_PREC_FP64 = 0, /* complex128 factors, FFT and apply */
  HZ_PREC_C64 = 1   /* complex64 factors (FP32 arithmetic), FFT and apply */
};
/* Krylov options */
enum { HZ_DRIVER_FGMRES = 0, HZ_DRIVER_GMRES_REF = 1 };
/* Arnoldi orthogonalisation.  GMRES_REF: REF_CONJ is the reference's w^H V_i
 * coefficient (FGMRES.cpp:511), MGS the textbook V_i^H w.  FGMRES: MGS is
 * modified Gram-Schmidt in its inverse compact-WY form (coefficients
 * (I + L)^{-1} V^H w, L = strict lower part of V^H V: two passes over the
 * basis per step; restarts m <= 64, the sequential form beyond), MGS_SEQ the
 * sequential form (one pass per basis vector). */
enum { HZ_ORTHO_REF_CONJ = 0, HZ_ORTHO_MGS = 1, HZ_ORTHO_MGS_SEQ = 2 };
enum { HZ_MASK_FULL = 0, HZ_MASK_PHYS = 1 };

/* Problem descriptor.  Replaces the per-axis size_m structs (definitions.h:56-68)
 * and the compile-time macros omega, c_z, nu, HOMO (definitions.h:163-329). */
typedef struct {
  int64_t n[3];     /* nodes per axis including layers */
  double h[3];      /* spacing per axis (node i at (i+1) h) */
  double l[3];      /* domain length per axis (z length sets the spectral symbol) */
  int32_t pml[3];   /* PML nodes per side (alpha profile, functions.cpp:1839-1878) */
  int32_t spg[3];   /* sponge nodes per side (REF: z only; delta-L taper and phys mask) */
  double omega;     /* angular frequency */
  double beta;      /* complex shift of the preconditioner (beta_eq) */
  double pml_c;     /* d(x) = pml_c x^2 (1000 in functions.cpp:893-906) */
  int32_t mode;     /* HZ_MODE_* */
  int32_t kz_symbol;/* HZ_KZ_* */
  int32_t c0_rule;  /* HZ_C0_* */
  int32_t skip_ratio; /* > 0: planes with z symbol lambda_k >= skip_ratio * k0^2 are not factored
                         and the preconditioner is zero on them (SAVE_MEM, FGMRES.cpp:196-206;
                         the reference's ratio is 5, or 15 at nu = 2); 0: every plane */
  double c_bg;      /* > 0: constant background speed of the planes (reference HOMO) */
  double skip_k0;   /* wavenumber k0 of the skip test (the reference's omega / c_z); 0: omega / c_bg,
                       or omega over the slowest background speed when c_bg = 0 */
} hz_desc;

typedef struct hz_problem hz_problem;

typedef struct {
  int32_t driver;       /* HZ_DRIVER_* */
  int32_t m;            /* restart length (FGMRES niter, main.cpp:150) */
  int32_t max_restarts; /* restart count (main.cpp:151) */
  int32_t ortho;        /* HZ_ORTHO_* */
  int32_t mask;         /* HZ_MASK_* residual box */
  int32_t monitor;      /* GMRES_REF: the reference's per-iteration u = L0^{-1} x (FGMRES.cpp:601) */
  double tol;           /* relative residual target */
  double breakdown;     /* |h_{j+1,j}| threshold (FGMRES.cpp:522) */
  double* hist;         /* optional host array of per-iteration residuals (may be NULL) */
  int32_t hist_cap;
  int32_t prec;         /* HZ_PREC_* used by the driver's preconditioner applies */
  int32_t batch;        /* FGMRES: right-hand sides advanced in lockstep, sharing each batched
                           preconditioner and operator apply (0: all of them, at most 16) */
  int32_t graph;        /* FGMRES: 1 = replay each chunk of Arnoldi steps (preconditioner, operator,
                           Gram-Schmidt, normalisation, Givens update) from a captured HIP graph */
  int32_t autotune;     /* FGMRES: 1 = run hz_autotune for (prec, batch size) before the first step */
  int32_t chunk;        /* FGMRES: Arnoldi steps between host reads of the convergence words
                           (0: 1 step, or 5 with graph) */
} hz_krylov_opts;

typedef struct {
  int32_t iters;     /* inner iterations (one preconditioned operator apply each) */
  int32_t restarts;  /* restart cycles started */
  int32_t converged;
  int32_t breakdown;
  double rel_res;    /* final true relative residual over the mask */
  double res;        /* final absolute residual over the mask */
  double diff_sol;   /* GMRES_REF monitor ||u_k - u_{k-1}||/||u_{k-1}|| (FGMRES.cpp:644) */
  int32_t n_hist;
  int32_t reserved;
  double t_solve_s;  /* wall time of the Krylov loop */
} hz_krylov_stats;

/* Launch choices picked by hz_autotune for one (precision, right-hand-side count). */
enum {
  HZ_SOLVE_DEFAULT = 0, /* streaming multi-RHS solve where instantiated, else chunked */
  HZ_SOLVE_STREAM = 1,  /* factor tiles streamed through an LDS ring, up to 8 RHS per pass */
  HZ_SOLVE_CHUNKED = 2, /* chunked GEMV kernel, any width */
  HZ_SOLVE_STREAM_1 = 3, /* single-RHS ring with an LDS cross-group reduction (200-wide planes) */
  HZ_SOLVE_STREAM_MFMA = 4, /* streaming solve with the 8-RHS block product on FP32 MFMA (complex64) */
  HZ_SOLVE_SYM = 5,         /* symmetric packed factors (upper block triangle of G_j D_j), one RHS per pass */
  HZ_SOLVE_TWIST = 6        /* twisted packed factors (chains from both ends meeting at y-line ny/2), the
                               two planes of each factor slot per pass (complex64) */
};
typedef struct {
  int32_t solve_kernel; /* HZ_SOLVE_* */
  int32_t group_cap;    /* right-hand sides per streaming pass: 8, 4, 2 or 1 */
  int32_t fft_lines;    /* (x,y) lines per z-FFT workgroup: 8, 16 or 32 (LDS Stockham), 0 for the
                           register-resident kernel of nz = 64, 96, 128, 160, 192, 200 */
  int32_t from_cache;   /* 1 when read from the HZ_TUNE_CACHE file */
  double solve_ms;      /* plane solve of nrhs vectors with the choice */
  double fft_ms;        /* one forward z-FFT of nrhs vectors with the choice */
} hz_tune_info;

/* ---------------------------------------------------------------- library */
const char* hz_last_error(void);
int hz_version(void);

/* ---------------------------------------------------------------- problem */
/* Set up one frequency: PML tables (alpha, functions.cpp:1839-1878;
 * beta2D_pml / beta3D, functions.cpp:2242-2402), background c0 (SetSoundSpeed2D,
 * functions.cpp:1686-1738), delta-L (GenerateDeltaL, functions.cpp:1740-1824)
 * and the device copies of the medium.  c_host: N sound speeds (host memory).
 * Replaces the setup block of FGMRES, FGMRES.cpp:72-101. */
int hz_problem_create(const hz_desc* desc, const double* c_host, hz_problem** out);
int hz_problem_destroy(hz_problem* p);
/* Use an external HIP stream (hipStream_t) for every later call; NULL = own stream. */
int hz_problem_set_stream(hz_problem* p, void* hip_stream);
int hz_problem_sync(hz_problem* p);
/* The HIP stream (hipStream_t) every call of this problem is ordered on. */
void* hz_problem_stream(const hz_problem* p);
/* dims[0..3] = nx, ny, nz, N; *factor_bytes = bytes held by the plane factors. */
int hz_problem_info(const hz_problem* p, int64_t* dims, double* factor_bytes);
/* *nslot = factor slots (distinct z symbols among the solved planes: planes k
 * and nz - k share one), *nskip = planes that SAVE_MEM leaves unsolved.
 * Replaces nothing in the reference (its plane loop factors every kept plane,
 * FGMRES.cpp:196-206); reports what the plane map of hz_problem_create chose. */
int hz_problem_slots(const hz_problem* p, int* nslot, int* nskip);

/* ------------------------------------------------------- preconditioner */
/* Factor every kz plane M_k = D_xy + kappa - kz_k^2 as a block-tridiagonal
 * Schur recursion over y-lines with explicit nx x nx inverses (Gauss-Jordan,
 * partial pivoting).  Replaces the per-plane PARDISO phases 11+22,
 * FGMRES.cpp:176-298 (and DirFactFastDiagStructOnline,
 * HODLR/BinaryTreesHODLR.cpp:2870-2969). */
int hz_precond_factor(hz_problem* p, int prec);
/* out = L0^{-1} in = F_z^{-1} diag(M_k^{-1}) F_z in, in precision `prec`
 * (must have been factored).  Replaces Solve3DSparseUsingFT_CSR,
 * functions.cpp:5837-6061. */
int hz_precond_apply(hz_problem* p, int prec, int nrhs, const hz_c128* in, hz_c128* out);
/* hz_precond_apply without any device allocation: the z-transformed planes
 * live in a fixed 64 MiB device array of the library (one per device, held
 * until the process ends) instead of the problem's grown work buffer.  No
 * reference counterpart: for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops that forbid hipMalloc).  nrhs * N
 * must fit the array (8 Mi complex64 or 4 Mi complex128 values), else
 * HZ_ERR_ARG.  Calls that overlap on one device (several streams or host
 * threads) share the array and race. */
int hz_precond_apply_noalloc(hz_problem* p, int prec, int nrhs, const hz_c128* in, hz_c128* out);
/* out = (L0^{-1})^H in, with the conjugate-transposed plane factors
 * (no reference counterpart: adjoint solves of the misfit gradient). */
int hz_precond_apply_adjoint(hz_problem* p, int prec, int nrhs, const hz_c128* in, hz_c128* out);

/* HODLR compression of the plane factors (the reference's HODLR build:
 * SymRecCompressStruct / LowRankApproxStruct, HODLR/BinaryTreesHODLR.cpp:298-416,
 * used by DirFactFastDiagStructOnline :2870-2969 and applied by
 * DirSolveFastDiagStruct :2971-3018 inside Solve3DSparseUsingFT_HODLR,
 * functions.cpp:6064-6244).  Every block inverse of precision `prec` (already
 * factored) is compressed in its symmetric form K_j = G_j D_j: nodes of more
 * than `smallsize` rows split into n - ceil(n/2) and ceil(n/2) rows, the lower
 * off-diagonal block keeps the singular triplets with S[t] / S[0] >= eps
 * (0 <= eps < 1; 0 keeps every nonzero one), the upper one is its transpose.
 * Later forward preconditioner applies of that precision use the compressed
 * records; adjoint applies use the dense factors, which flags =
 * HZ_HODLR_FREE_DENSE releases.  nx <= 256, tree depth 3..6 (smallsize between
 * about nx/64 and nx/8), at most rank 16 per block.  eps < 0 drops the
 * compressed form (back to the dense factors if kept).  A new
 * hz_precond_factor drops it too. */
/* HZ_HODLR_TWISTED (complex64, where the twisted factors exist): compress the
 * twisted factors (chains from both ends of each y-line chain meeting at
 * ny/2) instead of the one-directional ones; the forward applies then run two
 * workgroups per plane over half chains.  No counterpart in the reference,
 * whose HODLR recursion is one-directional (BinaryTreesHODLR.cpp). */
enum { HZ_HODLR_FREE_DENSE = 1, HZ_HODLR_TWISTED = 2 };
int hz_precond_compress(hz_problem* p, int prec, double eps, int smallsize, int flags);
typedef struct {
  int32_t active;     /* 1 when the forward applies of prec use the compressed records */
  int32_t smallsize;
  int32_t levels;     /* depth of the row tree */
  int32_t nodes;      /* compressed off-diagonal blocks per block inverse */
  int32_t leaf_max;   /* rows of the largest leaf */
  int32_t max_rank;
  double eps;
  double mean_rank;   /* over every (slot, y-line, node) */
  double bytes;       /* device bytes of the compressed records */
  double dense_bytes; /* device bytes of the dense factors still held */
} hz_hodlr_info;
int hz_precond_hodlr_info(const hz_problem* p, int prec, hz_hodlr_info* out);
/* ranks of plane k's compressed blocks: out[j * nodes + q] for y-line j and
 * node q (breadth-first order of the row tree); cap >= ny * nodes.  Planes
 * that share a factor slot report the same ranks; a plane skipped by the
 * SAVE_MEM rule gives HZ_ERR_ARG. */
int hz_precond_hodlr_ranks(const hz_problem* p, int prec, int64_t k, int32_t* out, int64_t cap);

/* ---------------------------------------------------------- operators */
/* out = L0 in.  Replaces Multiply3DSparseUsingFT, functions.cpp:6247-6318. */
int hz_L0_apply(hz_problem* p, int nrhs, const hz_c128* in, hz_c128* out);
/* out = A in, the true operator: REF = L0 - diag(deltaL); S7 = 7-point PML
 * stencil + omega^2/c^2 (beta3D, functions.cpp:2242-2303; ComputeResidual,
 * functions.cpp:6446-6504). */
int hz_op_apply(hz_problem* p, int nrhs, const hz_c128* in, hz_c128* out);
/* out = A^H in. */
int hz_op_apply_adjoint(hz_problem* p, int nrhs, const hz_c128* in, hz_c128* out);
/* out = A L0^{-1} in (REF: in - deltaL * L0^{-1} in).  Replaces
 * ApplyCoeffMatrixA_CSR, functions.cpp:5554-5575. */
int hz_op_apply_preconditioned(hz_problem* p, int prec, int nrhs, const hz_c128* in, hz_c128* out);
/* r = f - A u and rel[q] = ||mask(r_q)|| / ||f_q|| (host array of nrhs). */
int hz_residual(hz_problem* p, int nrhs, const hz_c128* f, const hz_c128* u, hz_c128* r, int mask,
                double* rel);

/* ------------------------------------------------------ Krylov drivers */
/* Solve A u = f for nrhs right-hand sides (stats: host array of nrhs).
 *  HZ_DRIVER_FGMRES:    flexible GMRES, MGS, complex Givens, FP64 true-operator
 *                       matvec and residual, preconditioner in opts->prec; the
 *                       right-hand sides run in lockstep batches of opts->batch
 *                       (at most 64); every Arnoldi step, the Givens update
 *                       included, runs on the device, the host reads one
 *                       convergence word per right-hand side and step;
 *  HZ_DRIVER_GMRES_REF: the reference's restarted GMRES on (I - deltaL L0^{-1}),
 *                       zgels least squares, per-iteration true residual over
 *                       the mask, u = L0^{-1} x (FGMRES.cpp:6-725).
 * Replaces FGMRES (templates.h:59). */
int hz_fgmres(hz_problem* p, int nrhs, const hz_c128* f, hz_c128* u, const hz_krylov_opts* opts,
              hz_krylov_stats* stats);
/* Right-preconditioned BiCGStab with shadow residual 2 r0; stops when the
 * residual over the mask (absolute when tol_abs, else relative) < tol.
 * hist (host, optional): relative residual after each iteration of the first
 * right-hand side.  Replaces BCGSTAB (templates.h:60; BcGSTAB.cpp:6-666). */
int hz_bicgstab(hz_problem* p, int prec, int nrhs, const hz_c128* f, hz_c128* u, int maxit, double tol,
                int tol_abs, int mask, double* hist, int hist_cap, hz_krylov_stats* stats);

/* Time the launch variants of the preconditioner for nrhs right-hand sides
 * (plane-solve kernel family and right-hand-side group size, z-FFT lines per
 * workgroup) on this problem's factors and keep the fastest for later calls
 * with that precision and count.  Needs hz_precond_factor(prec).  With the
 * environment variable HZ_TUNE_CACHE naming a file, choices are read from and
 * appended to it (key: grid, precision, nrhs).  info (optional) receives the
 * choice.  No reference counterpart (PARDISO's iparm defaults are fixed). */
int hz_autotune(hz_problem* p, int prec, int nrhs, hz_tune_info* info);

/* ------------------------------------------------------ several problems */
/* Factor nprob problems (e.g. several frequencies on grids of different sizes)
 * concurrently, each on its own stream, queued from the calling thread.  No reference counterpart (the
 * reference factors one frequency per run, FGMRES.cpp:176-298). */
int hz_precond_factor_multi(int nprob, hz_problem* const* probs, int prec);
/* hz_fgmres on nprob distinct problems at once: problem i solves nrhs[i]
 * right-hand sides f[i] -> u[i] (device, nrhs[i]*N_i each) with stats[i] (host,
 * nrhs[i]); every problem runs its lockstep batches on its own stream, and the
 * calling thread queues the work of all problems (no other host threads).
 * opts->hist, when given, records problem 0's first right-hand side.  No
 * reference counterpart. */
int hz_fgmres_multi(int nprob, hz_problem* const* probs, const int* nrhs, const hz_c128* const* f, hz_c128* const* u,
                    const hz_krylov_opts* opts, hz_krylov_stats* const* stats);

/* ------------------------------------------------------------- adjoint */
/* Receiver misfit J = 1/2 sum_s ||P u_s - d_s||^2 (S7 mode) for nsrc sources
 * f (device, nsrc*N), receivers rec (host, nrec flat indices) and data d
 * (host, nsrc*nrec).  Forward fields are kept on the device when
 * store_fields > 0, recomputed by hz_misfit_backward when it is 0, and
 * chosen against the field budget with HZ_FIELDS_AUTO.  No reference
 * counterpart. */
int hz_misfit_forward(hz_problem* p, int nsrc, const hz_c128* f, int nrec, const int64_t* rec,
                      const hz_c128* data, const hz_krylov_opts* opts, int store_fields, double* J);
/* store_fields = HZ_FIELDS_AUTO: keep the nsrc forward fields (nsrc * N complex128) when they fit the
 * field budget, else recompute them in hz_misfit_backward.  The budget is the value of
 * hz_misfit_set_budget, or (0, the default) half of the device memory free when hz_misfit_forward
 * starts.  hz_misfit_fields_stored reports the choice of the last forward call. */
enum { HZ_FIELDS_AUTO = -1 };
int hz_misfit_set_budget(hz_problem* p, double bytes);
int hz_misfit_fields_stored(const hz_problem* p, int* stored);
/* grad (device, N doubles) = dJ/dm, m = 1/c^2:
 *   A^H lambda_s = P^T (P u_s - d_s),  grad = -omega^2 sum_s Re(conj(u_s) lambda_s). */
int hz_misfit_backward(hz_problem* p, double* grad);

/* ------------------------------------------------------ stage hooks */
/* z-DFT of every (x,y) line: sign -1 forward (unscaled), +1 backward (x 1/nz),
 * computed in `prec`; in/out device c128.  The DftiCompute loops of
 * functions.cpp:5876-5881, 6041-6046. */
int hz_fft_z(hz_problem* p, int prec, int sign, int nrhs, const hz_c128* in, hz_c128* out);
/* In-place plane solves M_k x_k = b_k (adjoint: M_k^H) on z-transformed data,
 * x: nrhs * N complex of the factor precision (complex128 for HZ_PREC_FP64,
 * complex64 for HZ_PREC_C64).  The per-plane PARDISO phase-33 loop of
 * functions.cpp:5941-6028. */
int hz_plane_solve(hz_problem* p, int prec, int adjoint, int nrhs, void* x);

/* ---- velocity models, solution and convergence files (host memory) ---- */

/* Reads a velocity model of the physical box (the nodes inside the pml + spg
 * layers of every axis) stored as float32 in [y][x][z] order, z fastest, into
 * c (host, N doubles, x fastest) and extends it into the layers: x and y layer
 * nodes copy the nearest physical node of their row / column
 * (HeteroSoundSpeed3DExtensionToPML, functions.cpp:2675-2716), z layer nodes
 * the nearest physical z node (the reference leaves them at zero).  Rejects a
 * short file and non-positive or non-finite speeds.  Replaces
 * SetSoundSpeed3DFromFile (functions.cpp:1458-1531). */
int hz_velocity_from_file(const char* path, const hz_desc* desc, double* c);
/* The reference's model file name for spacing h: "OT_small_at_%dm.bin" with (int) h
 * (functions.cpp:1471). */
int hz_velocity_file_name(double h, char* buf, int len);
/* Writes the physical box of u (host, N complex128) as one "%18.16lf %18.16lf" line per node,
 * x fastest, then y, then z (reducePML3D order, functions.cpp:238-263): the sol3D_N%d.dat
 * format of main.cpp:827-834. */
int hz_write_solution(const char* path, const hz_desc* desc, const hz_c128* u);
/* Reads that format into the physical box of u (host, N complex128; layer nodes set to zero);
 * rejects a file with fewer or more values than physical nodes. */
int hz_read_solution(const char* path, const hz_desc* desc, hz_c128* u);
/* Writes a convergence log, one "%d %e %e %lf" line per entry: iteration, residual, relative
 * residual, relative change of the physical-box solution (FGMRES.cpp:646). */
int hz_write_convergence(const char* path, int count, const int* iter, const double* res, const double* relres,
                         const double* diff);
/* The reference's convergence file name
 * "convergence_N%d_PML%d_Lx%d_FREQ%d_SPG%6.lf_BETA%5.3lf_FGMRES.dat" (FGMRES.cpp:50): physical
 * x nodes, x PML nodes, (int) x length, (int) frequency nu, z sponge width 2 spg h_z, beta. */
int hz_convergence_file_name(const hz_desc* desc, double length_x, double nu, double beta, char* buf, int len);

/* ---- accuracy studies over solution files (host memory) ---- */

/* Runge's ratio of three nested solutions (Runge, functions.cpp:932-1134, 3D branch): the files
 * hold cubic physical boxes of n_coarse, n_mid, n_fine nodes per side in the sol3D_N%d.dat
 * format, with (n_fine + 1) a multiple of (n_coarse + 1) and of (n_mid + 1) (the reference's
 * 49, 99, 199).  out[0] = h * sum over the coarse box of |u_fine - u_coarse|, out[1] the same
 * over the middle box (IntegralNorm3D, functions.cpp:1348-1360), out[2] = out[0] / out[1]; the
 * reference prints log2(out[2]) as the order (main.cpp:375-376).  Rejects boxes that are not
 * nested and files with fewer or more values than the box. */
int hz_runge_ratio(const char* coarse, const char* mid, const char* fine, int n_coarse, int n_mid, int n_fine,
                   double h, double* out);
/* Beta3D's shell norms (Beta3D, functions.cpp:1136-1229; compute_and_print_circle_norm_nopml,
 * :1290-1313; check_norm_circle_3D_nopml, :629-674): coarse holds a box of grid - 1 nodes per
 * side, fine one of 2 grid - 1; node (i, j, k) of the coarse box sits at ((i+1) h, (j+1) h,
 * (k+1) h).  For e = 2 .. grid - 2: radius[e-2] = e h and norm[e-2] = ||u_coarse - u_fine|| /
 * ||u_fine|| (Frobenius) over the coarse nodes with 2h <= |x - source| <= e h; count must be
 * grid - 3. */
int hz_beta3d_norms(const char* coarse, const char* fine, int grid, double h, const double* source, int count,
                    double* radius, double* norm);
/* Writes those norms as "%lf %lf" lines (functions.cpp:1295-1302). */
int hz_write_beta_norms(const char* path, int count, const double* radius, const double* norm);
/* The norms file name "beta_3D_norm2_Nx%dNy%d_SPG%d.dat" for n physical nodes per side and
 * spg z sponge nodes (functions.cpp:1295). */
int hz_beta_norm_file_name(int n, int spg, char* buf, int len);

#ifdef __cplusplus
}
#endif
#endif /* HZ_H */
