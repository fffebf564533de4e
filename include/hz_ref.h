// hz_ref.h -- reference-compatible C++ entry points over the hz C ABI.
//
// Same names, argument lists and types as the reference's hot-path functions
// (templates.h:59-60, 219-225; structs of definitions.h:44-75), so a caller of
// the reference links against libhz.so instead of the MKL PARDISO build.
// The PARDISO handle arguments (iparm, perm, pt, D2csr) are accepted and
// ignored: plane factors live inside the library, cached per grid and shift.
// Physical constants that the reference takes from compile-time macros
// (nu, c_z, HOMO, OVERTRUST, REL_RES_EXIT, RES_EXIT; definitions.h:192-303) are
// set once with hz_ref_configure().
//
// The medium of the drop-in solvers (FGMRES, BCGSTAB, Solve3DSparseUsingFT_CSR,
// ApplyCoeffMatrixA_CSR, Multiply3DSparseUsingFT) is, in this order of
// precedence: the array given to hz_ref_set_sound_speed, the OverThrust file of
// hz_ref_config.velocity_file (SetSoundSpeed3DFromFile, FGMRES.cpp:80), or the
// reference's built-in model (HOMO c_z, else the linear model of
// SetSoundSpeed3D).  A caller that computes deltaL from its own medium must
// configure the same medium, or L0 and deltaL describe different media.
#ifndef HZ_REF_H
#define HZ_REF_H

#include <complex>

#include "hz.h"

typedef std::complex<double> dtype;

struct size_m {
  double l;
  int n;
  double h;
  int pml_pts;
  int spg_pts;
  int n_nopml;
  double ta;
  double tc;
  double td;
  double c0, c1, c2, c3;
};

struct point {
  double x;
  double y;
  double z;
};

struct MatrixCSRComplex {
  int* ia = nullptr;
  int* ja = nullptr;
  dtype* values = nullptr;
  int non_zeros = 0;
  int solve = 0;
};
typedef struct MatrixCSRComplex zcsr;

extern "C" {
typedef struct {
  double nu;            // omega = 2 pi nu                 (definitions.h:293, 302)
  double c_z;           // homogeneous speed               (definitions.h:294)
  int homo;             // 1: HOMO build, 0: linear hetero (definitions.h:193, functions.cpp:1371-1378)
  double rel_res_exit;  // FGMRES stop                     (definitions.h:199)
  double res_exit;      // BiCGStab stop (absolute)        (definitions.h:198)
  double pml_c;         // d(x) = pml_c x^2                (functions.cpp:897)
  int prec;             // HZ_PREC_* of the preconditioner (reference: FP64)
  const char* velocity_file;  // OVERTRUST model file (float32 [y][x][z] over the physical box), or NULL;
                              // the path is copied  (definitions.h:192, functions.cpp:1458-1531)
} hz_ref_config;

int hz_ref_configure(const hz_ref_config* cfg);
// Medium of the drop-in solvers: n sound speeds over the whole grid (layers
// included, x fastest), copied; NULL clears it.  The solvers fail (error
// message, NaN output) on a grid of another size.
int hz_ref_set_sound_speed(const double* c, int64_t n);
// Status of the last drop-in call on this thread (HZ_OK or an HZ_ERR_* code):
// the reference's functions return void.
int hz_ref_last_status(void);
void hz_ref_release(void);
}

// ---- set-up functions of the reference path (host memory, no device work).
// PML stretching factor at (possibly half-integer) node i of one axis, with
// omega = 2 pi nu and d(x) = pml_c x^2 of the configuration.
// functions.cpp:1839 / templates.h:160
dtype alpha(size_m xyz, double i);
// Coefficient of the 5-point PML plane operator at node (i, j): diag_case 0
// the diagonal (plus kwave_beta2), +-1 the x neighbours, +-2 the y neighbours.
// functions.cpp:2347 / templates.h:72
dtype beta2D_pml(size_m x, size_m y, int diag_case, dtype kwave_beta2, int i, int j);
// Coefficient of the 7-point PML operator at node (i, j, k): diag_case 0 the
// diagonal, +-1, +-2, +-3 the x, y, z neighbours.  functions.cpp:2242 (not
// declared in the reference's headers)
dtype beta3D(size_m x, size_m y, size_m z, int diag_case, int i, int j, int k);
// CSR (1-based) of one plane: w == -1 assembles the pattern and the values of
// beta2D_pml with kwave_beta2 and records the diagonal positions in freqs;
// w >= 0 adds kwave_beta2 to the recorded diagonal entries.  Acsr->ia, ja,
// values sized by the caller, Acsr->non_zeros = 5 nx ny - 2 nx - 2 ny.
// functions.cpp:3270 / templates.h:186
void GenSparseMatrixOnline2DwithPMLFast(int w, size_m x, size_m y, zcsr* Acsr, dtype kwave_beta2, int* freqs);
// 9-point plane operators (DIAG_PATTERN 9, definitions.h:188-189).  diag_case
// as the reference's DIAG9 (functions.cpp:3373-3412): 0 the node, +-1 x
// neighbours, +-3 y neighbours, 4 / -4 the (x+1, y+1) / (x-1, y-1) corners,
// 2 / -2 the (x-1, y+1) / (x+1, y-1) corners.
// functions.cpp:2404 / templates.h:73 (weight sigma moved from x to the corners)
dtype beta2D_pml_9pts(size_m x, size_m y, int diag_case, dtype kwave_beta2, int i, int j, double sigma);
// functions.cpp:2467 / templates.h:74 (stencil weights a, c, d)
dtype beta2D_pml_9pts_opt(size_m x, size_m y, int diag_case, dtype kwave_beta2, int i, int j, double a, double c,
                          double d);
// CSR (1-based) of one 9-point plane with beta2D_pml_9pts_opt at the
// reference's weights a = 1, c = 1, d = 0 (functions.cpp:3553-3555; its corner
// coefficients are then zero and the values those of the 5-point plane):
// w == -1 pattern + values + diagonal positions, w >= 0 adds kwave_beta2 to the
// diagonal.  Acsr->non_zeros = 9 nx ny - 6 nx - 6 ny + 4.
// functions.cpp:3532 / templates.h:71
void GenSparseMatrixOnline2DwithPMLand9Pts(int w, size_m x, size_m y, zcsr* Acsr, dtype kwave_beta2, int* freqs);
// The same with beta2D_pml_9pts at weight sigma.  functions.cpp:3414 / templates.h:69
void GenSparseMatrixOnline2DwithPMLand9Points(int w, size_m x, size_m y, size_m z, zcsr* Acsr, dtype kwave_beta2,
                                              int* freqs, double sigma);
// Sound speed of the built-in model on the physical x/y nodes of every plane
// (HOMO: c_z; else c = c1 x + c2 y + c3 with c(0,0) = 600, c(0,Ly) = 1500,
// c(Lx,Ly) = 1800).  functions.cpp:1362 / templates.h:172
void SetSoundSpeed3D(size_m x, size_m y, size_m z, dtype* sound3D, point source);
// OverThrust model from the file hz_velocity_file_name(x.h) in the working
// directory (or hz_ref_config.velocity_file when set), physical box only.
// functions.cpp:1458 / templates.h:86
void SetSoundSpeed3DFromFile(size_m x, size_m y, size_m z, dtype* sound3D, point source);
// x/y layer nodes copy the nearest physical node of their row / column.
// functions.cpp:2675 / templates.h:83
void HeteroSoundSpeed3DExtensionToPML(size_m x, size_m y, size_m z, dtype* sound3D);
// Background c0 of the planes: the running-sum mean of 1/c^2 over the physical
// z range (SetSoundSpeed2D's accumulator is never cleared), physical x/y only.
// functions.cpp:1686 / templates.h:173
void SetSoundSpeed2D(size_m x, size_m y, size_m z, dtype* sound3D, dtype* sound2D, point source);
// functions.cpp:2636 / templates.h:178
void HeteroSoundSpeed2DExtensionToPML(size_m x, size_m y, dtype* sound2D);
// deltaL = omega^2 ((1 + i beta)/c0^2 - 1/c^2) on planes spg..nz-spg, Gaussian
// taper (sigma 0.3) in the sponge planes.  functions.cpp:1740 / templates.h:177
void GenerateDeltaL(size_m x, size_m y, size_m z, dtype* sound3D, dtype* sound2D, dtype* deltaL, double beta_eq);

// FGMRES.cpp:6 / templates.h:59
void FGMRES(size_m x, size_m y, size_m z, int m, int rest, const point source, dtype* x_sol, dtype* x_orig,
            const dtype* f, double thresh, double& diff_sol, double beta_eq);
// BcGSTAB.cpp:6 / templates.h:60
void BCGSTAB(size_m x, size_m y, size_m z, int m, const point source, dtype* x_sol, dtype* x_orig, const dtype* f,
             double thresh, double& diff_sol, double beta_eq);
// functions.cpp:5837 / templates.h:219
void Solve3DSparseUsingFT_CSR(size_m x, size_m y, size_m z, int* iparm, int* perm, size_t* pt, zcsr**& D2csr,
                              const dtype* f, dtype* x_sol, double beta_eq, double thresh);
// functions.cpp:5554 / templates.h:220
void ApplyCoeffMatrixA_CSR(size_m x, size_m y, size_m z, int* iparm, int* perm, size_t* pt, zcsr**& D2csr,
                           const dtype* w, const dtype* deltaL, dtype* g, double beta_eq, double thresh);
// functions.cpp:6247 / templates.h:225
void Multiply3DSparseUsingFT(size_m x, size_m y, size_m z, int* iparm, int* perm, size_t* pt, zcsr**& D2csr,
                             const dtype* u, dtype* f_sol, double thresh);

#endif  // HZ_REF_H
