// hz_ref.h -- reference-compatible C++ entry points over the hz C ABI.
//
// Same names, argument lists and types as the reference's hot-path functions
// (templates.h:59-60, 219-225; structs of definitions.h:44-75), so a caller of
// the reference links against libhz.so instead of the MKL PARDISO build.
// The PARDISO handle arguments (iparm, perm, pt, D2csr) are accepted and
// ignored: plane factors live inside the library, cached per grid and shift.
// Physical constants that the reference takes from compile-time macros
// (nu, c_z, HOMO, REL_RES_EXIT, RES_EXIT; definitions.h:193-303) are set once
// with hz_ref_configure().
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
} hz_ref_config;

int hz_ref_configure(const hz_ref_config* cfg);
void hz_ref_release(void);
}

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
