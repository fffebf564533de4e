// ref_link_check.cpp -- stands where the reference's Source/functions.cpp
// stands in a build of the reference against libhz.so: a translation unit
// that defines every function libhz.so replaces from functions.cpp (same
// signatures).  Built with the renaming flags of the integration recipe
// (-D<name>=ref_<name>), its definitions get other names and the callers bind
// to libhz.so's; built without them, they interpose and the first call aborts.
#include <cstdio>
#include <cstdlib>

#include "hz_ref.h"

[[noreturn]] static void bound(const char* name) {
  fprintf(stderr, "the translation unit's own %s was called instead of libhz.so's\n", name);
  abort();
}

void Solve3DSparseUsingFT_CSR(size_m, size_m, size_m, int*, int*, size_t*, zcsr**&, const dtype*, dtype*, double,
                              double) {
  bound("Solve3DSparseUsingFT_CSR");
}
void ApplyCoeffMatrixA_CSR(size_m, size_m, size_m, int*, int*, size_t*, zcsr**&, const dtype*, const dtype*, dtype*,
                           double, double) {
  bound("ApplyCoeffMatrixA_CSR");
}
void Multiply3DSparseUsingFT(size_m, size_m, size_m, int*, int*, size_t*, zcsr**&, const dtype*, dtype*, double) {
  bound("Multiply3DSparseUsingFT");
}
dtype alpha(size_m, double) { bound("alpha"); }
dtype beta3D(size_m, size_m, size_m, int, int, int, int) { bound("beta3D"); }
dtype beta2D_pml(size_m, size_m, int, dtype, int, int) { bound("beta2D_pml"); }
void GenSparseMatrixOnline2DwithPMLFast(int, size_m, size_m, zcsr*, dtype, int*) {
  bound("GenSparseMatrixOnline2DwithPMLFast");
}
void SetSoundSpeed3D(size_m, size_m, size_m, dtype*, point) { bound("SetSoundSpeed3D"); }
void SetSoundSpeed3DFromFile(size_m, size_m, size_m, dtype*, point) { bound("SetSoundSpeed3DFromFile"); }
void SetSoundSpeed2D(size_m, size_m, size_m, dtype*, dtype*, point) { bound("SetSoundSpeed2D"); }
void GenerateDeltaL(size_m, size_m, size_m, dtype*, dtype*, dtype*, double) { bound("GenerateDeltaL"); }
void HeteroSoundSpeed3DExtensionToPML(size_m, size_m, size_m, dtype*) { bound("HeteroSoundSpeed3DExtensionToPML"); }
void HeteroSoundSpeed2DExtensionToPML(size_m, size_m, dtype*) { bound("HeteroSoundSpeed2DExtensionToPML"); }
dtype beta2D_pml_9pts(size_m, size_m, int, dtype, int, int, double) { bound("beta2D_pml_9pts"); }
dtype beta2D_pml_9pts_opt(size_m, size_m, int, dtype, int, int, double, double, double) {
  bound("beta2D_pml_9pts_opt");
}
void GenSparseMatrixOnline2DwithPMLand9Pts(int, size_m, size_m, zcsr*, dtype, int*) {
  bound("GenSparseMatrixOnline2DwithPMLand9Pts");
}
void GenSparseMatrixOnline2DwithPMLand9Points(int, size_m, size_m, size_m, zcsr*, dtype, int*, double) {
  bound("GenSparseMatrixOnline2DwithPMLand9Points");
}
