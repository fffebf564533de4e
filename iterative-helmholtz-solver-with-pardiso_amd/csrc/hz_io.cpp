// Velocity-model files and the reference's solution / convergence text files
// (host memory; no device work).
//
//   hz_velocity_from_file   SetSoundSpeed3DFromFile (functions.cpp:1458-1531)
//                           + HeteroSoundSpeed3DExtensionToPML (functions.cpp:2675-2716)
//   hz_write_solution       sol3D_N%d.dat writer (main.cpp:827-834) over the physical box
//                           in reducePML3D order (functions.cpp:238-263)
//   hz_read_solution        the same format read back (Beta3D / Runge inputs, functions.cpp:932-1239)
//   hz_write_convergence    convergence_*.dat lines (FGMRES.cpp:50, 646)
//   hz_runge_ratio          Runge's ratio of three nested solution files (functions.cpp:932-1134)
//   hz_beta3d_norms         Beta3D's shell norms of two nested solution files
//                           (functions.cpp:1136-1229, 629-674, 1290-1313)
//   hz_write_beta_norms     beta_3D_norm2_*.dat lines (functions.cpp:1295-1302)
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hz.h"
#include "hz_internal.h"

namespace {

// layer nodes per side of axis a: the reference's physical box excludes
// pml_pts in x/y and spg_pts in z (reducePML3D); the descriptor carries both
int64_t layers(const hz_desc* d, int a) { return (int64_t)d->pml[a] + d->spg[a]; }

void check_desc(const hz_desc* d) {
  HZ_CHECK(d, HZ_ERR_ARG, "null descriptor");
  for (int a = 0; a < 3; ++a) {
    HZ_CHECK(d->n[a] >= 1, HZ_ERR_ARG, "axis %d: n=%lld", a, (long long)d->n[a]);
    HZ_CHECK(2 * layers(d, a) < d->n[a], HZ_ERR_ARG, "axis %d: layers (%d + %d per side) leave no physical nodes", a,
             d->pml[a], d->spg[a]);
  }
}

struct File {
  FILE* f = nullptr;
  File(const char* path, const char* mode) {
    HZ_CHECK(path, HZ_ERR_ARG, "null path");
    f = std::fopen(path, mode);
    HZ_CHECK(f, HZ_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
  }
  ~File() {
    if (f) std::fclose(f);
  }
};

// a sol3D_N%d.dat file of an n x n x n physical box (x fastest), exactly n^3 values
std::vector<hz_c128> read_box(const char* path, int64_t n) {
  HZ_CHECK(n >= 1, HZ_ERR_ARG, "box of %lld nodes per side", (long long)n);
  File f(path, "r");
  std::vector<hz_c128> u((size_t)(n * n * n));
  for (size_t l = 0; l < u.size(); ++l)
    HZ_CHECK(std::fscanf(f.f, "%lf %lf", &u[l].re, &u[l].im) == 2, HZ_ERR_ARG,
             "%s: %zu of %zu values (box of %lld nodes per side)", path, l, u.size(), (long long)n);
  double extra;
  HZ_CHECK(std::fscanf(f.f, "%lf", &extra) != 1, HZ_ERR_ARG, "%s: more values than a box of %lld nodes per side",
           path, (long long)n);
  return u;
}

// value at coarse node (i, j, k) of a box of nc nodes per side, taken from the
// box of nf = s (nc + 1) - 1 nodes per side: fine node s (i + 1) - 1 is at the
// same position (functions.cpp:1102-1112)
inline const hz_c128& restrict_at(const std::vector<hz_c128>& uf, int64_t nf, int64_t s, int64_t i, int64_t j,
                                  int64_t k) {
  return uf[(size_t)((s * (i + 1) - 1) + nf * (s * (j + 1) - 1) + nf * nf * (s * (k + 1) - 1))];
}

int64_t step_between(int64_t nc, int64_t nf) {
  HZ_CHECK(nc >= 1 && nf >= nc && (nf + 1) % (nc + 1) == 0, HZ_ERR_ARG,
           "a box of %lld nodes per side is not nested in one of %lld ((n + 1) must divide)", (long long)nc,
           (long long)nf);
  return (nf + 1) / (nc + 1);
}

}  // namespace

#define HZ_API_BEGIN try {
#define HZ_API_END                                    \
  }                                                   \
  catch (const hz::Error& e) {                        \
    return e.code;                                    \
  }                                                   \
  catch (const std::exception& e) {                   \
    hz::set_error("%s", e.what());                    \
    return HZ_ERR_ARG;                                \
  }                                                   \
  return HZ_OK;

extern "C" {

int hz_velocity_file_name(double h, char* buf, int len) {
  HZ_API_BEGIN
  HZ_CHECK(buf && len > 0, HZ_ERR_ARG, "null buffer");
  const int w = std::snprintf(buf, (size_t)len, "OT_small_at_%dm.bin", (int)h);
  HZ_CHECK(w > 0 && w < len, HZ_ERR_ARG, "buffer of %d bytes too short", len);
  HZ_API_END
}

int hz_velocity_from_file(const char* path, const hz_desc* d, double* c) {
  HZ_API_BEGIN
  check_desc(d);
  HZ_CHECK(c, HZ_ERR_ARG, "null output");
  const int64_t nx = d->n[0], ny = d->n[1], nz = d->n[2];
  const int64_t lx = layers(d, 0), ly = layers(d, 1), lz = layers(d, 2);
  const int64_t n1 = nx - 2 * lx, n2 = ny - 2 * ly, n3 = nz - 2 * lz;  // physical x, y, z
  const size_t count = (size_t)(n1 * n2 * n3);
  std::vector<float> data(count);
  {
    File f(path, "rb");
    const size_t got = std::fread(data.data(), sizeof(float), count, f.f);
    HZ_CHECK(got == count, HZ_ERR_ARG, "%s: %zu of %zu float32 values (physical box %lld x %lld x %lld)", path, got,
             count, (long long)n1, (long long)n2, (long long)n3);
  }
  const int64_t nxy = nx * ny;
  // physical box: file order [y][x][z], z fastest
  for (int64_t j = 0; j < n2; ++j)
    for (int64_t i = 0; i < n1; ++i)
      for (int64_t k = 0; k < n3; ++k) {
        const float v = data[(size_t)((j * n1 + i) * n3 + k)];
        HZ_CHECK(std::isfinite(v) && v > 0.0f, HZ_ERR_ARG, "%s: speed %g at (x %lld, y %lld, z %lld)", path, (double)v,
                 (long long)i, (long long)j, (long long)k);
        c[(k + lz) * nxy + (j + ly) * nx + (i + lx)] = (double)v;
      }
  // z layers (sponge / PML): the nearest physical z node of the same column.
  // The reference leaves them at zero, which its delta-L cannot use.
  for (int64_t k = 0; k < nz; ++k) {
    if (k >= lz && k < nz - lz) continue;
    const int64_t ks = k < lz ? lz : nz - lz - 1;
    for (int64_t j = ly; j < ny - ly; ++j)
      for (int64_t i = lx; i < nx - lx; ++i) c[k * nxy + j * nx + i] = c[ks * nxy + j * nx + i];
  }
  // x layers copy the first / last physical x node of the row, then the y
  // layers copy the first / last physical row (corners included), every k
  // (the reference's x pass also visits the y-layer rows, which its y pass
  // then overwrites: the result is the same)
  for (int64_t k = 0; k < nz; ++k) {
    double* ck = c + k * nxy;
    for (int64_t j = ly; j < ny - ly; ++j) {
      for (int64_t i = 0; i < lx; ++i) ck[j * nx + i] = ck[j * nx + lx];
      for (int64_t i = nx - lx; i < nx; ++i) ck[j * nx + i] = ck[j * nx + nx - lx - 1];
    }
    for (int64_t i = 0; i < nx; ++i) {
      for (int64_t j = 0; j < ly; ++j) ck[j * nx + i] = ck[ly * nx + i];
      for (int64_t j = ny - ly; j < ny; ++j) ck[j * nx + i] = ck[(ny - ly - 1) * nx + i];
    }
  }
  HZ_API_END
}

int hz_write_solution(const char* path, const hz_desc* d, const hz_c128* u) {
  HZ_API_BEGIN
  check_desc(d);
  HZ_CHECK(u, HZ_ERR_ARG, "null solution");
  File f(path, "w");
  const int64_t nx = d->n[0], ny = d->n[1], nz = d->n[2];
  const int64_t lx = layers(d, 0), ly = layers(d, 1), lz = layers(d, 2);
  for (int64_t k = lz; k < nz - lz; ++k)
    for (int64_t j = ly; j < ny - ly; ++j)
      for (int64_t i = lx; i < nx - lx; ++i) {
        const hz_c128 v = u[(k * ny + j) * nx + i];
        HZ_CHECK(std::fprintf(f.f, "%18.16lf %18.16lf\n", v.re, v.im) > 0, HZ_ERR_ARG, "%s: write failed", path);
      }
  HZ_API_END
}

int hz_read_solution(const char* path, const hz_desc* d, hz_c128* u) {
  HZ_API_BEGIN
  check_desc(d);
  HZ_CHECK(u, HZ_ERR_ARG, "null output");
  File f(path, "r");
  const int64_t nx = d->n[0], ny = d->n[1], nz = d->n[2];
  const int64_t lx = layers(d, 0), ly = layers(d, 1), lz = layers(d, 2);
  std::memset(u, 0, sizeof(hz_c128) * (size_t)(nx * ny * nz));
  for (int64_t k = lz; k < nz - lz; ++k)
    for (int64_t j = ly; j < ny - ly; ++j)
      for (int64_t i = lx; i < nx - lx; ++i) {
        hz_c128 v;
        HZ_CHECK(std::fscanf(f.f, "%lf %lf", &v.re, &v.im) == 2, HZ_ERR_ARG,
                 "%s: physical node (%lld, %lld, %lld) missing", path, (long long)(i - lx), (long long)(j - ly),
                 (long long)(k - lz));
        u[(k * ny + j) * nx + i] = v;
      }
  double extra;
  HZ_CHECK(std::fscanf(f.f, "%lf", &extra) != 1, HZ_ERR_ARG, "%s: more values than physical nodes", path);
  HZ_API_END
}

int hz_write_convergence(const char* path, int count, const int* iter, const double* res, const double* relres,
                         const double* diff) {
  HZ_API_BEGIN
  HZ_CHECK(count >= 0 && (count == 0 || (iter && res && relres && diff)), HZ_ERR_ARG, "bad convergence arrays");
  File f(path, "w");
  for (int q = 0; q < count; ++q)
    HZ_CHECK(std::fprintf(f.f, "%d %e %e %lf\n", iter[q], res[q], relres[q], diff[q]) > 0, HZ_ERR_ARG,
             "%s: write failed", path);
  HZ_API_END
}

int hz_convergence_file_name(const hz_desc* d, double length_x, double nu, double beta, char* buf, int len) {
  HZ_API_BEGIN
  check_desc(d);
  HZ_CHECK(buf && len > 0, HZ_ERR_ARG, "null buffer");
  const int64_t nxp = d->n[0] - 2 * layers(d, 0);
  const int w = std::snprintf(buf, (size_t)len, "convergence_N%d_PML%d_Lx%d_FREQ%d_SPG%6.lf_BETA%5.3lf_FGMRES.dat",
                              (int)nxp, d->pml[0], (int)length_x, (int)nu, d->h[2] * 2 * d->spg[2], beta);
  HZ_CHECK(w > 0 && w < len, HZ_ERR_ARG, "buffer of %d bytes too short", len);
  HZ_API_END
}

int hz_runge_ratio(const char* coarse, const char* mid, const char* fine, int n_coarse, int n_mid, int n_fine,
                   double h, double* out) {
  HZ_API_BEGIN
  HZ_CHECK(out, HZ_ERR_ARG, "null output");
  HZ_CHECK(std::isfinite(h) && h > 0.0, HZ_ERR_ARG, "h = %g", h);
  const int64_t s1 = step_between(n_coarse, n_fine), s2 = step_between(n_mid, n_fine);
  const std::vector<hz_c128> u1 = read_box(coarse, n_coarse), u2 = read_box(mid, n_mid), u3 = read_box(fine, n_fine);
  // h * sum over the coarser box of |u_fine - u| (IntegralNorm3D, functions.cpp:1348-1360)
  auto dist = [&](const std::vector<hz_c128>& u, int64_t n, int64_t s) {
    double sum = 0.0;
    for (int64_t k = 0; k < n; ++k)
      for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
          const hz_c128& f = restrict_at(u3, n_fine, s, i, j, k);
          const hz_c128& c = u[(size_t)(i + n * j + n * n * k)];
          sum += std::hypot(f.re - c.re, f.im - c.im);
        }
    return sum * h;
  };
  out[0] = dist(u1, n_coarse, s1);
  out[1] = dist(u2, n_mid, s2);
  out[2] = out[0] / out[1];
  HZ_API_END
}

int hz_beta3d_norms(const char* coarse, const char* fine, int grid, double h, const double* source, int count,
                    double* radius, double* norm) {
  HZ_API_BEGIN
  HZ_CHECK(grid >= 4 && count == grid - 3, HZ_ERR_ARG, "grid %d: %d shells requested, the study has grid - 3", grid,
           count);
  HZ_CHECK(source && radius && norm, HZ_ERR_ARG, "null argument");
  HZ_CHECK(std::isfinite(h) && h > 0.0, HZ_ERR_ARG, "h = %g", h);
  const int64_t nc = grid - 1, nf = 2 * (int64_t)grid - 1;
  const std::vector<hz_c128> uc = read_box(coarse, nc), uf = read_box(fine, nf);
  // Shell e (2 <= e <= grid - 2) holds the coarse nodes with 2h <= r <= e h
  // (check_norm_circle_3D_nopml, functions.cpp:654-666), so each node is binned
  // at the first shell that holds it and the shells are prefix sums.  The
  // comparisons are the reference's own double expressions.
  const double r0 = 2 * h;
  std::vector<double> dsum((size_t)grid, 0.0), osum((size_t)grid, 0.0);
  for (int64_t k = 0; k < nc; ++k)
    for (int64_t j = 0; j < nc; ++j)
      for (int64_t i = 0; i < nc; ++i) {
        const double x = (i + 1) * h - source[0];
        const double y = (j + 1) * h - source[1];
        const double z = (k + 1) * h - source[2];
        const double r = std::sqrt(x * x + y * y + z * z);
        if (!(r >= r0)) continue;
        int64_t e = std::max<int64_t>(2, (int64_t)std::ceil(r / h));
        while (e > 2 && r <= (e - 1) * h) --e;
        while (e <= grid - 2 && !(r <= e * h)) ++e;
        if (e > grid - 2) continue;
        const hz_c128& f = restrict_at(uf, nf, 2, i, j, k);
        const hz_c128& c = uc[(size_t)(i + nc * j + nc * nc * k)];
        const double dr = c.re - f.re, di = c.im - f.im;
        dsum[(size_t)e] += dr * dr + di * di;
        osum[(size_t)e] += f.re * f.re + f.im * f.im;
      }
  // ||u_coarse - u_fine|| / ||u_fine|| over each shell (RelError, Frobenius)
  double dc = 0.0, oc = 0.0;
  for (int e = 2; e <= grid - 2; ++e) {
    dc += dsum[(size_t)e];
    oc += osum[(size_t)e];
    radius[e - 2] = e * h;
    norm[e - 2] = std::sqrt(dc) / std::sqrt(oc);
  }
  HZ_API_END
}

int hz_write_beta_norms(const char* path, int count, const double* radius, const double* norm) {
  HZ_API_BEGIN
  HZ_CHECK(count >= 0 && (count == 0 || (radius && norm)), HZ_ERR_ARG, "bad norm arrays");
  File f(path, "w");
  for (int q = 0; q < count; ++q)
    HZ_CHECK(std::fprintf(f.f, "%lf %lf\n", radius[q], norm[q]) > 0, HZ_ERR_ARG, "%s: write failed", path);
  HZ_API_END
}

int hz_beta_norm_file_name(int n, int spg, char* buf, int len) {
  HZ_API_BEGIN
  HZ_CHECK(buf && len > 0, HZ_ERR_ARG, "null buffer");
  const int w = std::snprintf(buf, (size_t)len, "beta_3D_norm2_Nx%dNy%d_SPG%d.dat", n, n, spg);
  HZ_CHECK(w > 0 && w < len, HZ_ERR_ARG, "buffer of %d bytes too short", len);
  HZ_API_END
}

}  // extern "C"
