// hz_tune.cpp -- the autotuner: times the preconditioner's launch variants on
// a problem's own factors and keeps the fastest per (precision, right-hand
// sides).  Variants: the plane-solve kernel family (streaming multi-RHS or
// chunked), the streaming kernel's right-hand-side group cap (one pass over the
// factors serves up to 8 vectors; smaller groups put more workgroups on the
// chip when nz is below the CU count) and the z-FFT lines per workgroup.
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <mutex>
#include <sstream>
#include <string>

#include "hz_internal.h"

namespace hz {
namespace {

// average milliseconds of `reps` runs of f on the problem's stream (after one
// untimed run)
double time_ms(hz_problem* p, const std::function<void()>& f, int reps) {
  f();
  hipEvent_t e0, e1;
  HZ_HIP(hipEventCreate(&e0));
  HZ_HIP(hipEventCreate(&e1));
  HZ_HIP(hipEventRecord(e0, p->stream));
  for (int r = 0; r < reps; ++r) f();
  HZ_HIP(hipEventRecord(e1, p->stream));
  HZ_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  HZ_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

std::string cache_key(const hz_problem* p, int prec, int nrhs) {
  hipDeviceProp_t prop;
  std::string arch = "gpu";
  if (hipGetDeviceProperties(&prop, p->device) == hipSuccess) arch = prop.gcnArchName;
  const size_t colon = arch.find(':');
  if (colon != std::string::npos) arch = arch.substr(0, colon);
  std::ostringstream k;
  k << arch << " " << p->nx << " " << p->ny << " " << p->nz << " " << prec << " " << nrhs;
  return k.str();
}

std::mutex g_cache_mu;

bool cache_read(const std::string& path, const std::string& key, Tuning& t) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  std::ifstream in(path);
  std::string line;
  bool found = false;
  while (std::getline(in, line)) {
    // "<arch> nx ny nz prec nrhs | family cap lines"
    const size_t bar = line.find('|');
    if (bar == std::string::npos) continue;
    std::string k = line.substr(0, bar);
    while (!k.empty() && k.back() == ' ') k.pop_back();
    if (k != key) continue;
    std::istringstream v(line.substr(bar + 1));
    Tuning c;
    if (v >> c.solve_family >> c.group_cap >> c.fft_lines) {
      t = c;
      found = true;  // the last entry for a key wins
    }
  }
  return found;
}

void cache_append(const std::string& path, const std::string& key, const Tuning& t) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  std::ofstream out(path, std::ios::app);
  if (out) out << key << " | " << t.solve_family << " " << t.group_cap << " " << t.fft_lines << "\n";
}

}  // namespace

void autotune(hz_problem* p, int prec, int nrhs, hz_tune_info* info) {
  HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
  HZ_CHECK(nrhs >= 1, HZ_ERR_ARG, "bad right-hand-side count %d", nrhs);
  HZ_CHECK(p->fac[prec].ready, HZ_ERR_STATE, "hz_precond_factor(prec=%d) has not run", prec);
  ensure_dense(p, prec, p->stream);  // every candidate timed on complete factors
  const char* cpath = getenv("HZ_TUNE_CACHE");
  const std::string key = cache_key(p, prec, nrhs);
  Tuning best;
  bool from_cache = cpath && cpath[0] && cache_read(cpath, key, best);

  // work vectors: complex128 input of the forward FFT and the plane-solve buffer
  const size_t vb = (size_t)p->N * nrhs * sizeof(double2);
  if (p->w_a.bytes < vb) p->w_a.alloc(vb);
  if (p->w_b.bytes < vb) p->w_b.alloc(vb);
  launch_fill((int64_t)p->N * nrhs, make_double2(1.0, 0.0), p->w_b.as<double2>(), p->stream);
  launch_fill((int64_t)p->N * nrhs, make_double2(0.0, 0.0), p->w_a.as<double2>(), p->stream);
  const int reps = 3;
  double best_solve = 0, best_fft = 0;

  if (!from_cache) {
    // plane solve: streaming kernel with each group cap, chunked kernel
    std::vector<Tuning> cand;
    if (stream_solve_available(prec, (int)p->nx)) {
      for (int cap : {8, 4, 2, 1}) {
        if (prec == HZ_PREC_FP64 && cap > 4) continue;
        if (cap > 1 && cap > nrhs) continue;
        Tuning t;
        t.solve_family = HZ_SOLVE_STREAM;
        t.group_cap = cap;
        cand.push_back(t);
      }
    }
    if (nrhs <= 4 && p->fac[prec].sym_ready) {
      Tuning t;
      t.solve_family = HZ_SOLVE_SYM;
      cand.push_back(t);
    }
    if (nrhs < 8 && twist_solve_available(p, prec)) {
      Tuning t;
      t.solve_family = HZ_SOLVE_TWIST;
      cand.push_back(t);
    }
    if (nrhs >= 8 && mfma_solve_available(prec, (int)p->nx)) {
      Tuning t;
      t.solve_family = HZ_SOLVE_STREAM_MFMA;
      cand.push_back(t);
    }
    if (nrhs == 1 && single_ring_available(prec, (int)p->nx)) {
      Tuning t;
      t.solve_family = HZ_SOLVE_STREAM_1;
      cand.push_back(t);
    }
    {
      Tuning t;
      t.solve_family = HZ_SOLVE_CHUNKED;
      cand.push_back(t);
    }
    best_solve = -1;
    for (const Tuning& t : cand) {
      const double ms = time_ms(p, [&] { launch_plane_solve_with(p, prec, nrhs, p->w_a.ptr, false, t, p->stream); },
                                reps);
      if (best_solve < 0 || ms < best_solve) {
        best_solve = ms;
        best.solve_family = t.solve_family;
        best.group_cap = t.group_cap;
      }
    }
    // forward z-FFT (complex128 in, the precision's plane layout out)
    best_fft = -1;
    for (int lines : {0, 8, 16, 32}) {
      if (!fft_lines_fit(prec, (int)p->nz, lines)) continue;
      const double ms = time_ms(p, [&] {
        launch_fft_z_lines(p, lines, prec, -1, nrhs, p->w_b.ptr, 0, p->w_a.ptr, prec == HZ_PREC_C64 ? 1 : 0, 0,
                           nullptr, nullptr, p->stream);
      }, reps);
      if (best_fft < 0 || ms < best_fft) {
        best_fft = ms;
        best.fft_lines = lines;
      }
    }
    if (cpath && cpath[0]) cache_append(cpath, key, best);
  } else {
    best_solve = time_ms(p, [&] { launch_plane_solve_with(p, prec, nrhs, p->w_a.ptr, false, best, p->stream); }, 1);
    best_fft = time_ms(p, [&] {
      launch_fft_z_lines(p, best.fft_lines, prec, -1, nrhs, p->w_b.ptr, 0, p->w_a.ptr, prec == HZ_PREC_C64 ? 1 : 0,
                         0, nullptr, nullptr, p->stream);
    }, 1);
  }
  p->tuned[{prec, nrhs}] = best;
  if (info) {
    info->solve_kernel = best.solve_family;
    info->group_cap = best.group_cap;
    info->fft_lines = best.fft_lines;
    info->from_cache = from_cache ? 1 : 0;
    info->solve_ms = best_solve;
    info->fft_ms = best_fft;
  }
}

}  // namespace hz
