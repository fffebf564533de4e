// hz_noalloc.hip -- the preconditioner apply for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops that forbid
// hipMalloc).  The z-transformed planes live in a fixed device array of this
// library instead of the problem's grown work buffer.
//
// Every device holds its own copy of the array from module load to process
// end; its address on each device comes from hipGetSymbolAddress, taken once
// per device on the first call with that device current.  All calls on one
// device share the array: calls that overlap on one device (different
// streams, or several host threads) race on it.
#include <mutex>

#include "hz_internal.h"

namespace hz {
namespace {

constexpr size_t SCRATCH_BYTES = size_t(64) << 20;  // 64 MiB: 8 Mi complex64 or 4 Mi complex128 plane values
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};

// device address of `scratch` on `dev` (made current for the lookup)
void* scratch_on(int dev) {
  HZ_CHECK(dev >= 0 && dev < MAX_DEVICES, HZ_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    int cur = -1;
    HZ_HIP(hipGetDevice(&cur));
    if (cur != dev) HZ_HIP(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    HZ_HIP(e);
    HZ_CHECK(ptr != nullptr, HZ_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch[dev] = ptr;
  }
  return g_scratch[dev];
}

}  // namespace
}  // namespace hz

using namespace hz;

extern "C" int hz_precond_apply_noalloc(hz_problem* p, int prec, int nrhs, const hz_c128* in, hz_c128* out) {
  try {
    HZ_CHECK(p && in && out && nrhs >= 1, HZ_ERR_ARG, "bad arguments");
    HZ_CHECK(prec == HZ_PREC_FP64 || prec == HZ_PREC_C64, HZ_ERR_ARG, "bad precision %d", prec);
    const size_t elem = prec == HZ_PREC_FP64 ? sizeof(double2) : sizeof(float2);
    const size_t need = (size_t)p->N * (size_t)nrhs * elem;
    HZ_CHECK(need <= SCRATCH_BYTES, HZ_ERR_ARG,
             "hz_precond_apply_noalloc: %d right-hand sides of %lld points need %zu bytes of plane storage, "
             "more than its fixed %zu-byte array",
             nrhs, (long long)p->N, need, SCRATCH_BYTES);
    void* work = scratch_on(p->device);
    precond_apply_ws(p, prec, nrhs, (const double2*)in, (double2*)out, false, false, work);
  } catch (const Error& e) {
    return e.code;
  }
  return HZ_OK;
}
