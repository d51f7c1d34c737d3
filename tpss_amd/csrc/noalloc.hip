// noalloc.hip -- the allocation-free variant of tpss_precond_vmult.
//
// The preconditioner's additive slot buffer and the residual of the
// multiplicative sweeps that update it live in device memory that the
// smoother allocates on first use.  Callers that cannot allocate (code under
// HIP graph capture, serving loops, harnesses that forbid hipMalloc) use
// tpss_precond_vmult_noalloc instead: it takes that scratch from one
// file-scope device array of TPSS_NOALLOC_SCRATCH_BYTES, which every GPU the
// library runs on holds until the process ends, and calls no allocating
// routine.  Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.hpp"
#include "smoother.hpp"

namespace tpss
{
namespace
{
constexpr size_t kScratchBytes = TPSS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
void *scratch_ptr[kMaxDevices] = {};

// device address of the current device's scratch array, looked up once per
// device with that device current (each GPU has its own copy at its own address)
void *device_scratch()
{
  int dev = 0;
  TPSS_HIP(hipGetDevice(&dev));
  TPSS_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if(!scratch_ptr[dev])
  {
    void *p = nullptr;
    TPSS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TPSS_REQUIRE(p != nullptr, "no device address for the scratch array");
    scratch_ptr[dev] = p;
  }
  return scratch_ptr[dev];
}
}  // namespace
}  // namespace tpss

using namespace tpss;

extern "C" {

int tpss_noalloc_init(void)
{
  return guarded([&] { (void)device_scratch(); });
}

int tpss_precond_vmult_noalloc(tpss_smoother *sm, void *dst, const void *src, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && dst && src, "null argument");
    TPSS_REQUIRE(dst != src, "dst and src must not alias");
    const size_t need = precond_scratch_bytes(sm);
    if(need > kScratchBytes)
      throw Error(TPSS_E_INVALID_ARG,
                  "tpss_precond_vmult_noalloc: this smoother needs " + std::to_string(need) +
                    " bytes of scratch, the allocation-free entry point has " +
                    std::to_string(kScratchBytes));
    report_nonfinite(sm);
    char *ptr = need ? static_cast<char *>(device_scratch()) : nullptr;
    // the scratch is written before it is read in every pass (the slot buffer
    // by the patch kernel, the residual by the residual kernel of each color),
    // so the previous call's contents need no clearing; section timers (which
    // create events) are not recorded in this call
    const bool timing = sm->timing;
    sm->timing = false;
    sm->scratch = ptr;
    struct Restore
    {
      tpss_smoother *sm;
      bool timing;
      ~Restore()
      {
        sm->scratch = nullptr;
        sm->timing = timing;
      }
    } restore{sm, timing};
    PassCheck armed(sm);
    precond_vmult_any(sm, dst, src, as_stream(stream), false, 1.0);
    check_finite(sm, dst, as_stream(stream));
  });
}

}  // extern "C"
