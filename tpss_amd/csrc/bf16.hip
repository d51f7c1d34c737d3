// bf16.hip -- patch solves with BF16 MFMA contractions and FP32 accumulation.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "ops.hpp"
#include "smoother.hpp"

namespace tpss
{
template<typename VT>
void precond_vmult_bf16(tpss_smoother *, VT *, const VT *, hipStream_t, bool, double)
{
  throw Error(TPSS_E_UNSUPPORTED, "BF16 patch solves are not built yet");
}
template void precond_vmult_bf16<double>(tpss_smoother *, double *, const double *, hipStream_t,
                                         bool, double);
template void precond_vmult_bf16<float>(tpss_smoother *, float *, const float *, hipStream_t, bool,
                                        double);
}  // namespace tpss
