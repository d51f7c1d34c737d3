// ops.hpp -- host-side launch helpers shared by the translation units of libtpss_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace tpss
{
template<int V>
using ic = std::integral_constant<int, V>;

// calls f(ic<dim>, ic<np>) for the supported (dim, p+1) pairs
template<typename F>
void dispatch_dim_np(int dim, int np, F &&f)
{
  switch(dim * 10 + np)
  {
    case 22: f(ic<2>{}, ic<2>{}); return;
    case 23: f(ic<2>{}, ic<3>{}); return;
    case 24: f(ic<2>{}, ic<4>{}); return;
    case 25: f(ic<2>{}, ic<5>{}); return;
    case 32: f(ic<3>{}, ic<2>{}); return;
    case 33: f(ic<3>{}, ic<3>{}); return;
    case 34: f(ic<3>{}, ic<4>{}); return;
    case 35: f(ic<3>{}, ic<5>{}); return;
    default: throw Error(TPSS_E_UNSUPPORTED, "unsupported (dim, degree)");
  }
}

inline unsigned blocks_for(int64_t n, int threads) { return unsigned((n + threads - 1) / threads); }

void operator_apply_any(const tpss_level *lv, void *y, const void *x, const void *b,
                        const uint32_t *cells, int64_t n_list, hipStream_t s);
// y = A x in FP64 whatever the level's vector precision (setup of fitted local solvers)
void operator_apply_f64(const tpss_level *lv, double *y, const double *x, const uint32_t *cells,
                        int64_t n_list, hipStream_t s);
// y = A x and the per-block partial sums of x.y in partial[0, nb) (at most cap
// of them) on FP64 3D Q3 Cartesian levels: returns nb, or 0 (nothing launched)
// on other levels
int operator_apply_dot_f64(const tpss_level *lv, double *y, const double *x, double *partial,
                           int64_t cap, hipStream_t s);
void fill_zero(void *x, int prec, int64_t n, hipStream_t s);
void axpby(void *y, const void *x, int prec, int64_t n, double a, double b, hipStream_t s);
void convert(void *dst, int dp, const void *src, int sp, int64_t n, hipStream_t s);
size_t elem_size(int prec);

// deterministic FP64 dot product into a device scalar
struct DotWorkspace
{
  DeviceArray<double> partial;
  static constexpr int NB = 1024;
  DotWorkspace() : partial(NB) {}
  // blocks of 256 threads of a grid-stride pass over n entries
  static int blocks(int64_t n) { return int(std::min<int64_t>(NB, std::max<int64_t>(1, (n + 255) / 256))); }
  template<typename T>
  void dot(const T *a, const T *b, int64_t n, double *out, hipStream_t s)
  {
    const int nb = blocks(n);
    hipLaunchKernelGGL((k_dot_partial<T>), dim3(nb), dim3(256), 0, s, a, b, n, partial.ptr);
    finish(nb, out, s);
  }
  // out = the sum of the first nb partial sums (written by a pass of nb blocks)
  void finish(int nb, double *out, hipStream_t s)
  {
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, partial.ptr, nb, out);
    TPSS_HIP(hipGetLastError());
  }
};

}  // namespace tpss
