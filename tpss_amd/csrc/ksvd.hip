// ksvd.hip -- batched Kronecker SVD in FP64 on the device (compute_ksvd,
// kroneckersvd.h:215-296): the best Kronecker-rank-r approximation
//   A ~ sum_i B1_i (x) B0_i,  A of size (m1 m0) x (n1 n0),
// from the leading singular triplets of the rearrangement
//   R(A)[i1*n1 + j1, i0*n0 + j0] = A[i1*m0 + i0, j1*n0 + j0]   (reshuffle_full, :92-117).
// The input is either A itself or a sum of Kronecker products
//   A = sum_{i < in_rank} Abig_i (x) Bsmall_i,  R(A) = sum_i vec(Abig_i) vec(Bsmall_i)^T
// (compute_ksvd for std::vector<std::array<Table, 2>>, :306-402, whose Lanczos
// products rankk_vector_multiplication never form A).
// Golub-Kahan-Lanczos bidiagonalisation of R(A) started from e_0 with full
// reorthogonalisation, r^2 + 10 steps (or until breakdown), then the SVD of
// the small bidiagonal matrix (one-sided Jacobi) and the factors scaled by
// sqrt(sigma_i) (:286-295).  Sign convention: the first non-negligible entry
// of B0_i is positive.  One block per matrix; Lanczos vectors in a global
// workspace.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "common.hpp"
#include "ops.hpp"

namespace tpss
{
namespace
{
constexpr int KS_T = 256;
constexpr int KS_MAXIT = 40;

__device__ double block_sum(double v, double *sh)
{
  sh[threadIdx.x] = v;
  __syncthreads();
  for(int w = KS_T / 2; w > 0; w >>= 1)
  {
    if(int(threadIdx.x) < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// element (row, col) of R(A)
__device__ __forceinline__ double rearranged(const double *A, int m1, int n1, int m0, int n0,
                                             int row, int col)
{
  const int i1 = row / n1, j1 = row % n1, i0 = col / n0, j0 = col % n0;
  return A[size_t(i1 * m0 + i0) * (n1 * n0) + j1 * n0 + j0];
}

constexpr int KS_MAXIN = 16;

// SUM: Aall = the big factors (batch x in_rank x Rn), Ball = the small ones
// (batch x in_rank x Cn); otherwise Aall = the dense matrices
template<bool SUM>
__global__ void __launch_bounds__(KS_T)
k_ksvd(const double *__restrict__ Aall, const double *__restrict__ Ball, const int in_rank,
       const int m1, const int n1, const int m0, const int n0, const int rank, const int iters,
       double *__restrict__ work, double *__restrict__ B1, double *__restrict__ B0,
       double *__restrict__ sigma)
{
  __shared__ double sh[KS_T];
  __shared__ double alpha[KS_MAXIT + 1], beta[KS_MAXIT + 2];
  __shared__ double dots[KS_MAXIN];
  __shared__ int steps;
  const int Rn = m1 * n1, Cn = m0 * n0;
  const size_t b = blockIdx.x;
  const double *A = SUM ? Aall + b * size_t(in_rank) * Rn : Aall + b * size_t(Rn) * Cn;
  const double *Bsm = SUM ? Ball + b * size_t(in_rank) * Cn : nullptr;
  // workspace per matrix: U[(iters+1) x Rn] (u_0 = 0), V[iters x Cn], P[Cn], Q[Rn]
  double *U = work + b * (size_t(iters + 1) * Rn + size_t(iters) * Cn + Cn + Rn);
  double *V = U + size_t(iters + 1) * Rn;
  double *P = V + size_t(iters) * Cn;
  double *Q = P + Cn;
  for(int i = threadIdx.x; i < Rn; i += KS_T) U[i] = 0.0;
  for(int i = threadIdx.x; i < Cn; i += KS_T) P[i] = i == 0 ? 1.0 : 0.0;
  if(threadIdx.x == 0)
  {
    beta[0] = 1.0;
    steps = 0;
  }
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  int it = 0;
  for(; it < iters; ++it)
  {
    if(!(fabs(beta[it]) > eps)) break;
    // beta_{it+1} = ||p||, v_it = p / beta
    double s = 0.0;
    for(int i = threadIdx.x; i < Cn; i += KS_T) s += P[i] * P[i];
    const double bn = sqrt(block_sum(s, sh));
    if(threadIdx.x == 0) beta[it + 1] = bn;
    double *v = V + size_t(it) * Cn;
    for(int i = threadIdx.x; i < Cn; i += KS_T) v[i] = bn > 0.0 ? P[i] / bn : 0.0;
    __syncthreads();
    // r = R v - beta u_it, orthogonalised against the previous u's
    if constexpr(SUM)
    {
      for(int i = 0; i < in_rank; ++i)
      {
        double d = 0.0;
        for(int col = threadIdx.x; col < Cn; col += KS_T) d += Bsm[size_t(i) * Cn + col] * v[col];
        d = block_sum(d, sh);
        if(threadIdx.x == 0) dots[i] = d;
      }
      __syncthreads();
      for(int row = threadIdx.x; row < Rn; row += KS_T)
      {
        double acc = 0.0;
        for(int i = 0; i < in_rank; ++i) acc += A[size_t(i) * Rn + row] * dots[i];
        Q[row] = acc - bn * U[size_t(it) * Rn + row];
      }
    }
    else
      for(int row = threadIdx.x; row < Rn; row += KS_T)
      {
        double acc = 0.0;
        for(int col = 0; col < Cn; ++col) acc += rearranged(A, m1, n1, m0, n0, row, col) * v[col];
        Q[row] = acc - bn * U[size_t(it) * Rn + row];
      }
    __syncthreads();
    for(int j = 1; j <= it; ++j)
    {
      const double *uj = U + size_t(j) * Rn;
      double d = 0.0;
      for(int i = threadIdx.x; i < Rn; i += KS_T) d += Q[i] * uj[i];
      d = block_sum(d, sh);
      for(int i = threadIdx.x; i < Rn; i += KS_T) Q[i] -= d * uj[i];
      __syncthreads();
    }
    s = 0.0;
    for(int i = threadIdx.x; i < Rn; i += KS_T) s += Q[i] * Q[i];
    const double an = sqrt(block_sum(s, sh));
    if(threadIdx.x == 0) alpha[it] = an;
    double *un = U + size_t(it + 1) * Rn;
    for(int i = threadIdx.x; i < Rn; i += KS_T) un[i] = an > 0.0 ? Q[i] / an : 0.0;
    __syncthreads();
    // p = R^T u - alpha v, orthogonalised against the v's
    if constexpr(SUM)
    {
      for(int i = 0; i < in_rank; ++i)
      {
        double d = 0.0;
        for(int row = threadIdx.x; row < Rn; row += KS_T) d += A[size_t(i) * Rn + row] * un[row];
        d = block_sum(d, sh);
        if(threadIdx.x == 0) dots[i] = d;
      }
      __syncthreads();
      for(int col = threadIdx.x; col < Cn; col += KS_T)
      {
        double acc = 0.0;
        for(int i = 0; i < in_rank; ++i) acc += Bsm[size_t(i) * Cn + col] * dots[i];
        P[col] = acc - an * v[col];
      }
    }
    else
      for(int col = threadIdx.x; col < Cn; col += KS_T)
      {
        double acc = 0.0;
        for(int row = 0; row < Rn; ++row) acc += rearranged(A, m1, n1, m0, n0, row, col) * un[row];
        P[col] = acc - an * v[col];
      }
    __syncthreads();
    for(int j = 0; j <= it; ++j)
    {
      const double *vj = V + size_t(j) * Cn;
      double d = 0.0;
      for(int i = threadIdx.x; i < Cn; i += KS_T) d += P[i] * vj[i];
      d = block_sum(d, sh);
      for(int i = threadIdx.x; i < Cn; i += KS_T) P[i] -= d * vj[i];
      __syncthreads();
    }
    if(threadIdx.x == 0) steps = it + 1;
    __syncthreads();
    if(!(an > eps)) break;
  }
  // small bidiagonal matrix Bd (steps x steps): diag alpha, super diag beta[2..]
  __shared__ double W[KS_MAXIT * KS_MAXIT], Vs[KS_MAXIT * KS_MAXIT], sv[KS_MAXIT];
  __shared__ int order[KS_MAXIT];
  const int k = steps;
  if(threadIdx.x == 0)
  {
    for(int i = 0; i < k * k; ++i)
    {
      W[i] = 0.0;
      Vs[i] = 0.0;
    }
    for(int i = 0; i < k; ++i)
    {
      W[i * k + i] = alpha[i];
      if(i + 1 < k) W[i * k + i + 1] = beta[i + 2];
      Vs[i * k + i] = 1.0;
    }
    // one-sided Jacobi on the columns of W
    for(int sweep = 0; sweep < 60; ++sweep)
    {
      double off = 0.0;
      for(int p = 0; p < k - 1; ++p)
        for(int q = p + 1; q < k; ++q)
        {
          double a = 0, bb = 0, g = 0;
          for(int i = 0; i < k; ++i)
          {
            a += W[i * k + p] * W[i * k + p];
            bb += W[i * k + q] * W[i * k + q];
            g += W[i * k + p] * W[i * k + q];
          }
          if(fabs(g) <= 1e-300 || fabs(g) <= 1e-15 * sqrt(a * bb)) continue;
          off = fmax(off, fabs(g) / sqrt(a * bb));
          const double zeta = (bb - a) / (2.0 * g);
          const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
          for(int i = 0; i < k; ++i)
          {
            const double wp = W[i * k + p], wq = W[i * k + q];
            W[i * k + p] = cs * wp - sn * wq;
            W[i * k + q] = sn * wp + cs * wq;
            const double vp = Vs[i * k + p], vq = Vs[i * k + q];
            Vs[i * k + p] = cs * vp - sn * vq;
            Vs[i * k + q] = sn * vp + cs * vq;
          }
        }
      if(off < 1e-15) break;
    }
    for(int j = 0; j < k; ++j)
    {
      double s2 = 0.0;
      for(int i = 0; i < k; ++i) s2 += W[i * k + j] * W[i * k + j];
      sv[j] = sqrt(s2);
      order[j] = j;
    }
    for(int i = 1; i < k; ++i)
    {
      const int t = order[i];
      int j = i - 1;
      while(j >= 0 && sv[order[j]] < sv[t])
      {
        order[j + 1] = order[j];
        --j;
      }
      order[j + 1] = t;
    }
  }
  __syncthreads();
  // factors: left_i = U_k (W[:,j]/s_j), right_i = V_k Vs[:,j]
  for(int r = 0; r < rank; ++r)
  {
    const bool have = r < k;
    const int j = have ? order[r] : 0;
    const double sj = have ? sv[j] : 0.0;
    const double sq = sqrt(sj);
    double *b1 = B1 + (b * rank + r) * size_t(Rn);
    double *b0 = B0 + (b * rank + r) * size_t(Cn);
    for(int row = threadIdx.x; row < Rn; row += KS_T)
    {
      double acc = 0.0;
      if(have && sj > 0.0)
        for(int i = 0; i < k; ++i) acc += U[size_t(i + 1) * Rn + row] * W[i * k + j];
      b1[row] = have && sj > 0.0 ? acc / sj * sq : 0.0;
    }
    for(int col = threadIdx.x; col < Cn; col += KS_T)
    {
      double acc = 0.0;
      if(have)
        for(int i = 0; i < k; ++i) acc += V[size_t(i) * Cn + col] * Vs[i * k + j];
      b0[col] = acc * sq;
    }
    __syncthreads();
    // sign: first non-negligible entry of B0 positive
    double mx = 0.0;
    for(int col = threadIdx.x; col < Cn; col += KS_T) mx = fmax(mx, fabs(b0[col]));
    sh[threadIdx.x] = mx;
    __syncthreads();
    for(int w = KS_T / 2; w > 0; w >>= 1)
    {
      if(int(threadIdx.x) < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
      __syncthreads();
    }
    const double bmax = sh[0];
    __syncthreads();
    if(threadIdx.x == 0)
    {
      int f = 0;
      while(f < Cn - 1 && fabs(b0[f]) <= 1e-12 * bmax) ++f;
      sh[0] = b0[f] < 0.0 ? -1.0 : 1.0;
      sigma[b * rank + r] = sj;
    }
    __syncthreads();
    const double sgn = sh[0];
    for(int row = threadIdx.x; row < Rn; row += KS_T) b1[row] *= sgn;
    for(int col = threadIdx.x; col < Cn; col += KS_T) b0[col] *= sgn;
    __syncthreads();
  }
}
}  // namespace
}  // namespace tpss

using namespace tpss;

extern "C" {

int tpss_ksvd_batched(const double *A, int batch, int m1, int n1, int m0, int n0, int rank,
                      double *B1, double *B0, double *sigma, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(A && B1 && B0 && sigma, "null argument");
    TPSS_REQUIRE(batch >= 0 && m1 > 0 && n1 > 0 && m0 > 0 && n0 > 0 && rank >= 1,
                 "invalid dimensions");
    const int iters = std::min(KS_MAXIT, std::min(rank * rank + 10, std::min(m1 * n1, m0 * n0)));
    TPSS_REQUIRE(rank <= iters, "rank exceeds the Lanczos space");
    if(batch == 0) return;
    const size_t Rn = size_t(m1) * n1, Cn = size_t(m0) * n0;
    const size_t per = (iters + 1) * Rn + iters * Cn + Cn + Rn;
    DeviceArray<double> work{per * size_t(batch)};
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_ksvd<false>, dim3(unsigned(batch)), dim3(KS_T), 0, s, A, nullptr, 0, m1,
                       n1, m0, n0, rank, iters, work.ptr, B1, B0, sigma);
    TPSS_HIP(hipGetLastError());
    TPSS_HIP(hipStreamSynchronize(s));
  });
}

int tpss_ksvd_sum_batched(const double *Abig, const double *Bsmall, int batch, int in_rank, int m1,
                          int n1, int m0, int n0, int rank, double *B1, double *B0, double *sigma,
                          void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(Abig && Bsmall && B1 && B0 && sigma, "null argument");
    TPSS_REQUIRE(batch >= 0 && m1 > 0 && n1 > 0 && m0 > 0 && n0 > 0 && rank >= 1,
                 "invalid dimensions");
    TPSS_REQUIRE(in_rank >= 1 && in_rank <= KS_MAXIN, "in_rank must be in 1..16");
    const int iters = std::min(KS_MAXIT, std::min(rank * rank + 10, std::min(m1 * n1, m0 * n0)));
    TPSS_REQUIRE(rank <= iters, "rank exceeds the Lanczos space");
    if(batch == 0) return;
    const size_t Rn = size_t(m1) * n1, Cn = size_t(m0) * n0;
    const size_t per = (iters + 1) * Rn + iters * Cn + Cn + Rn;
    DeviceArray<double> work{per * size_t(batch)};
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_ksvd<true>, dim3(unsigned(batch)), dim3(KS_T), 0, s, Abig, Bsmall,
                       in_rank, m1, n1, m0, n0, rank, iters, work.ptr, B1, B0, sigma);
    TPSS_HIP(hipGetLastError());
    TPSS_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
