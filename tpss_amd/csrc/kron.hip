// kron.hip -- generic Kronecker-sum matrices and the CP decomposition on the
// device, batched over independent matrices (FP64).
//
// * tpss_kron_vmult: A = sum_{r < rank} A_{r,order-1} (x) ... (x) A_{r,0}
//   (dimension 0 rightmost, vector index i = sum_d i_d prod_{e<d} size_e) with
//   rectangular, direction-dependent factor sizes m_d x n_d: vmult / Tvmult /
//   vmult_add of Tensors::TensorProductMatrix in the "basic" state
//   (tensor_product_matrix.h:136-304, tensor_product_matrix.templates.h:200-291),
//   as a sequence of mode-d products (internal::My::contract_general_impl with
//   collapsed n_pre / n_post, move_to_deal_ii.h:701-799).
// * tpss_kron_table: the dense matrix (as_table / as_transpose_table).
// * tpss_kron_pinv: the inverse of the basic state, InverseTable
//   (matrix_helper.h:21-68): LAPACKFullMatrix::compute_inverse_svd(threshold)
//   (singular values s_i with s_i / s_max < threshold dropped) or
//   compute_inverse_svd_with_kernel(kernel_size) (the kernel_size smallest
//   dropped); here by one-sided Jacobi SVD.
// * tpss_kcp_batched: compute_kcp (kroneckersvd.h:501-585), the rank-R CP
//   decomposition of a sum of 3-fold Kronecker products by alternating least
//   squares, initialised with the first R input terms, R^2 + 10 sweeps, each
//   mode solved with the pseudo-inverse of the Hadamard product of the other
//   modes' Gram matrices (alignedlinalg.h:214-237, 471-485).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "ops.hpp"

namespace tpss
{
namespace
{
constexpr int KR_T = 256;
constexpr int KR_MAX_ORDER = 3;
// largest intermediate tensor of one mode-product sequence held in LDS
constexpr int KR_MAX_SIZE = 4096;

struct KronShape
{
  int order, rank;
  int m[KR_MAX_ORDER], n[KR_MAX_ORDER];
  int fac_off[KR_MAX_ORDER];  // offset of factor d inside one rank-1 term
  int term;                   // values per rank-1 term
};

// y (sizes out) = mode-d product of x (sizes in) with the m_d x n_d matrix F
// (transposed: F^T, n_d x m_d); sizes of the other dimensions unchanged
__device__ void mode_product(const double *F, const int md, const int nd, const bool transpose,
                             const int d, const int *sz_in, const int order, const double *x,
                             double *y)
{
  int pre = 1, post = 1;
  for(int e = 0; e < d; ++e) pre *= sz_in[e];
  for(int e = d + 1; e < order; ++e) post *= sz_in[e];
  const int rows = transpose ? nd : md, cols = transpose ? md : nd;
  const int total = pre * rows * post;
  for(int o = threadIdx.x; o < total; o += blockDim.x)
  {
    const int a = o % pre, i = (o / pre) % rows, c = o / (pre * rows);
    double s = 0.0;
    for(int j = 0; j < cols; ++j)
    {
      const double f = transpose ? F[j * nd + i] : F[i * nd + j];
      s += f * x[a + pre * (j + cols * c)];
    }
    y[o] = s;
  }
}

__global__ void __launch_bounds__(KR_T)
k_kron_vmult(const KronShape sh, const double *__restrict__ factors, const double *__restrict__ src,
             double *__restrict__ dst, const int transpose, const int add)
{
  __shared__ double buf[2][KR_MAX_SIZE];
  const size_t b = blockIdx.x;
  int nin = 1, nout = 1;
  for(int d = 0; d < sh.order; ++d)
  {
    nin *= transpose ? sh.m[d] : sh.n[d];
    nout *= transpose ? sh.n[d] : sh.m[d];
  }
  const double *x = src + b * size_t(nin);
  double *y = dst + b * size_t(nout);
  const double *fb = factors + b * size_t(sh.rank) * sh.term;
  double acc[KR_MAX_SIZE / KR_T];
  for(int q = 0; q < KR_MAX_SIZE / KR_T; ++q) acc[q] = 0.0;
  for(int r = 0; r < sh.rank; ++r)
  {
    const double *fr = fb + size_t(r) * sh.term;
    int sz[KR_MAX_ORDER];
    for(int d = 0; d < sh.order; ++d) sz[d] = transpose ? sh.m[d] : sh.n[d];
    const double *in = x;
    int cur = 0;
    for(int d = 0; d < sh.order; ++d)
    {
      mode_product(fr + sh.fac_off[d], sh.m[d], sh.n[d], transpose != 0, d, sz, sh.order, in,
                   buf[cur]);
      __syncthreads();
      sz[d] = transpose ? sh.n[d] : sh.m[d];
      in = buf[cur];
      cur ^= 1;
    }
    for(int o = threadIdx.x, q = 0; o < nout; o += KR_T, ++q) acc[q] += in[o];
    __syncthreads();
  }
  for(int o = threadIdx.x, q = 0; o < nout; o += KR_T, ++q) y[o] = add ? y[o] + acc[q] : acc[q];
}

// dense matrix (M x N, row-major) of every batch item, or its transpose
__global__ void k_kron_table(const KronShape sh, const double *__restrict__ factors,
                             const int transpose, double *__restrict__ out, const int64_t total)
{
  int M = 1, N = 1;
  for(int d = 0; d < sh.order; ++d)
  {
    M *= sh.m[d];
    N *= sh.n[d];
  }
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(g >= total) return;
  const int64_t per = int64_t(M) * N;
  const int64_t b = g / per, e = g % per;
  const int row = int(transpose ? e % M : e / N), col = int(transpose ? e / M : e % N);
  const double *fb = factors + b * size_t(sh.rank) * sh.term;
  double s = 0.0;
  for(int r = 0; r < sh.rank; ++r)
  {
    double p = 1.0;
    int ri = row, ci = col;
    for(int d = 0; d < sh.order; ++d)
    {
      const int id = ri % sh.m[d], jd = ci % sh.n[d];
      ri /= sh.m[d];
      ci /= sh.n[d];
      p *= fb[size_t(r) * sh.term + sh.fac_off[d] + id * sh.n[d] + jd];
    }
    s += p;
  }
  out[g] = s;
}

__device__ double kr_block_sum(double v, double *sh)
{
  sh[threadIdx.x] = v;
  __syncthreads();
  for(int w = KR_T / 2; w > 0; w >>= 1)
  {
    if(int(threadIdx.x) < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Pseudo-inverse of n x n matrices by one-sided Jacobi (Hestenes): the columns
// of W = A V are made orthogonal by plane rotations, V accumulating them; then
// A = U S V^T with S = column norms of W, U = W S^{-1}, and
// pinv = V S^+ U^T = sum_{kept i} v_i w_i^T / s_i^2.  One block per matrix,
// W and V in the workspace (n^2 each), the n/2 disjoint column pairs of a
// round-robin step rotated by the waves of the block.
__global__ void __launch_bounds__(KR_T)
k_pinv_jacobi(const double *__restrict__ A, const int n, const double threshold,
              const int kernel_size, double *__restrict__ work, double *__restrict__ out)
{
  __shared__ double red[KR_T];
  __shared__ double sv[256];
  __shared__ int perm[256];
  __shared__ int rotated;
  const size_t b = blockIdx.x;
  const double *a = A + b * size_t(n) * n;
  double *W = work + b * size_t(2) * n * n;
  double *V = W + size_t(n) * n;
  double *P = out + b * size_t(n) * n;
  for(int e = threadIdx.x; e < n * n; e += KR_T)
  {
    W[e] = a[e];
    V[e] = (e / n == e % n) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64, nw = KR_T / 64;
  const int np = n + (n & 1);  // round-robin over an even number of slots
  for(int sweep = 0; sweep < 60; ++sweep)
  {
    if(threadIdx.x == 0) rotated = 0;
    __syncthreads();
    for(int step = 0; step < np - 1; ++step)
    {
      // pairs of the round-robin schedule: slot 0 fixed, the others rotate
      for(int k = wave; k < np / 2; k += nw)
      {
        auto slot = [&](int s) { return s == 0 ? 0 : 1 + (s - 1 + step) % (np - 1); };
        int p = slot(k), q = slot(np - 1 - k);
        if(p > q)
        {
          const int t = p;
          p = q;
          q = t;
        }
        if(q >= n) continue;  // the dummy slot of an odd n
        double al = 0.0, be = 0.0, ga = 0.0;
        for(int i = lane; i < n; i += 64)
        {
          const double wp = W[size_t(i) * n + p], wq = W[size_t(i) * n + q];
          al += wp * wp;
          be += wq * wq;
          ga += wp * wq;
        }
        for(int off = 32; off > 0; off >>= 1)
        {
          al += __shfl_xor(al, off);
          be += __shfl_xor(be, off);
          ga += __shfl_xor(ga, off);
        }
        if(fabs(ga) <= 1e-300 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        for(int i = lane; i < n; i += 64)
        {
          const double wp = W[size_t(i) * n + p], wq = W[size_t(i) * n + q];
          W[size_t(i) * n + p] = cs * wp - sn * wq;
          W[size_t(i) * n + q] = sn * wp + cs * wq;
          const double vp = V[size_t(i) * n + p], vq = V[size_t(i) * n + q];
          V[size_t(i) * n + p] = cs * vp - sn * vq;
          V[size_t(i) * n + q] = sn * vp + cs * vq;
        }
        if(lane == 0) rotated = 1;
      }
      __syncthreads();
    }
    if(!rotated) break;
    __syncthreads();
  }
  // singular values and their order (descending)
  for(int j = 0; j < n; ++j)
  {
    double s = 0.0;
    for(int i = threadIdx.x; i < n; i += KR_T) s += W[size_t(i) * n + j] * W[size_t(i) * n + j];
    s = kr_block_sum(s, red);
    if(threadIdx.x == 0)
    {
      sv[j] = sqrt(s);
      perm[j] = j;
    }
  }
  __syncthreads();
  if(threadIdx.x == 0)
  {
    for(int i = 1; i < n; ++i)
    {
      const int t = perm[i];
      int j = i - 1;
      while(j >= 0 && sv[perm[j]] < sv[t])
      {
        perm[j + 1] = perm[j];
        --j;
      }
      perm[j + 1] = t;
    }
  }
  __syncthreads();
  const double smax = sv[perm[0]];
  const int keep = kernel_size >= 0 ? max(0, n - kernel_size) : n;
  for(int e = threadIdx.x; e < n * n; e += KR_T)
  {
    const int row = e / n, col = e % n;
    double s = 0.0;
    for(int r = 0; r < keep; ++r)
    {
      const int j = perm[r];
      const double sj = sv[j];
      if(!(sj > 0.0) || (kernel_size < 0 && sj < threshold * smax)) continue;
      s += V[size_t(row) * n + j] * W[size_t(col) * n + j] / (sj * sj);
    }
    P[e] = s;
  }
}

// --------------------------------------------------------------- CP-ALS
constexpr int CP_MAXR = 8;      // output rank
constexpr int CP_MAXIN = 8;     // input rank
constexpr int CP_MAXLEN = 144;  // values of one factor matrix (up to 12 x 12)

// pseudo-inverse of the symmetric positive semi-definite R x R matrix G
// (cyclic Jacobi eigendecomposition; eigenvalues of modulus <= machine epsilon
// dropped, invert_safe of alignedlinalg.h:103-109), single thread
__device__ void spsd_pinv(int R, const double *G, double *Pi)
{
  double Aw[CP_MAXR * CP_MAXR], Q[CP_MAXR * CP_MAXR];
  for(int i = 0; i < R * R; ++i)
  {
    Aw[i] = G[i];
    Q[i] = (i / R == i % R) ? 1.0 : 0.0;
  }
  for(int sweep = 0; sweep < 50; ++sweep)
  {
    double off = 0.0;
    for(int p = 0; p < R; ++p)
      for(int q = p + 1; q < R; ++q) off += Aw[p * R + q] * Aw[p * R + q];
    if(off < 1e-32) break;
    for(int p = 0; p < R; ++p)
      for(int q = p + 1; q < R; ++q)
      {
        const double apq = Aw[p * R + q];
        if(fabs(apq) < 1e-300) continue;
        const double th = (Aw[q * R + q] - Aw[p * R + p]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for(int k = 0; k < R; ++k)
        {
          const double akp = Aw[k * R + p], akq = Aw[k * R + q];
          Aw[k * R + p] = c * akp - s * akq;
          Aw[k * R + q] = s * akp + c * akq;
        }
        for(int k = 0; k < R; ++k)
        {
          const double apk = Aw[p * R + k], aqk = Aw[q * R + k];
          Aw[p * R + k] = c * apk - s * aqk;
          Aw[q * R + k] = s * apk + c * aqk;
        }
        for(int k = 0; k < R; ++k)
        {
          const double qkp = Q[k * R + p], qkq = Q[k * R + q];
          Q[k * R + p] = c * qkp - s * qkq;
          Q[k * R + q] = s * qkp + c * qkq;
        }
      }
  }
  for(int i = 0; i < R; ++i)
    for(int j = 0; j < R; ++j)
    {
      double s = 0.0;
      for(int k = 0; k < R; ++k)
      {
        const double l = Aw[k * R + k];
        if(fabs(l) > 2.220446049250313e-16) s += Q[i * R + k] * Q[j * R + k] / l;
      }
      Pi[i * R + j] = s;
    }
}

// One block per batch item.  Modes named as in compute_kcp: mode 0 holds the
// vectorised in[i][2], mode 1 in[i][1], mode 2 in[i][0] (row-major vec).
// Each update A_k <- Ahat_k [hadamard_{l != k} (Ahat_l^T A_l)] pinv(V_k),
// V_k = hadamard_{l != k} (A_l^T A_l): the product of the mode-k unfolding of
// the input tensor with the Khatri-Rao product of the other modes, written
// with the input's own factors (the unfolding of a sum of rank-1 tensors).
__global__ void __launch_bounds__(KR_T)
k_kcp(const double *__restrict__ in, const int in_rank, const int R, const int len0,
      const int len1, const int len2, const int iters, double *__restrict__ out)
{
  __shared__ double Ah[3][CP_MAXIN * CP_MAXLEN];  // input factors, [term][entry]
  __shared__ double Ac[3][CP_MAXR * CP_MAXLEN];   // current factors, [col][entry]
  __shared__ double Gio[3][CP_MAXIN * CP_MAXR];   // Ahat_l^T A_l
  __shared__ double Goo[3][CP_MAXR * CP_MAXR];    // A_l^T A_l
  __shared__ double Wm[CP_MAXIN * CP_MAXR], Pi[CP_MAXR * CP_MAXR], Vm[CP_MAXR * CP_MAXR];
  const int len[3] = {len0, len1, len2};
  const size_t b = blockIdx.x;
  // input layout per batch item and term: in[i][0] (len2), in[i][1] (len1), in[i][2] (len0)
  const int term = len0 + len1 + len2;
  const double *ib = in + b * size_t(in_rank) * term;
  for(int i = 0; i < in_rank; ++i)
    for(int k = 0; k < 3; ++k)
    {
      const int off = k == 0 ? len2 + len1 : k == 1 ? len2 : 0;
      for(int e = threadIdx.x; e < len[k]; e += KR_T)
      {
        Ah[k][i * CP_MAXLEN + e] = ib[size_t(i) * term + off + e];
        if(i < R) Ac[k][i * CP_MAXLEN + e] = ib[size_t(i) * term + off + e];
      }
    }
  __syncthreads();
  auto gram = [&](int l) {
    // Gio[l] = Ahat_l^T A_l (in_rank x R), Goo[l] = A_l^T A_l (R x R)
    for(int t = threadIdx.x; t < in_rank * R + R * R; t += KR_T)
    {
      double s = 0.0;
      if(t < in_rank * R)
      {
        const int i = t / R, r = t % R;
        for(int e = 0; e < len[l]; ++e) s += Ah[l][i * CP_MAXLEN + e] * Ac[l][r * CP_MAXLEN + e];
        Gio[l][t] = s;
      }
      else
      {
        const int u = t - in_rank * R, r1 = u / R, r2 = u % R;
        for(int e = 0; e < len[l]; ++e) s += Ac[l][r1 * CP_MAXLEN + e] * Ac[l][r2 * CP_MAXLEN + e];
        Goo[l][u] = s;
      }
    }
  };
  for(int l = 0; l < 3; ++l) gram(l);
  __syncthreads();
  for(int it = 0; it < iters; ++it)
    for(int k = 0; k < 3; ++k)
    {
      const int l1 = k == 0 ? 1 : 0, l2 = k == 2 ? 1 : 2;
      for(int t = threadIdx.x; t < in_rank * R + R * R; t += KR_T)
      {
        if(t < in_rank * R)
          Wm[t] = Gio[l1][t] * Gio[l2][t];
        else
        {
          const int u = t - in_rank * R;
          Vm[u] = Goo[l1][u] * Goo[l2][u];
        }
      }
      __syncthreads();
      if(threadIdx.x == 0) spsd_pinv(R, Vm, Pi);
      __syncthreads();
      // A_k = Ahat_k (W Pi)
      for(int t = threadIdx.x; t < len[k] * R; t += KR_T)
      {
        const int e = t / R, r = t % R;
        double s = 0.0;
        for(int i = 0; i < in_rank; ++i)
        {
          double wp = 0.0;
          for(int q = 0; q < R; ++q) wp += Wm[i * R + q] * Pi[q * R + r];
          s += Ah[k][i * CP_MAXLEN + e] * wp;
        }
        Ac[k][r * CP_MAXLEN + e] = s;
      }
      __syncthreads();
      gram(k);
      __syncthreads();
    }
  // output per batch item and rank-1 term: out[r][0] (mode 2), out[r][1] (mode 1), out[r][2] (mode 0)
  double *ob = out + b * size_t(R) * term;
  for(int r = 0; r < R; ++r)
    for(int k = 0; k < 3; ++k)
    {
      const int off = k == 0 ? len2 + len1 : k == 1 ? len2 : 0;
      for(int e = threadIdx.x; e < len[k]; e += KR_T)
        ob[size_t(r) * term + off + e] = Ac[k][r * CP_MAXLEN + e];
    }
}

KronShape make_shape_kron(int order, int rank, const int *m, const int *n)
{
  TPSS_REQUIRE(order >= 1 && order <= KR_MAX_ORDER, "order must be 1, 2 or 3");
  TPSS_REQUIRE(rank >= 1, "rank must be positive");
  TPSS_REQUIRE(m && n, "null argument");
  KronShape sh{};
  sh.order = order;
  sh.rank = rank;
  int off = 0;
  for(int d = 0; d < order; ++d)
  {
    TPSS_REQUIRE(m[d] > 0 && n[d] > 0, "factor sizes must be positive");
    sh.m[d] = m[d];
    sh.n[d] = n[d];
    sh.fac_off[d] = off;
    off += m[d] * n[d];
  }
  sh.term = off;
  return sh;
}
}  // namespace
}  // namespace tpss

using namespace tpss;

extern "C" {

int tpss_kron_vmult(int order, int rank, const int *m, const int *n, int batch,
                    const double *factors, const double *src, double *dst, int transpose, int add,
                    void *stream)
{
  return guarded([&] {
    const KronShape sh = make_shape_kron(order, rank, m, n);
    TPSS_REQUIRE(batch >= 0, "batch must be non-negative");
    if(batch == 0) return;
    TPSS_REQUIRE(factors && src && dst, "null argument");
    TPSS_REQUIRE(static_cast<const void *>(src) != static_cast<const void *>(dst),
                 "dst and src must not alias");
    // largest intermediate: after the first d mode products dimensions < d
    // have the output sizes, the others the input sizes
    for(int d = 0; d <= order; ++d)
    {
      int64_t sz = 1;
      for(int e = 0; e < order; ++e)
        sz *= (e < d) == (transpose == 0) ? m[e] : n[e];
      TPSS_REQUIRE(sz <= KR_MAX_SIZE, "Kronecker vectors of at most 4096 entries");
    }
    hipLaunchKernelGGL(k_kron_vmult, dim3(unsigned(batch)), dim3(KR_T), 0, as_stream(stream), sh,
                       factors, src, dst, transpose, add);
    TPSS_HIP(hipGetLastError());
  });
}

int tpss_kron_table(int order, int rank, const int *m, const int *n, int batch,
                    const double *factors, int transpose, double *out, void *stream)
{
  return guarded([&] {
    const KronShape sh = make_shape_kron(order, rank, m, n);
    TPSS_REQUIRE(batch >= 0, "batch must be non-negative");
    if(batch == 0) return;
    TPSS_REQUIRE(factors && out, "null argument");
    int64_t per = 1;
    for(int d = 0; d < order; ++d) per *= int64_t(m[d]) * n[d];
    const int64_t total = per * batch;
    TPSS_REQUIRE(total < (int64_t(1) << 31), "dense matrices too large");
    hipLaunchKernelGGL(k_kron_table, dim3(blocks_for(total, 256)), dim3(256), 0,
                       as_stream(stream), sh, factors, transpose, out, total);
    TPSS_HIP(hipGetLastError());
  });
}

int tpss_kron_pinv(const double *A, int n, int batch, double threshold, int kernel_size,
                   double *out, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(n >= 1 && n <= 256, "matrices of size 1..256");
    TPSS_REQUIRE(batch >= 0, "batch must be non-negative");
    TPSS_REQUIRE(threshold >= 0.0, "threshold must be non-negative");
    TPSS_REQUIRE(kernel_size >= -1 && kernel_size <= n, "kernel_size in -1..n");
    if(batch == 0) return;
    TPSS_REQUIRE(A && out && A != out, "null or aliased argument");
    hipStream_t s = as_stream(stream);
    DeviceArray<double> work{size_t(2) * n * n * size_t(batch)};
    hipLaunchKernelGGL(k_pinv_jacobi, dim3(unsigned(batch)), dim3(KR_T), 0, s, A, n, threshold,
                       kernel_size, work.ptr, out);
    TPSS_HIP(hipGetLastError());
    TPSS_HIP(hipStreamSynchronize(s));
  });
}

int tpss_kcp_batched(const double *in, int batch, int in_rank, const int *m, const int *n,
                     int rank, int iterations, double *out, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(m && n, "null argument");
    TPSS_REQUIRE(batch >= 0, "batch must be non-negative");
    TPSS_REQUIRE(rank >= 1 && rank <= CP_MAXR, "rank must be in 1..8");
    TPSS_REQUIRE(in_rank >= rank && in_rank <= CP_MAXIN,
                 "the initial guess takes the first `rank` input terms: rank <= in_rank <= 8");
    int len[3];
    for(int k = 0; k < 3; ++k)
    {
      TPSS_REQUIRE(m[k] > 0 && n[k] > 0, "factor sizes must be positive");
      len[k] = m[k] * n[k];
      TPSS_REQUIRE(len[k] <= CP_MAXLEN, "factor matrices of at most 144 entries");
    }
    if(batch == 0) return;
    TPSS_REQUIRE(in && out, "null argument");
    const int iters = iterations > 0 ? iterations : rank * rank + 10;
    // m[k], n[k]: sizes of in[i][k]; the kernel's mode 0 is in[i][2]
    hipLaunchKernelGGL(k_kcp, dim3(unsigned(batch)), dim3(KR_T), 0, as_stream(stream), in, in_rank,
                       rank, len[2], len[1], len[0], iters, out);
    TPSS_HIP(hipGetLastError());
  });
}

}  // extern "C"
