// mlp.hip -- learned per-patch damping (the hook of the relaxation factor at
// schwarz_preconditioner.templates.h:631-633, which the reference keeps scalar).
//
//  * features: the singular values of the Jacobian of every cell of a patch at
//    its centre, sorted descending and scaled by n (= 1/h), cells in patch
//    lexicographic order (3 x 8 = 24 inputs in 3D, 2 x 4 = 8 in 2D);
//  * target ("optimal local damping" computed by the method itself): for a
//    residual r = A e of a random error e and the local correction
//    d_j = R_j^T A_j^{-1} R_j r, the A-norm line search along d_j,
//        omega_j = (d_j^T r) / (d_j^T A d_j),
//    accumulated over several e (ratio of sums) and clamped to the output range;
//    it is 1 wherever the local solver is the exact patch inverse;
//  * network: in -> H (tanh) -> H (tanh) -> 1, omega = lo + (hi - lo) sigmoid(z),
//    weights stored in BF16, products accumulated in FP32; trained with Adam on
//    an FP32 master copy (forward with the BF16-rounded weights).
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "ops.hpp"
#include "smoother.hpp"

struct tpss_mlp
{
  int n_in = 0, n_hidden = 0;
  double lo = 0.2, hi = 1.2;
  size_t n_params = 0;
  tpss::DeviceArray<float> master;  // FP32 parameters (training copy)
  tpss::DeviceArray<__hip_bfloat16> wbf;  // BF16 parameters used by every forward
  tpss::DeviceArray<float> adam_m, adam_v, grad;
  long step = 0;
};

namespace tpss
{
namespace
{
// parameter layout: W1[H][IN], b1[H], W2[H][H], b2[H], W3[H], b3
struct Offs
{
  size_t w1, b1, w2, b2, w3, b3, total;
};
Offs offsets(int in, int h)
{
  Offs o;
  o.w1 = 0;
  o.b1 = o.w1 + size_t(h) * in;
  o.w2 = o.b1 + h;
  o.b2 = o.w2 + size_t(h) * h;
  o.w3 = o.b2 + h;
  o.b3 = o.w3 + h;
  o.total = o.b3 + 1;
  return o;
}

__global__ void k_to_bf16(const float *__restrict__ src, __hip_bfloat16 *__restrict__ dst, const size_t n)
{
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i < n) dst[i] = __float2bfloat16(src[i]);
}

// forward for a batch; one thread per sample, weights staged in LDS as FP32
// values of the BF16 parameters; optionally keeps the activations for backward
template<int IN, int H>
__global__ void __launch_bounds__(128)
k_mlp_forward(const __hip_bfloat16 *__restrict__ wb, const float *__restrict__ X, const int64_t B,
              const float lo, const float hi, float *__restrict__ out, float *__restrict__ h1s,
              float *__restrict__ h2s, float *__restrict__ zs)
{
  __shared__ float W1[H * IN], b1[H], W2[H * H], b2[H], W3[H], b3;
  const Offs o = {0, size_t(H) * IN, size_t(H) * IN + H, size_t(H) * IN + H + size_t(H) * H,
                  size_t(H) * IN + 2 * H + size_t(H) * H, size_t(H) * IN + 3 * H + size_t(H) * H, 0};
  for(int i = threadIdx.x; i < H * IN; i += blockDim.x) W1[i] = __bfloat162float(wb[o.w1 + i]);
  for(int i = threadIdx.x; i < H * H; i += blockDim.x) W2[i] = __bfloat162float(wb[o.w2 + i]);
  for(int i = threadIdx.x; i < H; i += blockDim.x)
  {
    b1[i] = __bfloat162float(wb[o.b1 + i]);
    b2[i] = __bfloat162float(wb[o.b2 + i]);
    W3[i] = __bfloat162float(wb[o.w3 + i]);
  }
  if(threadIdx.x == 0) b3 = __bfloat162float(wb[o.b3]);
  __syncthreads();
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(s >= B) return;
  float x[IN], a1[H], a2[H];
#pragma unroll
  for(int i = 0; i < IN; ++i) x[i] = X[s * IN + i];
#pragma unroll 4
  for(int k = 0; k < H; ++k)
  {
    float acc = b1[k];
#pragma unroll
    for(int i = 0; i < IN; ++i) acc += W1[k * IN + i] * x[i];
    a1[k] = tanhf(acc);
  }
#pragma unroll 4
  for(int k = 0; k < H; ++k)
  {
    float acc = b2[k];
#pragma unroll
    for(int i = 0; i < H; ++i) acc += W2[k * H + i] * a1[i];
    a2[k] = tanhf(acc);
  }
  float z = b3;
#pragma unroll
  for(int i = 0; i < H; ++i) z += W3[i] * a2[i];
  const float sg = 1.0f / (1.0f + expf(-z));
  out[s] = lo + (hi - lo) * sg;
  if(h1s)
  {
    for(int k = 0; k < H; ++k)
    {
      h1s[s * H + k] = a1[k];
      h2s[s * H + k] = a2[k];
    }
    zs[s] = z;
  }
}

// backward, part 1 (one thread per sample): output and hidden-layer deltas of
// L = 1/(2B) sum_s (omega_s - target_s)^2 ; loss terms per sample
template<int IN, int H>
__global__ void __launch_bounds__(128)
k_mlp_deltas(const __hip_bfloat16 *__restrict__ wb, const float *__restrict__ target,
             const float *__restrict__ outv, const float *__restrict__ h1s,
             const float *__restrict__ h2s, const float *__restrict__ zs, const int64_t B,
             const float lo, const float hi, float *__restrict__ D1, float *__restrict__ D2,
             float *__restrict__ DZ, float *__restrict__ lossv)
{
  __shared__ float W2[H * H], W3[H];
  const size_t ow2 = size_t(H) * IN + H, ow3 = ow2 + size_t(H) * H + H;
  for(int i = threadIdx.x; i < H * H; i += blockDim.x) W2[i] = __bfloat162float(wb[ow2 + i]);
  for(int i = threadIdx.x; i < H; i += blockDim.x) W3[i] = __bfloat162float(wb[ow3 + i]);
  __syncthreads();
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(s >= B) return;
  const float invB = 1.0f / float(B);
  const float sg = 1.0f / (1.0f + expf(-zs[s]));
  const float err = outv[s] - target[s];
  lossv[s] = 0.5f * err * err * invB;
  const float dz = err * invB * (hi - lo) * sg * (1.0f - sg);
  DZ[s] = dz;
  float d2[H];
#pragma unroll 8
  for(int k = 0; k < H; ++k)
  {
    const float a2 = h2s[s * H + k];
    d2[k] = dz * W3[k] * (1.0f - a2 * a2);
    D2[s * H + k] = d2[k];
  }
#pragma unroll 4
  for(int i = 0; i < H; ++i)
  {
    float acc = 0.f;
#pragma unroll 8
    for(int k = 0; k < H; ++k) acc += W2[k * H + i] * d2[k];
    const float a1 = h1s[s * H + i];
    D1[s * H + i] = acc * (1.0f - a1 * a1);
  }
}

// backward, part 2 (one thread per parameter): sums over the batch,
// deterministic order
template<int IN, int H>
__global__ void __launch_bounds__(256)
k_mlp_param_grad(const float *__restrict__ X, const float *__restrict__ h1s,
                 const float *__restrict__ h2s, const float *__restrict__ D1,
                 const float *__restrict__ D2, const float *__restrict__ DZ, const int64_t B,
                 float *__restrict__ grad)
{
  const size_t ob1 = size_t(H) * IN, ow2 = ob1 + H, ob2 = ow2 + size_t(H) * H, ow3 = ob2 + H,
               ob3 = ow3 + H, total = ob3 + 1;
  const size_t pidx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(pidx >= total) return;
  float acc = 0.f;
  if(pidx < ob1)
  {
    const int k = int(pidx / IN), i = int(pidx % IN);
    for(int64_t s = 0; s < B; ++s) acc += D1[s * H + k] * X[s * IN + i];
  }
  else if(pidx < ow2)
  {
    const int k = int(pidx - ob1);
    for(int64_t s = 0; s < B; ++s) acc += D1[s * H + k];
  }
  else if(pidx < ob2)
  {
    const int k = int((pidx - ow2) / H), i = int((pidx - ow2) % H);
    for(int64_t s = 0; s < B; ++s) acc += D2[s * H + k] * h1s[s * H + i];
  }
  else if(pidx < ow3)
  {
    const int k = int(pidx - ob2);
    for(int64_t s = 0; s < B; ++s) acc += D2[s * H + k];
  }
  else if(pidx < ob3)
  {
    const int i = int(pidx - ow3);
    for(int64_t s = 0; s < B; ++s) acc += DZ[s] * h2s[s * H + i];
  }
  else
  {
    for(int64_t s = 0; s < B; ++s) acc += DZ[s];
  }
  grad[pidx] = acc;
}

__global__ void k_sum_partials(const float *__restrict__ partial, const int nb, const size_t n,
                               float *__restrict__ out)
{
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= n) return;
  float s = 0.f;
  for(int b = 0; b < nb; ++b) s += partial[size_t(b) * n + i];
  out[i] = s;
}

__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, const size_t n, const float lr, const float b1c,
                       const float b2c)
{
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / b1c) / (sqrtf(vi / b2c) + eps);
}

// ---------------------------------------------------------------- features

__device__ void sym3_eigvals(double a[3][3], double *ev)
{
  // cyclic Jacobi on a symmetric 3x3 matrix
  for(int sweep = 0; sweep < 30; ++sweep)
  {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if(off < 1e-30 * (a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2] + 1e-300)) break;
    for(int p = 0; p < 2; ++p)
      for(int q = p + 1; q < 3; ++q)
      {
        const double apq = a[p][q];
        if(apq == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for(int r = 0; r < 3; ++r)
        {
          const double arp = a[r][p], arq = a[r][q];
          a[r][p] = c * arp - sn * arq;
          a[r][q] = sn * arp + c * arq;
        }
        for(int r = 0; r < 3; ++r)
        {
          const double apr = a[p][r], aqr = a[q][r];
          a[p][r] = c * apr - sn * aqr;
          a[q][r] = sn * apr + c * aqr;
        }
      }
  }
  ev[0] = a[0][0];
  ev[1] = a[1][1];
  ev[2] = a[2][2];
}

template<int DIM>
__global__ void k_patch_features(const double *__restrict__ vert, const uint32_t *__restrict__ vtx,
                                 const int P, const int n, float *__restrict__ feat)
{
  const int j = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(j >= P) return;
  int v[3];
  unpack_vertex(vtx[j], v);
  constexpr int NF = DIM * (1 << DIM);
  for(int cb = 0; cb < (1 << DIM); ++cb)
  {
    int cc[3] = {0, 0, 0};
    for(int e = 0; e < DIM; ++e) cc[e] = v[e] - 1 + ((cb >> e) & 1);
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if(vert)
    {
      for(int c = 0; c < (1 << DIM); ++c)
      {
        int64_t idx = 0, st = 1;
        for(int e = 0; e < DIM; ++e)
        {
          idx += int64_t(cc[e] + ((c >> e) & 1)) * st;
          st *= (n + 1);
        }
        for(int b = 0; b < DIM; ++b)
        {
          double w = 1.0;
          for(int e = 0; e < DIM; ++e)
          {
            const int bit = (c >> e) & 1;
            w *= e == b ? (bit ? 1.0 : -1.0) : 0.5;
          }
          for(int a = 0; a < DIM; ++a) J[a][b] += w * vert[idx * DIM + a];
        }
      }
    }
    else
      for(int a = 0; a < DIM; ++a) J[a][a] = 1.0 / n;
    double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for(int a = 0; a < DIM; ++a)
      for(int b = 0; b < DIM; ++b)
        for(int c = 0; c < DIM; ++c) G[a][b] += J[c][a] * J[c][b];
    if(DIM == 2) G[2][2] = 0.0;
    double ev[3];
    sym3_eigvals(G, ev);
    double sv[3];
    for(int a = 0; a < 3; ++a) sv[a] = sqrt(fmax(ev[a], 0.0)) * n;
    // descending
    for(int a = 0; a < 3; ++a)
      for(int b = a + 1; b < 3; ++b)
        if(sv[b] > sv[a])
        {
          const double t = sv[a];
          sv[a] = sv[b];
          sv[b] = t;
        }
    for(int a = 0; a < DIM; ++a) feat[size_t(j) * NF + cb * DIM + a] = float(sv[a]);
  }
}

// per patch: num += sum_i D[i] r[i], den += sum_i D[i] y[i] over the patch DoFs
template<int DIM, int K, typename VT>
__global__ void k_patch_dots(const VT *__restrict__ D, const VT *__restrict__ r,
                             const VT *__restrict__ y, const uint32_t *__restrict__ vtx,
                             const int *__restrict__ ids, const int count, const int n,
                             double *__restrict__ num, double *__restrict__ den)
{
  constexpr int NP = K / 2, NPD = cpow(K, DIM), NDC = cpow(NP, DIM);
  __shared__ double sa[64], sb[64];
  const int slot = blockIdx.x;
  if(slot >= count) return;
  const int j = ids[slot];
  int v[3];
  unpack_vertex(vtx[j], v);
  double a = 0, bsum = 0;
  for(int i = threadIdx.x; i < NPD; i += blockDim.x)
  {
    const int i0 = i % K, i1 = DIM > 1 ? (i / K) % K : 0, i2 = DIM > 2 ? i / (K * K) : 0;
    const int64_t g = patch_cell<DIM, NP>(v, n, i0, i1, i2) * NDC + patch_local<DIM, NP>(i0, i1, i2);
    a += double(D[g]) * double(r[g]);
    bsum += double(D[g]) * double(y[g]);
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = bsum;
  __syncthreads();
  for(int w = 32; w > 0; w >>= 1)
  {
    if(int(threadIdx.x) < w)
    {
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if(threadIdx.x == 0)
  {
    num[j] += sa[0];
    den[j] += sb[0];
  }
}

// D[patch DoFs of ids] = A_j^{-1} R_j r  (plain local solves, no damping)
template<int DIM, int K, typename VT>
__global__ void __launch_bounds__((PatchShape<DIM, K>::THREADS))
k_patch_solve_ids(const VT *__restrict__ r, VT *__restrict__ D, const uint32_t *__restrict__ vtx,
                  const uint32_t *__restrict__ fac, const VT *__restrict__ S_tab,
                  const VT *__restrict__ lam_tab, const int *__restrict__ ids, const int count,
                  const int n)
{
  using PS = PatchShape<DIM, K>;
  constexpr int NP = PS::NP, L = PS::L, NPD = PS::NPD, NDC = PS::NDC, PPB = PS::PPB;
  __shared__ VT buf[PPB][NPD];
  __shared__ VT Ss[PPB][DIM][K * K];
  __shared__ VT ls[PPB][DIM][K];
  const int pl = threadIdx.x / L, t = threadIdx.x % L;
  const int s0 = int(blockIdx.x) * PPB;
  const bool valid = pl < PPB && s0 + pl < count;
  const int plc = pl < PPB ? pl : 0;
  for(int idx = threadIdx.x; idx < PPB * DIM * K * K; idx += blockDim.x)
  {
    const int q = idx / (DIM * K * K), rr = idx % (DIM * K * K);
    const int d = rr / (K * K), e = rr % (K * K);
    Ss[q][d][e] = s0 + q < count ? S_tab[size_t(fac[size_t(ids[s0 + q]) * DIM + d]) * K * K + e] : VT(0);
  }
  for(int idx = threadIdx.x; idx < PPB * DIM * K; idx += blockDim.x)
  {
    const int q = idx / (DIM * K), rr = idx % (DIM * K);
    const int d = rr / K, e = rr % K;
    ls[q][d][e] = s0 + q < count ? lam_tab[size_t(fac[size_t(ids[s0 + q]) * DIM + d]) * K + e] : VT(1);
  }
  int v[3] = {1, 1, 1};
  if(valid) unpack_vertex(vtx[ids[s0 + pl]], v);
  __syncthreads();
  const int ty = DIM > 1 ? t % K : 0, tz = DIM > 2 ? t / K : 0;
  VT rr_[K], o[K];
#pragma unroll
  for(int i = 0; i < K; ++i)
    rr_[i] = valid ? r[patch_cell<DIM, NP>(v, n, i, ty, tz) * NDC + patch_local<DIM, NP>(i, ty, tz)] : VT(0);
  fd_lines<DIM, K, VT>(rr_, o, buf[plc], Ss[plc][0], Ss[plc][1], Ss[plc][DIM > 2 ? 2 : 1], ls[plc][0],
                       ls[plc][1], ls[plc][DIM > 2 ? 2 : 1], VT(1), t, pl < PPB);
  if(!valid) return;
#pragma unroll
  for(int a = 0; a < K; ++a)
    D[patch_cell<DIM, NP>(v, n, a, ty, tz) * NDC + patch_local<DIM, NP>(a, ty, tz)] = o[a];
}

__global__ void k_damping_from_ratio(const double *__restrict__ num, const double *__restrict__ den,
                                     const int P, const double lo, const double hi,
                                     float *__restrict__ out)
{
  const int j = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(j >= P) return;
  double w = den[j] > 0.0 ? num[j] / den[j] : 1.0;
  w = fmin(hi, fmax(lo, w));
  out[j] = float(w);
}

__global__ void k_lambda_from_ratio(const double *__restrict__ num, const double *__restrict__ den,
                                    const int *__restrict__ ids, const int count,
                                    double *__restrict__ out)
{
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= count) return;
  const int j = ids[i];
  out[j] = den[j] > 0.0 ? num[j] / den[j] : 0.0;
}

int red_black_color(int dim, const int *v)
{
  int mask = 0, sum = 0;
  for(int d = 0; d < dim; ++d)
  {
    const int x = v[d] - 1, s = x & 1;
    mask |= s << d;
    sum += (x + s) / 2;
  }
  return 2 * mask + (sum & 1);
}

template<typename F>
void dispatch_mlp(int in, int h, F &&f)
{
  if(in == 24 && h == 64) return f(ic<24>{}, ic<64>{});
  if(in == 8 && h == 64) return f(ic<8>{}, ic<64>{});
  if(in == 24 && h == 16) return f(ic<24>{}, ic<16>{});
  if(in == 8 && h == 16) return f(ic<8>{}, ic<16>{});
  throw Error(TPSS_E_UNSUPPORTED, "MLP shape must be (24|8) inputs and 64|16 hidden units");
}

void mlp_refresh_bf16(tpss_mlp *m, hipStream_t s)
{
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(int64_t(m->n_params), 256)), dim3(256), 0, s,
                     m->master.ptr, m->wbf.ptr, m->n_params);
  TPSS_HIP(hipGetLastError());
}

void mlp_forward(tpss_mlp *m, const float *X, float *out, int64_t B, hipStream_t s, float *h1,
                 float *h2, float *z)
{
  dispatch_mlp(m->n_in, m->n_hidden, [&](auto IN, auto H) {
    constexpr int I = decltype(IN)::value, HH = decltype(H)::value;
    hipLaunchKernelGGL((k_mlp_forward<I, HH>), dim3(blocks_for(B, 128)), dim3(128), 0, s,
                       m->wbf.ptr, X, B, float(m->lo), float(m->hi), out, h1, h2, z);
    TPSS_HIP(hipGetLastError());
  });
}

// gradient of the mean-squared loss over one batch into m->grad; returns the loss
double mlp_grad(tpss_mlp *m, const float *X, const float *target, int64_t B, hipStream_t s)
{
  const size_t H = size_t(m->n_hidden);
  DeviceArray<float> out{size_t(B)}, h1{size_t(B) * H}, h2{size_t(B) * H}, z{size_t(B)},
    D1{size_t(B) * H}, D2{size_t(B) * H}, DZ{size_t(B)}, lossv{size_t(B)};
  mlp_forward(m, X, out.ptr, B, s, h1.ptr, h2.ptr, z.ptr);
  dispatch_mlp(m->n_in, m->n_hidden, [&](auto IN, auto HC) {
    constexpr int I = decltype(IN)::value, HH = decltype(HC)::value;
    hipLaunchKernelGGL((k_mlp_deltas<I, HH>), dim3(blocks_for(B, 128)), dim3(128), 0, s, m->wbf.ptr,
                       target, out.ptr, h1.ptr, h2.ptr, z.ptr, B, float(m->lo), float(m->hi), D1.ptr,
                       D2.ptr, DZ.ptr, lossv.ptr);
    hipLaunchKernelGGL((k_mlp_param_grad<I, HH>), dim3(blocks_for(int64_t(m->n_params), 256)),
                       dim3(256), 0, s, X, h1.ptr, h2.ptr, D1.ptr, D2.ptr, DZ.ptr, B, m->grad.ptr);
    TPSS_HIP(hipGetLastError());
  });
  DeviceArray<float> lsum(1);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1), 0, s, lossv.ptr, int(B), size_t(1), lsum.ptr);
  TPSS_HIP(hipGetLastError());
  float l = 0.f;
  TPSS_HIP(hipMemcpyAsync(&l, lsum.ptr, sizeof(float), hipMemcpyDeviceToHost, s));
  TPSS_HIP(hipStreamSynchronize(s));
  return double(l);
}

template<int DIM, int K>
void patch_features_t(const tpss_smoother *sm, float *dev_out, hipStream_t s)
{
  const LocalSolvers &ls = *sm->ls;
  const tpss_level *lv = sm->level;
  hipLaunchKernelGGL((k_patch_features<DIM>), dim3(blocks_for(ls.P, 128)), dim3(128), 0, s,
                     lv->distorted ? lv->d_vertices.ptr : nullptr, ls.d_vtx.ptr, ls.P, ls.n, dev_out);
  TPSS_HIP(hipGetLastError());
}

void patch_features(const tpss_smoother *sm, float *dev_out, hipStream_t s)
{
  dispatch_dim_np(sm->ls->dim, sm->ls->np, [&](auto D, auto NPc) {
    patch_features_t<decltype(D)::value, 2 * decltype(NPc)::value>(sm, dev_out, s);
  });
}

// omega_j = sum_e d_j^T r / sum_e d_j^T A d_j (FP64 vectors and solves)
void optimal_damping(tpss_smoother *sm, int n_samples, uint64_t seed, double lo, double hi,
                     float *dev_out, hipStream_t s)
{
  const LocalSolvers &ls = *sm->ls;
  tpss_level *lv = sm->level;
  TPSS_REQUIRE(lv->vec_prec == TPSS_F64, "optimal damping needs an FP64 level");
  const int dim = ls.dim;
  // red-black groups of the smoother's patches (same-color patches are DoF
  // disjoint and not face coupled, so A (sum_j d_j) restricted to patch j is A d_j)
  std::vector<std::vector<int>> groups(2 << dim);
  for(int j = 0; j < ls.P; ++j)
    groups[red_black_color(dim, &ls.vtx_unpacked[size_t(j) * dim])].push_back(j);
  std::vector<std::unique_ptr<DeviceArray<int>>> gids;
  for(auto &gr : groups)
  {
    gids.emplace_back(new DeviceArray<int>());
    gids.back()->upload(gr);
  }
  const int64_t N = lv->N;
  DeviceArray<double> e{size_t(N)}, r{size_t(N)}, D{size_t(N)}, y{size_t(N)}, num{size_t(ls.P)},
    den{size_t(ls.P)};
  fill_zero(num.ptr, TPSS_F64, ls.P, s);
  fill_zero(den.ptr, TPSS_F64, ls.P, s);
  for(int smp = 0; smp < n_samples; ++smp)
  {
    hipLaunchKernelGGL((k_random<double>), dim3(blocks_for(N, 256)), dim3(256), 0, s, e.ptr, N,
                       seed + uint64_t(smp));
    operator_apply_any(lv, r.ptr, e.ptr, nullptr, nullptr, 0, s);
    for(size_t c = 0; c < groups.size(); ++c)
    {
      const int count = int(groups[c].size());
      if(count == 0) continue;
      fill_zero(D.ptr, TPSS_F64, N, s);
      dispatch_dim_np(dim, ls.np, [&](auto Dm, auto NPc) {
        constexpr int DIM = decltype(Dm)::value, K = 2 * decltype(NPc)::value;
        using PS = PatchShape<DIM, K>;
        // block-diagonal local solvers (elasticity): one solve per component
        // block, with that component's factor rows
        for(int comp = 0; comp < ls.ncomp; ++comp)
          hipLaunchKernelGGL((k_patch_solve_ids<DIM, K, double>),
                             dim3((count + PS::PPB - 1) / PS::PPB), dim3(PS::THREADS), 0, s,
                             r.ptr + comp * lv->Ns, D.ptr + comp * lv->Ns, ls.d_vtx.ptr,
                             ls.d_fac.ptr + size_t(comp) * ls.P * dim, ls.S64.ptr, ls.lam64.ptr,
                             gids[c]->ptr, count, ls.n);
        TPSS_HIP(hipGetLastError());
        operator_apply_any(lv, y.ptr, D.ptr, nullptr, nullptr, 0, s);
        // d_j^T r and d_j^T A d_j summed over the component blocks
        for(int comp = 0; comp < ls.ncomp; ++comp)
          hipLaunchKernelGGL((k_patch_dots<DIM, K, double>), dim3(count), dim3(64), 0, s,
                             D.ptr + comp * lv->Ns, r.ptr + comp * lv->Ns, y.ptr + comp * lv->Ns,
                             ls.d_vtx.ptr, gids[c]->ptr, count, ls.n, num.ptr, den.ptr);
        TPSS_HIP(hipGetLastError());
      });
    }
  }
  hipLaunchKernelGGL(k_damping_from_ratio, dim3(blocks_for(ls.P, 256)), dim3(256), 0, s, num.ptr,
                     den.ptr, ls.P, lo, hi, dev_out);
  TPSS_HIP(hipGetLastError());
  TPSS_HIP(hipStreamSynchronize(s));
}
// Largest eigenvalue of A_j^{-1}_local A_j (local solver vs exact local matrix)
// per patch by power iteration in the A_j inner product: v <- M_j^{-1} A_j v,
// lambda = (A v)^T M^{-1} (A v) / v^T A v (the Rayleigh quotient of M^{-1} A),
// all patches of one red-black color at once (not face coupled, so
// (A v)|_j = A_j v_j for v supported on the color's patches)
void local_lambda_max(tpss_smoother *sm, int n_iter, uint64_t seed, double *dev_lambda,
                      hipStream_t s)
{
  const LocalSolvers &ls = *sm->ls;
  tpss_level *lv = sm->level;
  TPSS_REQUIRE(lv->vec_prec == TPSS_F64, "local spectra need an FP64 level");
  const int dim = ls.dim;
  std::vector<std::vector<int>> groups(2 << dim);
  for(int j = 0; j < ls.P; ++j)
    groups[red_black_color(dim, &ls.vtx_unpacked[size_t(j) * dim])].push_back(j);
  const int64_t N = lv->N;
  DeviceArray<double> e{size_t(N)}, v{size_t(N)}, w{size_t(N)}, num{size_t(ls.P)},
    den{size_t(ls.P)};
  DeviceArray<int> gid;
  hipLaunchKernelGGL((k_random<double>), dim3(blocks_for(N, 256)), dim3(256), 0, s, e.ptr, N,
                     seed);
  for(size_t c = 0; c < groups.size(); ++c)
  {
    const int count = int(groups[c].size());
    if(count == 0) continue;
    TPSS_HIP(hipStreamSynchronize(s));  // the previous group is done with gid
    gid.upload(groups[c]);
    dispatch_dim_np(dim, ls.np, [&](auto Dm, auto NPc) {
      constexpr int DIM = decltype(Dm)::value, K = 2 * decltype(NPc)::value;
      using PS = PatchShape<DIM, K>;
      // start: v = M^{-1} e on the group's patches, zero elsewhere
      fill_zero(v.ptr, TPSS_F64, N, s);
      hipLaunchKernelGGL((k_patch_solve_ids<DIM, K, double>), dim3((count + PS::PPB - 1) / PS::PPB),
                         dim3(PS::THREADS), 0, s, e.ptr, v.ptr, ls.d_vtx.ptr, ls.d_fac.ptr,
                         ls.S64.ptr, ls.lam64.ptr, gid.ptr, count, ls.n);
      // the iterates alternate between v and e, both zero off the group's patches
      fill_zero(e.ptr, TPSS_F64, N, s);
      for(int it = 0; it < n_iter; ++it)
      {
        operator_apply_any(lv, w.ptr, v.ptr, nullptr, nullptr, 0, s);
        if(it + 1 == n_iter)
        {
          fill_zero(num.ptr, TPSS_F64, ls.P, s);
          fill_zero(den.ptr, TPSS_F64, ls.P, s);
        }
        // v <- M^{-1} (A v) on the patches (the previous v's entries are overwritten;
        // the rest of v stays zero)
        hipLaunchKernelGGL((k_patch_solve_ids<DIM, K, double>),
                           dim3((count + PS::PPB - 1) / PS::PPB), dim3(PS::THREADS), 0, s, w.ptr,
                           e.ptr, ls.d_vtx.ptr, ls.d_fac.ptr, ls.S64.ptr, ls.lam64.ptr, gid.ptr,
                           count, ls.n);
        if(it + 1 == n_iter)
          // num = w . M^{-1} w, den = w . v
          hipLaunchKernelGGL((k_patch_dots<DIM, K, double>), dim3(count), dim3(64), 0, s, w.ptr,
                             e.ptr, v.ptr, ls.d_vtx.ptr, gid.ptr, count, ls.n, num.ptr, den.ptr);
        std::swap(v.ptr, e.ptr);
        TPSS_HIP(hipGetLastError());
      }
      // e is overwritten by the iteration: a fresh random start for the next group
      hipLaunchKernelGGL((k_random<double>), dim3(blocks_for(N, 256)), dim3(256), 0, s, e.ptr, N,
                         seed + 1 + c);
      hipLaunchKernelGGL(k_lambda_from_ratio, dim3(blocks_for(count, 256)), dim3(256), 0, s,
                         num.ptr, den.ptr, gid.ptr, count, dev_lambda);
      TPSS_HIP(hipGetLastError());
    });
  }
  TPSS_HIP(hipStreamSynchronize(s));
}
}  // namespace
}  // namespace tpss

using namespace tpss;

extern "C" {

int tpss_mlp_create(int n_in, int n_hidden, double omega_min, double omega_max, uint64_t seed,
                    tpss_mlp **out)
{
  return guarded([&] {
    TPSS_REQUIRE(out, "null argument");
    TPSS_REQUIRE(omega_max > omega_min, "empty damping range");
    auto m = std::make_unique<tpss_mlp>();
    m->n_in = n_in;
    m->n_hidden = n_hidden;
    dispatch_mlp(n_in, n_hidden, [](auto, auto) {});
    m->lo = omega_min;
    m->hi = omega_max;
    const Offs o = offsets(n_in, n_hidden);
    m->n_params = o.total;
    // Glorot-uniform weights from SplitMix64, zero biases; output bias such
    // that the initial damping is 1 (inside [lo, hi])
    std::vector<float> h(o.total, 0.f);
    auto u = [&](uint64_t i) {
      uint64_t zz = seed * 0x100000000ull + i + 0x9E3779B97F4A7C15ull;
      zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
      zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
      zz ^= zz >> 31;
      return 2.0 * (double(zz >> 11) / 9007199254740992.0) - 1.0;
    };
    const double a1 = std::sqrt(6.0 / (n_in + n_hidden)), a2 = std::sqrt(6.0 / (2.0 * n_hidden)),
                 a3 = std::sqrt(6.0 / (n_hidden + 1));
    for(size_t i = 0; i < size_t(n_hidden) * n_in; ++i) h[o.w1 + i] = float(a1 * u(i));
    for(size_t i = 0; i < size_t(n_hidden) * n_hidden; ++i) h[o.w2 + i] = float(a2 * u(100000 + i));
    for(int i = 0; i < n_hidden; ++i) h[o.w3 + i] = float(a3 * u(200000 + i));
    const double s0 = (1.0 - omega_min) / (omega_max - omega_min);
    h[o.b3] = float(std::log(s0 / (1.0 - s0)));
    m->master.upload(h);
    m->wbf.resize(o.total);
    m->adam_m.upload(std::vector<float>(o.total, 0.f));
    m->adam_v.upload(std::vector<float>(o.total, 0.f));
    m->grad.resize(o.total);
    mlp_refresh_bf16(m.get(), nullptr);
    TPSS_HIP(hipDeviceSynchronize());
    *out = m.release();
  });
}

int tpss_mlp_destroy(tpss_mlp *m)
{
  delete m;
  return TPSS_OK;
}

int tpss_mlp_n_params(const tpss_mlp *m, size_t *n)
{
  return guarded([&] {
    TPSS_REQUIRE(m && n, "null argument");
    *n = m->n_params;
  });
}

int tpss_mlp_set_params(tpss_mlp *m, const float *host, size_t n)
{
  return guarded([&] {
    TPSS_REQUIRE(m && host && n == m->n_params, "parameter count mismatch");
    m->master.upload(std::vector<float>(host, host + n));
    mlp_refresh_bf16(m, nullptr);
    TPSS_HIP(hipDeviceSynchronize());
  });
}

int tpss_mlp_get_params(const tpss_mlp *m, float *host, size_t n)
{
  return guarded([&] {
    TPSS_REQUIRE(m && host && n == m->n_params, "parameter count mismatch");
    TPSS_HIP(hipMemcpy(host, m->master.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int tpss_mlp_forward(tpss_mlp *m, const float *X, float *out, int64_t batch, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(m && X && out && batch >= 0, "null argument");
    if(batch) mlp_forward(m, X, out, batch, as_stream(stream), nullptr, nullptr, nullptr);
  });
}

int tpss_mlp_loss_grad(tpss_mlp *m, const float *X, const float *target, int64_t batch,
                       float *grad_out, double *loss, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(m && X && target && batch > 0, "null argument");
    const double l = mlp_grad(m, X, target, batch, as_stream(stream));
    if(loss) *loss = l;
    if(grad_out)
      TPSS_HIP(hipMemcpy(grad_out, m->grad.ptr, m->n_params * sizeof(float), hipMemcpyDeviceToDevice));
  });
}

int tpss_mlp_train(tpss_mlp *m, const float *X, const float *target, int64_t n, int epochs,
                   int batch, double lr, uint64_t seed, double *final_loss)
{
  return guarded([&] {
    TPSS_REQUIRE(m && X && target && n > 0 && epochs >= 1 && batch >= 1, "invalid arguments");
    hipStream_t s = nullptr;
    const int64_t nb = (n + batch - 1) / batch;
    // deterministic batch order: a seeded rotation of the sample blocks per epoch
    double last = 0.0;
    for(int ep = 0; ep < epochs; ++ep)
    {
      double acc = 0.0;
      for(int64_t b = 0; b < nb; ++b)
      {
        const int64_t bb = (b + int64_t((seed + uint64_t(ep) * 7919u) % uint64_t(nb))) % nb;
        const int64_t off = bb * batch;
        const int64_t cnt = std::min<int64_t>(batch, n - off);
        acc += mlp_grad(m, X + off * m->n_in, target + off, cnt, s) * double(cnt);
        ++m->step;
        const float b1c = 1.f - std::pow(0.9f, float(m->step));
        const float b2c = 1.f - std::pow(0.999f, float(m->step));
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(int64_t(m->n_params), 256)), dim3(256), 0, s,
                           m->master.ptr, m->adam_m.ptr, m->adam_v.ptr, m->grad.ptr, m->n_params,
                           float(lr), b1c, b2c);
        TPSS_HIP(hipGetLastError());
        mlp_refresh_bf16(m, s);
      }
      last = acc / double(n);
    }
    TPSS_HIP(hipStreamSynchronize(s));
    if(final_loss) *final_loss = last;
  });
}

int tpss_smoother_patch_features(const tpss_smoother *sm, float *dev_out, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && dev_out, "null argument");
    patch_features(sm, dev_out, as_stream(stream));
  });
}

int tpss_smoother_optimal_damping(tpss_smoother *sm, int n_samples, uint64_t seed,
                                  double omega_min, double omega_max, float *dev_out)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && dev_out && n_samples >= 1, "invalid arguments");
    TPSS_REQUIRE(omega_max > omega_min, "empty damping range");
    optimal_damping(sm, n_samples, seed, omega_min, omega_max, dev_out, nullptr);
  });
}

int tpss_smoother_local_lambda_max(tpss_smoother *sm, int n_iter, uint64_t seed,
                                   double *dev_lambda)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && dev_lambda && n_iter >= 1, "invalid arguments");
    TPSS_REQUIRE(sm->level->ncomp == 1, "local spectra: Poisson levels only");
    local_lambda_max(sm, n_iter, seed, dev_lambda, nullptr);
  });
}

int tpss_smoother_set_damping_mlp(tpss_smoother *sm, tpss_mlp *m)
{
  return guarded([&] {
    TPSS_REQUIRE(sm, "null argument");
    if(!m)
    {
      sm->use_omega = false;
      return;
    }
    const int nf = sm->ls->dim * (1 << sm->ls->dim);
    TPSS_REQUIRE(nf == m->n_in, "MLP input size does not match the patch features");
    DeviceArray<float> feat{size_t(sm->ls->P) * nf};
    patch_features(sm, feat.ptr, nullptr);
    sm->d_omega.resize(sm->ls->P);
    mlp_forward(m, feat.ptr, sm->d_omega.ptr, sm->ls->P, nullptr, nullptr, nullptr, nullptr);
    TPSS_HIP(hipStreamSynchronize(nullptr));
    sm->use_omega = true;
  });
}

}  // extern "C"
