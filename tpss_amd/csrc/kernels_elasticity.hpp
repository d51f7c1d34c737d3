// kernels_elasticity.hpp -- SIPG linear elasticity operator on Cartesian meshes,
// the operator of the reference's LinElasticity model problem
// (LinElasticity::MW::MatrixIntegrator, apps/linelasticity_integrator.h:31-124,
// which calls deal.II's LocalIntegrators::Elasticity and ::GradDiv):
//   cell      2 mu eps(u):eps(v) + lambda div u div v
//   interior  2 mu (sigma [u].[v] - {eps(u) n}.[v] - {eps(v) n}.[u])
//             + lambda (sigma [u.n][v.n] - {div u}[v.n] - {div v}[u.n])
//   boundary  the same with the mirrored exterior state u_out = -u_in,
//             grad u_out = grad u_in (penalty 2 sigma, Dirichlet u = 0)
// sigma = ip * gamma / h, gamma = p (p + 1).
//
// Vectors are component-blocked: x[c * Ns + cell * NDC + local].  CPB cells per
// block (64 / NDC for small cells), one thread per cell DoF / quadrature point
// (QGauss(p+1)).  Faces are evaluated element-centrically (every output entry
// is written by exactly one thread): per face the traces of both sides are
// contracted along the normal for all tangential nodes at once, interpolated
// tangentially per face quadrature point, and the fluxes integrated back.
// The face direction is a compile-time constant, so no per-thread array is
// indexed at run time.  Arithmetic in FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels.hpp"
#include "kernels_general.hpp"

namespace tpss
{
// contraction of an LDS tensor along direction e: out(t) = sum_j M[o*NP + j] src(t|e->j),
// o = t_e; with TRANS the matrix is read as M[j*NP + o]
template<int DIM, int NP, bool TRANS>
__device__ __forceinline__ double contract(const double *src, const double *M, const int *tid,
                                           const int t, const int e)
{
  constexpr int ST[3] = {1, NP, NP * NP};
  const int o = tid[e];
  const int base = t - o * ST[e];
  double s = 0.0;
#pragma unroll
  for(int j = 0; j < NP; ++j) s += (TRANS ? M[j * NP + o] : M[o * NP + j]) * src[base + j * ST[e]];
  return s;
}

template<int NP>
struct ElastTab
{
  double V[NP * NP];   // V[q*NP + i] = phi_i(x_q)
  double G[NP * NP];   // G[q*NP + i] = phi_i'(x_q) / h
  double f[2][NP];     // phi_i(0), phi_i(1)
  double fd[2][NP];    // phi_i'(0) / h, phi_i'(1) / h
  double w[NP];        // h * Gauss weight
  double mu, lambda, sigma;
};

template<int DIM, int NP>
constexpr int elasticity_cells_per_block()
{
  return cpow(NP, DIM) >= 64 ? 1 : 64 / cpow(NP, DIM);
}

template<int DIM, int NP>
constexpr int elasticity_threads()
{
  return ((elasticity_cells_per_block<DIM, NP>() * cpow(NP, DIM) + 63) / 64) * 64;
}

template<int DIM, int NP, typename VT, bool RESIDUAL>
__global__ void __launch_bounds__((elasticity_threads<DIM, NP>()), 4)
k_elasticity(const VT *__restrict__ x, VT *__restrict__ y, const VT *__restrict__ b,
             const ElastTab<NP> tab, const int n, const int64_t Ns,
             const uint32_t *__restrict__ cells, const int64_t count)
{
  constexpr int NDC = cpow(NP, DIM);
  constexpr int NF = cpow(NP, DIM - 1);
  constexpr int CPB = elasticity_cells_per_block<DIM, NP>();
  constexpr int ST[3] = {1, NP, NP * NP};
  __shared__ double U[CPB][DIM][NDC], UN[CPB][DIM][NDC];
  __shared__ double A[CPB][NDC], B[CPB][NDC];
  __shared__ double S[CPB][DIM][DIM][NDC];
  __shared__ double PA[CPB][2][DIM][NF], PB[CPB][2][DIM][NF];
  __shared__ double FA[CPB][DIM][NF], FB[CPB][DIM][DIM][NF];
  __shared__ ElastTab<NP> tl;

  const int t = threadIdx.x;
  const int cs = t / NDC, ts0 = t % NDC;
  const int64_t slot = int64_t(blockIdx.x) * CPB + cs;
  const bool act = cs < CPB && slot < count;
  const int c0 = cs < CPB ? cs : 0;  // LDS slot of this thread (idle threads alias slot 0)
  const int ts = act ? ts0 : 0;
  const int64_t cell = act ? (cells ? int64_t(cells[slot]) : slot) : 0;
  int cc[3] = {0, 0, 0};
  {
    int64_t r = cell;
    for(int d = 0; d < DIM; ++d)
    {
      cc[d] = int(r % n);
      r /= n;
    }
  }
  int tid[3] = {0, 0, 0};
  {
    int r = ts;
    for(int d = 0; d < DIM; ++d)
    {
      tid[d] = r % NP;
      r /= NP;
    }
  }
  if(t == 0) tl = tab;
  if(act)
    for(int c = 0; c < DIM; ++c) U[c0][c][ts] = double(x[c * Ns + cell * NDC + ts]);
  __syncthreads();

  // ---------------- cell term: physical gradients at the Gauss points
  for(int c = 0; c < DIM; ++c)
    for(int a = 0; a < DIM; ++a)
    {
      const double *src = U[c0][c];
#pragma unroll
      for(int e = 0; e < DIM; ++e)
      {
        const double v = contract<DIM, NP, false>(src, e == a ? tl.G : tl.V, tid, ts, e);
        double *dst = (e % 2 == 0) ? A[c0] : B[c0];
        __syncthreads();
        if(act) dst[ts] = v;
        __syncthreads();
        src = dst;
      }
      const double g = src[ts];
      __syncthreads();
      if(act) S[c0][c][a][ts] = g;
    }
  __syncthreads();
  // stress times JxW: mu (grad u + grad u^T) + lambda div u I
  {
    double g[DIM][DIM], w = 1.0;
    for(int e = 0; e < DIM; ++e) w *= tl.w[tid[e]];
    double div = 0.0;
#pragma unroll
    for(int c = 0; c < DIM; ++c)
    {
#pragma unroll
      for(int a = 0; a < DIM; ++a) g[c][a] = S[c0][c][a][ts];
      div += g[c][c];
    }
    __syncthreads();
    if(act)
#pragma unroll
      for(int c = 0; c < DIM; ++c)
#pragma unroll
        for(int a = 0; a < DIM; ++a)
          S[c0][c][a][ts] = w * (tl.mu * (g[c][a] + g[a][c]) + (c == a ? tl.lambda * div : 0.0));
  }
  __syncthreads();
  double acc[DIM];
#pragma unroll
  for(int c = 0; c < DIM; ++c)
  {
    acc[c] = 0.0;
    for(int a = 0; a < DIM; ++a)
    {
      const double *src = S[c0][c][a];
#pragma unroll
      for(int e = 0; e < DIM; ++e)
      {
        const double v = contract<DIM, NP, true>(src, e == a ? tl.G : tl.V, tid, ts, e);
        double *dst = (e % 2 == 0) ? A[c0] : B[c0];
        __syncthreads();
        if(act) dst[ts] = v;
        __syncthreads();
        src = dst;
      }
      acc[c] += src[ts];
      __syncthreads();
    }
  }

  // ---------------- faces, direction D fixed at compile time
  auto face_dir = [&](auto Dc, int64_t cstride) {
    constexpr int D = decltype(Dc)::value;
    constexpr int E1 = D == 0 ? 1 : 0;
    constexpr int E2 = DIM == 3 ? (D == 2 ? 1 : 2) : E1;
#pragma unroll 1
    for(int side = 0; side < 2; ++side)
    {
      const bool interior = side == 1 ? cc[D] < n - 1 : cc[D] > 0;
      const int64_t nb = interior ? (side == 1 ? cell + cstride : cell - cstride) : cell;
      const int nside = 1 - side;
      if(interior && act)
        for(int c = 0; c < DIM; ++c) UN[c0][c][ts] = double(x[c * Ns + nb * NDC + ts]);
      __syncthreads();
      // contraction along the normal: value (PA) and normal derivative (PB) traces
      // of both sides for every component and tangential node
      if(act)
        for(int o = ts; o < 2 * DIM * NF; o += NDC)
        {
          const int sd = o / (DIM * NF), rem = o % (DIM * NF), c = rem / NF, m = rem % NF;
          if(sd == 1 && !interior) continue;
          const int i1 = m % NP, i2 = DIM == 3 ? m / NP : 0;
          const int base = i1 * ST[E1] + (DIM == 3 ? i2 * ST[E2] : 0);
          const double *src = sd == 0 ? U[c0][c] : UN[c0][c];
          const int fs = sd == 0 ? side : nside;
          double a = 0.0, bsum = 0.0;
#pragma unroll
          for(int id = 0; id < NP; ++id)
          {
            const double u = src[base + id * ST[D]];
            a += tl.f[fs][id] * u;
            bsum += tl.fd[fs][id] * u;
          }
          PA[c0][sd][c][m] = a;
          PB[c0][sd][c][m] = bsum;
        }
      __syncthreads();
      if(act && ts < NF)
      {
        const int l1 = ts % NP, l2 = DIM == 3 ? ts / NP : 0;
        // traces and physical gradients of this cell (o) and the exterior state (x)
        double uo[DIM], go[DIM][DIM], ux[DIM], gx[DIM][DIM];
#pragma unroll
        for(int c = 0; c < DIM; ++c)
        {
          uo[c] = ux[c] = 0.0;
#pragma unroll
          for(int a = 0; a < DIM; ++a) go[c][a] = gx[c][a] = 0.0;
        }
#pragma unroll 1
        for(int m = 0; m < NF; ++m)
        {
          const int i1 = m % NP, i2 = DIM == 3 ? m / NP : 0;
          const double t1v = tl.V[l1 * NP + i1], t1d = tl.G[l1 * NP + i1];
          const double t2v = DIM == 3 ? tl.V[l2 * NP + i2] : 1.0;
          const double t2d = DIM == 3 ? tl.G[l2 * NP + i2] : 0.0;
          const double wvv = t1v * t2v, wdv = t1d * t2v, wvd = t1v * t2d;
#pragma unroll
          for(int c = 0; c < DIM; ++c)
          {
            const double a = PA[c0][0][c][m];
            uo[c] += wvv * a;
            go[c][D] += wvv * PB[c0][0][c][m];
            go[c][E1] += wdv * a;
            if(DIM == 3) go[c][E2] += wvd * a;
            if(interior)
            {
              const double a1 = PA[c0][1][c][m];
              ux[c] += wvv * a1;
              gx[c][D] += wvv * PB[c0][1][c][m];
              gx[c][E1] += wdv * a1;
              if(DIM == 3) gx[c][E2] += wvd * a1;
            }
          }
        }
        if(!interior)
#pragma unroll
          for(int c = 0; c < DIM; ++c)
          {
            ux[c] = -uo[c];
#pragma unroll
            for(int a = 0; a < DIM; ++a) gx[c][a] = go[c][a];
          }
        const double nrm = side == 1 ? 1.0 : -1.0;
        const double dS = tl.w[l1] * (DIM == 3 ? tl.w[l2] : 1.0);
        double J[DIM], divo = 0.0, divx = 0.0;
#pragma unroll
        for(int c = 0; c < DIM; ++c)
        {
          J[c] = uo[c] - ux[c];
          divo += go[c][c];
          divx += gx[c][c];
        }
#pragma unroll
        for(int c = 0; c < DIM; ++c)
        {
          // {2 mu eps(u) n + lambda div(u) n}_c
          const double fo = tl.mu * (go[c][D] + go[D][c]) + (c == D ? tl.lambda * divo : 0.0);
          const double fx = tl.mu * (gx[c][D] + gx[D][c]) + (c == D ? tl.lambda * divx : 0.0);
          const double F = 0.5 * (fo + fx) * nrm;
          FA[c0][c][ts] = dS * (-F + 2.0 * tl.mu * tl.sigma * J[c] +
                                (c == D ? tl.lambda * tl.sigma * J[D] : 0.0));
          // -1/2 (2 mu eps(v) n . [u] + lambda div v [u].n)
          double fb[DIM];
#pragma unroll
          for(int a = 0; a < DIM; ++a) fb[a] = 0.0;
          fb[D] += -0.5 * tl.mu * nrm * J[c];
          if(c == D)
#pragma unroll
            for(int a = 0; a < DIM; ++a) fb[a] += -0.5 * tl.mu * nrm * J[a];
          fb[c] += -0.5 * tl.lambda * nrm * J[D];
#pragma unroll
          for(int a = 0; a < DIM; ++a) FB[c0][c][a][ts] = dS * fb[a];
        }
      }
      __syncthreads();
      if(act)
      {
        // test functions of this cell: phi_i = f(i_D) * tangential factors
        const int id = tid[D], i1 = tid[E1], i2 = DIM == 3 ? tid[E2] : 0;
        const double fv = tl.f[side][id], fdv = tl.fd[side][id];
#pragma unroll 1
        for(int l = 0; l < NF; ++l)
        {
          const int l1 = l % NP, l2 = DIM == 3 ? l / NP : 0;
          const double t1v = tl.V[l1 * NP + i1], t1d = tl.G[l1 * NP + i1];
          const double t2v = DIM == 3 ? tl.V[l2 * NP + i2] : 1.0;
          const double t2d = DIM == 3 ? tl.G[l2 * NP + i2] : 0.0;
          const double wv = fv * t1v * t2v, wn = fdv * t1v * t2v, w1 = fv * t1d * t2v,
                       w2 = fv * t1v * t2d;
#pragma unroll
          for(int c = 0; c < DIM; ++c)
          {
            double s = wv * FA[c0][c][l] + wn * FB[c0][c][D][l] + w1 * FB[c0][c][E1][l];
            if(DIM == 3) s += w2 * FB[c0][c][E2][l];
            acc[c] += s;
          }
        }
      }
      __syncthreads();
    }
  };
  const int64_t nn = int64_t(n) * n;
  face_dir(std::integral_constant<int, 0>{}, int64_t(1));
  face_dir(std::integral_constant<int, 1>{}, int64_t(n));
  if constexpr(DIM == 3) face_dir(std::integral_constant<int, DIM == 3 ? 2 : 1>{}, nn);

  if(act)
#pragma unroll
    for(int c = 0; c < DIM; ++c)
    {
      const int64_t i = c * Ns + cell * NDC + ts;
      if(RESIDUAL)
        y[i] = VT(double(b[i]) - acc[c]);
      else
        y[i] = VT(acc[c]);
    }
}

}  // namespace tpss
