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
// Vectors are component-blocked: x[c * Ns + cell * NDC + local].  One block per
// cell, one thread per cell DoF / quadrature point (QGauss(p+1)); faces are
// evaluated element-centrically like the Poisson kernel for moved vertices, so
// every output entry is written by exactly one thread.  Arithmetic in FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "kernels_general.hpp"

namespace tpss
{
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

template<int DIM, int NP, typename VT, bool RESIDUAL>
__global__ void __launch_bounds__(128)
k_elasticity(const VT *__restrict__ x, VT *__restrict__ y, const VT *__restrict__ b,
             const ElastTab<NP> tab, const int n, const int64_t Ns,
             const uint32_t *__restrict__ cells)
{
  constexpr int NDC = cpow(NP, DIM);
  constexpr int NF = cpow(NP, DIM - 1);
  __shared__ double U[DIM][NDC], UN[DIM][NDC], A[NDC], B[NDC];
  __shared__ double S[DIM][DIM][NDC];
  __shared__ double FA[DIM][NF], FB[DIM][DIM][NF];

  const int64_t cell = cells ? int64_t(cells[blockIdx.x]) : int64_t(blockIdx.x);
  const int t = threadIdx.x;
  const bool act = t < NDC;
  const int ts = act ? t : 0;
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
  if(act)
    for(int c = 0; c < DIM; ++c) U[c][t] = double(x[c * Ns + cell * NDC + t]);
  __syncthreads();

  // ---------------- cell term: physical gradients at the Gauss points
  for(int c = 0; c < DIM; ++c)
    for(int a = 0; a < DIM; ++a)
    {
      const double *src = U[c];
#pragma unroll
      for(int e = 0; e < DIM; ++e)
      {
        const double v = contract<DIM, NP, false>(src, e == a ? tab.G : tab.V, tid, ts, e);
        double *dst = (e % 2 == 0) ? A : B;
        __syncthreads();
        if(act) dst[t] = v;
        __syncthreads();
        src = dst;
      }
      const double g = src[ts];
      __syncthreads();
      if(act) S[c][a][t] = g;
    }
  __syncthreads();
  // stress times JxW: mu (grad u + grad u^T) + lambda div u I
  {
    double g[DIM][DIM], w = 1.0;
    for(int e = 0; e < DIM; ++e) w *= tab.w[tid[e]];
    double div = 0.0;
    for(int c = 0; c < DIM; ++c)
    {
      for(int a = 0; a < DIM; ++a) g[c][a] = S[c][a][ts];
      div += g[c][c];
    }
    __syncthreads();
    if(act)
      for(int c = 0; c < DIM; ++c)
        for(int a = 0; a < DIM; ++a)
          S[c][a][t] = w * (tab.mu * (g[c][a] + g[a][c]) + (c == a ? tab.lambda * div : 0.0));
  }
  __syncthreads();
  double acc[DIM];
  for(int c = 0; c < DIM; ++c)
  {
    acc[c] = 0.0;
    for(int a = 0; a < DIM; ++a)
    {
      const double *src = S[c][a];
#pragma unroll
      for(int e = 0; e < DIM; ++e)
      {
        const double v = contract<DIM, NP, true>(src, e == a ? tab.G : tab.V, tid, ts, e);
        double *dst = (e % 2 == 0) ? A : B;
        __syncthreads();
        if(act) dst[t] = v;
        __syncthreads();
        src = dst;
      }
      acc[c] += src[ts];
      __syncthreads();
    }
  }

  // ---------------- faces
  int64_t cstride = 1;
  for(int d = 0; d < DIM; ++d)
  {
    int e1 = -1, e2 = -1;
    for(int e = 0; e < DIM; ++e)
      if(e != d)
      {
        if(e1 < 0)
          e1 = e;
        else
          e2 = e;
      }
    for(int side = 0; side < 2; ++side)
    {
      const bool interior = side == 1 ? cc[d] < n - 1 : cc[d] > 0;
      const int64_t nb = interior ? (side == 1 ? cell + cstride : cell - cstride) : cell;
      const int nside = 1 - side;
      if(interior && act)
        for(int c = 0; c < DIM; ++c) UN[c][t] = double(x[c * Ns + nb * NDC + t]);
      __syncthreads();
      if(t < NF)
      {
        const int l1 = t % NP, l2 = DIM == 3 ? t / NP : 0;
        // traces and physical gradients of this cell (o) and the exterior state (x)
        double uo[DIM], go[DIM][DIM], ux[DIM], gx[DIM][DIM];
        for(int c = 0; c < DIM; ++c)
        {
          uo[c] = ux[c] = 0.0;
          for(int a = 0; a < DIM; ++a) go[c][a] = gx[c][a] = 0.0;
        }
        for(int k = 0; k < NDC; ++k)
        {
          int kk[3] = {0, 0, 0};
          int r = k;
          for(int e = 0; e < DIM; ++e)
          {
            kk[e] = r % NP;
            r /= NP;
          }
          const double t1v = tab.V[l1 * NP + kk[e1]], t1d = tab.G[l1 * NP + kk[e1]];
          const double t2v = DIM == 3 ? tab.V[l2 * NP + kk[e2 < 0 ? 0 : e2]] : 1.0;
          const double t2d = DIM == 3 ? tab.G[l2 * NP + kk[e2 < 0 ? 0 : e2]] : 0.0;
          {
            const double fv = tab.f[side][kk[d]], fdv = tab.fd[side][kk[d]];
            const double wv = fv * t1v * t2v, wn = fdv * t1v * t2v, w1 = fv * t1d * t2v,
                         w2 = fv * t1v * t2d;
            for(int c = 0; c < DIM; ++c)
            {
              const double val = U[c][k];
              uo[c] += wv * val;
              go[c][d] += wn * val;
              go[c][e1] += w1 * val;
              if(DIM == 3) go[c][e2] += w2 * val;
            }
          }
          if(interior)
          {
            const double fv = tab.f[nside][kk[d]], fdv = tab.fd[nside][kk[d]];
            const double wv = fv * t1v * t2v, wn = fdv * t1v * t2v, w1 = fv * t1d * t2v,
                         w2 = fv * t1v * t2d;
            for(int c = 0; c < DIM; ++c)
            {
              const double val = UN[c][k];
              ux[c] += wv * val;
              gx[c][d] += wn * val;
              gx[c][e1] += w1 * val;
              if(DIM == 3) gx[c][e2] += w2 * val;
            }
          }
        }
        if(!interior)
          for(int c = 0; c < DIM; ++c)
          {
            ux[c] = -uo[c];
            for(int a = 0; a < DIM; ++a) gx[c][a] = go[c][a];
          }
        const double nrm = side == 1 ? 1.0 : -1.0;
        const double dS = tab.w[l1] * (DIM == 3 ? tab.w[l2] : 1.0);
        double J[DIM], divo = 0.0, divx = 0.0;
        for(int c = 0; c < DIM; ++c)
        {
          J[c] = uo[c] - ux[c];
          divo += go[c][c];
          divx += gx[c][c];
        }
        for(int c = 0; c < DIM; ++c)
        {
          // {2 mu eps(u) n + lambda div(u) n}_c
          const double fo = tab.mu * (go[c][d] + go[d][c]) + (c == d ? tab.lambda * divo : 0.0);
          const double fx = tab.mu * (gx[c][d] + gx[d][c]) + (c == d ? tab.lambda * divx : 0.0);
          const double F = 0.5 * (fo + fx) * nrm;
          FA[c][t] = dS * (-F + 2.0 * tab.mu * tab.sigma * J[c] +
                           (c == d ? tab.lambda * tab.sigma * J[d] : 0.0));
          // -1/2 (2 mu eps(v) n . [u] + lambda div v [u].n)
          double fb[DIM];
          for(int a = 0; a < DIM; ++a) fb[a] = 0.0;
          fb[d] += -0.5 * tab.mu * nrm * J[c];
          if(c == d)
            for(int a = 0; a < DIM; ++a) fb[a] += -0.5 * tab.mu * nrm * J[a];
          fb[c] += -0.5 * tab.lambda * nrm * J[d];
          for(int a = 0; a < DIM; ++a) FB[c][a][t] = dS * fb[a];
        }
      }
      __syncthreads();
      if(act)
      {
        // test functions of this cell: phi_i = f(i_d) * tangential factors
        const int id = tid[d], i1 = tid[e1], i2 = DIM == 3 ? tid[e2 < 0 ? 0 : e2] : 0;
        const double fv = tab.f[side][id], fdv = tab.fd[side][id];
        for(int l = 0; l < NF; ++l)
        {
          const int l1 = l % NP, l2 = DIM == 3 ? l / NP : 0;
          const double t1v = tab.V[l1 * NP + i1], t1d = tab.G[l1 * NP + i1];
          const double t2v = DIM == 3 ? tab.V[l2 * NP + i2] : 1.0;
          const double t2d = DIM == 3 ? tab.G[l2 * NP + i2] : 0.0;
          const double wv = fv * t1v * t2v, wn = fdv * t1v * t2v, w1 = fv * t1d * t2v,
                       w2 = fv * t1v * t2d;
          for(int c = 0; c < DIM; ++c)
          {
            double s = wv * FA[c][l] + wn * FB[c][d][l] + w1 * FB[c][e1][l];
            if(DIM == 3) s += w2 * FB[c][e2 < 0 ? 0 : e2][l];
            acc[c] += s;
          }
        }
      }
      __syncthreads();
    }
    cstride *= n;
  }
  if(act)
    for(int c = 0; c < DIM; ++c)
    {
      const int64_t i = c * Ns + cell * NDC + t;
      if(RESIDUAL)
        y[i] = VT(double(b[i]) - acc[c]);
      else
        y[i] = VT(acc[c]);
    }
}

}  // namespace tpss
