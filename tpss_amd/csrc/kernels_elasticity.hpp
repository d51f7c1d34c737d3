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
// (QGauss(p+1)) handling all components; arithmetic in the vector type.
//   cell term: the gradients of all components by sum factorisation (x, y, z
//     stages sharing their 1D interpolations), the stress times JxW, and its
//     integration against the test gradients (z, y, x stages);
//   faces, one direction at a time (both faces of the cell): traces of both
//     sides contracted along the normal (the neighbour's straight from global
//     memory; the face neighbours' cell indices are formed once per block),
//     fluxes per face point, fluxes projected on the tangential test
//     functions (3D: sum-factorised, one tangential direction after the other,
//     through the cell-stage arrays, which the faces leave free), and every
//     thread adds the terms of its test function.
// Faces are evaluated element-centrically (every output entry is written by
// exactly one thread).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels.hpp"

namespace tpss
{
template<int NP, typename T>
struct ElastTab
{
  T V[NP * NP];   // V[q*NP + i] = phi_i(x_q)
  T G[NP * NP];   // G[q*NP + i] = phi_i'(x_q) / h
  T f[2][NP];     // phi_i(0), phi_i(1)
  T fd[2][NP];    // phi_i'(0) / h, phi_i'(1) / h
  T w[NP];        // h * Gauss weight
  T mu, lambda, sigma;
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
__global__ void __launch_bounds__((elasticity_threads<DIM, NP>()))
k_elasticity(const VT *__restrict__ x, VT *__restrict__ y, const VT *__restrict__ b,
             const ElastTab<NP, VT> tab, const int n, const int64_t Ns,
             const uint32_t *__restrict__ cells, const int64_t count)
{
  using T = VT;
  constexpr int NDC = cpow(NP, DIM);
  constexpr int NF = cpow(NP, DIM - 1);
  constexpr int CPB = elasticity_cells_per_block<DIM, NP>();
  constexpr int NT = elasticity_threads<DIM, NP>();
  constexpr int ST[3] = {1, NP, NP * NP};
  constexpr int NS = DIM * DIM;  // stage arrays per cell
  __shared__ T U[CPB][DIM][NDC];
  __shared__ T R1[CPB][NS][NDC], R2[CPB][NS][NDC];  // stage arrays, used in turn
  // normal traces [face side][cell side][component][tangential node]; later the
  // projections of the fluxes on the tangential test functions ([.][0][.][.])
  __shared__ T PA[CPB][2][2][DIM][NF], PB[CPB][2][2][DIM][NF];
  __shared__ T FA[CPB][2][DIM][NF], FB[CPB][2][DIM][DIM][NF];
  __shared__ int64_t nbc[CPB][2 * DIM];  // face neighbours [2 D + side] (-1: boundary face)
  __shared__ ElastTab<NP, T> tl;

  const int t = threadIdx.x;
  const int cs = t / NDC, ts0 = t % NDC;
  const int64_t slot = int64_t(blockIdx.x) * CPB + cs;
  const bool act = cs < CPB && slot < count;
  const int c0 = cs < CPB ? cs : 0;  // LDS slot of this thread (idle threads alias slot 0)
  const int ts = act ? ts0 : 0;
  auto cell_of = [&](int s) -> int64_t {
    const int64_t sl = int64_t(blockIdx.x) * CPB + s;
    return sl < count ? (cells ? int64_t(cells[sl]) : sl) : int64_t(-1);
  };
  const int64_t cell = act ? cell_of(cs) : 0;
  // slot s of the block holds a cell (no load of the cell list)
  auto has_cell = [&](int s) { return int64_t(blockIdx.x) * CPB + s < count; };
  const int ix = ts % NP, iy = (ts / NP) % NP, iz = DIM == 3 ? ts / (NP * NP) : 0;
  const int ti[3] = {ix, iy, iz};

  {
    const T *src = reinterpret_cast<const T *>(&tab);
    T *dst = reinterpret_cast<T *>(&tl);
    constexpr int NTAB = int(sizeof(ElastTab<NP, T>) / sizeof(T));
    for(int i = t; i < NTAB; i += NT) dst[i] = src[i];
  }
  if(t < CPB * 2 * DIM)
  {
    const int s = t / (2 * DIM), f = t % (2 * DIM), D = f >> 1, side = f & 1;
    const int64_t ce = cell_of(s);
    int64_t nb = -1;
    if(ce >= 0)
    {
      const int64_t cstride = D == 0 ? int64_t(1) : D == 1 ? int64_t(n) : int64_t(n) * n;
      const int cd = int((ce / cstride) % n);
      if(side == 1 ? cd < n - 1 : cd > 0) nb = side == 1 ? ce + cstride : ce - cstride;
    }
    nbc[s][f] = nb;
  }
  if(act)
#pragma unroll
    for(int c = 0; c < DIM; ++c) U[c0][c][ts] = x[c * Ns + cell * NDC + ts];
  __syncthreads();

  // ---------------- cell term
  T acc[DIM];
#pragma unroll
  for(int c = 0; c < DIM; ++c) acc[c] = T(0);
  {
    // x-stage: R1[2c] = V_x u_c, R1[2c+1] = G_x u_c
    if(act)
    {
      const int base = ts - ix;
#pragma unroll
      for(int c = 0; c < DIM; ++c)
      {
        T v = 0, g = 0;
#pragma unroll
        for(int j = 0; j < NP; ++j)
        {
          const T u = U[c0][c][base + j];
          v += tl.V[ix * NP + j] * u;
          g += tl.G[ix * NP + j] * u;
        }
        R1[c0][2 * c][ts] = v;
        R1[c0][2 * c + 1][ts] = g;
      }
    }
    __syncthreads();
    T g[DIM][DIM];  // g[c][a] = d u_c / d x_a at the quadrature point ts
    if constexpr(DIM == 3)
    {
      // y-stage: R2[3c] = V_y G_x u_c, R2[3c+1] = G_y V_x u_c, R2[3c+2] = V_y V_x u_c
      if(act)
      {
        const int base = ts - iy * NP;
#pragma unroll
        for(int c = 0; c < DIM; ++c)
        {
          T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
          for(int j = 0; j < NP; ++j)
          {
            const T av = R1[c0][2 * c][base + j * NP], ag = R1[c0][2 * c + 1][base + j * NP];
            const T vy = tl.V[iy * NP + j];
            s0 += vy * ag;
            s1 += tl.G[iy * NP + j] * av;
            s2 += vy * av;
          }
          R2[c0][3 * c][ts] = s0;
          R2[c0][3 * c + 1][ts] = s1;
          R2[c0][3 * c + 2][ts] = s2;
        }
      }
      __syncthreads();
      // z-stage into registers
      const int base = ts - iz * NP * NP;
#pragma unroll
      for(int c = 0; c < DIM; ++c)
      {
        T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for(int j = 0; j < NP; ++j)
        {
          const T vz = tl.V[iz * NP + j];
          s0 += vz * R2[c0][3 * c][base + j * NP * NP];
          s1 += vz * R2[c0][3 * c + 1][base + j * NP * NP];
          s2 += tl.G[iz * NP + j] * R2[c0][3 * c + 2][base + j * NP * NP];
        }
        g[c][0] = s0;
        g[c][1] = s1;
        g[c][2] = s2;
      }
    }
    else
    {
      const int base = ts - iy * NP;
#pragma unroll
      for(int c = 0; c < DIM; ++c)
      {
        T s0 = 0, s1 = 0;
#pragma unroll
        for(int j = 0; j < NP; ++j)
        {
          s0 += tl.V[iy * NP + j] * R1[c0][2 * c + 1][base + j * NP];
          s1 += tl.G[iy * NP + j] * R1[c0][2 * c][base + j * NP];
        }
        g[c][0] = s0;
        g[c][1] = s1;
      }
    }
    // stress times JxW: mu (grad u + grad u^T) + lambda div u I, into the array
    // not read in the last stage (3D: R1, 2D: R2)
    {
      T w = T(1), div = T(0);
#pragma unroll
      for(int e = 0; e < DIM; ++e) w *= tl.w[ti[e]];
#pragma unroll
      for(int c = 0; c < DIM; ++c) div += g[c][c];
      T(*SG)[NDC] = DIM == 3 ? R1[c0] : R2[c0];
      if(act)
#pragma unroll
        for(int c = 0; c < DIM; ++c)
#pragma unroll
          for(int a = 0; a < DIM; ++a)
            SG[c * DIM + a][ts] =
              w * (tl.mu * (g[c][a] + g[a][c]) + (c == a ? tl.lambda * div : T(0)));
    }
    __syncthreads();
    if constexpr(DIM == 3)
    {
      // z-integration: R2[3c] = V_z^T s_c0, R2[3c+1] = V_z^T s_c1, R2[3c+2] = G_z^T s_c2
      if(act)
      {
        const int base = ts - iz * NP * NP;
#pragma unroll
        for(int c = 0; c < DIM; ++c)
        {
          T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
          for(int q = 0; q < NP; ++q)
          {
            const T vz = tl.V[q * NP + iz];
            s0 += vz * R1[c0][c * 3][base + q * NP * NP];
            s1 += vz * R1[c0][c * 3 + 1][base + q * NP * NP];
            s2 += tl.G[q * NP + iz] * R1[c0][c * 3 + 2][base + q * NP * NP];
          }
          R2[c0][3 * c][ts] = s0;
          R2[c0][3 * c + 1][ts] = s1;
          R2[c0][3 * c + 2][ts] = s2;
        }
      }
      __syncthreads();
      // y-integration: R1[2c] = V_y^T (.)0, R1[2c+1] = G_y^T (.)1 + V_y^T (.)2
      if(act)
      {
        const int base = ts - iy * NP;
#pragma unroll
        for(int c = 0; c < DIM; ++c)
        {
          T s0 = 0, s1 = 0;
#pragma unroll
          for(int q = 0; q < NP; ++q)
          {
            const T vy = tl.V[q * NP + iy];
            s0 += vy * R2[c0][3 * c][base + q * NP];
            s1 += tl.G[q * NP + iy] * R2[c0][3 * c + 1][base + q * NP] +
                  vy * R2[c0][3 * c + 2][base + q * NP];
          }
          R1[c0][2 * c][ts] = s0;
          R1[c0][2 * c + 1][ts] = s1;
        }
      }
    }
    else
    {
      // y-integration: R1[2c] = V_y^T s_c0, R1[2c+1] = G_y^T s_c1
      if(act)
      {
        const int base = ts - iy * NP;
#pragma unroll
        for(int c = 0; c < DIM; ++c)
        {
          T s0 = 0, s1 = 0;
#pragma unroll
          for(int q = 0; q < NP; ++q)
          {
            s0 += tl.V[q * NP + iy] * R2[c0][c * 2][base + q * NP];
            s1 += tl.G[q * NP + iy] * R2[c0][c * 2 + 1][base + q * NP];
          }
          R1[c0][2 * c][ts] = s0;
          R1[c0][2 * c + 1][ts] = s1;
        }
      }
    }
    __syncthreads();
    // x-integration
    if(act)
    {
      const int base = ts - ix;
#pragma unroll
      for(int c = 0; c < DIM; ++c)
      {
        T s = 0;
#pragma unroll
        for(int q = 0; q < NP; ++q)
          s += tl.G[q * NP + ix] * R1[c0][2 * c][base + q] +
               tl.V[q * NP + ix] * R1[c0][2 * c + 1][base + q];
        acc[c] = s;
      }
    }
  }

  // ---------------- faces, direction D fixed at compile time (both faces)
  auto face_dir = [&](auto Dc) {
    constexpr int D = decltype(Dc)::value;
    constexpr int E1 = D == 0 ? 1 : 0;
    constexpr int E2 = DIM == 3 ? (D == 2 ? 1 : 2) : E1;
    __syncthreads();  // PA / PB of the previous direction have been read
    // traces along the normal: value (PA) and normal derivative (PB) of both
    // cell sides for every component and tangential node; the cell's own side
    // (from LDS) and the neighbour's (from global memory) in separate passes, so
    // that a wave runs one of the two branches per iteration
#pragma unroll
    for(int sd = 0; sd < 2; ++sd)
      for(int o = t; o < CPB * 2 * DIM * NF; o += NT)
      {
        const int s = o / (2 * DIM * NF), side = (o / (DIM * NF)) % 2;
        const int c = (o / NF) % DIM, m = o % NF;
        if(!has_cell(s)) continue;
        const bool inner = nbc[s][2 * D + side] >= 0;
        if(sd == 1 && !inner) continue;
        const int i1 = m % NP, i2 = DIM == 3 ? m / NP : 0;
        const int base = i1 * ST[E1] + (DIM == 3 ? i2 * ST[E2] : 0);
        const int fs = sd == 0 ? side : 1 - side;
        T a = 0, bs = 0;
        if(sd == 0)
        {
#pragma unroll
          for(int id = 0; id < NP; ++id)
          {
            const T u = U[s][c][base + id * ST[D]];
            a += tl.f[fs][id] * u;
            bs += tl.fd[fs][id] * u;
          }
        }
        else
        {
          const VT *src = x + c * Ns + nbc[s][2 * D + side] * NDC + base;
#pragma unroll
          for(int id = 0; id < NP; ++id)
          {
            const T u = src[id * ST[D]];
            a += tl.f[fs][id] * u;
            bs += tl.fd[fs][id] * u;
          }
        }
        PA[s][side][sd][c][m] = a;
        PB[s][side][sd][c][m] = bs;
      }
    __syncthreads();
    // fluxes at every face point
    for(int o = t; o < CPB * 2 * NF; o += NT)
    {
      const int s = o / (2 * NF), side = (o / NF) % 2, l = o % NF;
      if(!has_cell(s)) continue;
      const bool inner = nbc[s][2 * D + side] >= 0;
      const int l1 = l % NP, l2 = DIM == 3 ? l / NP : 0;
      // traces and physical gradients of this cell (o) and the exterior state (x)
      T uo[DIM], go[DIM][DIM], ux[DIM], gx[DIM][DIM];
#pragma unroll
      for(int c = 0; c < DIM; ++c)
      {
        uo[c] = ux[c] = T(0);
#pragma unroll
        for(int a = 0; a < DIM; ++a) go[c][a] = gx[c][a] = T(0);
      }
#pragma unroll
      for(int m = 0; m < NF; ++m)
      {
        const int i1 = m % NP, i2 = DIM == 3 ? m / NP : 0;
        const T t1v = tl.V[l1 * NP + i1], t1d = tl.G[l1 * NP + i1];
        const T t2v = DIM == 3 ? tl.V[l2 * NP + i2] : T(1);
        const T t2d = DIM == 3 ? tl.G[l2 * NP + i2] : T(0);
        const T wvv = t1v * t2v, wdv = t1d * t2v, wvd = t1v * t2d;
#pragma unroll
        for(int c = 0; c < DIM; ++c)
        {
          const T a = PA[s][side][0][c][m];
          uo[c] += wvv * a;
          go[c][D] += wvv * PB[s][side][0][c][m];
          go[c][E1] += wdv * a;
          if(DIM == 3) go[c][E2] += wvd * a;
          if(inner)
          {
            const T a1 = PA[s][side][1][c][m];
            ux[c] += wvv * a1;
            gx[c][D] += wvv * PB[s][side][1][c][m];
            gx[c][E1] += wdv * a1;
            if(DIM == 3) gx[c][E2] += wvd * a1;
          }
        }
      }
      if(!inner)
#pragma unroll
        for(int c = 0; c < DIM; ++c)
        {
          ux[c] = -uo[c];
#pragma unroll
          for(int a = 0; a < DIM; ++a) gx[c][a] = go[c][a];
        }
      const T nrm = side == 1 ? T(1) : T(-1);
      const T dS = tl.w[l1] * (DIM == 3 ? tl.w[l2] : T(1));
      T J[DIM], divo = 0, divx = 0;
#pragma unroll
      for(int c = 0; c < DIM; ++c)
      {
        J[c] = uo[c] - ux[c];
        divo += go[c][c];
        divx += gx[c][c];
      }
      const T hm = T(-0.5) * tl.mu * nrm, hl = T(-0.5) * tl.lambda * nrm;
#pragma unroll
      for(int c = 0; c < DIM; ++c)
      {
        // {2 mu eps(u) n + lambda div(u) n}_c
        const T fo = tl.mu * (go[c][D] + go[D][c]) + (c == D ? tl.lambda * divo : T(0));
        const T fx = tl.mu * (gx[c][D] + gx[D][c]) + (c == D ? tl.lambda * divx : T(0));
        const T F = T(0.5) * (fo + fx) * nrm;
        FA[s][side][c][l] = dS * (-F + T(2) * tl.mu * tl.sigma * J[c] +
                                  (c == D ? tl.lambda * tl.sigma * J[D] : T(0)));
        // -1/2 (2 mu eps(v) n . [u] + lambda div v [u].n)
        T fb[DIM];
#pragma unroll
        for(int a = 0; a < DIM; ++a) fb[a] = T(0);
        fb[D] += hm * J[c];
        if(c == D)
#pragma unroll
          for(int a = 0; a < DIM; ++a) fb[a] += hm * J[a];
        fb[c] += hl * J[D];
#pragma unroll
        for(int a = 0; a < DIM; ++a) FB[s][side][c][a][l] = dS * fb[a];
      }
    }
    __syncthreads();
    // projections on the tangential test functions (into PA / PB [.][0])
    if constexpr(DIM == 3)
    {
      // sum-factorised: along e1 into the cell-stage arrays (free during the faces),
      // SA = sum_l1 phi_i1 FA + phi'_i1 FB_E1, SB = sum_l1 phi_i1 FB_E2,
      // SC = sum_l1 phi_i1 FB_D at (i1, l2); then along e2
      constexpr int NQ = CPB * 2 * DIM * NF;
      static_assert(3 * NQ <= CPB * NS * NDC, "projection stage fits the stage arrays");
      T *SA = &R2[0][0][0], *SB = SA + NQ, *SC = SA + 2 * NQ;
      for(int o = t; o < NQ; o += NT)
      {
        const int s = o / (2 * DIM * NF), side = (o / (DIM * NF)) % 2;
        const int c = (o / NF) % DIM, m = o % NF;
        if(!has_cell(s)) continue;
        const int i1 = m % NP, l2 = m / NP;
        T sa = 0, sb = 0, sc = 0;
#pragma unroll
        for(int l1 = 0; l1 < NP; ++l1)
        {
          const int l = l1 + NP * l2;
          const T v = tl.V[l1 * NP + i1], g = tl.G[l1 * NP + i1];
          sa += v * FA[s][side][c][l] + g * FB[s][side][c][E1][l];
          sb += v * FB[s][side][c][E2][l];
          sc += v * FB[s][side][c][D][l];
        }
        SA[o] = sa;
        SB[o] = sb;
        SC[o] = sc;
      }
      __syncthreads();
      for(int o = t; o < NQ; o += NT)
      {
        const int s = o / (2 * DIM * NF), side = (o / (DIM * NF)) % 2;
        const int c = (o / NF) % DIM, m = o % NF;
        if(!has_cell(s)) continue;
        const int i1 = m % NP, i2 = m / NP, base = o - m;
        T pv = 0, pd = 0;
#pragma unroll
        for(int l2 = 0; l2 < NP; ++l2)
        {
          const int r = base + i1 + NP * l2;
          const T v = tl.V[l2 * NP + i2], g = tl.G[l2 * NP + i2];
          pv += v * SA[r] + g * SB[r];
          pd += v * SC[r];
        }
        PA[s][side][0][c][m] = pv;
        PB[s][side][0][c][m] = pd;
      }
    }
    else
    {
      // 2D: one tangential direction (NF = NP)
      for(int o = t; o < CPB * 2 * DIM * NF; o += NT)
      {
        const int s = o / (2 * DIM * NF), side = (o / (DIM * NF)) % 2;
        const int c = (o / NF) % DIM, i1 = o % NF;
        if(!has_cell(s)) continue;
        T pv = 0, pd = 0;
#pragma unroll
        for(int l = 0; l < NF; ++l)
        {
          const T v = tl.V[l * NP + i1];
          pv += v * FA[s][side][c][l] + tl.G[l * NP + i1] * FB[s][side][c][E1][l];
          pd += v * FB[s][side][c][D][l];
        }
        PA[s][side][0][c][i1] = pv;
        PB[s][side][0][c][i1] = pd;
      }
    }
    __syncthreads();
    if(act)
    {
      // test functions of this cell: phi_i = f(i_D) * tangential factors
      const int id = ti[D];
      const int m = ti[E1] + (DIM == 3 ? NP * ti[E2] : 0);
#pragma unroll
      for(int side = 0; side < 2; ++side)
      {
        const T fv = tl.f[side][id], fdv = tl.fd[side][id];
#pragma unroll
        for(int c = 0; c < DIM; ++c)
          acc[c] += fv * PA[c0][side][0][c][m] + fdv * PB[c0][side][0][c][m];
      }
    }
  };
  face_dir(std::integral_constant<int, 0>{});
  face_dir(std::integral_constant<int, 1>{});
  if constexpr(DIM == 3) face_dir(std::integral_constant<int, DIM == 3 ? 2 : 1>{});

  if(act)
#pragma unroll
    for(int c = 0; c < DIM; ++c)
    {
      const int64_t i = c * Ns + cell * NDC + ts;
      if(RESIDUAL)
        y[i] = b[i] - acc[c];
      else
        y[i] = acc[c];
    }
}

}  // namespace tpss
