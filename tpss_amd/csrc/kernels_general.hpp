// kernels_general.hpp -- SIPG Poisson operator on meshes with moved vertices
// ((bi/tri)linear cells), matrix-free with QGauss(p+1) on cells and faces:
// Laplace::DG::MF::Operator::{apply_cell, apply_face, apply_boundary}
// (apps/laplace_integrator.h:613-743).  The geometry (Jacobians at quadrature
// points) is recomputed from the cell vertices on the fly, so the only HBM
// traffic is the vectors and the (n+1)^dim vertex array.
//
// One block per cell, one thread per cell DoF / quadrature point (nq = np).
// Faces are evaluated element-centrically: every cell integrates its own test
// functions against the flux of each of its 2*dim faces (interior faces are
// thus evaluated from both sides, which is what makes the kernel write-race free).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace tpss
{
template<int NP>
struct GeneralTab
{
  double V[NP * NP];   // V[q*NP + i]  = phi_i(x_q), Gauss points x_q
  double Dv[NP * NP];  // Dv[q*NP + i] = phi_i'(x_q)
  double qx[NP], qw[NP];
  double f[2][NP];     // phi_i(0), phi_i(1)
  double fd[2][NP];    // phi_i'(0), phi_i'(1)
};

// J[a][b] = d x_a / d xi_b of the multilinear map with corner coordinates X
template<int DIM>
__device__ __forceinline__ void map_jacobian(const double (*X)[3], const double *xi,
                                             double (*J)[3])
{
  for(int a = 0; a < DIM; ++a)
    for(int b = 0; b < DIM; ++b) J[a][b] = 0.0;
  for(int c = 0; c < (1 << DIM); ++c)
  {
    for(int b = 0; b < DIM; ++b)
    {
      double w = 1.0;
      for(int e = 0; e < DIM; ++e)
      {
        const int bit = (c >> e) & 1;
        if(e == b)
          w *= bit ? 1.0 : -1.0;
        else
          w *= bit ? xi[e] : 1.0 - xi[e];
      }
      for(int a = 0; a < DIM; ++a) J[a][b] += w * X[c][a];
    }
  }
}

// Ji = J^{-1}, returns det J
template<int DIM>
__device__ __forceinline__ double invert(const double (*J)[3], double (*Ji)[3])
{
  if(DIM == 2)
  {
    const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    Ji[0][0] = J[1][1] / det;
    Ji[0][1] = -J[0][1] / det;
    Ji[1][0] = -J[1][0] / det;
    Ji[1][1] = J[0][0] / det;
    return det;
  }
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double id = 1.0 / det;
  Ji[0][0] = c00 * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = c01 * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = c02 * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  return det;
}

template<int DIM>
__device__ __forceinline__ void load_corners(const double *__restrict__ vert, const int n,
                                             const int *cc, double (*X)[3], int lane)
{
  // lanes 0 .. 2^DIM*DIM-1 load one coordinate each
  if(lane < (1 << DIM) * DIM)
  {
    const int c = lane / DIM, a = lane % DIM;
    int64_t idx = 0, st = 1;
    for(int e = 0; e < DIM; ++e)
    {
      idx += int64_t(cc[e] + ((c >> e) & 1)) * st;
      st *= (n + 1);
    }
    X[c][a] = vert[idx * DIM + a];
  }
}

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

template<int DIM, int NP, typename VT, bool RESIDUAL>
__global__ void __launch_bounds__(128, 2)
k_sipg_general(const VT *__restrict__ x, VT *__restrict__ y, const VT *__restrict__ b,
               const double *__restrict__ vert, const GeneralTab<NP> tab, const int n,
               const double ip, const int p, const uint32_t *__restrict__ cells)
{
  constexpr int NDC = cpow(NP, DIM);
  constexpr int NF = cpow(NP, DIM - 1);
  constexpr int NFACE = 2 * DIM;
  __shared__ double U[NDC], A[NDC], B[NDC];
  __shared__ double UN[NFACE][NDC];  // the face neighbours (face f = 2 d + side)
  __shared__ double F[DIM][NDC];
  __shared__ double X[8][3], XN[NFACE][8][3];
  __shared__ double FV[NFACE][NF], FH[NFACE][DIM][NF];
  // traces contracted along each face normal: value (PA) and normal derivative
  // (PB) of the own cell [0] and the neighbour [1] on the NF tangential nodes
  __shared__ double PA[NFACE][2][NF], PB[NFACE][2][NF];
  __shared__ double HP[NFACE][2];  // penalty lengths of the own cell and the neighbour
  __shared__ GeneralTab<NP> tl;  // the 1D tables, indexed at run time

  const int64_t cell = cells ? int64_t(cells[blockIdx.x]) : int64_t(blockIdx.x);
  const int t = threadIdx.x;
  if(t == 0) tl = tab;
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
  if(act) U[t] = double(x[cell * NDC + t]);
  load_corners<DIM>(vert, n, cc, X, t);
  __syncthreads();

  // ---------------- cell term: reference gradient at the quadrature point t
  double g[3] = {0, 0, 0};
#pragma unroll
  for(int a = 0; a < DIM; ++a)
  {
    // contract along every direction e with Dv (e == a) or V (e != a)
    const double *src = U;
#pragma unroll
    for(int e = 0; e < DIM; ++e)
    {
      const double v = contract<DIM, NP, false>(src, e == a ? tl.Dv : tl.V, tid, ts, e);
      double *dst = (e % 2 == 0) ? A : B;
      __syncthreads();
      if(act) dst[t] = v;
      __syncthreads();
      src = dst;
    }
    g[a] = src[ts];
    __syncthreads();
  }
  {
    double xi[3], J[3][3], Ji[3][3];
    double w = 1.0;
    for(int e = 0; e < DIM; ++e)
    {
      xi[e] = tl.qx[tid[e]];
      w *= tl.qw[tid[e]];
    }
    map_jacobian<DIM>(X, xi, J);
    const double det = invert<DIM>(J, Ji);
    const double jxw = fabs(det) * w;
    // flux in reference coordinates: Ji Ji^T g |det J| w
    double ph[3];
    for(int c = 0; c < DIM; ++c)
    {
      double s = 0.0;
      for(int bb = 0; bb < DIM; ++bb) s += Ji[bb][c] * g[bb];
      ph[c] = s;  // physical gradient
    }
    for(int a = 0; a < DIM; ++a)
    {
      double s = 0.0;
      for(int c = 0; c < DIM; ++c) s += Ji[a][c] * ph[c];
      if(act) F[a][t] = s * jxw;
    }
  }
  __syncthreads();
  double acc = 0.0;
#pragma unroll
  for(int a = 0; a < DIM; ++a)
  {
    const double *src = F[a];
#pragma unroll
    for(int e = 0; e < DIM; ++e)
    {
      const double v = contract<DIM, NP, true>(src, e == a ? tl.Dv : tl.V, tid, ts, e);
      double *dst = (e % 2 == 0) ? A : B;
      __syncthreads();
      if(act) dst[t] = v;
      __syncthreads();
      src = dst;
    }
    acc += src[ts];
    __syncthreads();
  }

  // ---------------- face terms, all 2 DIM faces of the cell at once
  const double gamma = p == 0 ? 1.0 : double(p) * (p + 1);
  auto face_interior = [&](int f) {
    const int d = f / 2;
    return (f % 2) == 1 ? cc[d] < n - 1 : cc[d] > 0;
  };
  {
    int64_t st = 1;
    for(int f = 0; f < NFACE; ++f)
    {
      const int d = f / 2, side = f % 2;
      if(d > 0 && side == 0) st *= n;
      if(!face_interior(f)) continue;
      const int64_t nb = side == 1 ? cell + st : cell - st;
      int cn[3] = {cc[0], cc[1], cc[2]};
      cn[d] += side == 1 ? 1 : -1;
      if(act) UN[f][t] = double(x[nb * NDC + t]);
      load_corners<DIM>(vert, n, cn, XN[f], t);
    }
  }
  __syncthreads();
  constexpr int ST[3] = {1, NP, NP * NP};
  // contraction along the normal: every face, both sides, all tangential nodes
  for(int o = t; o < NFACE * 2 * NF; o += blockDim.x)
  {
    const int f = o / (2 * NF), sd = (o / NF) % 2, m = o % NF;
    if(sd == 1 && !face_interior(f)) continue;
    const int d = f / 2, side = f % 2;
    const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
    const int i1 = m % NP, i2 = DIM == 3 ? m / NP : 0;
    const int base = i1 * ST[e1] + (DIM == 3 ? i2 * ST[e2] : 0);
    const double *src = sd == 0 ? U : UN[f];
    const int fs = sd == 0 ? side : 1 - side;
    double a = 0.0, bsum = 0.0;
#pragma unroll
    for(int id = 0; id < NP; ++id)
    {
      const double u = src[base + id * ST[d]];
      a += tl.f[fs][id] * u;
      bsum += tl.fd[fs][id] * u;
    }
    PA[f][sd][m] = a;
    PB[f][sd][m] = bsum;
  }
  // penalty lengths h = 1/|(J^{-1} n)_d| at the first face quadrature point of
  // both cells (apps/laplace_integrator.h:651-656, 692-693), one thread per face
  if(t < NFACE)
  {
    const int f = t, d = f / 2, side = f % 2;
    double xi0[3], J0[3][3], Ji0[3][3], n0[3], nn0 = 0.0;
    for(int e = 0; e < DIM; ++e) xi0[e] = e == d ? double(side) : tl.qx[0];
    map_jacobian<DIM>(X, xi0, J0);
    invert<DIM>(J0, Ji0);
    for(int a = 0; a < DIM; ++a)
    {
      n0[a] = Ji0[d][a] * (side == 0 ? -1.0 : 1.0);
      nn0 += n0[a] * n0[a];
    }
    nn0 = sqrt(nn0);
    double sdot = 0.0;
    for(int a = 0; a < DIM; ++a)
    {
      n0[a] /= nn0;
      sdot += Ji0[d][a] * n0[a];
    }
    HP[f][0] = 1.0 / fabs(sdot);
    HP[f][1] = 1.0;
    if(face_interior(f))
    {
      xi0[d] = double(1 - side);
      map_jacobian<DIM>(XN[f], xi0, J0);
      invert<DIM>(J0, Ji0);
      sdot = 0.0;
      for(int a = 0; a < DIM; ++a) sdot += Ji0[d][a] * n0[a];
      HP[f][1] = 1.0 / fabs(sdot);
    }
  }
  __syncthreads();
  // fluxes at every face quadrature point
  for(int pt = t; pt < NFACE * NF; pt += blockDim.x)
  {
    const int f = pt / NF, l = pt % NF;
    const int d = f / 2, side = f % 2, nside = 1 - side;
    const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
    const bool interior = face_interior(f);
    const int l1 = l % NP, l2 = DIM == 3 ? l / NP : 0;
    double xio[3], xin[3];
    const double wq = tl.qw[l1] * (DIM == 3 ? tl.qw[l2] : 1.0);
    for(int e = 0; e < DIM; ++e)
    {
      const double qe = e == e1 ? tl.qx[l1] : tl.qx[l2];
      xio[e] = e == d ? double(side) : qe;
      xin[e] = e == d ? double(nside) : qe;
    }
    // traces: value, normal and the two tangential reference derivatives
    double uo = 0, un = 0, gon = 0, got1 = 0, got2 = 0, gnn = 0, gnt1 = 0, gnt2 = 0;
    for(int m = 0; m < NF; ++m)
    {
      const int i1 = m % NP, i2 = DIM == 3 ? m / NP : 0;
      const double t1v = tl.V[l1 * NP + i1], t1d = tl.Dv[l1 * NP + i1];
      const double t2v = DIM == 3 ? tl.V[l2 * NP + i2] : 1.0;
      const double t2d = DIM == 3 ? tl.Dv[l2 * NP + i2] : 0.0;
      const double wvv = t1v * t2v, wdv = t1d * t2v, wvd = t1v * t2d;
      const double a0 = PA[f][0][m];
      uo += wvv * a0;
      gon += wvv * PB[f][0][m];
      got1 += wdv * a0;
      got2 += wvd * a0;
      if(interior)
      {
        const double a1 = PA[f][1][m];
        un += wvv * a1;
        gnn += wvv * PB[f][1][m];
        gnt1 += wdv * a1;
        gnt2 += wvd * a1;
      }
    }
    double J[3][3], Jio[3][3];
    map_jacobian<DIM>(X, xio, J);
    const double deto = invert<DIM>(J, Jio);
    // outward unit normal of this cell and surface element (Nanson)
    double nv[3] = {0, 0, 0}, nn = 0.0;
    for(int a = 0; a < DIM; ++a)
    {
      nv[a] = Jio[d][a] * (side == 0 ? -1.0 : 1.0);
      nn += nv[a] * nv[a];
    }
    nn = sqrt(nn);
    for(int a = 0; a < DIM; ++a) nv[a] /= nn;
    const double dS = fabs(deto) * nn * wq;
    // reference-gradient coefficients of the normal derivative: co = J^{-1} n
    double co[3] = {0, 0, 0};
    for(int bb = 0; bb < DIM; ++bb)
    {
      double s2 = 0.0;
      for(int a = 0; a < DIM; ++a) s2 += Jio[bb][a] * nv[a];
      co[bb] = s2;
    }
    const double dno = gon * co[d] + got1 * co[e1] + (DIM == 3 ? got2 * co[e2] : 0.0);
    double dnn = 0.0;
    if(interior)
    {
      double Jn[3][3], Jin[3][3];
      map_jacobian<DIM>(XN[f], xin, Jn);
      invert<DIM>(Jn, Jin);
      double cn[3] = {0, 0, 0};
      for(int bb = 0; bb < DIM; ++bb)
      {
        double s2 = 0.0;
        for(int a = 0; a < DIM; ++a) s2 += Jin[bb][a] * nv[a];
        cn[bb] = s2;
      }
      dnn = gnn * cn[d] + gnt1 * cn[e1] + (DIM == 3 ? gnt2 * cn[e2] : 0.0);
    }
    double fval, fnd;
    if(interior)
    {
      const double sigma = ip * 0.5 * 2.0 * gamma * (0.5 / HP[f][0] + 0.5 / HP[f][1]);
      const double jump = uo - un, avg = 0.5 * (dno + dnn);
      fval = dS * (jump * sigma - avg);
      fnd = dS * (-0.5 * jump);
    }
    else
    {
      const double sigma = ip * 2.0 * gamma / HP[f][0];
      fval = dS * (uo * sigma - dno);
      fnd = dS * (-uo);
    }
    FV[f][l] = fval;
    for(int bb = 0; bb < DIM; ++bb) FH[f][bb][l] = fnd * co[bb];
  }
  __syncthreads();
  if(act)
  {
    // test functions of this cell: phi_i = f(i_d) * tangential factors, all faces
    for(int f = 0; f < NFACE; ++f)
    {
      const int d = f / 2, side = f % 2;
      const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
      const int id = tid[d], i1 = tid[e1], i2 = DIM == 3 ? tid[e2] : 0;
      const double fv = tl.f[side][id], fdv = tl.fd[side][id];
      double s = 0.0;
      for(int l = 0; l < NF; ++l)
      {
        const int l1 = l % NP, l2 = DIM == 3 ? l / NP : 0;
        const double t1v = tl.V[l1 * NP + i1], t1d = tl.Dv[l1 * NP + i1];
        const double t2v = DIM == 3 ? tl.V[l2 * NP + i2] : 1.0;
        const double t2d = DIM == 3 ? tl.Dv[l2 * NP + i2] : 0.0;
        s += fv * t1v * t2v * FV[f][l];
        s += fdv * t1v * t2v * FH[f][d][l];
        s += fv * t1d * t2v * FH[f][e1][l];
        if(DIM == 3) s += fv * t1v * t2d * FH[f][e2][l];
      }
      acc += s;
    }
  }
  if(act)
  {
    if(RESIDUAL)
      y[cell * NDC + t] = VT(double(b[cell * NDC + t]) - acc);
    else
      y[cell * NDC + t] = VT(acc);
  }
}

}  // namespace tpss
