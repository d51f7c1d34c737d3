// kernels_general.hpp -- SIPG Poisson operator on meshes with moved vertices
// ((bi/tri)linear cells), matrix-free with QGauss(p+1) on cells and faces:
// Laplace::DG::MF::Operator::{apply_cell, apply_face, apply_boundary}
// (apps/laplace_integrator.h:613-743).  The geometry (Jacobians at quadrature
// points) is recomputed on the fly from the multilinear coefficients of each
// cell's map, so the only HBM traffic is the vectors and the (n+1)^dim vertex
// array.
//
// One block per cell, one thread per cell DoF / quadrature point (nq = np);
// arithmetic in the vector type (FP32 levels compute in FP32, FP64 levels in
// FP64; the map coefficients are formed in FP64 from the vertices).  Faces are
// evaluated element-centrically: every cell integrates its own test functions
// against the flux of each of its 2*dim faces (interior faces are evaluated
// from both sides, which keeps the kernel free of write races).
//
// Every contraction is sum-factorised, and independent work shares barriers:
//   P0 load u, the face neighbours, the corners of the cell and the far corners
//      of the face neighbours
//   P1 x-interpolation of u (value, derivative); map coefficients of the cell;
//      per face the tangent and normal-direction derivatives of both maps as
//      polynomials in the face coordinates
//   P2 y-interpolation; face traces along the normals (value, normal derivative)
//   P3 z-interpolation -> reference gradient -> geometry -> reference flux;
//      face traces along the first tangential direction
//   P4 z-integration; face traces along the second tangential direction,
//      face geometry and fluxes per face point
//   P5 y-integration; face fluxes projected on the tangential test functions
//      (3D: sum-factorised, first along e1, then -- after a barrier -- along e2)
//   P6 x-integration; add the face contributions of this test function
//
// Face geometry without 3x3 inverses.  The two cells of a face share the face's
// vertices, so on the face their maps have the same tangent vectors t1, t2 (the
// Jacobian columns of the tangential directions) and differ only in the column
// J_d of the normal direction.  With w = t1 x t2 (2D: t1 turned by 90 degrees),
// the outward unit normal is n = +-w/|w| (the sign makes n . J_d point away from
// the cell), the surface element is |det J| |J^{-T} e_d| = |w|, and the
// reference-gradient coefficients c = J^{-1} n of the normal derivative follow
// from J c = n: c_d = 1/(J_d . n) (n is orthogonal to t1, t2), and the tangential
// part solves the 2x2 Gram system G c_t = -c_d (t1 . J_d, t2 . J_d), G = (ti . tj);
// the neighbour's coefficients use its own J_d with the same t1, t2, n and G.
// The penalty length h = 1/|c_d| = |J_d . n| needs no division.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace tpss
{
template<int NP, typename T>
struct GeneralTab
{
  T V[NP * NP];   // V[q*NP + i]  = phi_i(x_q), Gauss points x_q
  T Dv[NP * NP];  // Dv[q*NP + i] = phi_i'(x_q)
  T qx[NP], qw[NP];
  T f[2][NP];     // phi_i(0), phi_i(1)
  T fd[2][NP];    // phi_i'(0), phi_i'(1)
};

// Jacobian J[a][b] = d x_a / d xi_b of the multilinear map
//   x(xi) = sum_S A[S] prod_{e in S} xi_e   (S: subsets of the directions, bit masks)
template<int DIM, typename T>
__device__ __forceinline__ void coef_jacobian(const T (*A)[3], const T *xi, T (*J)[3])
{
#pragma unroll
  for(int a = 0; a < DIM; ++a)
#pragma unroll
    for(int b = 0; b < DIM; ++b) J[a][b] = T(0);
#pragma unroll
  for(int S = 1; S < (1 << DIM); ++S)
#pragma unroll
    for(int b = 0; b < DIM; ++b)
    {
      if(!((S >> b) & 1)) continue;
      T w = T(1);
#pragma unroll
      for(int e = 0; e < DIM; ++e)
        if(e != b && ((S >> e) & 1)) w *= xi[e];
#pragma unroll
      for(int a = 0; a < DIM; ++a) J[a][b] += w * A[S][a];
    }
}

// 1/a: the hardware reciprocal in FP32 (1 ulp), an exact division in FP64
template<typename T>
__device__ __forceinline__ T recip(T a)
{
  if constexpr(std::is_same<T, float>::value)
    return __builtin_amdgcn_rcpf(a);
  else
    return T(1) / a;
}

template<typename T>
__device__ __forceinline__ T root(T a)
{
  if constexpr(std::is_same<T, float>::value)
    return __builtin_amdgcn_sqrtf(a);
  else
    return sqrt(a);
}

// Ji = J^{-1}, returns det J
template<int DIM, typename T>
__device__ __forceinline__ T invert(const T (*J)[3], T (*Ji)[3])
{
  if constexpr(DIM == 2)
  {
    const T det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const T id = recip(det);
    Ji[0][0] = J[1][1] * id;
    Ji[0][1] = -J[0][1] * id;
    Ji[1][0] = -J[1][0] * id;
    Ji[1][1] = J[0][0] * id;
    return det;
  }
  else
  {
    const T c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const T c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const T c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const T det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const T id = recip(det);
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
}

template<int DIM, int NP, typename VT, bool RESIDUAL>
__global__ void __launch_bounds__((cpow(NP, DIM) <= 64 ? 64 : 128), (std::is_same<VT, float>::value ? 5 : 1))
k_sipg_general(const VT *__restrict__ x, VT *__restrict__ y, const VT *__restrict__ b,
               const double *__restrict__ vert, const GeneralTab<NP, VT> tab, const int n,
               const double ip, const int p, const uint32_t *__restrict__ cells)
{
  using T = VT;
  constexpr int NDC = cpow(NP, DIM);
  constexpr int NF = cpow(NP, DIM - 1);
  constexpr int NFACE = 2 * DIM;
  constexpr int NCORN = 1 << DIM;
  constexpr int NT = NDC <= 64 ? 64 : 128;  // block size
  // lexicographic stride of direction e (by selection: a constant array indexed at
  // run time would be read from memory)
  auto stride = [](int e) { return e == 0 ? 1 : (e == 1 ? NP : NP * NP); };
  constexpr int NP2 = DIM == 3 ? NP : 1;    // tangential nodes along the second face direction

  __shared__ T U[NDC];
  __shared__ T A0[NDC], A1[NDC], A2[NDC];   // x-stage (gradient) / F (flux) / y-stage (integration)
  __shared__ T B0[NDC], B1[NDC], B2[NDC];   // y-stage (gradient) / z-stage (integration)
  __shared__ T PA[NFACE][2][NF], PB[NFACE][2][NF];  // normal traces; later the projections
  constexpr int NHC = NCORN / 2;            // corners per face
  constexpr int NFV = DIM == 3 ? 11 : 5;    // face polynomials per face
  __shared__ T C[NCORN][3];                 // map coefficients of the cell
  // per face: tangents (3D: t1 = [0] + q2 [2], t2 = [1] + q1 [2]; 2D: t1 = [0]) and
  // the normal-direction Jacobian columns of the cell (3D [3..6], 2D [1..2]) and of
  // the neighbour (3D [7..10], 2D [3..4]) as bilinear (2D: linear) polynomials in
  // the face coordinates (q1, q2)
  __shared__ T FG[NFACE][NFV][3];
  __shared__ T HI[NFACE][2];                // 1/h of the cell and of the neighbour
  __shared__ GeneralTab<NP, T> tl;
  // storage shared by arrays with disjoint lifetimes: the face neighbours (P0-P2)
  // with the face fluxes (P4-P5); the corners (P0-P1) with the tangential traces (P3-P4)
  constexpr int NUN = NFACE * NDC, NFLUX = NFACE * NF * (2 + DIM);
  constexpr int NTR = 3 * NFACE * 2 * NF, NXD = (NCORN + NFACE * NHC) * 3;
  __shared__ T RU[NUN > NFLUX ? NUN : NFLUX];
  __shared__ double RX[(NTR * int(sizeof(T)) + 7) / 8 > NXD ? (NTR * int(sizeof(T)) + 7) / 8 : NXD];
  T(*UN)[NDC] = reinterpret_cast<T(*)[NDC]>(RU);
  T(*JV)[NF] = reinterpret_cast<T(*)[NF]>(RU);
  T(*AV)[NF] = reinterpret_cast<T(*)[NF]>(RU + NFACE * NF);
  T(*FH)[DIM][NF] = reinterpret_cast<T(*)[DIM][NF]>(RU + 2 * NFACE * NF);
  // X[c]: corner c of the cell (c < NCORN), X[NCORN + f * NHC + c']: the corner of
  // the neighbour across face f opposite face corner c'
  double(*X)[3] = reinterpret_cast<double(*)[3]>(RX);
  T(*VA)[2][NF] = reinterpret_cast<T(*)[2][NF]>(RX);
  T(*DA)[2][NF] = reinterpret_cast<T(*)[2][NF]>(reinterpret_cast<T *>(RX) + NFACE * 2 * NF);
  T(*VB)[2][NF] = reinterpret_cast<T(*)[2][NF]>(reinterpret_cast<T *>(RX) + 2 * NFACE * 2 * NF);

  const int64_t cell = cells ? int64_t(cells[blockIdx.x]) : int64_t(blockIdx.x);
  const int t = threadIdx.x;
  const bool act = t < NDC;
  const int ts = act ? t : 0;
  // faces with a neighbour (bit f = 2 d + side) and the index of the cell's first
  // vertex (vertex indices fit in 32 bits: (n + 1)^3 < 2^31 for n <= 1000)
  int imask = 0, vbase = 0;
  {
    int r = int(cell);  // cells < n^3 <= 1000^3 < 2^31
    int vst = 1;
#pragma unroll
    for(int d = 0; d < DIM; ++d)
    {
      const int cd = r % n;
      r /= n;
      imask |= (cd > 0 ? 1 : 0) << (2 * d);
      imask |= (cd < n - 1 ? 1 : 0) << (2 * d + 1);
      vbase += cd * vst;
      vst *= n + 1;
    }
  }
  const int ix = ts % NP, iy = DIM > 1 ? (ts / NP) % NP : 0, iz = DIM > 2 ? ts / (NP * NP) : 0;
  auto interior = [&](int f) { return ((imask >> f) & 1) != 0; };
  auto nb_stride = [&](int d) {
    int64_t st = 1;
    for(int e = 0; e < d; ++e) st *= n;
    return st;
  };

  // ---------------- P0: loads
  {
    const T *src = reinterpret_cast<const T *>(&tab);
    T *dst = reinterpret_cast<T *>(&tl);
    constexpr int NTAB = int(sizeof(GeneralTab<NP, T>) / sizeof(T));
    for(int i = t; i < NTAB; i += NT) dst[i] = src[i];
  }
  if(act)
  {
    U[t] = x[cell * NDC + t];
#pragma unroll
    for(int f = 0; f < NFACE; ++f)
      if(interior(f))
      {
        const int64_t st = nb_stride(f / 2);
        UN[f][t] = x[(f % 2 == 1 ? cell + st : cell - st) * NDC + t];
      }
  }
  {
    int vs[3] = {1, n + 1, (n + 1) * (n + 1)};
    const int vb = vbase;
    for(int it = t; it < (NCORN + NFACE * NHC) * DIM; it += NT)
    {
      const int c = it / DIM, a = it % DIM;
      int idx = vb;
      if(c < NCORN)
      {
#pragma unroll
        for(int e = 0; e < DIM; ++e) idx += ((c >> e) & 1) * vs[e];
      }
      else
      {
        const int f = (c - NCORN) / NHC, cp = (c - NCORN) % NHC;
        if(!interior(f)) continue;
        const int d = f / 2, e1 = d == 0 ? 1 : 0;
        // offsets: 2 or -1 along d, the face corner's bits along e1 < e2
#pragma unroll
        for(int e = 0; e < DIM; ++e)
          idx += (e == d ? (f % 2 == 1 ? 2 : -1) : (e == e1 ? (cp & 1) : (cp >> 1))) * vs[e];
      }
      X[c][a] = vert[int64_t(idx) * DIM + a];
    }
  }
  __syncthreads();

  // ---------------- P1: x-interpolation; map coefficients
  if(act)
  {
    T v = 0, g = 0;
    const int base = ts - ix;
#pragma unroll
    for(int j = 0; j < NP; ++j)
    {
      const T u = U[base + j];
      v += tl.V[ix * NP + j] * u;
      g += tl.Dv[ix * NP + j] * u;
    }
    A0[t] = v;
    A1[t] = g;
  }
  for(int it = t; it < (NFACE + 1) * DIM; it += NT)
  {
    if(it >= NFACE * DIM)
    {
      // coordinate a of A_S = sum over subsets R of S of (-1)^{|S \ R|} X_R: one
      // subtraction stage per direction
      const int a = it - NFACE * DIM;
      double c[NCORN];
#pragma unroll
      for(int R = 0; R < NCORN; ++R) c[R] = X[R][a];
#pragma unroll
      for(int e = 0; e < DIM; ++e)
#pragma unroll
        for(int S = 0; S < NCORN; ++S)
          if((S >> e) & 1) c[S] -= c[S ^ (1 << e)];
#pragma unroll
      for(int S = 0; S < NCORN; ++S) C[S][a] = T(c[S]);
      continue;
    }
    // face f, coordinate a: Y = the face's corners, E / En = the edges along the
    // normal direction d of the cell / of the neighbour at those corners (pointing
    // to increasing x_d), all formed in FP64
    const int f = it / DIM, a = it % DIM;
    const int d = f / 2, side = f % 2, e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
    const bool inner = interior(f);
    double Y[NHC], E[NHC], En[NHC];
#pragma unroll
    for(int cp = 0; cp < NHC; ++cp)
    {
      const int cy = (side << d) | ((cp & 1) << e1) | (DIM == 3 ? ((cp >> 1) << e2) : 0);
      Y[cp] = X[cy][a];
      const double z = X[cy ^ (1 << d)][a];
      E[cp] = side ? Y[cp] - z : z - Y[cp];
      const double w = inner ? X[NCORN + f * NHC + cp][a] : Y[cp];
      En[cp] = side ? w - Y[cp] : Y[cp] - w;
    }
    if constexpr(DIM == 3)
    {
      FG[f][0][a] = T(Y[1] - Y[0]);
      FG[f][1][a] = T(Y[2] - Y[0]);
      FG[f][2][a] = T(Y[3] - Y[2] - Y[1] + Y[0]);
      FG[f][3][a] = T(E[0]);
      FG[f][4][a] = T(E[1] - E[0]);
      FG[f][5][a] = T(E[2] - E[0]);
      FG[f][6][a] = T(E[3] - E[2] - E[1] + E[0]);
      FG[f][7][a] = T(En[0]);
      FG[f][8][a] = T(En[1] - En[0]);
      FG[f][9][a] = T(En[2] - En[0]);
      FG[f][10][a] = T(En[3] - En[2] - En[1] + En[0]);
    }
    else
    {
      FG[f][0][a] = T(Y[1] - Y[0]);
      FG[f][1][a] = T(E[0]);
      FG[f][2][a] = T(E[1] - E[0]);
      FG[f][3][a] = T(En[0]);
      FG[f][4][a] = T(En[1] - En[0]);
    }
  }
  __syncthreads();

  // ---------------- P2: y-interpolation; traces along the face normals
  T g0 = 0, g1 = 0, g2 = 0;
  if(act)
  {
    const int base = ts - iy * NP;
    T b0 = 0, b1 = 0, b2 = 0;
#pragma unroll
    for(int j = 0; j < NP; ++j)
    {
      const T a0 = A0[base + j * NP], a1 = A1[base + j * NP];
      const T vy = tl.V[iy * NP + j], dy = tl.Dv[iy * NP + j];
      b0 += vy * a1;
      b1 += dy * a0;
      b2 += vy * a0;
    }
    if constexpr(DIM == 2)
    {
      g0 = b0;
      g1 = b1;
    }
    else
    {
      B0[t] = b0;
      B1[t] = b1;
      B2[t] = b2;
    }
  }
  {
    auto normal_trace = [&](int d, int side, int sd, int m) {
      const int f = 2 * d + side;
      if(sd == 1 && !interior(f)) return;
      const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
      const int i1 = m % NP, i2 = DIM == 3 ? m / NP : 0;
      const int base = i1 * stride(e1) + (DIM == 3 ? i2 * stride(e2) : 0);
      const T *src = sd == 0 ? U : UN[f];
      const int fs = sd == 0 ? side : 1 - side;
      T a = 0, bs = 0;
#pragma unroll
      for(int id = 0; id < NP; ++id)
      {
        const T u = src[base + id * stride(d)];
        a += tl.f[fs][id] * u;
        bs += tl.fd[fs][id] * u;
      }
      PA[f][sd][m] = a;
      PB[f][sd][m] = bs;
    };
    if constexpr(DIM == 3 && 4 * NF <= NT && NT < 8 * NF)
    {
      // (3D Q3, Q4) the items (side, sd, m) of one normal direction fill a pass:
      // one pass per direction, which makes d and the strides constants
      if(t < 4 * NF)
      {
        const int side = t / (2 * NF), sd = (t / NF) % 2, m = t % NF;
#pragma unroll
        for(int d = 0; d < DIM; ++d) normal_trace(d, side, sd, m);
      }
    }
    else
      for(int o = t; o < NFACE * 2 * NF; o += NT)
      {
        const int f = o / (2 * NF);
        normal_trace(f / 2, f % 2, (o / NF) % 2, o % NF);
      }
  }
  __syncthreads();

  // ---------------- P3: z-interpolation, geometry, reference flux; tangential
  // traces along the first face direction
  if(act)
  {
    if constexpr(DIM == 3)
    {
      const int base = ts - iz * NP * NP;
#pragma unroll
      for(int j = 0; j < NP; ++j)
      {
        const T vz = tl.V[iz * NP + j], dz = tl.Dv[iz * NP + j];
        g0 += vz * B0[base + j * NP * NP];
        g1 += vz * B1[base + j * NP * NP];
        g2 += dz * B2[base + j * NP * NP];
      }
    }
    T xi[3] = {tl.qx[ix], tl.qx[iy], DIM == 3 ? tl.qx[iz] : T(0)};
    T J[3][3], Ji[3][3];
    coef_jacobian<DIM>(C, xi, J);
    const T det = invert<DIM>(J, Ji);
    const T jxw = fabs(det) * tl.qw[ix] * tl.qw[iy] * (DIM == 3 ? tl.qw[iz] : T(1));
    const T g[3] = {g0, g1, g2};
    T ph[3];
#pragma unroll
    for(int c = 0; c < DIM; ++c)
    {
      T s = 0;
#pragma unroll
      for(int bb = 0; bb < DIM; ++bb) s += Ji[bb][c] * g[bb];
      ph[c] = s;  // physical gradient
    }
    T F[3];
#pragma unroll
    for(int a = 0; a < DIM; ++a)
    {
      T s = 0;
#pragma unroll
      for(int c = 0; c < DIM; ++c) s += Ji[a][c] * ph[c];
      F[a] = s * jxw;
    }
    // F_a into the x-stage arrays (last read in P2)
    A0[t] = F[0];
    A1[t] = F[1];
    if(DIM == 3) A2[t] = F[2];
  }
  for(int o = t; o < NFACE * 2 * NF; o += NT)
  {
    // (f, side, l1, i2): along e1 from the nodes i1 to the point l1
    const int f = o / (2 * NF), sd = (o / NF) % 2, r = o % NF;
    if(sd == 1 && !interior(f)) continue;
    const int l1 = r % NP, i2 = r / NP;
    T va = 0, da = 0, vb = 0;
#pragma unroll
    for(int i1 = 0; i1 < NP; ++i1)
    {
      const T a = PA[f][sd][i1 + NP * i2], pb = PB[f][sd][i1 + NP * i2];
      const T v = tl.V[l1 * NP + i1];
      va += v * a;
      da += tl.Dv[l1 * NP + i1] * a;
      vb += v * pb;
    }
    VA[f][sd][r] = va;
    DA[f][sd][r] = da;
    VB[f][sd][r] = vb;
  }
  __syncthreads();

  // ---------------- P4: z-integration; face fluxes at every face point
  if(act)
  {
    if constexpr(DIM == 3)
    {
      const int base = ts - iz * NP * NP;
      T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
      for(int q = 0; q < NP; ++q)
      {
        const T vz = tl.V[q * NP + iz], dz = tl.Dv[q * NP + iz];
        s0 += vz * A0[base + q * NP * NP];
        s1 += vz * A1[base + q * NP * NP];
        s2 += dz * A2[base + q * NP * NP];
      }
      B0[t] = s0;
      B1[t] = s1;
      B2[t] = s2;
    }
  }
  const T gamma = p == 0 ? T(1) : T(p) * T(p + 1);
  for(int pt = t; pt < NFACE * NF; pt += NT)
  {
    const int f = pt / NF, l = pt % NF;
    const int side = f % 2;
    const bool inner = interior(f);
    const int l1 = l % NP, l2 = DIM == 3 ? l / NP : 0;
    // traces: value, normal and tangential reference derivatives of both sides
    T tr[2][4];  // [side][value, d_normal, d_e1, d_e2]
#pragma unroll
    for(int sd = 0; sd < 2; ++sd)
    {
      T u = 0, gn = 0, gt1 = 0, gt2 = 0;
      if(sd == 0 || inner)
      {
#pragma unroll
        for(int i2 = 0; i2 < NP2; ++i2)
        {
          const int r = l1 + NP * i2;
          const T v2 = DIM == 3 ? tl.V[l2 * NP + i2] : T(1);
          const T d2 = DIM == 3 ? tl.Dv[l2 * NP + i2] : T(0);
          const T va = VA[f][sd][r];
          u += v2 * va;
          gn += v2 * VB[f][sd][r];
          gt1 += v2 * DA[f][sd][r];
          gt2 += d2 * va;
        }
      }
      tr[sd][0] = u;
      tr[sd][1] = gn;
      tr[sd][2] = gt1;
      tr[sd][3] = gt2;
    }
    const T q1 = tl.qx[l1], q2 = DIM == 3 ? tl.qx[l2] : T(0);
    // tangents and the normal-direction Jacobian columns of both cells
    T t1[3], t2[3] = {0, 0, 0}, jd[3], jn[3], w[3];
#pragma unroll
    for(int a = 0; a < DIM; ++a)
    {
      if constexpr(DIM == 3)
      {
        t1[a] = FG[f][0][a] + q2 * FG[f][2][a];
        t2[a] = FG[f][1][a] + q1 * FG[f][2][a];
        jd[a] = FG[f][3][a] + q1 * FG[f][4][a] + q2 * (FG[f][5][a] + q1 * FG[f][6][a]);
        jn[a] = FG[f][7][a] + q1 * FG[f][8][a] + q2 * (FG[f][9][a] + q1 * FG[f][10][a]);
      }
      else
      {
        t1[a] = FG[f][0][a];
        jd[a] = FG[f][1][a] + q1 * FG[f][2][a];
        jn[a] = FG[f][3][a] + q1 * FG[f][4][a];
      }
    }
    if constexpr(DIM == 3)
    {
      w[0] = t1[1] * t2[2] - t1[2] * t2[1];
      w[1] = t1[2] * t2[0] - t1[0] * t2[2];
      w[2] = t1[0] * t2[1] - t1[1] * t2[0];
    }
    else
    {
      w[0] = t1[1];
      w[1] = -t1[0];
    }
    T ww = 0, wj = 0;
#pragma unroll
    for(int a = 0; a < DIM; ++a)
    {
      ww += w[a] * w[a];
      wj += w[a] * jd[a];
    }
    const T nw = root(ww), inw = recip(nw);
    // outward: n . J_d > 0 on the face x_d = 1 of the cell, < 0 on x_d = 0
    const T sg = ((wj > T(0)) == (side == 1)) ? inw : -inw;
    T nv[3];
#pragma unroll
    for(int a = 0; a < DIM; ++a) nv[a] = sg * w[a];
    const T wq = tl.qw[l1] * (DIM == 3 ? tl.qw[l2] : T(1));
    const T dS = nw * wq;
    // Gram matrix of the tangents (2D: t1 . t1) and its inverse
    T g11 = 0, g12 = 0, g22 = 0;
#pragma unroll
    for(int a = 0; a < DIM; ++a)
    {
      g11 += t1[a] * t1[a];
      g12 += t1[a] * t2[a];
      g22 += t2[a] * t2[a];
    }
    const T iG = DIM == 3 ? recip(g11 * g22 - g12 * g12) : recip(g11);
    // c = J^{-1} n in the face frame: [0] normal direction, [1] e1, [2] e2
    auto coeffs = [&](const T *jc, T *c) {
      T jnv = 0, r1 = 0, r2 = 0;
#pragma unroll
      for(int a = 0; a < DIM; ++a)
      {
        jnv += jc[a] * nv[a];
        r1 += t1[a] * jc[a];
        r2 += t2[a] * jc[a];
      }
      const T cd = recip(jnv);
      c[0] = cd;
      if constexpr(DIM == 3)
      {
        c[1] = -cd * (g22 * r1 - g12 * r2) * iG;
        c[2] = -cd * (g11 * r2 - g12 * r1) * iG;
      }
      else
      {
        c[1] = -cd * r1 * iG;
        c[2] = T(0);
      }
    };
    T co[3];
    coeffs(jd, co);
    const T dno = tr[0][1] * co[0] + tr[0][2] * co[1] + (DIM == 3 ? tr[0][3] * co[2] : T(0));
    T jv, av, fnd, cnd = T(1);
    if(inner)
    {
      T cn[3];
      coeffs(jn, cn);
      cnd = cn[0];
      const T dnn =
        tr[1][1] * cn[0] + tr[1][2] * cn[1] + (DIM == 3 ? tr[1][3] * cn[2] : T(0));
      const T jump = tr[0][0] - tr[1][0];
      jv = dS * jump;
      av = -dS * T(0.5) * (dno + dnn);
      fnd = dS * T(-0.5) * jump;
    }
    else
    {
      jv = dS * tr[0][0];
      av = -dS * dno;
      fnd = -dS * tr[0][0];
    }
    JV[f][l] = jv;
    AV[f][l] = av;
#pragma unroll
    for(int bb = 0; bb < DIM; ++bb) FH[f][bb][l] = fnd * co[bb];  // face frame
    // inverse penalty lengths 1/h = |(J^{-1} n)_d| at the first face quadrature
    // point of both cells (apps/laplace_integrator.h:651-656, 692-693)
    if(l == 0)
    {
      HI[f][0] = fabs(co[0]);
      HI[f][1] = inner ? fabs(cnd) : T(1);
    }
  }
  __syncthreads();

  // ---------------- P5: y-integration; face projections on the tangential
  // test functions (into PA / PB, last read in P3)
  if(act)
  {
    if constexpr(DIM == 3)
    {
      const int base = ts - iy * NP;
      T s0 = 0, s12 = 0;
#pragma unroll
      for(int q = 0; q < NP; ++q)
      {
        const T vy = tl.V[q * NP + iy], dy = tl.Dv[q * NP + iy];
        s0 += vy * B0[base + q * NP];
        s12 += dy * B1[base + q * NP] + vy * B2[base + q * NP];
      }
      A2[t] = s0;    // A2 was last read in P4
      U[t] = s12;    // U was last read in P2
    }
    else
    {
      const int base = ts - iy * NP;
      T s0 = 0, s1 = 0;
#pragma unroll
      for(int q = 0; q < NP; ++q)
      {
        s0 += tl.V[q * NP + iy] * A0[base + q * NP];
        s1 += tl.Dv[q * NP + iy] * A1[base + q * NP];
      }
      A2[t] = s0;
      U[t] = s1;
    }
  }
  if constexpr(DIM == 3)
  {
    // projections sum-factorised over the two face coordinates, first over l1
    // (into the trace storage, last read in P4): at (i1, l2)
    //   SA = sum_l1 phi_i1(l1) (sigma JV + AV) + phi'_i1(l1) FH_e1
    //   SB = sum_l1 phi_i1(l1) FH_e2,  SC = sum_l1 phi_i1(l1) FH_n
    T(*SA)[NF] = reinterpret_cast<T(*)[NF]>(RX);
    T(*SB)[NF] = SA + NFACE;
    T(*SC)[NF] = SA + 2 * NFACE;
    for(int it = t; it < NFACE * NF; it += NT)
    {
      const int f = it / NF, m = it % NF;
      const int i1 = m % NP, l2 = m / NP;
      const T sigma = interior(f) ? T(ip) * gamma * T(0.5) * (HI[f][0] + HI[f][1])
                                  : T(ip) * T(2) * gamma * HI[f][0];
      T sa = 0, sb = 0, sc = 0;
#pragma unroll
      for(int l1 = 0; l1 < NP; ++l1)
      {
        const int l = l1 + NP * l2;
        const T v = tl.V[l1 * NP + i1], dv = tl.Dv[l1 * NP + i1];
        sa += v * (sigma * JV[f][l] + AV[f][l]) + dv * FH[f][1][l];
        sb += v * FH[f][2][l];
        sc += v * FH[f][0][l];
      }
      SA[f][m] = sa;
      SB[f][m] = sb;
      SC[f][m] = sc;
    }
    __syncthreads();
    // then over l2: the projections on phi_i1 phi_i2 (value) and on the normal
    // derivative's coefficient
    for(int it = t; it < NFACE * NF; it += NT)
    {
      const int f = it / NF, m = it % NF;
      const int i1 = m % NP, i2 = m / NP;
      T pv = 0, pd = 0;
#pragma unroll
      for(int l2 = 0; l2 < NP; ++l2)
      {
        const int r = i1 + NP * l2;
        const T v = tl.V[l2 * NP + i2], dv = tl.Dv[l2 * NP + i2];
        pv += v * SA[f][r] + dv * SB[f][r];
        pd += v * SC[f][r];
      }
      PA[f][0][m] = pv;
      PB[f][0][m] = pd;
    }
  }
  else
  {
    for(int it = t; it < NFACE * NF; it += NT)
    {
      const int f = it / NF, i1 = it % NF;
      const T sigma = interior(f) ? T(ip) * gamma * T(0.5) * (HI[f][0] + HI[f][1])
                                  : T(ip) * T(2) * gamma * HI[f][0];
      T pv = 0, pd = 0;
#pragma unroll
      for(int l = 0; l < NF; ++l)
      {
        const T v = tl.V[l * NP + i1];
        pv += v * (sigma * JV[f][l] + AV[f][l]) + tl.Dv[l * NP + i1] * FH[f][1][l];
        pd += v * FH[f][0][l];
      }
      PA[f][0][i1] = pv;
      PB[f][0][i1] = pd;
    }
  }
  __syncthreads();

  // ---------------- P6: x-integration and the face terms of this test function
  if(act)
  {
    const int base = ts - ix;
    T acc = 0;
#pragma unroll
    for(int q = 0; q < NP; ++q)
      acc += tl.Dv[q * NP + ix] * A2[base + q] + tl.V[q * NP + ix] * U[base + q];
    const int ti[3] = {ix, iy, iz};
#pragma unroll
    for(int f = 0; f < NFACE; ++f)
    {
      const int d = f / 2, side = f % 2;
      const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
      const int m = ti[e1] + (DIM == 3 ? NP * ti[e2] : 0);
      acc += tl.f[side][ti[d]] * PA[f][0][m] + tl.fd[side][ti[d]] * PB[f][0][m];
    }
    if(RESIDUAL)
      y[cell * NDC + t] = b[cell * NDC + t] - acc;
    else
      y[cell * NDC + t] = acc;
  }
}

}  // namespace tpss
