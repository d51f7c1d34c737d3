// kernels_mvp_split.hpp -- the residual-free color pass of kernels_mvp.hpp
// (3D, K = 8) with one patch spread over the four waves of a 256-thread block.
//
// On the coarse levels of a multigrid hierarchy a color holds a few to a few
// thousand patches, far fewer than the GPU holds waves, and the pass takes as
// long as one patch takes from its first load to its last store.  Spreading a
// patch over four waves cuts the arithmetic each wave runs in sequence to a
// quarter; the values it computes are those of k_patch_mvp_fused with full
// (K x K) products in every direction (the even-odd blocks are a throughput
// device; the full products give the same result to rounding).
//
// Thread t: x-line l = t / 4 of the patch, (y, z) = (l % 8, l / 8), and the
// entries e = 2 (t % 4) and e + 1 of that line (x-cell t % 4 / 2, cell-local
// x 2 (t % 2) and 2 (t % 2) + 1).  In the y and z stages of the solve the
// same thread takes entries 2 (t % 4), 2 (t % 4) + 1 of the y- or z-line
// t / 4.  The four threads of a line are consecutive lanes of one wave (a
// quad): where a stage needs the whole line it takes it from the quad with DPP
// quad permutations (the x-line entering the solve, the z-line between the
// forward and backward z products), otherwise a line is read and rewritten in
// place by its quad; lines pass between waves through LDS behind block
// barriers (six per patch with the outer-face coupling).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tpss
{
template<typename VT, typename CT>
__device__ __forceinline__ void load_pair(const VT *p, CT *o)
{
  o[0] = CT(p[0]);
  o[1] = CT(p[1]);
}

// value of lane sel_i of the lane's quad (four consecutive lanes) for a DPP
// quad_perm control CTRL = sel_0 | sel_1 << 2 | sel_2 << 4 | sel_3 << 6
template<int CTRL>
__device__ __forceinline__ float quad_perm(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template<int CTRL>
__device__ __forceinline__ double quad_perm(double v)
{
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// the 8 entries of a line held as two per lane by the four lanes of a quad
template<typename CT>
__device__ __forceinline__ void quad_gather(const CT *mine, CT *all)
{
  all[0] = quad_perm<0x00>(mine[0]);
  all[1] = quad_perm<0x00>(mine[1]);
  all[2] = quad_perm<0x55>(mine[0]);
  all[3] = quad_perm<0x55>(mine[1]);
  all[4] = quad_perm<0xAA>(mine[0]);
  all[5] = quad_perm<0xAA>(mine[1]);
  all[6] = quad_perm<0xFF>(mine[0]);
  all[7] = quad_perm<0xFF>(mine[1]);
}

template<typename VT, typename CT>
__global__ void __launch_bounds__(256)
k_patch_mvp_split(const VT *__restrict__ b, VT *__restrict__ x, const OpTab<CT, 4> tab,
                  const uint32_t *__restrict__ vtx, const uint32_t *__restrict__ fac,
                  const CT *__restrict__ S_tab, const CT *__restrict__ lam_tab,
                  const float *__restrict__ omega, const CT relaxation, const int patch_begin,
                  const int n, const int x_is_zero, int *__restrict__ nonfinite)
{
  constexpr int K = 8, NP = 4, NDC = 64;
  constexpr int SY = 8, SZ = 72;     // solve buffer (x, y, z) -> x + SY y + SZ z
  constexpr int CSY = 12, CSZ = 96;  // coupling buffers
  __shared__ CT A1[K * CSZ];
  __shared__ CT A2[K * CSZ];  // also the solve buffer
  __shared__ CT A3[K * CSZ];
  __shared__ CT Ss[3][K * K];
  __shared__ CT ls[3][K];

  const int t = threadIdx.x;
  const int j = patch_begin + int(blockIdx.x);  // one patch per block, grid = patches
  // The pass takes as long as one patch's chain of memory round trips: the
  // patch's vertex and factor rows (block-uniform), then every other load of
  // the pass -- the factor tables, b, the outer neighbour rows -- in one batch,
  // with the tables staged in LDS after the neighbour contractions.  Absent
  // neighbours read the patch's own cells (in bounds) and are discarded by
  // selection, so that no load waits behind a branch.
  int v[3] = {1, 1, 1};
  unpack_vertex(vtx[j], v);
  uint32_t fw[3];
#pragma unroll
  for(int d = 0; d < 3; ++d) fw[d] = fac[size_t(j) * 3 + d];

  const int line = t >> 2, q = t & 3, hq = q & 1;
  const int ty = line & 7, tz = line >> 3;
  const int ly = ty & 3, lz = tz & 3, by = ty >> 2, bz = tz >> 2;
  const int cx = q >> 1, lx0 = 2 * hq, e0 = 2 * q;
  const int64_t nn = int64_t(n) * n;
  const int64_t c_lo = patch_cell<3, NP>(v, n, 0, ty, tz);
  const int64_t cell = c_lo + cx;
  const int lyz = NP * ly + NP * NP * lz;
  CT sv = CT(0), lv = CT(1);
  if(t < 3 * K * K) sv = S_tab[size_t(fw[t / (K * K)]) * K * K + t % (K * K)];
  if(t < 3 * K) lv = lam_tab[size_t(fw[t / K]) * K + t % K];
  CT g[2];
  load_pair(b + cell * NDC + lyz + lx0, g);

  CT z0[2] = {CT(0), CT(0)}, t1[2] = {CT(0), CT(0)}, t2[2] = {CT(0), CT(0)};
  if(!x_is_zero)
  {
    // z_0: the x-neighbour of this thread's x-cell (left of cell 0, right of cell 1)
    const bool hx = cx == 0 ? (v[0] - 1 > 0) : (v[0] + 1 < n);
    CT xr[NP];
    load_row<NP>(x + (hx ? (cx == 0 ? c_lo - 1 : c_lo + 2) : cell) * NDC + lyz, xr);
    // the rows along the face normal of the y- and z-neighbour planes at this
    // line's position (two entries each)
    bool has[2];
    CT rows[2][NP][2];
#pragma unroll
    for(int dir = 1; dir < 3; ++dir)
    {
      const int bside = dir == 1 ? by : bz;
      has[dir - 1] = bside == 0 ? (v[dir] - 1 > 0) : (v[dir] + 1 < n);
      const int64_t step = dir == 1 ? int64_t(n) : nn;
      const VT *pn = x + (cell + (has[dir - 1] ? (bside == 0 ? -step : step) : 0)) * NDC +
                     (dir == 1 ? NP * NP * lz : NP * ly);
      const int rs = dir == 1 ? NP : NP * NP;  // stride of the rows along the normal
#pragma unroll
      for(int jj = 0; jj < NP; ++jj) load_pair(pn + rs * jj + lx0, rows[dir - 1][jj]);
    }
    {
      const CT *F = cx == 0 ? tab.Fpm : tab.Fmp;
#pragma unroll
      for(int k = 0; k < 2; ++k)
      {
        CT s = 0;
#pragma unroll
        for(int jj = 0; jj < NP; ++jj) s += F[(lx0 + k) * NP + jj] * xr[jj];
        z0[k] = hx ? s : CT(0);
      }
    }
    // z_1 (y) and z_2 (z): the neighbour plane contracted along the face
    // normal at this line's position, then t = M_x z over the x-cell (the
    // other two entries of the cell from the partner lane t ^ 1)
#pragma unroll
    for(int dir = 1; dir < 3; ++dir)
    {
      const int bside = dir == 1 ? by : bz, lpos = dir == 1 ? ly : lz;
      const CT *F = bside == 0 ? tab.Fpm : tab.Fmp;
      CT z[2] = {CT(0), CT(0)};
#pragma unroll
      for(int jj = 0; jj < NP; ++jj)
      {
        const CT f = F[lpos * NP + jj];
        z[0] += f * rows[dir - 1][jj][0];
        z[1] += f * rows[dir - 1][jj][1];
      }
      if(!has[dir - 1]) z[0] = z[1] = CT(0);
      const CT zo0 = quad_perm<0xB1>(z[0]), zo1 = quad_perm<0xB1>(z[1]);  // lane ^ 1
      CT zc[NP];
#pragma unroll
      for(int jj = 0; jj < NP; ++jj)
      {
        const bool mine = (jj >> 1) == hq;
        zc[jj] = (jj & 1) ? (mine ? z[1] : zo1) : (mine ? z[0] : zo0);
      }
#pragma unroll
      for(int k = 0; k < 2; ++k)
      {
        CT s = 0;
#pragma unroll
        for(int jj = 0; jj < NP; ++jj) s += tab.M[(lx0 + k) * NP + jj] * zc[jj];
        if(dir == 1)
          t1[k] = s;
        else
          t2[k] = s;
      }
    }
  }
  // the patch tables, published by the next barrier
  if(t < 3 * K * K) (&Ss[0][0])[t] = sv;
  if(t < 3 * K) (&ls[0][0])[t] = lv;
  const CT w = omega ? CT(omega[j]) : relaxation;

  if(!x_is_zero)
  {
    // cross-line mass contractions:  g -= M_y M_x z_2 + M_z (M_x z_1 + M_y z_0)
    CT *Z = A1;
    CT *T2 = A2;
#pragma unroll
    for(int k = 0; k < 2; ++k)
    {
      Z[e0 + k + CSY * ty + CSZ * tz] = z0[k];
      T2[e0 + k + CSY * ty + CSZ * tz] = t2[k];
    }
    __syncthreads();
    CT a[2], bb[2];
    const int y0 = by * NP;
#pragma unroll
    for(int k = 0; k < 2; ++k)
    {
      CT s = t1[k], s2 = 0;
#pragma unroll
      for(int jj = 0; jj < NP; ++jj)
      {
        s += tab.M[ly * NP + jj] * Z[e0 + k + CSY * (y0 + jj) + CSZ * tz];
        s2 += tab.M[ly * NP + jj] * T2[e0 + k + CSY * (y0 + jj) + CSZ * tz];
      }
      a[k] = s;
      bb[k] = s2;
    }
#pragma unroll
    for(int k = 0; k < 2; ++k) A3[e0 + k + CSY * ty + CSZ * tz] = a[k];
    __syncthreads();
    const int z00 = bz * NP;
#pragma unroll
    for(int k = 0; k < 2; ++k)
    {
      CT s = bb[k];
#pragma unroll
      for(int kk = 0; kk < NP; ++kk) s += tab.M[lz * NP + kk] * A3[e0 + k + CSY * ty + CSZ * (z00 + kk)];
      g[k] -= s;
    }
  }

  // fast-diagonalization solve on the patch in B (the coupling's reads of A2
  // precede its last barrier)
  CT *B = A2;
  CT rr[K];
  CT o[2];
  if(x_is_zero) __syncthreads();  // publishes Ss / ls (else the coupling's barriers do)
  // forward x: the x-line (ty, tz), gathered from the quad
  quad_gather(g, rr);
#pragma unroll
  for(int k = 0; k < 2; ++k)
  {
    CT s = 0;
#pragma unroll
    for(int a = 0; a < K; ++a) s += Ss[0][a * K + e0 + k] * rr[a];
    o[k] = s;
  }
#pragma unroll
  for(int k = 0; k < 2; ++k) B[e0 + k + SY * ty + SZ * tz] = o[k];
  __syncthreads();
  // forward y: the y-line (x0, z0) = (line % 8, line / 8)
  const int p0 = line & 7, p1 = line >> 3;
#pragma unroll
  for(int a = 0; a < K; ++a) rr[a] = B[p0 + SY * a + SZ * p1];
#pragma unroll
  for(int k = 0; k < 2; ++k)
  {
    CT s = 0;
#pragma unroll
    for(int a = 0; a < K; ++a) s += Ss[1][a * K + e0 + k] * rr[a];
    o[k] = s;
  }
#pragma unroll
  for(int k = 0; k < 2; ++k) B[p0 + SY * (e0 + k) + SZ * p1] = o[k];
  __syncthreads();
  // forward z and the eigenvalue scaling: the z-line (x0, y0) = (line % 8, line / 8)
#pragma unroll
  for(int a = 0; a < K; ++a) rr[a] = B[p0 + SY * p1 + SZ * a];
  {
    const CT lxy = ls[0][p0] + ls[1][p1];
#pragma unroll
    for(int k = 0; k < 2; ++k)
    {
      CT s = 0;
#pragma unroll
      for(int a = 0; a < K; ++a) s += Ss[2][a * K + e0 + k] * rr[a];
      o[k] = eig_scale(s, w, lxy + ls[2][e0 + k]);
    }
  }
  // backward z: the scaled z-line gathered from the quad
  quad_gather(o, rr);
#pragma unroll
  for(int k = 0; k < 2; ++k)
  {
    CT s = 0;
#pragma unroll
    for(int i = 0; i < K; ++i) s += Ss[2][(e0 + k) * K + i] * rr[i];
    o[k] = s;
  }
#pragma unroll
  for(int k = 0; k < 2; ++k) B[p0 + SY * p1 + SZ * (e0 + k)] = o[k];
  __syncthreads();
  // backward y
#pragma unroll
  for(int a = 0; a < K; ++a) rr[a] = B[p0 + SY * a + SZ * p1];
#pragma unroll
  for(int k = 0; k < 2; ++k)
  {
    CT s = 0;
#pragma unroll
    for(int i = 0; i < K; ++i) s += Ss[1][(e0 + k) * K + i] * rr[i];
    o[k] = s;
  }
#pragma unroll
  for(int k = 0; k < 2; ++k) B[p0 + SY * (e0 + k) + SZ * p1] = o[k];
  __syncthreads();
  // backward x: this thread's two entries of the x-line (ty, tz)
#pragma unroll
  for(int a = 0; a < K; ++a) rr[a] = B[a + SY * ty + SZ * tz];
#pragma unroll
  for(int k = 0; k < 2; ++k)
  {
    CT s = 0;
#pragma unroll
    for(int i = 0; i < K; ++i) s += Ss[0][(e0 + k) * K + i] * rr[i];
    o[k] = s;
  }
  VT *px = x + cell * NDC + lyz + lx0;
  if(!x_is_zero && w != CT(1))
  {
    CT old[2];
    load_pair(px, old);
#pragma unroll
    for(int k = 0; k < 2; ++k) o[k] += (CT(1) - w) * old[k];
  }
  if(nonfinite && !(__builtin_isfinite(o[0]) && __builtin_isfinite(o[1]))) *nonfinite = 1;
  px[0] = VT(o[0]);
  px[1] = VT(o[1]);
}

}  // namespace tpss
