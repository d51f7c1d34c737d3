// kernels_mvp.hpp -- residual-free multiplicative color pass on Cartesian meshes.
//
// For a patch j of the current color the reference computes
//   r = b - A x,  u_j = A_j^{-1} R_j r,  x += omega R_j^T u_j
// (schwarz_preconditioner.templates.h:585-662, 691-732).  On tensor-product
// cells A_j is exactly the patch block R_j A R_j^T, so with B_j the coupling of
// the patch to the cells just outside it,
//   R_j r = b_j - A_j x_j - B_j x_out   and
//   x_j <- (1 - omega) x_j + omega A_j^{-1} (b_j - B_j x_out).
// B_j x_out is the face coupling of the SIPG operator through the patch's outer
// faces: per patch cell and direction d (one outer face each) the neighbour
// term T_d(F) x_nbr of the cell operator, evaluated here in sum-factorised form
//   M_z (M_y z_0 + M_x z_1) + M_y M_x z_2,  z_d = F_d x_nbr(d) along d.
// Valid when same-color patches are not face-coupled (red-black coloring), so
// no patch of the pass reads cells another patch of the pass writes.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tpss
{
typedef short tpss_bf16x4 __attribute__((__vector_size__(4 * sizeof(short))));
typedef float tpss_floatx4 __attribute__((__vector_size__(4 * sizeof(float))));

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short bf16_bits(float f)
{
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return short(u >> 16);
}

// Fast-diagonalization inverse of the two patches of a block (3D, K = 8) on
// the matrix cores: every 1D contraction along direction d is one
// v_mfma_f32_16x16x16_bf16 per chunk of 16 lines, with A = diag(S_a^T, S_b^T)
// (16 x 16, block diagonal over the two patches) and B the 16 contraction
// values (8 per patch) of 16 lines; products in BF16, sums in FP32.  The
// patch data (FP32) lives in B2[2][512] (x fastest); the two waves of the
// block take two chunks each.  Lane layout of the 16x16x16 MFMA (wave64):
// A: row l%16, k = 4(l/16)+0..3; B: column l%16, k = 4(l/16)+0..3;
// D: column l%16, rows 4(l/16)+0..3.
__device__ __forceinline__ int lds_index(int d, int k, int line)
{
  // K = 8: line enumerates the other two indices (first fastest)
  return d == 0 ? k + 8 * line : d == 1 ? (line & 7) + 8 * k + 64 * (line >> 3) : line + 64 * k;
}

__device__ __forceinline__ void fd_mfma_pair(float (*B2)[512], const float (*S)[3][64],
                                             const float (*lam)[3][8], const float *w2)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n16 = lane & 15, q = lane >> 4;       // column (line) / k-quarter
  const int pq = q >> 1, kofs = 4 * (q & 1);      // patch of the B/D rows, first row index
  const int arow = lane & 15, pa = arow >> 3, ao = arow & 7;  // A row: patch, output index
  for(int stage = 0; stage < 6; ++stage)
  {
    const int d = stage < 3 ? stage : 5 - stage;
    const bool fwd = stage < 3;
    // A operand: S_p^T (forward) or S_p (backward) of patch pa if the k block
    // belongs to the same patch, else zero
    tpss_bf16x4 a;
#pragma unroll
    for(int e = 0; e < 4; ++e)
    {
      const int kk = kofs + e;
      float v = 0.f;
      if(pa == pq) v = fwd ? S[pa][d][kk * 8 + ao] : S[pa][d][ao * 8 + kk];
      a[e] = bf16_bits(v);
    }
#pragma unroll
    for(int cc = 0; cc < 2; ++cc)
    {
      const int line = 16 * (2 * wave + cc) + n16;
      tpss_bf16x4 b;
#pragma unroll
      for(int e = 0; e < 4; ++e) b[e] = bf16_bits(B2[pq][lds_index(d, kofs + e, line)]);
      tpss_floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
      if(stage == 2)
      {
        // diagonal scaling after the last forward contraction (z): line = (x, y)
        const int x0 = line & 7, y0 = line >> 3;
        const float lxy = lam[pq][0][x0] + lam[pq][1][y0];
#pragma unroll
        for(int e = 0; e < 4; ++e) acc[e] *= w2[pq] / (lxy + lam[pq][2][kofs + e]);
      }
#pragma unroll
      for(int e = 0; e < 4; ++e) B2[pq][lds_index(d, kofs + e, line)] = acc[e];
    }
    __syncthreads();
  }
}

// FD = 0: vector FMA contractions in CT; FD = 1: BF16 MFMA contractions with
// FP32 accumulation (3D, K = 8, CT = float)
// Waves per SIMD the register allocation must allow.  Up to K = 8: FP32 patch
// solves 5 (<= 96 VGPRs) so that more patches keep their loads in flight while
// others compute; longer lines (K = 10) hold more values per thread: 3.  FP64
// solves unconstrained (3D K = 8: 186 VGPRs, 2 waves; forcing 3 spills and
// measured 20 % slower).

template<int K, typename CT, int FD>
constexpr int mvp_min_waves()
{
  return FD == 1 || !std::is_same<CT, float>::value ? 1 : K <= 8 ? 5 : 3;
}

// Launch shape of the fused pass: one patch per block where a patch fills a
// wave (3D, K = 8, vector solves), so that the block is a single wave, its
// barriers cost nothing and no patch waits for another; otherwise the shape of
// the other patch kernels (the BF16-MFMA solve pairs two patches per block).
template<int DIM, int K, int FD>
struct MvpShape
{
  using PS = PatchShape<DIM, K>;
  static constexpr int PPB = (FD == 0 && PS::L == 64) ? 1 : PS::PPB;
  static constexpr int THREADS = ((PPB * PS::L + 63) / 64) * 64;
};

template<int DIM, int K, typename VT, typename CT, int FD = 0>
__global__ void __launch_bounds__((MvpShape<DIM, K, FD>::THREADS), (mvp_min_waves<K, CT, FD>()))
k_patch_mvp_fused(const VT *__restrict__ b, VT *__restrict__ x, const OpTab<CT, K / 2> tab,
                  const uint32_t *__restrict__ vtx, const uint32_t *__restrict__ fac,
                  const CT *__restrict__ S_tab, const CT *__restrict__ lam_tab,
                  const uint32_t *__restrict__ eo_tab, const float *__restrict__ omega,
                  const CT relaxation, const int patch_begin,
                  const int patch_end, const int n, const int x_is_zero,
                  int *__restrict__ nonfinite)
{
  using PS = PatchShape<DIM, K>;
  using MS = MvpShape<DIM, K, FD>;
  constexpr int NP = PS::NP, L = PS::L, NPD = PS::NPD, NDC = PS::NDC, PPB = MS::PPB;
  // LDS strides (y, z) of the patch buffers, padded for 3D K = 8 so that the
  // line transposes of the solve (FSY, FSZ) and the cross-line reads of the
  // coupling stage (CSY, CSZ) are free of bank conflicts.  The padded solve
  // buffer reuses the coupling buffer A2, whose last read precedes a barrier
  // that every thread passes before the solve writes.
  constexpr bool PAD = FD == 0 && DIM == 3 && K == 8;
  constexpr int FSY = K, FSZ = PAD ? 72 : K * K;
  constexpr int CSY = PAD ? 12 : K, CSZ = PAD ? 96 : K * K;
  static_assert(!PAD || K * FSZ <= K * CSZ, "solve buffer must fit in the coupling buffer");
  __shared__ CT buf[PPB][PAD ? 1 : NPD];
  __shared__ CT A1[PPB][DIM > 2 ? K * CSZ : NPD];
  __shared__ CT A2[PPB][DIM > 2 ? K * CSZ : 1];
  __shared__ CT Ss[PPB][DIM][K * K];
  __shared__ CT ls[PPB][DIM][K];

  const int pl = threadIdx.x / L, t = threadIdx.x % L;
  const int j0 = patch_begin + int(blockIdx.x) * PPB;
  const int j = j0 + pl;
  const bool valid = pl < PPB && j < patch_end;
  const int plc = pl < PPB ? pl : 0;
  const bool owner = pl < PPB;

  // FP32 solves with one patch per wave (3D, K = 8): the eigenvector matrices
  // of the wave's patch are read with wave-uniform addresses straight into
  // scalar registers (operands of the contraction FMAs) instead of being
  // staged through LDS, which takes most LDS reads off the pass.  (FP64
  // matrices would need twice the scalar registers and spill.)
  constexpr bool WAVE_PATCH = FD == 0 && L == 64 && MS::THREADS == PPB * 64;
  constexpr bool S_SCALAR = WAVE_PATCH && std::is_same<CT, float>::value;
  // One patch per wave, 3D, K = 8 (the fine levels): every global load of the
  // pass is issued in two batches -- the patch's vertex and factor rows, then
  // its factor tables, b and the outer neighbour rows -- before the first wait,
  // so that a wave waits for two memory round trips, not for one per table and
  // neighbour.  Each lane loads one row of the y- and z-neighbour planes and
  // takes the other three rows of a plane from the lanes that loaded them: the
  // lanes sharing a y-plane form a quad (DPP), those sharing a z-plane sit 8
  // lanes apart (their rows pass through LDS: 4 writes and 16 reads of 16 bytes
  // per lane instead of 64 ds_bpermute).  The patch tables are staged
  // unconditionally (the full matrices are unused by the even-odd products).
  constexpr bool BATCH = WAVE_PATCH && DIM == 3 && K == 8 && !std::is_same<CT, float>::value;
  int v[3] = {1, 1, 1};
  uint32_t fw[3] = {0, 0, 0};  // factor rows of the wave's patch (BATCH)
  CT w;
  if constexpr(BATCH)
  {
    // a starting wave issues its loads ahead of the other wave's arithmetic on
    // the SIMD (raised priority until the last load of the batch is issued)
    __builtin_amdgcn_s_setprio(3);
    const int jw = min(j0, patch_end - 1);  // wave-uniform
    const uint32_t pv = vtx[jw];
#pragma unroll
    for(int d = 0; d < 3; ++d) fw[d] = fac[size_t(jw) * 3 + d];
    unpack_vertex(pv, v);
  }
  else
  {
    // one wave per patch with the matrices not in scalar registers (FP64): the
    // even-odd products take the half matrices straight from scalar loads, so
    // only patches without mirror symmetry stage the full matrices in LDS
    bool stage_S = true;
    if constexpr(WAVE_PATCH && !S_SCALAR)
    {
      const int jw = min(j0, patch_end - 1);
      stage_S = !(eo_tab && (eo_tab[fac[size_t(jw) * DIM]] & eo_tab[fac[size_t(jw) * DIM + 1]] &
                             eo_tab[fac[size_t(jw) * DIM + 2]]) != 0u);
    }
    if constexpr(!S_SCALAR)
      for(int idx = threadIdx.x; stage_S && idx < PPB * DIM * K * K; idx += blockDim.x)
      {
        const int q = idx / (DIM * K * K), r = idx % (DIM * K * K);
        const int d = r / (K * K), e = r % (K * K);
        const int jj = j0 + q;
        Ss[q][d][e] = jj < patch_end ? S_tab[size_t(fac[size_t(jj) * DIM + d]) * K * K + e] : CT(0);
      }
    for(int idx = threadIdx.x; idx < PPB * DIM * K; idx += blockDim.x)
    {
      const int q = idx / (DIM * K), r = idx % (DIM * K);
      const int d = r / K, e = r % K;
      const int jj = j0 + q;
      ls[q][d][e] = jj < patch_end ? lam_tab[size_t(fac[size_t(jj) * DIM + d]) * K + e] : CT(1);
    }
    if(valid) unpack_vertex(vtx[j], v);
    w = valid ? (omega ? CT(omega[j]) : relaxation) : CT(0);
  }

  const int ty = DIM > 1 ? t % K : 0;
  const int tz = DIM > 2 ? t / K : 0;
  const int ly = ty >= NP ? ty - NP : ty, lz = tz >= NP ? tz - NP : tz;
  const int by = ty >= NP ? 1 : 0, bz = tz >= NP ? 1 : 0;
  const int64_t nn = int64_t(n) * n;

  // the x-line (ty, tz) of the patch: row lyz of the cells c_lo and c_lo + 1
  const int64_t c_lo = patch_cell<DIM, NP>(v, n, 0, ty, tz);
  const int lyz = patch_local<DIM, NP>(0, ty, tz);
  CT g[K], r[K], o[K];
  // BATCH: this lane's entries of the patch tables, staged in LDS after the
  // coupling stage (their loads go out with those of b and the neighbours)
  CT sv[3], lv = CT(1);
  if constexpr(BATCH)
  {
    if constexpr(!S_SCALAR)
#pragma unroll
      for(int d = 0; d < 3; ++d) sv[d] = S_tab[size_t(fw[d]) * K * K + t];
    if(t < 3 * K) lv = lam_tab[size_t(fw[t / K]) * K + t % K];
    load_row<NP>(b + c_lo * NDC + lyz, g);
    load_row<NP>(b + (c_lo + 1) * NDC + lyz, g + NP);
  }
  else if(valid)
  {
    load_row<NP>(b + c_lo * NDC + lyz, g);
    load_row<NP>(b + (c_lo + 1) * NDC + lyz, g + NP);
  }
  else
#pragma unroll
    for(int i = 0; i < K; ++i) g[i] = CT(0);

  if(!x_is_zero)
  {
    // z_0: x-neighbours of the two x-cells of this line
    CT z0[K], t1[K], t2[K];
#pragma unroll
    for(int i = 0; i < K; ++i)
    {
      z0[i] = 0;
      t1[i] = 0;
      t2[i] = 0;
    }
    if constexpr(BATCH)
    {
      // the outer neighbour rows of this lane: the x-neighbours of its line and
      // one row of each of the two y- and z-neighbour planes; absent neighbours
      // read the patch's own cells (in bounds) and are discarded below
      const bool hl = v[0] - 1 > 0, hr = v[0] + 1 < n;
      const bool hy = by == 0 ? (v[1] - 1 > 0) : (v[1] + 1 < n);
      const bool hz = bz == 0 ? (v[2] - 1 > 0) : (v[2] + 1 < n);
      CT xl[NP], xr[NP], ry[2][NP], rz[2][NP];
      load_row<NP>(x + (hl ? c_lo - 1 : c_lo) * NDC + lyz, xl);
      load_row<NP>(x + (hr ? c_lo + 2 : c_lo + 1) * NDC + lyz, xr);
      const int64_t sy = hy ? (by == 0 ? -int64_t(n) : int64_t(n)) : 0;
      const int64_t sz = hz ? (bz == 0 ? -nn : nn) : 0;
#pragma unroll
      for(int cx = 0; cx < 2; ++cx)
      {
        // y-plane lz, row ly; z-plane ly, row lz
        load_row<NP>(x + (c_lo + cx + sy) * NDC + NP * NP * lz + NP * ly, ry[cx]);
        load_row<NP>(x + (c_lo + cx + sz) * NDC + NP * ly + NP * NP * lz, rz[cx]);
      }
      __builtin_amdgcn_s_setprio(0);
      // z_0 = F x_out along x, z_1 / z_2 = the y- / z-neighbour planes contracted
      // along the face normal, from the rows loaded above.  The contractions run
      // for absent neighbours too and are discarded by selection: a branch would
      // let the compiler move the loads into it, behind the waits of the others.
      const int lane0 = int(threadIdx.x) & ~(3 << 3);
#pragma unroll
      for(int i = 0; i < NP; ++i)
      {
        CT s = 0, s2 = 0;
#pragma unroll
        for(int jj = 0; jj < NP; ++jj)
        {
          s += tab.Fpm[i * NP + jj] * xl[jj];
          s2 += tab.Fmp[i * NP + jj] * xr[jj];
        }
        z0[i] = hl ? s : CT(0);
        z0[NP + i] = hr ? s2 : CT(0);
      }
      {
        // the quad's four lanes (ly = 0..3) hold the four rows of the plane
        const CT *F = by == 0 ? tab.Fpm : tab.Fmp;
        CT z1[K];
#pragma unroll
        for(int cx = 0; cx < 2; ++cx)
        {
          CT xn[NP * NP];
#pragma unroll
          for(int jj = 0; jj < NP; ++jj)
#pragma unroll
            for(int lx = 0; lx < NP; ++lx) xn[lx + NP * jj] = quad_bcast_rt(ry[cx][lx], jj);
#pragma unroll
          for(int lx = 0; lx < NP; ++lx)
          {
            CT s = 0;
#pragma unroll
            for(int jj = 0; jj < NP; ++jj) s += F[ly * NP + jj] * xn[lx + NP * jj];
            z1[cx * NP + lx] = s;
          }
        }
#pragma unroll
        for(int cx = 0; cx < 2; ++cx)
#pragma unroll
          for(int i = 0; i < NP; ++i)
          {
            CT s = 0;
#pragma unroll
            for(int jj = 0; jj < NP; ++jj) s += tab.M[i * NP + jj] * z1[cx * NP + jj];
            t1[cx * NP + i] = hy ? s : CT(0);
          }
      }
      {
        // the lanes of line (ty, NP bz + kk), kk = 0..3, hold the rows kk: exchanged
        // through LDS (A2 is free until the coupling stage writes T2 there)
        const CT *F = bz == 0 ? tab.Fpm : tab.Fmp;
        CT z2[K];
        CT *RZ = A2[0];
#pragma unroll
        for(int cx = 0; cx < 2; ++cx)
#pragma unroll
          for(int lx = 0; lx < NP; ++lx) RZ[(t * 2 + cx) * NP + lx] = rz[cx][lx];
        __syncthreads();
#pragma unroll
        for(int cx = 0; cx < 2; ++cx)
        {
          CT xn[NP * NP];
#pragma unroll
          for(int kk = 0; kk < NP; ++kk)
#pragma unroll
            for(int lx = 0; lx < NP; ++lx) xn[lx + NP * kk] = RZ[((lane0 | (kk << 3)) * 2 + cx) * NP + lx];
#pragma unroll
          for(int lx = 0; lx < NP; ++lx)
          {
            CT s = 0;
#pragma unroll
            for(int kk = 0; kk < NP; ++kk) s += F[lz * NP + kk] * xn[lx + NP * kk];
            z2[cx * NP + lx] = s;
          }
        }
#pragma unroll
        for(int cx = 0; cx < 2; ++cx)
#pragma unroll
          for(int i = 0; i < NP; ++i)
          {
            CT s = 0;
#pragma unroll
            for(int jj = 0; jj < NP; ++jj) s += tab.M[i * NP + jj] * z2[cx * NP + jj];
            t2[cx * NP + i] = hz ? s : CT(0);
          }
      }
    }
    else
    {
      if(valid && v[0] - 1 > 0)
      {
        CT xl[NP];
        load_row<NP>(x + (c_lo - 1) * NDC + lyz, xl);
#pragma unroll
        for(int i = 0; i < NP; ++i)
        {
          CT s = 0;
#pragma unroll
          for(int jj = 0; jj < NP; ++jj) s += tab.Fpm[i * NP + jj] * xl[jj];
          z0[i] = s;
        }
      }
      if(valid && v[0] + 1 < n)
      {
        CT xr[NP];
        load_row<NP>(x + (c_lo + 2) * NDC + lyz, xr);
#pragma unroll
        for(int i = 0; i < NP; ++i)
        {
          CT s = 0;
#pragma unroll
          for(int jj = 0; jj < NP; ++jj) s += tab.Fmp[i * NP + jj] * xr[jj];
          z0[NP + i] = s;
        }
      }
      // z_1 (y-neighbour outside the patch) then t1 = M_x z_1 per x-cell
      {
        const bool has = by == 0 ? (v[1] - 1 > 0) : (v[1] + 1 < n);
        const CT *F = by == 0 ? tab.Fpm : tab.Fmp;
        if(valid && has)
        {
          const int64_t shift = by == 0 ? -int64_t(n) : int64_t(n);
          CT z1[K];
#pragma unroll
          for(int cx = 0; cx < 2; ++cx)
          {
            // the x-y plane lz of the neighbour cell: NP rows, contiguous
            CT xn[NP * NP];
            const VT *pn = x + (c_lo + cx + shift) * NDC + NP * NP * lz;
#pragma unroll
            for(int jj = 0; jj < NP; ++jj) load_row<NP>(pn + NP * jj, xn + NP * jj);
#pragma unroll
            for(int lx = 0; lx < NP; ++lx)
            {
              CT s = 0;
#pragma unroll
              for(int jj = 0; jj < NP; ++jj) s += F[ly * NP + jj] * xn[lx + NP * jj];
              z1[cx * NP + lx] = s;
            }
          }
#pragma unroll
          for(int cx = 0; cx < 2; ++cx)
#pragma unroll
            for(int i = 0; i < NP; ++i)
            {
              CT s = 0;
#pragma unroll
              for(int jj = 0; jj < NP; ++jj) s += tab.M[i * NP + jj] * z1[cx * NP + jj];
              t1[cx * NP + i] = s;
            }
        }
      }
      if(DIM > 2)
      {
        const bool has = bz == 0 ? (v[2] - 1 > 0) : (v[2] + 1 < n);
        const CT *F = bz == 0 ? tab.Fpm : tab.Fmp;
        if(valid && has)
        {
          const int64_t shift = bz == 0 ? -nn : nn;
          CT z2[K];
#pragma unroll
          for(int cx = 0; cx < 2; ++cx)
          {
            // the x-z plane ly of the neighbour cell: NP rows at stride NP^2
            CT xn[NP * NP];
            const VT *pn = x + (c_lo + cx + shift) * NDC + NP * ly;
#pragma unroll
            for(int kk = 0; kk < NP; ++kk) load_row<NP>(pn + NP * NP * kk, xn + NP * kk);
#pragma unroll
            for(int lx = 0; lx < NP; ++lx)
            {
              CT s = 0;
#pragma unroll
              for(int kk = 0; kk < NP; ++kk) s += F[lz * NP + kk] * xn[lx + NP * kk];
              z2[cx * NP + lx] = s;
            }
          }
#pragma unroll
          for(int cx = 0; cx < 2; ++cx)
#pragma unroll
            for(int i = 0; i < NP; ++i)
            {
              CT s = 0;
#pragma unroll
              for(int jj = 0; jj < NP; ++jj) s += tab.M[i * NP + jj] * z2[cx * NP + jj];
              t2[cx * NP + i] = s;
            }
        }
      }
    }
    if constexpr(BATCH) __syncthreads();  // the z-plane rows in A2 are read
    // cross-line mass contractions through LDS
    CT *Z = A1[plc];
    CT *T2 = A2[plc];
#pragma unroll
    for(int i = 0; i < K; ++i)
      if(owner)
      {
        Z[i + CSY * ty + CSZ * tz] = z0[i];
        if(DIM > 2) T2[i + CSY * ty + CSZ * tz] = t2[i];
      }
    __syncthreads();
    CT a[K], bb[K];
    const int y0 = by * NP;
#pragma unroll
    for(int i = 0; i < K; ++i)
    {
      CT s = t1[i], s2 = 0;
#pragma unroll
      for(int jj = 0; jj < NP; ++jj)
      {
        s += tab.M[ly * NP + jj] * Z[i + CSY * (y0 + jj) + CSZ * tz];
        if(DIM > 2) s2 += tab.M[ly * NP + jj] * T2[i + CSY * (y0 + jj) + CSZ * tz];
      }
      a[i] = s;
      bb[i] = s2;
    }
    if(DIM == 2)
    {
#pragma unroll
      for(int i = 0; i < K; ++i) g[i] -= a[i];
    }
    else
    {
      __syncthreads();
#pragma unroll
      for(int i = 0; i < K; ++i)
        if(owner) Z[i + CSY * ty + CSZ * tz] = a[i];
      __syncthreads();
      const int z00 = bz * NP;
#pragma unroll
      for(int i = 0; i < K; ++i)
      {
        CT s = bb[i];
#pragma unroll
        for(int kk = 0; kk < NP; ++kk) s += tab.M[lz * NP + kk] * Z[i + CSY * ty + CSZ * (z00 + kk)];
        g[i] -= s;
      }
    }
  }

  if constexpr(BATCH)
  {
    if constexpr(!S_SCALAR)
#pragma unroll
      for(int d = 0; d < 3; ++d) Ss[0][d][t] = sv[d];
    if(t < 3 * K) (&ls[0][0][0])[t] = lv;
    w = valid ? (omega ? CT(omega[min(j0, patch_end - 1)]) : relaxation) : CT(0);
    __syncthreads();
  }
  // the factor tables in LDS are complete (without the coupling stage no
  // barrier has passed since they were written)
  else if(x_is_zero)
    __syncthreads();
  if constexpr(FD == 1)
  {
    static_assert(DIM == 3 && K == 8 && PPB == 2, "MFMA patch solve: 3D, K = 8, two patches per block");
    __shared__ float w2[2];
    if(t == 0 && owner) w2[pl] = float(w);
    __syncthreads();  // the coupling stage above may still read A1/A2; buf is free
#pragma unroll
    for(int i = 0; i < K; ++i)
      if(owner) buf[plc][i + K * ty + K * K * tz] = float(g[i]);
    __syncthreads();
    fd_mfma_pair(reinterpret_cast<float(*)[512]>(&buf[0][0]),
                 reinterpret_cast<const float(*)[3][64]>(&Ss[0][0][0]),
                 reinterpret_cast<const float(*)[3][8]>(&ls[0][0][0]), w2);
#pragma unroll
    for(int i = 0; i < K; ++i) o[i] = buf[plc][i + K * ty + K * K * tz];
  }
  else
  {
#pragma unroll
    for(int i = 0; i < K; ++i) r[i] = g[i];
    if constexpr(WAVE_PATCH)
    {
      // wave index as a wave-uniform value; the last patch of the pass stands
      // in for a missing one (its results are discarded)
      const int pw = __builtin_amdgcn_readfirstlane(int(threadIdx.x)) / 64;
      const int jw = min(j0 + pw, patch_end - 1);
      const uint32_t f0 = fac[size_t(jw) * DIM + 0], f1 = fac[size_t(jw) * DIM + 1],
                     f2 = fac[size_t(jw) * DIM + 2];
      const CT *G0 = S_tab + size_t(f0) * K * K, *G1 = S_tab + size_t(f1) * K * K,
               *G2 = S_tab + size_t(f2) * K * K;
      // even-odd products where all three 1D problems are mirror-symmetric
      // (interior patches of a uniform mesh); a wave-uniform branch.  They read
      // only the first K/2 rows of each matrix, which fit the scalar registers
      // in FP64 too.
      if(eo_tab && (eo_tab[f0] & eo_tab[f1] & eo_tab[f2]) != 0u)
        fd_lines<DIM, K, CT, FSY, FSZ, true>(r, o, PAD ? A2[plc] : buf[plc], G0, G1, G2,
                                             ls[plc][0], ls[plc][1], ls[plc][2], w, t, owner);
      else
        fd_lines<DIM, K, CT, FSY, FSZ, false>(
          r, o, PAD ? A2[plc] : buf[plc], S_SCALAR ? G0 : Ss[plc][0], S_SCALAR ? G1 : Ss[plc][1],
          S_SCALAR ? G2 : Ss[plc][2], ls[plc][0], ls[plc][1], ls[plc][2], w, t, owner);
    }
    else
      fd_lines<DIM, K, CT, FSY, FSZ>(r, o, PAD ? A2[plc] : buf[plc], Ss[plc][0], Ss[plc][1],
                                     Ss[plc][DIM > 2 ? 2 : 1], ls[plc][0], ls[plc][1],
                                     ls[plc][DIM > 2 ? 2 : 1], w, t, owner);
  }
  if(!valid) return;
  if(!x_is_zero && w != CT(1))
  {
    CT old[K];
    load_row<NP>(x + c_lo * NDC + lyz, old);
    load_row<NP>(x + (c_lo + 1) * NDC + lyz, old + NP);
#pragma unroll
    for(int a = 0; a < K; ++a) o[a] += (CT(1) - w) * old[a];
  }
  if(nonfinite)
  {
    // the non-finite check of the caller's output, on the values stored here
    bool bad = false;
#pragma unroll
    for(int a = 0; a < K; ++a) bad = bad || !__builtin_isfinite(o[a]);
    if(bad) *nonfinite = 1;
  }
  store_row<NP>(x + c_lo * NDC + lyz, o);
  store_row<NP>(x + (c_lo + 1) * NDC + lyz, o + NP);
}

// BF16-MFMA patch solves of the block-diagonal elasticity smoother (3D Q2, k = 6),
// residual-update form (MODE 0 of k_patch_fd): dst += omega_j R_j^T A_jc^{-1} R_j r
// for one component block c, two patches per 128-thread block.  The 6 x 6 x 6
// block is embedded in the 8 x 8 x 8 layout of fd_mfma_pair and zero-padded: the
// padded rows and columns of the eigenvector matrices are zero, so padded entries
// stay zero through every contraction, and the padded eigenvalues are one (no
// division by zero).  src, dst, fac: the component's block and factor rows.
template<typename VT>
__global__ void __launch_bounds__(128)
k_patch_fd_bf16_q2(const VT *__restrict__ src, VT *__restrict__ dst,
                   const uint32_t *__restrict__ vtx, const uint32_t *__restrict__ fac,
                   const float *__restrict__ S_tab, const float *__restrict__ lam_tab,
                   const float *__restrict__ omega, const float relaxation,
                   const int patch_begin, const int patch_end, const int n)
{
  constexpr int K = 6, NP = 3, NDC = 27;
  __shared__ float B2[2][512];
  __shared__ float Ss[2][3][64];
  __shared__ float ls[2][3][8];
  __shared__ float w2[2];
  const int j0 = patch_begin + 2 * int(blockIdx.x);
  for(int idx = threadIdx.x; idx < 2 * 3 * 64; idx += 128)
  {
    const int q = idx / 192, r = idx % 192, d = r / 64, e = r % 64, a = e / 8, i = e % 8;
    const int j = j0 + q;
    float v = 0.f;
    if(j < patch_end && a < K && i < K) v = S_tab[size_t(fac[size_t(j) * 3 + d]) * K * K + a * K + i];
    Ss[q][d][e] = v;
  }
  for(int idx = threadIdx.x; idx < 2 * 3 * 8; idx += 128)
  {
    const int q = idx / 24, r = idx % 24, d = r / 8, i = r % 8;
    const int j = j0 + q;
    ls[q][d][i] = (j < patch_end && i < K) ? lam_tab[size_t(fac[size_t(j) * 3 + d]) * K + i] : 1.f;
  }
  if(threadIdx.x < 2)
  {
    const int j = j0 + int(threadIdx.x);
    w2[threadIdx.x] = j < patch_end ? (omega ? omega[j] : relaxation) : 0.f;
  }
  // global index of padded entry e (x fastest) of patch slot q, or -1
  auto entry = [&](int q, int e) -> int64_t {
    const int j = j0 + q, x = e % 8, y = (e / 8) % 8, z = e / 64;
    if(j >= patch_end || x >= K || y >= K || z >= K) return -1;
    int v[3];
    unpack_vertex(vtx[j], v);
    return patch_cell<3, NP>(v, n, x, y, z) * NDC + patch_local<3, NP>(x, y, z);
  };
  for(int idx = threadIdx.x; idx < 2 * 512; idx += 128)
  {
    const int64_t g = entry(idx / 512, idx % 512);
    B2[idx / 512][idx % 512] = g >= 0 ? float(src[g]) : 0.f;
  }
  __syncthreads();
  fd_mfma_pair(B2, Ss, ls, w2);  // ends with a barrier: B2 holds the corrections
  for(int idx = threadIdx.x; idx < 2 * 512; idx += 128)
  {
    const int64_t g = entry(idx / 512, idx % 512);
    if(g >= 0) dst[g] += VT(B2[idx / 512][idx % 512]);
  }
}

}  // namespace tpss
