// kernels.hpp -- device kernels of libtpss_amd (gfx950).
//
// Data layout in HBM (shared by every kernel):
//   level vector   x[cell * (p+1)^dim + local], cell = sum_d c_d n^d (x fastest),
//                  local lexicographic (x fastest) on Gauss-Lobatto nodes;
//   patch tables   vtx[j] = v_0 | v_1 << 10 | v_2 << 20 (interior vertex),
//                  fac[j*dim + d] = row of the 1D factor table for direction d;
//   factor table   S[f][a][i] (eigenvector i in column i), lam[f][i];
//   AVP buffer     U[(cell * 2^dim + slot) * (p+1)^dim + local], slot = position
//                  of the cell inside the patch that produced the value.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace tpss
{
constexpr __host__ __device__ int cpow(int b, int e) { return e == 0 ? 1 : b * cpow(b, e - 1); }

// ----------------------------------------------------------------------------
// SIPG operator on the Cartesian mesh (apps/laplace_integrator.h:578-743 for
// tensor-product cells): per cell and direction d the 1D line operator
//   z = D_d x_c + F_pm x_{c-e_d} + F_mp x_{c+e_d}   (along d)
// followed by the cell mass along every other direction, summed over d.
// One thread per cell DoF, CPB cells per block, lines staged in LDS.
// ----------------------------------------------------------------------------
template<typename T, int NP>
struct OpTab
{
  T M[NP * NP];
  T D[4][NP * NP];
  T Fpm[NP * NP];
  T Fmp[NP * NP];
};

// Contiguous rows of a cell (NP values along x) moved with 16-byte vector
// accesses; the types only claim element alignment, so any element offset is
// valid (global memory accesses need no 16-byte alignment on CDNA).
typedef float tpss_f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double tpss_d2u __attribute__((ext_vector_type(2), aligned(8)));

template<int NP, typename VT, typename CT>
__device__ __forceinline__ void load_row(const VT *__restrict__ p, CT *out)
{
  if constexpr(std::is_same<VT, float>::value && NP % 4 == 0)
  {
#pragma unroll
    for(int i = 0; i < NP; i += 4)
    {
      const tpss_f4u v = *reinterpret_cast<const tpss_f4u *>(p + i);
      out[i] = CT(v.x), out[i + 1] = CT(v.y), out[i + 2] = CT(v.z), out[i + 3] = CT(v.w);
    }
  }
  else if constexpr(std::is_same<VT, double>::value && NP % 2 == 0)
  {
#pragma unroll
    for(int i = 0; i < NP; i += 2)
    {
      const tpss_d2u v = *reinterpret_cast<const tpss_d2u *>(p + i);
      out[i] = CT(v.x), out[i + 1] = CT(v.y);
    }
  }
  else
  {
#pragma unroll
    for(int i = 0; i < NP; ++i) out[i] = CT(p[i]);
  }
}

template<int NP, typename VT, typename CT>
__device__ __forceinline__ void store_row(VT *__restrict__ p, const CT *in)
{
  if constexpr(std::is_same<VT, float>::value && NP % 4 == 0)
  {
#pragma unroll
    for(int i = 0; i < NP; i += 4)
    {
      tpss_f4u v;
      v.x = float(in[i]), v.y = float(in[i + 1]), v.z = float(in[i + 2]), v.w = float(in[i + 3]);
      *reinterpret_cast<tpss_f4u *>(p + i) = v;
    }
  }
  else if constexpr(std::is_same<VT, double>::value && NP % 2 == 0)
  {
#pragma unroll
    for(int i = 0; i < NP; i += 2)
    {
      tpss_d2u v;
      v.x = double(in[i]), v.y = double(in[i + 1]);
      *reinterpret_cast<tpss_d2u *>(p + i) = v;
    }
  }
  else
  {
#pragma unroll
    for(int i = 0; i < NP; ++i) p[i] = VT(in[i]);
  }
}

// value of v in lane SRC of this lane's quad (DPP quad_perm broadcast; every
// lane has a source, so no old value has to be set first)
template<int SRC>
__device__ __forceinline__ int quad_bcast_i(int v)
{
  return __builtin_amdgcn_mov_dpp(v, SRC * 0x55, 0xF, 0xF, true);
}
template<typename CT>
__device__ __forceinline__ CT quad_bcast_rt(CT v, int src)
{
  if constexpr(std::is_same<CT, float>::value)
  {
    const int i = __float_as_int(v);
    const int r = src == 0 ? quad_bcast_i<0>(i)
                : src == 1 ? quad_bcast_i<1>(i)
                : src == 2 ? quad_bcast_i<2>(i)
                           : quad_bcast_i<3>(i);
    return __int_as_float(r);
  }
  else
  {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    int rl, rh;
    if(src == 0)
      rl = quad_bcast_i<0>(lo), rh = quad_bcast_i<0>(hi);
    else if(src == 1)
      rl = quad_bcast_i<1>(lo), rh = quad_bcast_i<1>(hi);
    else if(src == 2)
      rl = quad_bcast_i<2>(lo), rh = quad_bcast_i<2>(hi);
    else
      rl = quad_bcast_i<3>(lo), rh = quad_bcast_i<3>(hi);
    return __hiloint2double(rh, rl);
  }
}

// value of v in lane `lane` of the wave (ds_bpermute)
template<typename CT>
__device__ __forceinline__ CT lane_read(CT v, int lane)
{
  if constexpr(std::is_same<CT, float>::value)
    return __int_as_float(__builtin_amdgcn_ds_bpermute(lane << 2, __float_as_int(v)));
  else
  {
    const int lo = __builtin_amdgcn_ds_bpermute(lane << 2, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(lane << 2, __double2hiint(v));
    return __hiloint2double(hi, lo);
  }
}

template<int DIM, int NP, typename VT, typename CT, int CPB, bool RESIDUAL>
__global__ void __launch_bounds__(256)
k_sipg_cartesian(const VT *__restrict__ x, VT *__restrict__ y, const VT *__restrict__ b,
                 const OpTab<CT, NP> tab, const int n, const int64_t ncell,
                 const uint32_t *__restrict__ cells, const int64_t n_list)
{
  constexpr int NDC = cpow(NP, DIM);
  __shared__ CT xs[CPB][NDC];
  __shared__ CT zs[CPB][NDC];
  const int cs = threadIdx.x / NDC, li = threadIdx.x % NDC;
  const int64_t slot = int64_t(blockIdx.x) * CPB + cs;
  const int64_t count = cells ? n_list : ncell;
  const bool active = cs < CPB && slot < count;
  const int64_t cell = active ? (cells ? int64_t(cells[slot]) : slot) : 0;
  const int csafe = cs < CPB ? cs : 0;

  int lid[DIM], cid[DIM];
  {
    int t = li;
    int c = int(cell);  // cells < n^3 <= 1000^3 < 2^31
    for(int d = 0; d < DIM; ++d)
    {
      lid[d] = t % NP;
      t /= NP;
      cid[d] = c % n;
      c /= n;
    }
  }
  if(cs < CPB) xs[cs][li] = active ? CT(x[cell * NDC + li]) : CT(0);
  __syncthreads();

  CT acc = 0;
  int64_t cstride = 1;
  int lstride = 1;
#pragma unroll
  for(int d = 0; d < DIM; ++d)
  {
    const bool has_l = cid[d] > 0, has_r = cid[d] < n - 1;
    const CT *D = tab.D[(has_l ? 1 : 0) | (has_r ? 2 : 0)];
    const int i = lid[d];
    const int base = li - i * lstride;
    CT z = 0;
#pragma unroll
    for(int j = 0; j < NP; ++j) z += D[i * NP + j] * xs[csafe][base + j * lstride];
    if(active && has_l)
    {
      const VT *xl = x + (cell - cstride) * NDC + base;
#pragma unroll
      for(int j = 0; j < NP; ++j) z += tab.Fpm[i * NP + j] * CT(xl[j * lstride]);
    }
    if(active && has_r)
    {
      const VT *xr = x + (cell + cstride) * NDC + base;
#pragma unroll
      for(int j = 0; j < NP; ++j) z += tab.Fmp[i * NP + j] * CT(xr[j * lstride]);
    }
    int estride = 1;
#pragma unroll
    for(int e = 0; e < DIM; ++e)
    {
      if(e != d)
      {
        __syncthreads();
        if(cs < CPB) zs[cs][li] = z;
        __syncthreads();
        const int ie = lid[e];
        const int be = li - ie * estride;
        CT w = 0;
#pragma unroll
        for(int j = 0; j < NP; ++j) w += tab.M[ie * NP + j] * zs[csafe][be + j * estride];
        z = w;
      }
      estride *= NP;
    }
    acc += z;
    cstride *= n;
    lstride *= NP;
  }
  if(active)
  {
    if(RESIDUAL)
      y[cell * NDC + li] = VT(CT(b[cell * NDC + li]) - acc);
    else
      y[cell * NDC + li] = VT(acc);
  }
}

// Row-tiled variant: a block owns TX consecutive cells of one x-row and stages
// that row (with its two x-halo cells) and the four neighbouring rows (+-y,
// +-z) in LDS with contiguous loads.  With all neighbours local the operator is
// evaluated in sum-factorised form
//   y = M_z (M_y z_0 + M_x z_1) + M_y M_x z_2    (3D)
//   y = M_y z_0 + M_x z_1                        (2D)
// where z_d is the 1D line operator along d (own cell + face neighbours).
template<int DIM, int NP>
constexpr int row_tile()
{
  return cpow(NP, DIM) >= 256 ? 1 : 256 / cpow(NP, DIM);
}

template<int DIM, int NP, typename VT, typename CT, bool RESIDUAL>
__global__ void __launch_bounds__(256)
k_sipg_rows(const VT *__restrict__ x, VT *__restrict__ y, const VT *__restrict__ b,
            const OpTab<CT, NP> tab, const int n)
{
  constexpr int NDC = cpow(NP, DIM);
  constexpr int TX = row_tile<DIM, NP>();
  constexpr int NROW = 2 * DIM - 1;  // own row + (2 DIM - 2) neighbour rows
  constexpr int TXH = TX + 2;
  __shared__ CT xs[NROW][TXH][NDC];
  __shared__ CT zs[DIM][TX][NDC];
  __shared__ CT as[2][TX][NDC];

  const int chunks = (n + TX - 1) / TX;
  int64_t rid = blockIdx.x;
  const int chunk = int(rid % chunks);
  rid /= chunks;
  const int c1 = int(rid % n);
  const int c2 = DIM == 3 ? int(rid / n) : 0;
  const int c0s = chunk * TX;
  const int64_t nn = int64_t(n) * n;
  const int64_t row_base = int64_t(c1) * n + (DIM == 3 ? int64_t(c2) * nn : 0);

  // ---- stage rows in LDS (contiguous cells along x)
  for(int idx = threadIdx.x; idx < NROW * TXH * NDC; idx += blockDim.x)
  {
    const int r = idx / (TXH * NDC);
    const int rem = idx % (TXH * NDC);
    const int hc = rem / NDC, l = rem % NDC;
    const int c0 = c0s - 1 + hc;
    int64_t base = -1;
    if(c0 >= 0 && c0 < n)
    {
      if(r == 0)
        base = row_base + c0;
      else if(r == 1 && c1 > 0)
        base = row_base - n + c0;
      else if(r == 2 && c1 < n - 1)
        base = row_base + n + c0;
      else if(r == 3 && c2 > 0)
        base = row_base - nn + c0;
      else if(r == 4 && c2 < n - 1)
        base = row_base + nn + c0;
    }
    xs[r][hc][l] = base >= 0 ? CT(x[base * NDC + l]) : CT(0);
  }
  __syncthreads();

  const int cs = threadIdx.x / NDC, li = threadIdx.x % NDC;
  const bool in_block = cs < TX;
  const int csafe = in_block ? cs : 0;
  const int c0 = c0s + csafe;
  const bool active = in_block && c0 < n;
  int lid[3] = {0, 0, 0};
  {
    int t = li;
    for(int d = 0; d < DIM; ++d)
    {
      lid[d] = t % NP;
      t /= NP;
    }
  }
  const int cid[3] = {c0, c1, c2};
  constexpr int LS[3] = {1, NP, NP * NP};

  // ---- 1D line operators z_d
#pragma unroll
  for(int d = 0; d < DIM; ++d)
  {
    const bool has_l = cid[d] > 0, has_r = cid[d] < n - 1;
    const CT *D = tab.D[(has_l ? 1 : 0) | (has_r ? 2 : 0)];
    const int i = lid[d];
    const int base = li - i * LS[d];
    const CT *own = xs[0][csafe + 1];
    const CT *lft = d == 0 ? xs[0][csafe] : xs[2 * d - 1][csafe + 1];
    const CT *rgt = d == 0 ? xs[0][csafe + 2] : xs[2 * d][csafe + 1];
    CT z = 0;
#pragma unroll
    for(int j = 0; j < NP; ++j)
    {
      const int o = base + j * LS[d];
      z += D[i * NP + j] * own[o];
      // outside the domain the staged rows hold zeros
      z += tab.Fpm[i * NP + j] * lft[o] + tab.Fmp[i * NP + j] * rgt[o];
    }
    if(in_block) zs[d][cs][li] = z;
  }
  __syncthreads();

  CT out = 0;
  if(DIM == 2)
  {
    const CT *z0 = zs[0][csafe], *z1 = zs[DIM > 1 ? 1 : 0][csafe];
    const int b1 = li - lid[1] * NP, b0 = li - lid[0];
#pragma unroll
    for(int j = 0; j < NP; ++j)
      out += tab.M[lid[1] * NP + j] * z0[b1 + j * NP] + tab.M[lid[0] * NP + j] * z1[b0 + j];
  }
  else
  {
    const CT *z0 = zs[0][csafe], *z1 = zs[1][csafe], *z2 = zs[DIM - 1][csafe];
    const int b0 = li - lid[0], b1 = li - lid[1] * NP;
    CT a = 0, bb = 0;
#pragma unroll
    for(int j = 0; j < NP; ++j)
    {
      a += tab.M[lid[1] * NP + j] * z0[b1 + j * NP] + tab.M[lid[0] * NP + j] * z1[b0 + j];
      bb += tab.M[lid[0] * NP + j] * z2[b0 + j];
    }
    if(in_block)
    {
      as[0][cs][li] = a;
      as[1][cs][li] = bb;
    }
    __syncthreads();
    const CT *a0 = as[0][csafe], *a1 = as[1][csafe];
    const int b2 = li - lid[2] * NP * NP;
#pragma unroll
    for(int j = 0; j < NP; ++j)
      out += tab.M[lid[2] * NP + j] * a0[b2 + j * NP * NP] + tab.M[lid[1] * NP + j] * a1[b1 + j * NP];
  }
  if(active)
  {
    const int64_t g = (row_base + c0) * NDC + li;
    if(RESIDUAL)
      y[g] = VT(CT(b[g]) - out);
    else
      y[g] = VT(out);
  }
}

// Line-tiled variant (3D): one thread per x-line (ly, lz) of a cell, its NP
// values in registers.  The x-direction line operator and the x mass
// contractions stay in registers; the y and z line operators read the cell's
// own planes from LDS and the face neighbours' planes (NP contiguous rows)
// from global memory, shared through the cache by the NP threads of a plane:
//   y = M_z (M_y z_0 + M_x z_1) + M_y M_x z_2,  z_d = line operator along d.
template<int NP>
constexpr int lines_cells_per_block()
{
  return 256 / (NP * NP);
}

// At least 4 waves per SIMD: FP64 Q3 fits 128 VGPRs without spills (130 unbounded,
// 3 waves; 4 waves measured 5 % faster at 64^3).
template<int NP, typename VT, typename CT, bool RESIDUAL>
__global__ void __launch_bounds__(256, 4)
k_sipg_lines(const VT *__restrict__ x, VT *__restrict__ y, const VT *__restrict__ b,
             const OpTab<CT, NP> tab, const int n, const int64_t ncell)
{
  constexpr int NL = NP * NP, NDC = NL * NP, CPB = lines_cells_per_block<NP>();
  __shared__ CT xs[CPB][NDC];  // own cells
  __shared__ CT zs[CPB][NDC];  // z_0, later M_y z_0 + M_x z_1
  __shared__ CT ws[CPB][NDC];  // M_x z_2
  const int cs = threadIdx.x / NL, ln = threadIdx.x % NL;
  const int ly = ln % NP, lz = ln / NP;
  const bool in = cs < CPB;
  const int csafe = in ? cs : 0;
  const int64_t cell = int64_t(blockIdx.x) * CPB + csafe;
  const bool act = in && cell < ncell;
  const int64_t nn = int64_t(n) * n;
  // 32-bit coordinates (cells < n^3 <= 1000^3 < 2^31): 64-bit divisions by a
  // run-time n are long instruction sequences
  const int ci = int(cell);
  const int cx = ci % n, cy = (ci / n) % n, cz = ci / (n * n);
  const int lofs = NP * ly + NL * lz;  // this line within the cell
  CT xo[NP];
  if(act)
    load_row<NP>(x + cell * NDC + lofs, xo);
  else
#pragma unroll
    for(int i = 0; i < NP; ++i) xo[i] = CT(0);
  if(in)
#pragma unroll
    for(int i = 0; i < NP; ++i) xs[cs][lofs + i] = xo[i];

  // z_0: along x, own line and the x-neighbours' lines
  CT z0[NP];
  {
    const bool hl = cx > 0, hr = cx < n - 1;
    const CT *D = tab.D[(hl ? 1 : 0) | (hr ? 2 : 0)];
    CT xl[NP], xr[NP];
#pragma unroll
    for(int i = 0; i < NP; ++i) xl[i] = xr[i] = CT(0);
    if(act && hl) load_row<NP>(x + (cell - 1) * NDC + lofs, xl);
    if(act && hr) load_row<NP>(x + (cell + 1) * NDC + lofs, xr);
#pragma unroll
    for(int i = 0; i < NP; ++i)
    {
      CT s = 0;
#pragma unroll
      for(int j = 0; j < NP; ++j)
        s += D[i * NP + j] * xo[j] + tab.Fpm[i * NP + j] * xl[j] + tab.Fmp[i * NP + j] * xr[j];
      z0[i] = s;
    }
  }
  __syncthreads();  // own cells staged

  // z_1 (along y) and z_2 (along z) for every x-node of this line
  CT z1[NP], z2[NP];
  {
    const bool hl = cy > 0, hr = cy < n - 1;
    const CT *D = tab.D[(hl ? 1 : 0) | (hr ? 2 : 0)];
#pragma unroll
    for(int i = 0; i < NP; ++i)
    {
      CT s = 0;
#pragma unroll
      for(int j = 0; j < NP; ++j) s += D[ly * NP + j] * xs[csafe][i + NP * j + NL * lz];
      z1[i] = s;
    }
#pragma unroll
    for(int side = 0; side < 2; ++side)
    {
      if(!(act && (side == 0 ? hl : hr))) continue;
      const CT *F = side == 0 ? tab.Fpm : tab.Fmp;
      const VT *pn = x + (side == 0 ? cell - n : cell + n) * NDC + NL * lz;
#pragma unroll
      for(int j = 0; j < NP; ++j)
      {
        CT row[NP];
        load_row<NP>(pn + NP * j, row);
        const CT f = F[ly * NP + j];
#pragma unroll
        for(int i = 0; i < NP; ++i) z1[i] += f * row[i];
      }
    }
  }
  {
    const bool hl = cz > 0, hr = cz < n - 1;
    const CT *D = tab.D[(hl ? 1 : 0) | (hr ? 2 : 0)];
#pragma unroll
    for(int i = 0; i < NP; ++i)
    {
      CT s = 0;
#pragma unroll
      for(int k = 0; k < NP; ++k) s += D[lz * NP + k] * xs[csafe][i + NP * ly + NL * k];
      z2[i] = s;
    }
#pragma unroll
    for(int side = 0; side < 2; ++side)
    {
      if(!(act && (side == 0 ? hl : hr))) continue;
      const CT *F = side == 0 ? tab.Fpm : tab.Fmp;
      const VT *pn = x + (side == 0 ? cell - nn : cell + nn) * NDC + NP * ly;
#pragma unroll
      for(int k = 0; k < NP; ++k)
      {
        CT row[NP];
        load_row<NP>(pn + NL * k, row);
        const CT f = F[lz * NP + k];
#pragma unroll
        for(int i = 0; i < NP; ++i) z2[i] += f * row[i];
      }
    }
  }
  // x mass contractions in registers
  CT m1[NP], w2[NP];
#pragma unroll
  for(int i = 0; i < NP; ++i)
  {
    CT s1 = 0, s2 = 0;
#pragma unroll
    for(int j = 0; j < NP; ++j)
    {
      s1 += tab.M[i * NP + j] * z1[j];
      s2 += tab.M[i * NP + j] * z2[j];
    }
    m1[i] = s1;
    w2[i] = s2;
  }
  if(in)
#pragma unroll
    for(int i = 0; i < NP; ++i)
    {
      zs[cs][lofs + i] = z0[i];
      ws[cs][lofs + i] = w2[i];
    }
  __syncthreads();
  CT a[NP], bb[NP];
#pragma unroll
  for(int i = 0; i < NP; ++i)
  {
    CT s = m1[i], s2 = 0;
#pragma unroll
    for(int j = 0; j < NP; ++j)
    {
      const CT m = tab.M[ly * NP + j];
      s += m * zs[csafe][i + NP * j + NL * lz];
      s2 += m * ws[csafe][i + NP * j + NL * lz];
    }
    a[i] = s;
    bb[i] = s2;
  }
  __syncthreads();
  if(in)
#pragma unroll
    for(int i = 0; i < NP; ++i) zs[cs][lofs + i] = a[i];
  __syncthreads();
  CT out[NP];
#pragma unroll
  for(int i = 0; i < NP; ++i)
  {
    CT s = bb[i];
#pragma unroll
    for(int k = 0; k < NP; ++k) s += tab.M[lz * NP + k] * zs[csafe][i + NP * ly + NL * k];
    out[i] = s;
  }
  if(!act) return;
  if(RESIDUAL)
  {
    CT bv[NP];
    load_row<NP>(b + cell * NDC + lofs, bv);
#pragma unroll
    for(int i = 0; i < NP; ++i) out[i] = bv[i] - out[i];
  }
  store_row<NP>(y + cell * NDC + lofs, out);
}

// Q3 form of k_sipg_lines (NP = 4: a cell's 16 lines are 16 lanes of one wave).
// The generic form loads the NP rows of every y- and z-neighbour plane in each of
// the plane's NP lanes, behind a branch per side (about 20 row loads per lane,
// one memory round trip after another).  Here every lane loads seven rows at
// once -- its line of the cell and the row at the same position in each of the
// six face neighbours (absent neighbours read the cell itself and are zeroed by
// selection) -- and takes the other rows of a neighbour plane from the lanes
// that loaded them: those of a y-plane are the lanes of its quad (DPP
// broadcast), those of a z-plane sit 4 lanes apart (staged in LDS, laid out
// as the neighbour cell).  The 1D
// tables indexed by a lane-varying row (mass, face coupling, boundary-dependent
// cell matrices) are read from LDS.  Same arithmetic, in the same order, as
// k_sipg_lines.  DOT (operator form): also the block's part of x.y in
// partial[blockIdx.x] (the CG's p.Ap without a second pass over p and Ap).
template<typename VT, typename CT, bool RESIDUAL, bool DOT = false>
__global__ void __launch_bounds__(256, 4)
k_sipg_lines_q3(const VT *__restrict__ x, VT *__restrict__ y, const VT *__restrict__ b,
                const OpTab<CT, 4> tab, const int n, const int64_t ncell,
                double *__restrict__ partial = nullptr)
{
  static_assert(!(DOT && RESIDUAL), "the dot is formed by the operator form");
  constexpr int NP = 4, NL = 16, NDC = 64, CPB = 16;
  static_assert(CPB * NL == 256, "16 cells of 16 lines per block");
  __shared__ CT xs[CPB][NDC];  // own cells
  __shared__ CT zs[CPB][NDC];  // z_0, later M_y z_0 + M_x z_1
  __shared__ CT ws[CPB][NDC];  // M_x z_2
  // the 1D tables in the layout of OpTab: M, D[0..3], F+-, F-+
  constexpr int TB = int(sizeof(OpTab<CT, 4>) / sizeof(CT));
  static_assert(TB == 7 * NL, "OpTab<CT, 4> holds seven 4 x 4 tables");
  __shared__ CT tb[256];  // every thread stores (a guarded store would pull its load behind the rows)
  const CT *tM = tb, *tFpm = tb + 5 * NL, *tFmp = tb + 6 * NL;
  const int cs = threadIdx.x / NL, ln = threadIdx.x % NL;
  const int ly = ln % NP, lz = ln / NP;
  const int64_t cell = int64_t(blockIdx.x) * CPB + cs;
  const bool act = cell < ncell;
  const int64_t cc = act ? cell : ncell - 1;  // idle lanes load an existing cell
  const int64_t nn = int64_t(n) * n;
  const int ci = int(cc);
  const int cx = ci % n, cy = (ci / n) % n, cz = ci / (n * n);
  const bool hxl = cx > 0, hxr = cx < n - 1, hyl = cy > 0, hyr = cy < n - 1;
  const bool hzl = cz > 0, hzr = cz < n - 1;
  const int lofs = NP * ly + NL * lz;  // this line within the cell, and the row
                                       // this lane loads from every neighbour
  // this thread's table entry, loaded with the rows (staged in LDS below)
  const CT tv = reinterpret_cast<const CT *>(&tab)[min(int(threadIdx.x), TB - 1)];
  CT xo[NP], xl[NP], xr[NP], ya[NP], yb[NP], za[NP], zb[NP], bv[NP];
  load_row<NP>(x + cc * NDC + lofs, xo);
  load_row<NP>(x + (hxl ? cc - 1 : cc) * NDC + lofs, xl);
  load_row<NP>(x + (hxr ? cc + 1 : cc) * NDC + lofs, xr);
  load_row<NP>(x + (hyl ? cc - n : cc) * NDC + lofs, ya);
  load_row<NP>(x + (hyr ? cc + n : cc) * NDC + lofs, yb);
  load_row<NP>(x + (hzl ? cc - nn : cc) * NDC + lofs, za);
  load_row<NP>(x + (hzr ? cc + nn : cc) * NDC + lofs, zb);
  if constexpr(RESIDUAL) load_row<NP>(b + cc * NDC + lofs, bv);
#pragma unroll
  for(int i = 0; i < NP; ++i)
  {
    xl[i] = hxl ? xl[i] : CT(0);
    xr[i] = hxr ? xr[i] : CT(0);
    ya[i] = hyl ? ya[i] : CT(0);
    yb[i] = hyr ? yb[i] : CT(0);
    za[i] = hzl ? za[i] : CT(0);
    zb[i] = hzr ? zb[i] : CT(0);
    xs[cs][lofs + i] = xo[i];
  }
  tb[threadIdx.x] = tv;
  __syncthreads();  // own cells and tables staged

  // z_0: along x, own line and the x-neighbours' lines
  CT z0[NP];
  {
    const CT *D = tb + NL * (1 + ((hxl ? 1 : 0) | (hxr ? 2 : 0)));
#pragma unroll
    for(int i = 0; i < NP; ++i)
    {
      CT s = 0;
#pragma unroll
      for(int j = 0; j < NP; ++j)
        s += D[i * NP + j] * xo[j] + tab.Fpm[i * NP + j] * xl[j] + tab.Fmp[i * NP + j] * xr[j];
      z0[i] = s;
    }
  }
  // z_1 (along y) and z_2 (along z) for every x-node of this line
  CT z1[NP], z2[NP];
  {
    const CT *D = tb + NL * (1 + ((hyl ? 1 : 0) | (hyr ? 2 : 0)));
#pragma unroll
    for(int i = 0; i < NP; ++i)
    {
      CT s = 0;
#pragma unroll
      for(int j = 0; j < NP; ++j) s += D[ly * NP + j] * xs[cs][i + NP * j + NL * lz];
      z1[i] = s;
    }
#pragma unroll
    for(int side = 0; side < 2; ++side)
    {
      const CT *F = side == 0 ? tFpm : tFmp;
#pragma unroll
      for(int j = 0; j < NP; ++j)
      {
        // row j of the neighbour's plane lz: loaded by lane j of this quad
        const CT f = F[ly * NP + j];
#pragma unroll
        for(int i = 0; i < NP; ++i) z1[i] += f * quad_bcast_rt(side == 0 ? ya[i] : yb[i], j);
      }
    }
  }
  {
    const CT *D = tb + NL * (1 + ((hzl ? 1 : 0) | (hzr ? 2 : 0)));
#pragma unroll
    for(int i = 0; i < NP; ++i)
    {
      CT s = 0;
#pragma unroll
      for(int k = 0; k < NP; ++k) s += D[lz * NP + k] * xs[cs][i + NP * ly + NL * k];
      z2[i] = s;
    }
    // the z-neighbours' rows of every lane through LDS (zs / ws are free until
    // z_0 and M_x z_2 are staged there): each neighbour cell laid out like xs
#pragma unroll
    for(int i = 0; i < NP; ++i)
    {
      zs[cs][lofs + i] = za[i];
      ws[cs][lofs + i] = zb[i];
    }
    __syncthreads();
#pragma unroll
    for(int side = 0; side < 2; ++side)
    {
      const CT *F = side == 0 ? tFpm : tFmp;
      const CT *nb = side == 0 ? zs[cs] : ws[cs];
#pragma unroll
      for(int k = 0; k < NP; ++k)
      {
        // row (ly, k) of the neighbour
        const CT f = F[lz * NP + k];
#pragma unroll
        for(int i = 0; i < NP; ++i) z2[i] += f * nb[i + NP * ly + NL * k];
      }
    }
    __syncthreads();  // every lane has read the rows before zs / ws are rewritten
  }
  // x mass contractions in registers
  CT m1[NP], w2[NP];
#pragma unroll
  for(int i = 0; i < NP; ++i)
  {
    CT s1 = 0, s2 = 0;
#pragma unroll
    for(int j = 0; j < NP; ++j)
    {
      s1 += tab.M[i * NP + j] * z1[j];
      s2 += tab.M[i * NP + j] * z2[j];
    }
    m1[i] = s1;
    w2[i] = s2;
  }
#pragma unroll
  for(int i = 0; i < NP; ++i)
  {
    zs[cs][lofs + i] = z0[i];
    ws[cs][lofs + i] = w2[i];
  }
  __syncthreads();
  CT a[NP], bb[NP];
#pragma unroll
  for(int i = 0; i < NP; ++i)
  {
    CT s = m1[i], s2 = 0;
#pragma unroll
    for(int j = 0; j < NP; ++j)
    {
      const CT m = tM[ly * NP + j];
      s += m * zs[cs][i + NP * j + NL * lz];
      s2 += m * ws[cs][i + NP * j + NL * lz];
    }
    a[i] = s;
    bb[i] = s2;
  }
  __syncthreads();
#pragma unroll
  for(int i = 0; i < NP; ++i) zs[cs][lofs + i] = a[i];
  __syncthreads();
  CT out[NP];
#pragma unroll
  for(int i = 0; i < NP; ++i)
  {
    CT s = bb[i];
#pragma unroll
    for(int k = 0; k < NP; ++k) s += tM[lz * NP + k] * zs[cs][i + NP * ly + NL * k];
    out[i] = s;
  }
  if constexpr(DOT)
  {
    __shared__ double wsum[4];
    double d = 0.0;
    if(act)
#pragma unroll
      for(int i = 0; i < NP; ++i) d += double(xo[i]) * double(out[i]);
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off);
    if((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = d;
    __syncthreads();
    if(threadIdx.x == 0) partial[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
  }
  if(!act) return;
  if(RESIDUAL)
#pragma unroll
    for(int i = 0; i < NP; ++i) out[i] = bv[i] - out[i];
  store_row<NP>(y + cell * NDC + lofs, out);
}

// ----------------------------------------------------------------------------
// Fused vertex-patch local solve (A2/A3/A4/A5/A12):
//   gather r_j = R_j src  ->  u_j = S diag(omega_j / sum lambda) S^T r_j  ->
//   MODE 0: dst += R_j^T u_j   (multiplicative color pass; same-color patches
//                               are DoF-disjoint so the update is race free)
//   MODE 1: U[cell, slot] = u_j (additive; summed per DoF afterwards, no atomics)
// Thread = one line of the patch tensor; the line direction rotates through
// LDS between the per-direction contractions.
// ----------------------------------------------------------------------------
template<int DIM, int K>
struct PatchShape
{
  static constexpr int NP = K / 2;
  static constexpr int L = cpow(K, DIM - 1);   // lines per patch
  static constexpr int NPD = cpow(K, DIM);     // DoFs per patch
  static constexpr int NDC = cpow(NP, DIM);    // DoFs per cell
  static constexpr int PPB = (128 / L) > 0 ? (128 / L) : 1;
  static constexpr int THREADS = ((PPB * L + 63) / 64) * 64;  // launched block size
};

inline __device__ __forceinline__ void unpack_vertex(uint32_t w, int *v)
{
  v[0] = int(w & 1023u);
  v[1] = int((w >> 10) & 1023u);
  v[2] = int((w >> 20) & 1023u);
}

// Patch-local multi-index (i0, i1, i2), i_d in [0, 2(p+1)), of the patch at
// interior vertex v -> global cell, cell-local DoF and the cell's position in
// the patch.  The DGQ 1D map i_d = cell_no*(p+1) + local (dof_info.h:358-369)
// with cells of the patch in lexicographic order (patch_info.templates.h:401-402).
template<int DIM, int NP>
__device__ __forceinline__ int64_t patch_cell(const int *v, int n, int i0, int i1, int i2)
{
  int64_t c = v[0] - 1 + (i0 >= NP ? 1 : 0);
  if(DIM > 1) c += int64_t(n) * (v[1] - 1 + (i1 >= NP ? 1 : 0));
  if(DIM > 2) c += int64_t(n) * n * (v[2] - 1 + (i2 >= NP ? 1 : 0));
  return c;
}

template<int DIM, int NP>
__device__ __forceinline__ int patch_local(int i0, int i1, int i2)
{
  int l = i0 >= NP ? i0 - NP : i0;
  if(DIM > 1) l += NP * (i1 >= NP ? i1 - NP : i1);
  if(DIM > 2) l += NP * NP * (i2 >= NP ? i2 - NP : i2);
  return l;
}

template<int DIM, int NP>
__device__ __forceinline__ int patch_slot(int i0, int i1, int i2)
{
  int s = (i0 >= NP ? 1 : 0);
  if(DIM > 1) s |= (i1 >= NP ? 2 : 0);
  if(DIM > 2) s |= (i2 >= NP ? 4 : 0);
  return s;
}

// the gather/scatter index table: out[j * K^dim + i] = global DoF of patch DoF i
template<int DIM, int K>
__global__ void k_patch_dof_indices(const uint32_t *__restrict__ vtx, const int P, const int n,
                                    uint32_t *__restrict__ out)
{
  constexpr int NP = K / 2, NPD = cpow(K, DIM), NDC = cpow(NP, DIM);
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(g >= int64_t(P) * NPD) return;
  const int j = int(g / NPD), i = int(g % NPD);
  int v[3];
  unpack_vertex(vtx[j], v);
  const int i0 = i % K, i1 = DIM > 1 ? (i / K) % K : 0, i2 = DIM > 2 ? i / (K * K) : 0;
  out[g] = uint32_t(patch_cell<DIM, NP>(v, n, i0, i1, i2) * NDC + patch_local<DIM, NP>(i0, i1, i2));
}

// s * w / lam.  FP32: the hardware reciprocal (1 ulp), well inside the FP32 patch
// solve's rounding.  FP64: the hardware reciprocal refined by two Newton steps
// r <- r + r (1 - lam r) (each doubles the correct bits; within an ulp or two of
// 1/lam), five instructions instead of the correctly rounded division's scale,
// refinement and fix-up sequence
template<typename CT>
__device__ __forceinline__ CT eig_scale(CT s, CT w, CT lam)
{
  if constexpr(std::is_same<CT, float>::value)
    return s * w * __builtin_amdgcn_rcpf(lam);
  else
  {
    double r = __builtin_amdgcn_rcp(lam);
    double e = fma(-lam, r, 1.0);
    r = fma(r, e, r);
    e = fma(-lam, r, 1.0);
    r = fma(r, e, r);
    return s * w * r;
  }
}

// o = S^T r (forward) and r = S o (backward) along one line, S[a * K + i] =
// component a of eigenvector i.  even_odd: the first K/2 eigenvectors are
// symmetric, the last K/2 antisymmetric under the reversal a -> K-1-a (the 1D
// problems of patches whose two cells have equal lengths and no boundary
// face), so each product takes two (K/2 x K/2) blocks of S on the folded line
// r_a +- r_{K-1-a}: half the multiply-adds.  (A compile-time choice: a
// run-time one keeps the line arrays out of registers.)
template<int K, bool EVEN_ODD, typename CT>
__device__ __forceinline__ void s_forward(const CT *S, const CT *r, CT *o)
{
  constexpr int P = K / 2;
  if constexpr(EVEN_ODD)
  {
    CT rp[P], rm[P];
#pragma unroll
    for(int a = 0; a < P; ++a)
    {
      rp[a] = r[a] + r[K - 1 - a];
      rm[a] = r[a] - r[K - 1 - a];
    }
#pragma unroll
    for(int k = 0; k < P; ++k)
    {
      CT sp = 0, sm = 0;
#pragma unroll
      for(int a = 0; a < P; ++a)
      {
        sp += S[a * K + k] * rp[a];
        sm += S[a * K + P + k] * rm[a];
      }
      o[k] = sp;
      o[P + k] = sm;
    }
  }
  else
  {
#pragma unroll
    for(int i = 0; i < K; ++i)
    {
      CT s = 0;
#pragma unroll
      for(int a = 0; a < K; ++a) s += S[a * K + i] * r[a];
      o[i] = s;
    }
  }
}

template<int K, bool EVEN_ODD, typename CT>
__device__ __forceinline__ void s_backward(const CT *S, const CT *o, CT *r)
{
  constexpr int P = K / 2;
  if constexpr(EVEN_ODD)
  {
#pragma unroll
    for(int a = 0; a < P; ++a)
    {
      CT sp = 0, sm = 0;
#pragma unroll
      for(int k = 0; k < P; ++k)
      {
        sp += S[a * K + k] * o[k];
        sm += S[a * K + P + k] * o[P + k];
      }
      r[a] = sp + sm;
      r[K - 1 - a] = sp - sm;
    }
  }
  else
  {
#pragma unroll
    for(int a = 0; a < K; ++a)
    {
      CT s = 0;
#pragma unroll
      for(int i = 0; i < K; ++i) s += S[a * K + i] * o[i];
      r[a] = s;
    }
  }
}

// Fast-diagonalization inverse of one patch held as x-lines: thread t owns the
// x-line (y, z) = (t % K, t / K); on entry r is that line of the patch
// right-hand side, on exit o is the same line of
//   u = (S_2 x S_1 x S_0) diag(w / (lam_0 + lam_1 + lam_2)) (S_2^T x S_1^T x S_0^T) r.
// The line direction rotates x -> y -> z -> y -> x through the patch's LDS
// buffer B (entry (x, y, z) at x + SY y + SZ z: padded strides keep the
// transposes free of LDS bank conflicts); EO: even-odd products in every
// direction.  Contains __syncthreads (every thread of the block must call it).
template<int DIM, int K, typename CT, int SY = K, int SZ = K * K, bool EO = false>
__device__ __forceinline__ void fd_lines(CT *r, CT *o, CT *B, const CT *S0, const CT *S1,
                                         const CT *S2, const CT *l0, const CT *l1, const CT *l2,
                                         const CT w, const int t, const bool owner)
{
  const int ty = DIM > 1 ? t % K : 0;
  const int tz = DIM > 2 ? t / K : 0;
  s_forward<K, EO>(S0, r, o);
  if(DIM == 2)
  {
#pragma unroll
    for(int i = 0; i < K; ++i)
      if(owner) B[i + SY * ty] = o[i];
    __syncthreads();
    const int x0 = t;
#pragma unroll
    for(int jj = 0; jj < K; ++jj) r[jj] = B[x0 + SY * jj];
    s_forward<K, EO>(S1, r, o);
#pragma unroll
    for(int jj = 0; jj < K; ++jj) o[jj] = eig_scale(o[jj], w, l0[x0] + l1[jj]);
    s_backward<K, EO>(S1, o, r);
    __syncthreads();
#pragma unroll
    for(int b = 0; b < K; ++b)
      if(owner) B[x0 + SY * b] = r[b];
    __syncthreads();
#pragma unroll
    for(int i = 0; i < K; ++i) r[i] = B[i + SY * ty];
  }
  else
  {
#pragma unroll
    for(int i = 0; i < K; ++i)
      if(owner) B[i + SY * ty + SZ * tz] = o[i];
    __syncthreads();
    {
      const int x0 = t % K, z0 = t / K;
#pragma unroll
      for(int jj = 0; jj < K; ++jj) r[jj] = B[x0 + SY * jj + SZ * z0];
      s_forward<K, EO>(S1, r, o);
#pragma unroll
      for(int jj = 0; jj < K; ++jj)
        if(owner) B[x0 + SY * jj + SZ * z0] = o[jj];
    }
    __syncthreads();
    {
      const int x0 = t % K, y0 = t / K;
#pragma unroll
      for(int l = 0; l < K; ++l) r[l] = B[x0 + SY * y0 + SZ * l];
      const CT lxy = l0[x0] + l1[y0];
      s_forward<K, EO>(S2, r, o);
#pragma unroll
      for(int l = 0; l < K; ++l) o[l] = eig_scale(o[l], w, lxy + l2[l]);
      s_backward<K, EO>(S2, o, r);
#pragma unroll
      for(int c = 0; c < K; ++c)
        if(owner) B[x0 + SY * y0 + SZ * c] = r[c];
    }
    __syncthreads();
    {
      const int x0 = t % K, z0 = t / K;
#pragma unroll
      for(int jj = 0; jj < K; ++jj) r[jj] = B[x0 + SY * jj + SZ * z0];
      s_backward<K, EO>(S1, r, o);
#pragma unroll
      for(int b = 0; b < K; ++b)
        if(owner) B[x0 + SY * b + SZ * z0] = o[b];
    }
    __syncthreads();
#pragma unroll
    for(int i = 0; i < K; ++i) r[i] = B[i + SY * ty + SZ * tz];
  }
  s_backward<K, EO>(S0, r, o);
}

template<int DIM, int K, typename VT, typename CT, int MODE>
__global__ void __launch_bounds__((PatchShape<DIM, K>::THREADS))
k_patch_fd(const VT *__restrict__ src, VT *__restrict__ dst, CT *__restrict__ ubuf,
           const uint32_t *__restrict__ vtx, const uint32_t *__restrict__ fac,
           const CT *__restrict__ S_tab, const CT *__restrict__ lam_tab,
           const float *__restrict__ omega, const CT relaxation, const int patch_begin,
           const int patch_end, const int n)
{
  using PS = PatchShape<DIM, K>;
  constexpr int NP = PS::NP, L = PS::L, NPD = PS::NPD, NDC = PS::NDC, PPB = PS::PPB;
  __shared__ CT buf[PPB][NPD];
  __shared__ CT Ss[PPB][DIM][K * K];
  __shared__ CT ls[PPB][DIM][K];

  const int pl = threadIdx.x / L, t = threadIdx.x % L;
  const int j0 = patch_begin + int(blockIdx.x) * PPB;
  const int j = j0 + pl;
  const bool valid = pl < PPB && j < patch_end;
  const int plc = pl < PPB ? pl : 0;
  constexpr int NT = PS::THREADS;

  // Every load of the pass in three batches instead of one memory round trip
  // per table entry and line element: the block's factor rows and this lane's
  // vertex; the table entries this thread stages and its patch line (two cell
  // rows of b, in MODE 0 also of dst -- the cells of a color's patches are
  // disjoint, so no other patch of the launch writes them); then the LDS
  // stores.  Loads are predicated, not branched, so that the compiler issues
  // them together.
  constexpr int NSE = PPB * DIM * K * K, NSI = (NSE + NT - 1) / NT;  // eigenvector entries
  constexpr int NLE = PPB * DIM * K, NLI = (NLE + NT - 1) / NT;      // eigenvalues
  uint32_t fs[NSI], fl[NLI];
#pragma unroll
  for(int k = 0; k < NSI; ++k)
  {
    const int idx = int(threadIdx.x) + k * NT;
    const int q = idx / (DIM * K * K), d = (idx % (DIM * K * K)) / (K * K);
    const bool ok = idx < NSE && j0 + q < patch_end;
    fs[k] = fac[size_t(ok ? j0 + q : patch_begin) * DIM + (ok ? d : 0)];
  }
#pragma unroll
  for(int k = 0; k < NLI; ++k)
  {
    const int idx = int(threadIdx.x) + k * NT;
    const int q = idx / (DIM * K), d = (idx % (DIM * K)) / K;
    const bool ok = idx < NLE && j0 + q < patch_end;
    fl[k] = fac[size_t(ok ? j0 + q : patch_begin) * DIM + (ok ? d : 0)];
  }
  int v[3] = {1, 1, 1};
  unpack_vertex(vtx[valid ? j : patch_begin], v);
  const CT w = valid ? (omega ? CT(omega[j]) : relaxation) : CT(0);

  auto cell_of = [&](int i0, int i1, int i2) { return patch_cell<DIM, NP>(v, n, i0, i1, i2); };
  auto loc_of = [&](int i0, int i1, int i2) { return patch_local<DIM, NP>(i0, i1, i2); };
  auto slot_of = [&](int i0, int i1, int i2) { return patch_slot<DIM, NP>(i0, i1, i2); };

  CT sv[NSI], lv[NLI];
#pragma unroll
  for(int k = 0; k < NSI; ++k)
  {
    const int idx = int(threadIdx.x) + k * NT;
    sv[k] = S_tab[size_t(fs[k]) * K * K + idx % (K * K)];
  }
#pragma unroll
  for(int k = 0; k < NLI; ++k)
  {
    const int idx = int(threadIdx.x) + k * NT;
    lv[k] = lam_tab[size_t(fl[k]) * K + idx % K];
  }
  // this thread's x-line (ty, tz) of the patch: row lyz of the cells c_lo, c_lo + 1
  const int ty = DIM > 1 ? t % K : 0;
  const int tz = DIM > 2 ? t / K : 0;
  const int64_t c_lo = cell_of(0, ty, tz);
  const int lyz = loc_of(0, ty, tz);
  CT r[K], o[K], old[K];
  load_row<NP>(src + c_lo * NDC + lyz, r);
  load_row<NP>(src + (c_lo + 1) * NDC + lyz, r + NP);
  if constexpr(MODE == 0)
  {
    load_row<NP>(dst + c_lo * NDC + lyz, old);
    load_row<NP>(dst + (c_lo + 1) * NDC + lyz, old + NP);
  }
#pragma unroll
  for(int k = 0; k < NSI; ++k)
  {
    const int idx = int(threadIdx.x) + k * NT;
    const int q = idx / (DIM * K * K), rr = idx % (DIM * K * K);
    const bool ok = idx < NSE;
    if(ok) Ss[q][rr / (K * K)][rr % (K * K)] = j0 + q < patch_end ? sv[k] : CT(0);
  }
#pragma unroll
  for(int k = 0; k < NLI; ++k)
  {
    const int idx = int(threadIdx.x) + k * NT;
    const int q = idx / (DIM * K), rr = idx % (DIM * K);
    if(idx < NLE) ls[q][rr / K][rr % K] = j0 + q < patch_end ? lv[k] : CT(1);
  }
  if(!valid)
#pragma unroll
    for(int i = 0; i < K; ++i) r[i] = CT(0);
  __syncthreads();

  const CT *S0 = Ss[plc][0];
  const CT *S1 = Ss[plc][1];
  const CT *S2 = Ss[plc][DIM > 2 ? 2 : 1];
  CT *B = buf[plc];

  fd_lines<DIM, K, CT>(r, o, B, S0, S1, S2, ls[plc][0], ls[plc][1], ls[plc][DIM > 2 ? 2 : 1], w,
                       t, pl < PPB);
  if(!valid) return;
  if constexpr(MODE == 0)
  {
#pragma unroll
    for(int a = 0; a < K; ++a) o[a] += old[a];
    store_row<NP>(dst + c_lo * NDC + lyz, o);
    store_row<NP>(dst + (c_lo + 1) * NDC + lyz, o + NP);
  }
  else
  {
    const int s0 = slot_of(0, ty, tz);
    store_row<NP>(ubuf + (c_lo * (1 << DIM) + s0) * NDC + lyz, o);
    store_row<NP>(ubuf + ((c_lo + 1) * (1 << DIM) + (s0 | 1)) * NDC + lyz, o + NP);
  }
}

// Restricted additive Schwarz (rscatter_add, patch_transfer.templates.h:140-158):
// of the patches that contain DoF (cell c, local node l), the one that keeps it
// in the restricted prolongation -- among the patches at interior vertices of
// the cell, the one whose patch multi-index m of the DoF minimises
// (max_d floor|K/2 - 1/2 - m_d|, lexicographic index of m), the order of
// compute_restricted_dofs_impl (dof_info.templates.h:192-313; ties cannot remain,
// two patches never hold one DoF at the same multi-index).  Returns the DoF's
// slot in that patch: bit d set if the cell is the upper one in direction d.
template<int DIM, int NP>
__device__ __forceinline__ int ras_owner_slot(const int *c, const int *l, const int n)
{
  int best = -1;
  int bd = 1 << 30, bi = 1 << 30;
#pragma unroll
  for(int slot = 0; slot < (1 << DIM); ++slot)
  {
    bool ok = true;
    int dist = 0, flat = 0, st = 1;
#pragma unroll
    for(int d = 0; d < DIM; ++d)
    {
      const int bit = (slot >> d) & 1;
      const int vd = c[d] + 1 - bit;
      ok = ok && vd >= 1 && vd <= n - 1;
      // |NP - 1/2 - m_d| rounded down: NP - 1 - l_d (lower cell), l_d (upper cell)
      dist = max(dist, bit ? l[d] : NP - 1 - l[d]);
      flat += (bit * NP + l[d]) * st;
      st *= 2 * NP;
    }
    if(ok && (dist < bd || (dist == bd && flat < bi)))
    {
      best = slot;
      bd = dist;
      bi = flat;
    }
  }
  return best;
}

// dst[i] (=|+=) alpha * sum over the patches containing DoF i of their buffered
// contribution (additive Schwarz without atomics).  RAS = 1: the sum divided by
// the number of those patches (use_ras_weights, schwarz_preconditioner.templates.h:
// 636-641); RAS = 2: only the contribution of the patch that keeps the DoF
// (use_ras_boolean_weights, rscatter_add, :646-649).
template<int DIM, int NP, typename VT, typename CT, bool ACCUMULATE, int RAS = 0>
__global__ void __launch_bounds__(256)
k_avp_sum(const CT *__restrict__ ubuf, VT *__restrict__ dst, const int n, const int64_t N,
          const double alpha)
{
  constexpr int NDC = cpow(NP, DIM);
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= N) return;
  const int64_t cell = i / NDC;
  const int l = int(i % NDC);
  int c[3] = {0, 0, 0};
  {
    int t = int(cell);  // cells < 2^24
    for(int d = 0; d < DIM; ++d)
    {
      c[d] = t % n;
      t /= n;
    }
  }
  CT s = 0;
  if constexpr(RAS == 2)
  {
    int lm[3] = {0, 0, 0};
    for(int d = 0, t = l; d < DIM; ++d, t /= NP) lm[d] = t % NP;
    const int slot = ras_owner_slot<DIM, NP>(c, lm, n);
    s = ubuf[(cell * (1 << DIM) + slot) * NDC + l];
  }
  else
  {
    int count = 0;
#pragma unroll
    for(int slot = 0; slot < (1 << DIM); ++slot)
    {
      bool ok = true;
#pragma unroll
      for(int d = 0; d < DIM; ++d)
      {
        const int vd = c[d] + 1 - ((slot >> d) & 1);
        ok = ok && vd >= 1 && vd <= n - 1;
      }
      if(ok)
      {
        s += ubuf[(cell * (1 << DIM) + slot) * NDC + l];
        ++count;
      }
    }
    if constexpr(RAS == 1) s /= CT(count);
  }
  if(ACCUMULATE)
    dst[i] = VT(double(dst[i]) + alpha * double(s));
  else
    dst[i] = VT(alpha * double(s));
}

// restricted-DoF flags of every patch (DoFInfo::restricted_dof_flags_patchwise,
// dof_info.templates.h:192-313): out[j * K^DIM + i] = 1 if patch j does not keep
// its patch DoF i in the restricted prolongation
template<int DIM, int NP>
__global__ void k_restricted_flags(const uint32_t *__restrict__ vtx, const int P, const int n,
                                   uint8_t *__restrict__ out)
{
  constexpr int K = 2 * NP, NPD = cpow(K, DIM);
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(g >= int64_t(P) * NPD) return;
  const int j = int(g / NPD), i = int(g % NPD);
  int v[3] = {1, 1, 1};
  unpack_vertex(vtx[j], v);
  int c[3] = {0, 0, 0}, lm[3] = {0, 0, 0}, slot = 0;
  for(int d = 0, t = i; d < DIM; ++d, t /= K)
  {
    const int m = t % K, bit = m / NP;
    lm[d] = m % NP;
    c[d] = v[d] - 1 + bit;
    slot |= bit << d;
  }
  out[g] = uint8_t(ras_owner_slot<DIM, NP>(c, lm, n) != slot);
}

// ----------------------------------------------------------------------------
// vectors
// ----------------------------------------------------------------------------
template<typename T>
__global__ void k_fill(T *__restrict__ x, const int64_t n, const T value)
{
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i < n) x[i] = value;
}

// y = a*x + b*y with host scalars
template<typename T>
__global__ void k_axpby(T *__restrict__ y, const T *__restrict__ x, const int64_t n,
                        const double a, const double b)
{
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i < n) y[i] = T(a * double(x[i]) + b * double(y[i]));
}

template<typename TD, typename TS>
__global__ void k_convert(TD *__restrict__ dst, const TS *__restrict__ src, const int64_t n)
{
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i < n) dst[i] = TD(src[i]);
}

inline __device__ __forceinline__ double splitmix_uniform(uint64_t seed, uint64_t i)
{
  uint64_t z = seed * 0x100000000ull + i;
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return 2.0 * (double(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

template<typename T>
__global__ void k_random(T *__restrict__ x, const int64_t n, const uint64_t seed)
{
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i < n) x[i] = T(splitmix_uniform(seed, uint64_t(i)));
}

// block partial sums of a.b (FP64 accumulation), deterministic two-pass reduction
template<typename T>
__global__ void __launch_bounds__(256)
k_dot_partial(const T *__restrict__ a, const T *__restrict__ b, const int64_t n,
              double *__restrict__ partial)
{
  __shared__ double sh[256];
  double s = 0.0;
  for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
      i += int64_t(gridDim.x) * blockDim.x)
    s += double(a[i]) * double(b[i]);
  sh[threadIdx.x] = s;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1)
  {
    if(threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if(threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

static __global__ void __launch_bounds__(256)
k_dot_final(const double *__restrict__ partial, const int np, double *__restrict__ out)
{
  __shared__ double sh[256];
  double s = 0.0;
  for(int i = threadIdx.x; i < np; i += blockDim.x) s += partial[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1)
  {
    if(threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if(threadIdx.x == 0) *out = sh[0];
}

// *out = the sum of np partial sums, np up to tens of thousands (one per block
// of an operator pass): 1024 threads with four independent loads in flight each
static __global__ void __launch_bounds__(1024)
k_dot_final_wide(const double *__restrict__ partial, const int np, double *__restrict__ out)
{
  __shared__ double sh[16];
  constexpr int T = 1024;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = threadIdx.x;
  for(; i + 3 * T < np; i += 4 * T)
  {
    s0 += partial[i];
    s1 += partial[i + T];
    s2 += partial[i + 2 * T];
    s3 += partial[i + 3 * T];
  }
  for(; i < np; i += T) s0 += partial[i];
  double s = (s0 + s1) + (s2 + s3);
#pragma unroll
  for(int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if(threadIdx.x < 64)
  {
    double v = threadIdx.x < 16 ? sh[threadIdx.x] : 0.0;
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if(threadIdx.x == 0) *out = v;
  }
}

// non-finite detector (grid-stride): flag[0] = 1 if any entry is NaN/Inf;
// flag may point to pinned host memory
template<typename T>
__global__ void k_nonfinite(const T *__restrict__ x, const int64_t n, int *flag)
{
  bool bad = false;
  for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
      i += int64_t(gridDim.x) * blockDim.x)
    bad |= !isfinite(x[i]);
  if(bad) *flag = 1;
}

// sum of one value per thread over a block of 256 threads (the tree of
// k_dot_partial); the result is valid in thread 0
__device__ __forceinline__ double block_sum_256(double v, double *sh)
{
  sh[threadIdx.x] = v;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1)
  {
    if(threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  return sh[0];
}

// CG update with device scalars, alpha = rz / pAp:  x += alpha p,
// r -= alpha Ap, block partial sums of the updated r.r (the grid-stride
// order of k_dot_partial, so the sums match dot(r, r) bit for bit) and, if
// r_low is given, the updated r in FP32 (the right-hand side of an FP32 V-cycle)
template<bool LOW>
__global__ void __launch_bounds__(256)
k_cg_update_dot(double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                const double *__restrict__ Ap, const int64_t n, const double *__restrict__ rz,
                const double *__restrict__ pAp, double *__restrict__ partial,
                float *__restrict__ r_low)
{
  __shared__ double sh[256];
  const double alpha = *rz / *pAp;
  double acc = 0.0;
  for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
      i += int64_t(gridDim.x) * blockDim.x)
  {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    if(LOW) r_low[i] = float(ri);
    acc += ri * ri;
  }
  const double tot = block_sum_256(acc, sh);
  if(threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// z = FP64(v) with block partial sums of r.z (grid-stride order of k_dot_partial)
template<typename T>
__global__ void __launch_bounds__(256)
k_convert_dot(double *__restrict__ z, const T *__restrict__ v, const double *__restrict__ r,
              const int64_t n, double *__restrict__ partial)
{
  __shared__ double sh[256];
  double acc = 0.0;
  for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
      i += int64_t(gridDim.x) * blockDim.x)
  {
    const double zi = double(v[i]);
    z[i] = zi;
    acc += r[i] * zi;
  }
  const double tot = block_sum_256(acc, sh);
  if(threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// p = z + (rz_new / rz_old) p
static __global__ void k_cg_direction(double *__restrict__ p, const double *__restrict__ z,
                                      const int64_t n, const double *__restrict__ rz_new,
                                      const double *__restrict__ rz_old)
{
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const double beta = *rz_new / *rz_old;
  if(i < n) p[i] = z[i] + beta * p[i];
}

// ----------------------------------------------------------------------------
// grid transfer for DG Q_p (embedding and its transpose)
// ----------------------------------------------------------------------------
template<int NP>
struct TransferTab
{
  double P[2][NP * NP];  // P[child][i*NP + j] = phi_j((x_i + child)/2)
};

// Grid transfer per coarse cell: the cell and its 2^DIM children are viewed as
// the coarse tensor c[jx][jy][jz] (NP^DIM) and the fine tensor F[X][Y][Z]
// ((2 NP)^DIM, X = child_x * NP + fine node), linked by the 1D prolongation
// Pf[X][j] = P[X / NP][(X % NP) NP + j]:
//   prolongation F = (Pf x Pf x Pf) c,   restriction c = (Pf^T x Pf^T x Pf^T) F,
// one direction at a time through LDS (sum factorisation); one block per
// coarse cell, fine values moved child by child (contiguous cell blocks).
template<int DIM, int NP>
struct TransferShape
{
  static constexpr int K = 2 * NP;
  static constexpr int NDC = cpow(NP, DIM);
  static constexpr int NFT = cpow(K, DIM);
  static constexpr int THREADS = NFT >= 256 ? 256 : ((NFT + 63) / 64) * 64;
};

// fine tensor index (X, Y, Z) of entry `local` of child `child`
template<int DIM, int NP>
__device__ __forceinline__ int fine_tensor_index(int child, int local)
{
  constexpr int K = 2 * NP;
  int o = 0, st = 1;
#pragma unroll
  for(int d = 0; d < DIM; ++d)
  {
    o += (((child >> d) & 1) * NP + local % NP) * st;
    local /= NP;
    st *= K;
  }
  return o;
}

template<int DIM>
__device__ __forceinline__ int64_t child_cell(const int *cc, int nc, int child)
{
  int64_t f = 0, st = 1;
  const int nf = 2 * nc;
#pragma unroll
  for(int d = 0; d < DIM; ++d)
  {
    f += int64_t(2 * cc[d] + ((child >> d) & 1)) * st;
    st *= nf;
  }
  return f;
}

// coordinates of coarse cell ccell (32-bit: cells < 1000^3 < 2^31; 64-bit
// divisions by a run-time nc are long instruction sequences)
template<int DIM>
__device__ __forceinline__ void cell_coords(int64_t ccell, int nc, int *cc)
{
  int r = int(ccell);
#pragma unroll
  for(int d = 0; d < DIM; ++d)
  {
    cc[d] = r % nc;
    r /= nc;
  }
}

// xf += P xc
template<int DIM, int NP, typename T>
__global__ void __launch_bounds__((TransferShape<DIM, NP>::THREADS))
k_prolongate_add(T *__restrict__ xf, const T *__restrict__ xc, const TransferTab<NP> tab,
                 const int nc)
{
  using TS = TransferShape<DIM, NP>;
  constexpr int K = TS::K, NDC = TS::NDC, NFT = TS::NFT, NT = TS::THREADS;
  __shared__ T Pf[K * NP];
  __shared__ T S0[NFT], S1[NFT];
  const int64_t cc = blockIdx.x;
  int ccd[3] = {0, 0, 0};
  cell_coords<DIM>(cc, nc, ccd);
  const int t = threadIdx.x;
  for(int i = t; i < K * NP; i += NT) Pf[i] = T(tab.P[(i / NP) / NP][((i / NP) % NP) * NP + i % NP]);
  for(int i = t; i < NDC; i += NT) S0[i] = xc[cc * NDC + i];
  __syncthreads();
  // x: S1[X + K (jy + NP jz)] = sum_jx Pf[X][jx] c[jx][jy][jz]
  for(int o = t; o < K * NDC / NP; o += NT)
  {
    const int X = o % K, r = o / K;
    T s = 0;
#pragma unroll
    for(int j = 0; j < NP; ++j) s += Pf[X * NP + j] * S0[j + NP * r];
    S1[o] = s;
  }
  __syncthreads();
  const T *last = S1;
  if constexpr(DIM == 3)
  {
    // y: S0[X + K (Y + K jz)] = sum_jy Pf[Y][jy] S1[X + K (jy + NP jz)]
    for(int o = t; o < K * K * NP; o += NT)
    {
      const int X = o % K, Y = (o / K) % K, jz = o / (K * K);
      T s = 0;
#pragma unroll
      for(int j = 0; j < NP; ++j) s += Pf[Y * NP + j] * S1[X + K * (j + NP * jz)];
      S0[o] = s;
    }
    __syncthreads();
    last = S0;
  }
  // last direction, straight into the fine vector (child by child)
  for(int o = t; o < NFT; o += NT)
  {
    const int child = o / NDC, local = o % NDC;
    const int ft = fine_tensor_index<DIM, NP>(child, local);
    const int Z = ft / cpow(K, DIM - 1), rest = ft % cpow(K, DIM - 1);
    T s = 0;
#pragma unroll
    for(int j = 0; j < NP; ++j) s += Pf[Z * NP + j] * last[rest + cpow(K, DIM - 1) * j];
    T *dst = xf + child_cell<DIM>(ccd, nc, child) * NDC + local;
    *dst = *dst + s;
  }
}

// xc = P^T xf
template<int DIM, int NP, typename T>
__global__ void __launch_bounds__((TransferShape<DIM, NP>::THREADS))
k_restrict(T *__restrict__ xc, const T *__restrict__ xf, const TransferTab<NP> tab, const int nc)
{
  using TS = TransferShape<DIM, NP>;
  constexpr int K = TS::K, NDC = TS::NDC, NFT = TS::NFT, NT = TS::THREADS;
  constexpr int KL = cpow(K, DIM - 1);
  __shared__ T Pf[K * NP];
  __shared__ T S0[NFT], S1[NFT];
  const int64_t cc = blockIdx.x;
  int ccd[3] = {0, 0, 0};
  cell_coords<DIM>(cc, nc, ccd);
  const int t = threadIdx.x;
  for(int i = t; i < K * NP; i += NT) Pf[i] = T(tab.P[(i / NP) / NP][((i / NP) % NP) * NP + i % NP]);
  for(int o = t; o < NFT; o += NT)
  {
    const int child = o / NDC, local = o % NDC;
    S0[fine_tensor_index<DIM, NP>(child, local)] = xf[child_cell<DIM>(ccd, nc, child) * NDC + local];
  }
  __syncthreads();
  // last direction: S1[rest + KL jl] = sum_L Pf[L][jl] S0[rest + KL L]
  for(int o = t; o < KL * NP; o += NT)
  {
    const int rest = o % KL, jl = o / KL;
    T s = 0;
#pragma unroll
    for(int L = 0; L < K; ++L) s += Pf[L * NP + jl] * S0[rest + KL * L];
    S1[o] = s;
  }
  __syncthreads();
  const T *cur = S1;
  if constexpr(DIM == 3)
  {
    // y: S0[X + K (jy + NP jz)] = sum_Y Pf[Y][jy] S1[X + K (Y + K jz)]
    for(int o = t; o < K * NP * NP; o += NT)
    {
      const int X = o % K, jy = (o / K) % NP, jz = o / (K * NP);
      T s = 0;
#pragma unroll
      for(int Y = 0; Y < K; ++Y) s += Pf[Y * NP + jy] * S1[X + K * (Y + K * jz)];
      S0[o] = s;
    }
    __syncthreads();
    cur = S0;
  }
  // x: c[jx + NP r] = sum_X Pf[X][jx] cur[X + K r]
  for(int o = t; o < NDC; o += NT)
  {
    const int jx = o % NP, r = o / NP;
    T s = 0;
#pragma unroll
    for(int X = 0; X < K; ++X) s += Pf[X * NP + jx] * cur[X + K * r];
    xc[cc * NDC + o] = s;
  }
}

// y = A x with a dense row-major FP64 matrix (coarse-grid inverse); one
// 64-lane wave per row
static __global__ void __launch_bounds__(256)
k_dense_matvec(const double *__restrict__ A, const double *__restrict__ x, double *__restrict__ y,
               const int m)
{
  const int row = int(blockIdx.x) * 4 + int(threadIdx.x) / 64;
  const int lane = threadIdx.x % 64;
  if(row >= m) return;
  double s = 0.0;
  for(int c = lane; c < m; c += 64) s += A[size_t(row) * m + c] * x[c];
  for(int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if(lane == 0) y[row] = s;
}

}  // namespace tpss
