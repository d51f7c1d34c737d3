// fit_kernels.hpp -- separable Kronecker fit of exact local matrices (FP64):
// the local solver of a distorted patch as fast diagonalization of
//   A_j ~ sum_d M_{dim-1} x .. x L_d x .. x M_0   (dimension 0 rightmost),
// fitted to A_j = R_j A R_j^T by alternating least squares.  For direction d
// the pair (L_d, M_d) solves the rank-2 Kronecker least-squares problem
//   min || A_j - L_d (x)_d R1 - M_d (x)_d R2 ||_F,
//   R1 = (x)_{e != d} M_e,  R2 = sum_{e != d} (x)_{f != d} (L_f if f == e else M_f)
// (one 2x2 normal-equation system per entry (i_d, j_d)): the CP-ALS of the
// reference's Kronecker SVD module (kroneckersvd.h:501-585) with the factors
// tied to the separable form that TensorProductMatrixSymmetricSum inverts.
//
// The local matrices come column by column from the operator kernel applied
// to unit vectors, one column of every patch of a batch at once: the patches
// of a batch share one red-black color, so no two of them share a cell or a
// face and their columns do not mix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "setup_kernels.hpp"

namespace tpss
{
// x[global DoF of patch-local DoF col of every batch patch] = val
template<int DIM, int K>
__global__ void __launch_bounds__(256)
k_fit_unit(double *__restrict__ x, const uint32_t *__restrict__ vtx,
           const int *__restrict__ plist, const int nb, const int n, const int col,
           const double val)
{
  constexpr int NP = K / 2, NDC = cpow(NP, DIM);
  const int jb = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(jb >= nb) return;
  int v[3];
  unpack_vertex(vtx[plist[jb]], v);
  const int i0 = col % K, i1 = DIM > 1 ? (col / K) % K : 0, i2 = DIM > 2 ? col / (K * K) : 0;
  x[patch_cell<DIM, NP>(v, n, i0, i1, i2) * NDC + patch_local<DIM, NP>(i0, i1, i2)] = val;
}

// A[jb][col][row] = y[global DoF of patch-local row] for every batch patch: the
// local matrix is symmetric, so the column is stored as a row (the writes of one
// launch are contiguous per patch instead of a stride of n_pd apart)
template<int DIM, int K>
__global__ void __launch_bounds__(256)
k_fit_gather(const double *__restrict__ y, const uint32_t *__restrict__ vtx,
             const int *__restrict__ plist, const int nb, const int n, const int col,
             double *__restrict__ A)
{
  constexpr int NP = K / 2, NDC = cpow(NP, DIM), NPD = cpow(K, DIM);
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(g >= int64_t(nb) * NPD) return;
  const int jb = int(g / NPD), r = int(g % NPD);
  int v[3];
  unpack_vertex(vtx[plist[jb]], v);
  const int i0 = r % K, i1 = DIM > 1 ? (r / K) % K : 0, i2 = DIM > 2 ? r / (K * K) : 0;
  A[(size_t(jb) * NPD + col) * NPD + r] =
    y[patch_cell<DIM, NP>(v, n, i0, i1, i2) * NDC + patch_local<DIM, NP>(i0, i1, i2)];
}

// Mass M and SIPG Laplace L of every 1D problem (the surrogate matrices that
// initialise the fit), same assembly as the local-solver setup.
template<int NP>
__global__ void __launch_bounds__(64)
k_assemble_1d(const double *__restrict__ h, const uint8_t *__restrict__ bdry,
              const double *__restrict__ coef, const Shape1D<NP> sh, const int p, const double ip,
              const int nprob, double *__restrict__ M_out, double *__restrict__ L_out)
{
  constexpr int K = 2 * NP;
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= nprob) return;
  double M[K * K], L[K * K];
  assemble_patch_1d<NP>(sh, p, h[2 * i], h[2 * i + 1], bdry[2 * i] != 0, bdry[2 * i + 1] != 0,
                        ip, coef[2 * i], coef[2 * i + 1], M, L);
  for(int e = 0; e < K * K; ++e)
  {
    M_out[size_t(i) * K * K + e] = M[e];
    L_out[size_t(i) * K * K + e] = L[e];
  }
}

// One block per batch patch.  Each thread owns rows of A_j and accumulates,
// for the direction d being updated, the partial contractions
//   T1[i_d][j_d] += A[r][c] * prod_{e != d} M_e[i_e][j_e]
//   T2[i_d][j_d] += A[r][c] * sum_{e != d} L_e[i_e][j_e] prod_{f != d, e} M_f[i_f][j_f]
// in registers (j_d is a loop index), then the block sums them in LDS.
template<int DIM, int K>
__global__ void __launch_bounds__(256)
k_separable_fit(const double *__restrict__ Abuf, const int *__restrict__ plist, const int nb,
                double *__restrict__ Mg, double *__restrict__ Lg, const int sweeps,
                double *__restrict__ err)
{
  constexpr int NPD = cpow(K, DIM), KK = K * K;
  __shared__ double Ms[DIM][KK], Ls[DIM][KK];
  __shared__ double T1[KK], T2[KK];
  __shared__ double gram[3];
  __shared__ double red[2][256];
  const int jb = blockIdx.x;
  if(jb >= nb) return;
  const int pj = plist[jb];
  const double *A = Abuf + size_t(jb) * NPD * NPD;
  const int t = threadIdx.x, nt = blockDim.x;
  for(int e = t; e < DIM * KK; e += nt)
  {
    Ms[e / KK][e % KK] = Mg[size_t(pj) * DIM * KK + e];
    Ls[e / KK][e % KK] = Lg[size_t(pj) * DIM * KK + e];
  }
  __syncthreads();

  for(int s = 0; s < sweeps; ++s)
#pragma unroll
    for(int d = 0; d < DIM; ++d)
    {
      if(t == 0)
      {
        double nM[3], nL[3], ML[3];
        for(int e = 0; e < DIM; ++e)
        {
          nM[e] = nL[e] = ML[e] = 0.0;
          for(int q = 0; q < KK; ++q)
          {
            nM[e] += Ms[e][q] * Ms[e][q];
            nL[e] += Ls[e][q] * Ls[e][q];
            ML[e] += Ms[e][q] * Ls[e][q];
          }
        }
        double g11 = 1.0, g12 = 0.0, g22 = 0.0;
        for(int e = 0; e < DIM; ++e)
          if(e != d) g11 *= nM[e];
        for(int e = 0; e < DIM; ++e)
        {
          if(e == d) continue;
          double t12 = ML[e], t22 = nL[e];
          for(int f = 0; f < DIM; ++f)
            if(f != d && f != e)
            {
              t12 *= nM[f];
              t22 *= nM[f];
            }
          g12 += t12;
          g22 += t22;
          for(int e2 = 0; e2 < DIM; ++e2)
          {
            if(e2 == d || e2 == e) continue;
            double tt = ML[e] * ML[e2];
            for(int f = 0; f < DIM; ++f)
              if(f != d && f != e && f != e2) tt *= nM[f];
            g22 += tt;
          }
        }
        gram[0] = g11;
        gram[1] = g12;
        gram[2] = g22;
      }
      for(int q = t; q < KK; q += nt)
      {
        T1[q] = 0.0;
        T2[q] = 0.0;
      }
      __syncthreads();
      // rows of A_j: thread t takes rows t, t + nt, ...
      for(int r = t; r < NPD; r += nt)
      {
        int ri[3] = {r % K, DIM > 1 ? (r / K) % K : 0, DIM > 2 ? r / (K * K) : 0};
        double a1[K], a2[K];
#pragma unroll
        for(int q = 0; q < K; ++q) a1[q] = a2[q] = 0.0;
        const double *row = A + size_t(r) * NPD;
        for(int c = 0; c < NPD; c += K)
        {
          // the K columns c .. c + K - 1 differ in j_0 only
          const int cj1 = DIM > 1 ? (c / K) % K : 0, cj2 = DIM > 2 ? c / (K * K) : 0;
          double s1 = 0.0, s2 = 0.0;
#pragma unroll
          for(int q = 0; q < K; ++q)
          {
            const int cj[3] = {q, cj1, cj2};
            double w1 = 1.0, w2 = 0.0;
#pragma unroll
            for(int e = 0; e < DIM; ++e)
            {
              if(e == d) continue;
              const double m = Ms[e][ri[e] * K + cj[e]], l = Ls[e][ri[e] * K + cj[e]];
              w2 = w2 * m + w1 * l;
              w1 *= m;
            }
            const double a = row[c + q];
            if(d == 0)
            {
              a1[q] += a * w1;
              a2[q] += a * w2;
            }
            else
            {
              s1 += a * w1;
              s2 += a * w2;
            }
          }
          if(d != 0)
          {
            // register slot j_d of the block (unrolled select)
            const int jd = d == 1 ? cj1 : cj2;
#pragma unroll
            for(int z = 0; z < K; ++z)
              if(z == jd)
              {
                a1[z] += s1;
                a2[z] += s2;
              }
          }
        }
        const int id = ri[d];
#pragma unroll
        for(int z = 0; z < K; ++z)
        {
          atomicAdd(&T1[id * K + z], a1[z]);
          atomicAdd(&T2[id * K + z], a2[z]);
        }
      }
      __syncthreads();
      {
        const double g11 = gram[0], g12 = gram[1], g22 = gram[2];
        const double det = g11 * g22 - g12 * g12;
        for(int q = t; q < KK; q += nt)
        {
          Ls[d][q] = (g22 * T1[q] - g12 * T2[q]) / det;
          Ms[d][q] = (g11 * T2[q] - g12 * T1[q]) / det;
        }
      }
      __syncthreads();
    }

  // relative Frobenius error of the fit
  double num = 0.0, den = 0.0;
  for(int r = t; r < NPD; r += nt)
  {
    const int ri[3] = {r % K, DIM > 1 ? (r / K) % K : 0, DIM > 2 ? r / (K * K) : 0};
    const double *row = A + size_t(r) * NPD;
    for(int c = 0; c < NPD; ++c)
    {
      const int cj[3] = {c % K, DIM > 1 ? (c / K) % K : 0, DIM > 2 ? c / (K * K) : 0};
      double fit = 0.0;
#pragma unroll
      for(int dd = 0; dd < DIM; ++dd)
      {
        double tt = 1.0;
#pragma unroll
        for(int e = 0; e < DIM; ++e)
          tt *= (e == dd ? Ls[e][ri[e] * K + cj[e]] : Ms[e][ri[e] * K + cj[e]]);
        fit += tt;
      }
      const double a = row[c];
      num += (a - fit) * (a - fit);
      den += a * a;
    }
  }
  red[0][t] = num;
  red[1][t] = den;
  __syncthreads();
  for(int w = nt / 2; w > 0; w /= 2)
  {
    if(t < w)
    {
      red[0][t] += red[0][t + w];
      red[1][t] += red[1][t + w];
    }
    __syncthreads();
  }
  if(t == 0) err[pj] = sqrt(red[0][0] / red[1][0]);
  for(int e = t; e < DIM * KK; e += nt)
  {
    Mg[size_t(pj) * DIM * KK + e] = Ms[e / KK][e % KK];
    Lg[size_t(pj) * DIM * KK + e] = Ls[e / KK][e % KK];
  }
}

// FD factors of the fitted matrices, one thread per (patch, direction);
// ok[patch] stays 1 only if every direction's fitted mass is s.p.d.
template<int K>
__global__ void __launch_bounds__(64)
k_fit_eig(const double *__restrict__ Mg, const double *__restrict__ Lg, const int nprob,
          const int dim, double *__restrict__ S_out, double *__restrict__ lam_out,
          int *__restrict__ ok)
{
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= nprob) return;
  double M[K * K], L[K * K], S[K * K], lam[K];
  for(int e = 0; e < K * K; ++e)
  {
    M[e] = Mg[size_t(i) * K * K + e];
    L[e] = Lg[size_t(i) * K * K + e];
  }
  if(!gen_eig<K>(L, M, S, lam))
  {
    ok[i / dim] = 0;
    return;
  }
  for(int e = 0; e < K * K; ++e) S_out[size_t(i) * K * K + e] = S[e];
  for(int e = 0; e < K; ++e) lam_out[size_t(i) * K + e] = lam[e];
}

// A fit is kept only if it is s.p.d.: every direction's mass is s.p.d. (checked
// by k_fit_eig) and the smallest eigenvalue of the Kronecker sum,
// sum_d lambda_d,min, is positive relative to the largest.  One thread per patch.
template<int K>
__global__ void __launch_bounds__(64)
k_fit_validate(const double *__restrict__ lam_fit, const int P, const int dim,
               int *__restrict__ ok)
{
  const int pj = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(pj >= P || !ok[pj]) return;
  double lo = 0.0, hi = 0.0;
  for(int d = 0; d < dim; ++d)
  {
    lo += lam_fit[(size_t(pj) * dim + d) * K];
    hi += fabs(lam_fit[(size_t(pj) * dim + d) * K + K - 1]);
  }
  if(!(lo > 1e-12 * hi)) ok[pj] = 0;
}

// factors of patches whose fit succeeded replace the surrogate ones
template<int K>
__global__ void __launch_bounds__(64)
k_fit_select(const double *__restrict__ S_fit, const double *__restrict__ lam_fit,
             const int *__restrict__ ok, const int nprob, const int dim, double *__restrict__ S,
             double *__restrict__ lam, double *__restrict__ err)
{
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= nprob) return;
  const int pj = i / dim;
  if(!ok[pj])
  {
    if(i % dim == 0) err[pj] = -1.0;
    return;
  }
  for(int e = 0; e < K * K; ++e) S[size_t(i) * K * K + e] = S_fit[size_t(i) * K * K + e];
  for(int e = 0; e < K; ++e) lam[size_t(i) * K + e] = lam_fit[size_t(i) * K + e];
}

}  // namespace tpss
