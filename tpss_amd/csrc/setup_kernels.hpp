// setup_kernels.hpp -- FP64 device setup of the local solvers:
// surrogate patch lengths (A9), 1D patch mass / SIPG-Laplace matrices (A7)
// and their generalized eigendecompositions (A8).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace tpss
{
// unit-interval shape data at Gauss points and faces (from the host basis)
template<int NP>
struct Shape1D
{
  double val[NP * NP], grad[NP * NP];  // [i*NP + q]
  double fval[NP * 2], fgrad[NP * 2];  // [i*2 + face]
  double qw[NP];
};

// Mass M and SIPG Laplace L of one patch direction (two cells of lengths h0, h1):
// FD::L2 / FD::Laplace cell operations (apps/common_integrator.h:161-192, 302-338),
// SIPG face operation with average factor 1 at the physical boundary and 1/2
// otherwise (common_integrator.h:196-240, fd_evaluation.h:651-688), interface
// cell 0 | cell 1 (common_integrator.h:242-289), DGQ 1D index c*(p+1)+i.
// L = a_cons * (stiffness + consistency terms) + a_pen * (penalty terms): (1, 1)
// for Poisson; for component c of linear elasticity in direction d
// a_cons = 2 mu (1 if c == d else 1/2) + lambda [c == d], a_pen = 2 mu + lambda [c == d]
// (CellStrain, NitscheStrain, CellGradDiv, NitscheGradDiv,
// apps/linelasticity_integrator.h:753-1145, basic penalty variant).
template<int NP>
__device__ void assemble_patch_1d(const Shape1D<NP> &sh, const int p, const double h0,
                                  const double h1, const bool bl, const bool br, const double ip,
                                  const double a_cons, const double a_pen, double *M, double *L)
{
  constexpr int K = 2 * NP;
  const double gamma = p == 0 ? 1.0 : double(p) * (p + 1);
  for(int e = 0; e < K * K; ++e)
  {
    M[e] = 0.0;
    L[e] = 0.0;
  }
  for(int c = 0; c < 2; ++c)
  {
    const double h = c == 0 ? h0 : h1;
    for(int i = 0; i < NP; ++i)
      for(int jj = 0; jj < NP; ++jj)
      {
        double m = 0, l = 0;
        for(int q = 0; q < NP; ++q)
        {
          const double dx = h * sh.qw[q];
          m += sh.val[jj * NP + q] * sh.val[i * NP + q] * dx;
          l += (sh.grad[i * NP + q] / h) * (sh.grad[jj * NP + q] / h) * dx;
        }
        M[(c * NP + i) * K + c * NP + jj] += m;
        L[(c * NP + i) * K + c * NP + jj] += a_cons * l;
      }
    for(int f = 0; f < 2; ++f)
    {
      const bool outer = (c == 0 && f == 0) || (c == 1 && f == 1);
      const bool at_bdry = outer && (c == 0 ? bl : br);
      const double a = at_bdry ? 1.0 : 0.5;
      const double nrm = f == 0 ? -1.0 : 1.0;
      const double pen = ip * a * 2.0 * gamma / h;  // ip * a * 2 gamma (1/2h + 1/2h)
      for(int i = 0; i < NP; ++i)
      {
        const double vi = sh.fval[i * 2 + f], gi = sh.fgrad[i * 2 + f] / h;
        for(int jj = 0; jj < NP; ++jj)
        {
          const double uj = sh.fval[jj * 2 + f], guj = sh.fgrad[jj * 2 + f] / h;
          L[(c * NP + i) * K + c * NP + jj] +=
            -a_cons * a * (vi * nrm * guj + gi * uj * nrm) + a_pen * pen * vi * uj * nrm * nrm;
        }
      }
    }
  }
  const double pen = ip * 0.5 * 2.0 * gamma * (0.5 / h0 + 0.5 / h1);
  for(int i = 0; i < NP; ++i)
  {
    const double v0 = sh.fval[i * 2 + 1], gv0 = sh.fgrad[i * 2 + 1] / h0;
    const double v1 = sh.fval[i * 2 + 0], gv1 = sh.fgrad[i * 2 + 0] / h1;
    for(int jj = 0; jj < NP; ++jj)
    {
      const double u0 = sh.fval[jj * 2 + 1], gu0 = sh.fgrad[jj * 2 + 1] / h0;
      const double u1 = sh.fval[jj * 2 + 0], gu1 = sh.fgrad[jj * 2 + 0] / h1;
      // normals: +1 on cell 0, -1 on cell 1
      L[i * K + NP + jj] += -a_cons * 0.5 * (v0 * gu1 - gv0 * u1) - a_pen * pen * v0 * u1;
      L[(NP + i) * K + jj] += -a_cons * 0.5 * (-v1 * gu0 + gv1 * u0) - a_pen * pen * v1 * u0;
    }
  }
}

// L s = lambda M s: Cholesky M = C C^T, cyclic Jacobi on C^{-1} L C^{-T},
// S = C^{-T} Q, ascending eigenvalues.  Returns false if M is not s.p.d. or an
// eigenvalue is not finite.  A 1D eigenvalue may be negative (an SIPG matrix of
// a distorted patch direction can be indefinite); what must not vanish is the
// sum over directions, checked per patch by k_check_eigenvalue_sums.
template<int K>
__device__ bool gen_eig(const double *L, const double *M, double *S, double *lam)
{
  double C[K * K], G[K * K], Q[K * K];
  for(int e = 0; e < K * K; ++e) C[e] = 0.0;
  for(int jj = 0; jj < K; ++jj)
  {
    double s = M[jj * K + jj];
    for(int m = 0; m < jj; ++m) s -= C[jj * K + m] * C[jj * K + m];
    if(!(s > 0.0)) return false;
    C[jj * K + jj] = sqrt(s);
    for(int i = jj + 1; i < K; ++i)
    {
      double t = M[i * K + jj];
      for(int m = 0; m < jj; ++m) t -= C[i * K + m] * C[jj * K + m];
      C[i * K + jj] = t / C[jj * K + jj];
    }
  }
  // Q <- C^{-1} L  (columns), then G <- Q C^{-T}
  for(int col = 0; col < K; ++col)
    for(int i = 0; i < K; ++i)
    {
      double t = L[i * K + col];
      for(int m = 0; m < i; ++m) t -= C[i * K + m] * Q[m * K + col];
      Q[i * K + col] = t / C[i * K + i];
    }
  for(int row = 0; row < K; ++row)
    for(int i = 0; i < K; ++i)
    {
      double t = Q[row * K + i];
      for(int m = 0; m < i; ++m) t -= C[i * K + m] * G[row * K + m];
      G[row * K + i] = t / C[i * K + i];
    }
  for(int i = 0; i < K; ++i)
    for(int jj = i + 1; jj < K; ++jj)
    {
      const double s = 0.5 * (G[i * K + jj] + G[jj * K + i]);
      G[i * K + jj] = s;
      G[jj * K + i] = s;
    }
  for(int i = 0; i < K; ++i)
    for(int jj = 0; jj < K; ++jj) Q[i * K + jj] = i == jj ? 1.0 : 0.0;
  for(int sweep = 0; sweep < 100; ++sweep)
  {
    double off = 0.0, diag = 0.0;
    for(int i = 0; i < K; ++i)
    {
      diag += G[i * K + i] * G[i * K + i];
      for(int jj = i + 1; jj < K; ++jj) off += G[i * K + jj] * G[i * K + jj];
    }
    if(off <= 1e-32 * diag) break;
    for(int pp = 0; pp < K - 1; ++pp)
      for(int qq = pp + 1; qq < K; ++qq)
      {
        const double apq = G[pp * K + qq];
        if(apq == 0.0) continue;
        const double theta = (G[qq * K + qq] - G[pp * K + pp]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for(int r = 0; r < K; ++r)
        {
          const double arp = G[r * K + pp], arq = G[r * K + qq];
          G[r * K + pp] = cs * arp - sn * arq;
          G[r * K + qq] = sn * arp + cs * arq;
        }
        for(int r = 0; r < K; ++r)
        {
          const double apr = G[pp * K + r], aqr = G[qq * K + r];
          G[pp * K + r] = cs * apr - sn * aqr;
          G[qq * K + r] = sn * apr + cs * aqr;
        }
        for(int r = 0; r < K; ++r)
        {
          const double vrp = Q[r * K + pp], vrq = Q[r * K + qq];
          Q[r * K + pp] = cs * vrp - sn * vrq;
          Q[r * K + qq] = sn * vrp + cs * vrq;
        }
      }
  }
  int order[K];
  for(int i = 0; i < K; ++i) order[i] = i;
  for(int i = 1; i < K; ++i)
  {
    const int t = order[i];
    int jj = i - 1;
    while(jj >= 0 && G[order[jj] * K + order[jj]] > G[t * K + t])
    {
      order[jj + 1] = order[jj];
      --jj;
    }
    order[jj + 1] = t;
  }
  for(int c = 0; c < K; ++c)
  {
    const int oc = order[c];
    lam[c] = G[oc * K + oc];
    for(int i = K - 1; i >= 0; --i)
    {
      double t = Q[i * K + oc];
      for(int m = i + 1; m < K; ++m) t -= C[m * K + i] * S[m * K + c];
      S[i * K + c] = t / C[i * K + i];
    }
    if(!isfinite(lam[c])) return false;
  }
  return true;
}

// one thread per 1D problem (h0, h1, boundary bits)
template<int NP>
__global__ void __launch_bounds__(64)
k_setup_fd(const double *__restrict__ h, const uint8_t *__restrict__ bdry,
           const double *__restrict__ coef, const Shape1D<NP> sh, const int p, const double ip,
           const int nprob, double *__restrict__ S_out, double *__restrict__ lam_out,
           int *__restrict__ status)
{
  constexpr int K = 2 * NP;
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= nprob) return;
  double M[K * K], L[K * K], S[K * K], lam[K];
  assemble_patch_1d<NP>(sh, p, h[2 * i], h[2 * i + 1], bdry[2 * i] != 0, bdry[2 * i + 1] != 0,
                        ip, coef[2 * i], coef[2 * i + 1], M, L);
  const bool ok = gen_eig<K>(L, M, S, lam);
  if(!ok) status[0] = 1;
  for(int e = 0; e < K * K; ++e) S_out[size_t(i) * K * K + e] = S[e];
  for(int e = 0; e < K; ++e) lam_out[size_t(i) * K + e] = lam[e];
}

// No eigenvalue of a local solver (the sums lambda_{i_0} + .. + lambda_{i_{d-1}}
// over the directions) may be nearly zero: the check of
// TensorProductMatrix::get_eigenvalues (tensor_product_matrix.h:357-371), taken
// relative to the largest sum.  One thread per (component block, patch) row.
template<int DIM, int K>
__global__ void __launch_bounds__(64)
k_check_eigenvalue_sums(const uint32_t *__restrict__ fac, const double *__restrict__ lam,
                        const int nrows, int *__restrict__ status)
{
  const int r = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(r >= nrows) return;
  const double *l[3];
  for(int d = 0; d < DIM; ++d) l[d] = lam + size_t(fac[size_t(r) * DIM + d]) * K;
  double big = 0.0;
  for(int d = 0; d < DIM; ++d) big += fmax(fabs(l[d][0]), fabs(l[d][K - 1]));
  double small = 1e300;
  int idx[3] = {0, 0, 0};
  for(int it = 0; it < cpow(K, DIM); ++it)
  {
    double s = 0.0;
    for(int d = 0; d < DIM; ++d) s += l[d][idx[d]];
    small = fmin(small, fabs(s));
    for(int d = 0; d < DIM; ++d)
    {
      if(++idx[d] < K) break;
      idx[d] = 0;
    }
  }
  if(!(small > 1e-12 * big)) status[0] = 1;
}

// Eigenpairs of mirror-symmetric 1D problems (two cells of equal length, no
// boundary face: M and L commute with the reversal a -> K-1-a) are K/2
// symmetric and K/2 antisymmetric eigenvectors.  For those the pairs are
// reordered symmetric first (each group in ascending eigenvalue order), the
// symmetry is made exact, and eo[f] = 1, so that patch solves can take the
// even-odd products.  Any other problem (boundary face, unequal lengths, a
// degenerate pair mixing the two kinds) keeps its order and eo[f] = 0.  The
// inverse (S diag(1/lam) S^T in tensor form) does not depend on the order.
template<int K>
__global__ void __launch_bounds__(64)
k_even_odd_order(double *__restrict__ S_tab, double *__restrict__ lam_tab, const int nf,
                 uint32_t *__restrict__ eo)
{
  constexpr int P = K / 2;
  const int f = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(f >= nf) return;
  double *S = S_tab + size_t(f) * K * K;
  double *lam = lam_tab + size_t(f) * K;
  int kind[K];
  int nsym = 0, nanti = 0;
  for(int c = 0; c < K; ++c)
  {
    double nrm = 0.0, ds = 0.0, da = 0.0;
    for(int a = 0; a < K; ++a) nrm = fmax(nrm, fabs(S[a * K + c]));
    for(int a = 0; a < P; ++a)
    {
      ds = fmax(ds, fabs(S[a * K + c] - S[(K - 1 - a) * K + c]));
      da = fmax(da, fabs(S[a * K + c] + S[(K - 1 - a) * K + c]));
    }
    kind[c] = ds <= 1e-11 * nrm ? 1 : (da <= 1e-11 * nrm ? -1 : 0);
    nsym += kind[c] == 1 ? 1 : 0;
    nanti += kind[c] == -1 ? 1 : 0;
  }
  if(nsym != P || nanti != P)
  {
    eo[f] = 0u;
    return;
  }
  double Sn[K * K], ln[K];
  int m = 0;
  for(int pass = 0; pass < 2; ++pass)
    for(int c = 0; c < K; ++c)
    {
      const int sg = kind[c];
      if(sg != (pass == 0 ? 1 : -1)) continue;
      ln[m] = lam[c];
      for(int a = 0; a < P; ++a)
      {
        const double v = 0.5 * (S[a * K + c] + sg * S[(K - 1 - a) * K + c]);
        Sn[a * K + m] = v;
        Sn[(K - 1 - a) * K + m] = sg * v;
      }
      ++m;
    }
  for(int e = 0; e < K * K; ++e) S[e] = Sn[e];
  for(int c = 0; c < K; ++c) lam[c] = ln[c];
  eo[f] = 1u;
}

// Surrogate lengths (A9, mapping_info.templates.h:42-240): per patch, direction
// and 1D cell slot the average over the 2^(dim-1) cells of that slot of the
// Gauss-Lobatto weighted distance between opposite faces along reference lines
// (straight for trilinear cells, so the arc length is the end-point distance).
template<int DIM, int NQ>
__global__ void __launch_bounds__(128)
k_surrogate_h(const double *__restrict__ vert, const uint32_t *__restrict__ vtx, const int P,
              const int n, const double *__restrict__ gx, const double *__restrict__ gw,
              double *__restrict__ h_out, uint8_t *__restrict__ bdry_out)
{
  const int jp = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if(jp >= P) return;
  int v[3];
  unpack_vertex(vtx[jp], v);
  auto vertex = [&](const int *iv) -> const double * {
    int64_t idx = 0, st = 1;
    for(int e = 0; e < DIM; ++e)
    {
      idx += iv[e] * st;
      st *= (n + 1);
    }
    return vert + idx * DIM;
  };
  for(int d = 0; d < DIM; ++d)
  {
    for(int slot = 0; slot < 2; ++slot)
    {
      double acc = 0.0;
      int nc = 0;
      for(int cb = 0; cb < (1 << DIM); ++cb)
      {
        if(((cb >> d) & 1) != slot) continue;
        int cell[3];
        for(int e = 0; e < DIM; ++e) cell[e] = v[e] - 1 + ((cb >> e) & 1);
        constexpr int NT = DIM == 2 ? NQ : NQ * NQ;
        double len = 0.0;
        for(int tq = 0; tq < NT; ++tq)
        {
          double xr[3] = {0, 0, 0}, w = 1.0;
          int tt = tq;
          for(int e = 0; e < DIM; ++e)
          {
            if(e == d) continue;
            xr[e] = gx[tt % NQ];
            w *= gw[tt % NQ];
            tt /= NQ;
          }
          double pa[3] = {0, 0, 0}, pb[3] = {0, 0, 0};
          for(int corner = 0; corner < (1 << DIM); ++corner)
          {
            double wa = 1.0, wb = 1.0;
            int iv[3];
            for(int e = 0; e < DIM; ++e)
            {
              const int bb = (corner >> e) & 1;
              iv[e] = cell[e] + bb;
              if(e == d)
              {
                wa *= bb == 0 ? 1.0 : 0.0;
                wb *= bb == 1 ? 1.0 : 0.0;
              }
              else
              {
                const double s = bb ? xr[e] : 1.0 - xr[e];
                wa *= s;
                wb *= s;
              }
            }
            const double *X = vertex(iv);
            for(int e = 0; e < DIM; ++e)
            {
              pa[e] += wa * X[e];
              pb[e] += wb * X[e];
            }
          }
          double dist = 0.0;
          for(int e = 0; e < DIM; ++e) dist += (pb[e] - pa[e]) * (pb[e] - pa[e]);
          len += w * sqrt(dist);
        }
        acc += len;
        ++nc;
      }
      h_out[(size_t(jp) * DIM + d) * 2 + slot] = acc / nc;
    }
    bdry_out[(size_t(jp) * DIM + d) * 2 + 0] = v[d] == 1;
    bdry_out[(size_t(jp) * DIM + d) * 2 + 1] = v[d] == n - 1;
  }
}

}  // namespace tpss
