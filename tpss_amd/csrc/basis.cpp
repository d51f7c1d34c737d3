// basis.cpp -- host-side FP64 1D finite-element data: Gauss / Gauss-Lobatto
// rules, Lagrange basis on Gauss-Lobatto nodes, 1D SIPG cell and face
// matrices.  These feed the device kernels as small tables.
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.hpp"

namespace tpss
{
namespace
{
thread_local std::string g_last_error;

// P_n and P_n' on [-1,1]
void legendre_pd(int n, double x, double &pn, double &dpn)
{
  double p0 = 1.0, p1 = x;
  if(n == 0)
  {
    pn = 1.0;
    dpn = 0.0;
    return;
  }
  for(int k = 2; k <= n; ++k)
  {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  pn = p1;
  dpn = n * (x * p1 - p0) / (x * x - 1.0);
}

void gauss_legendre(int nq, double *x, double *w)
{
  for(int i = 0; i < nq; ++i)
  {
    double z = std::cos(M_PI * (i + 0.75) / (nq + 0.5));
    double pn = 0, dpn = 1;
    for(int it = 0; it < 100; ++it)
    {
      legendre_pd(nq, z, pn, dpn);
      const double dz = pn / dpn;
      z -= dz;
      if(std::fabs(dz) < 1e-16) break;
    }
    legendre_pd(nq, z, pn, dpn);
    x[nq - 1 - i] = 0.5 * (1.0 + z);
    w[nq - 1 - i] = 1.0 / ((1.0 - z * z) * dpn * dpn);
  }
}
}  // namespace

void set_last_error(const std::string &msg) { g_last_error = msg; }
const char *last_error_cstr() { return g_last_error.c_str(); }

// Gauss-Lobatto nodes on [0,1] and weights (w may be null)
void gauss_lobatto_nodes(int np, double *x, double *w)
{
  const int N = np - 1;
  if(np == 1)
  {
    x[0] = 0.5;
    if(w) w[0] = 1.0;
    return;
  }
  x[0] = 0.0;
  x[N] = 1.0;
  for(int i = 1; i < N; ++i)
  {
    double z = -std::cos(M_PI * i / N);
    for(int it = 0; it < 100; ++it)
    {
      double pn, dpn;
      legendre_pd(N, z, pn, dpn);
      const double d2 = (2.0 * z * dpn - N * (N + 1.0) * pn) / (1.0 - z * z);
      const double dz = dpn / d2;
      z -= dz;
      if(std::fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (1.0 + z);
  }
  if(w)
    for(int i = 0; i <= N; ++i)
    {
      double pn, dpn;
      const double z = 2.0 * x[i] - 1.0;
      if(i == 0)
        pn = (N % 2) ? -1.0 : 1.0;
      else if(i == N)
        pn = 1.0;
      else
        legendre_pd(N, z, pn, dpn);
      w[i] = 1.0 / (N * (N + 1.0) * pn * pn);
    }
}

void lagrange_eval(int np, const double *nodes, double x, double *val, double *der)
{
  for(int j = 0; j < np; ++j)
  {
    double v = 1.0, d = 0.0;
    for(int m = 0; m < np; ++m)
    {
      if(m == j) continue;
      const double den = nodes[j] - nodes[m];
      d = (d * (x - nodes[m]) + v) / den;
      v = v * (x - nodes[m]) / den;
    }
    val[j] = v;
    der[j] = d;
  }
}

Basis1D make_basis(int p)
{
  Basis1D b;
  b.p = p;
  b.np = p + 1;
  b.nq = p + 1;
  gauss_lobatto_nodes(b.np, b.nodes, nullptr);
  gauss_legendre(b.nq, b.qx, b.qw);
  double v[MAXNP], d[MAXNP];
  for(int q = 0; q < b.nq; ++q)
  {
    lagrange_eval(b.np, b.nodes, b.qx[q], v, d);
    for(int i = 0; i < b.np; ++i)
    {
      b.val[i * b.nq + q] = v[i];
      b.grad[i * b.nq + q] = d[i];
    }
  }
  for(int f = 0; f < 2; ++f)
  {
    lagrange_eval(b.np, b.nodes, double(f), v, d);
    for(int i = 0; i < b.np; ++i)
    {
      b.fval[i * 2 + f] = v[i];
      b.fgrad[i * 2 + f] = d[i];
    }
  }
  return b;
}

// 2 gamma (1/2 h_l^-1 + 1/2 h_r^-1), gamma = p(p+1)  (apps/common_integrator.h:30-37)
double nitsche_penalty(int p, double h_left, double h_right)
{
  const double gamma = p == 0 ? 1.0 : double(p) * (p + 1);
  return 2.0 * gamma * (0.5 / h_left + 0.5 / h_right);
}

Operator1D make_operator_1d(int p, double h, double ip)
{
  const Basis1D b = make_basis(p);
  const int np = b.np, nq = b.nq;
  Operator1D op;
  std::memset(&op, 0, sizeof(op));
  const double sig_i = ip * 0.5 * nitsche_penalty(p, h, h);
  const double sig_b = ip * nitsche_penalty(p, h, h);
  double K[MAXNP * MAXNP], Fmm[MAXNP * MAXNP], Fpp[MAXNP * MAXNP], Bl[MAXNP * MAXNP],
    Br[MAXNP * MAXNP];
  for(int i = 0; i < np; ++i)
    for(int j = 0; j < np; ++j)
    {
      double m = 0, l = 0;
      for(int q = 0; q < nq; ++q)
      {
        m += b.val[i * nq + q] * b.val[j * nq + q] * h * b.qw[q];
        l += b.grad[i * nq + q] * b.grad[j * nq + q] / h * b.qw[q];
      }
      const int ij = i * np + j;
      op.M[ij] = m;
      K[ij] = l;
      const double vi1 = b.fval[i * 2 + 1], gi1 = b.fgrad[i * 2 + 1] / h;
      const double vi0 = b.fval[i * 2 + 0], gi0 = b.fgrad[i * 2 + 0] / h;
      const double vj1 = b.fval[j * 2 + 1], gj1 = b.fgrad[j * 2 + 1] / h;
      const double vj0 = b.fval[j * 2 + 0], gj0 = b.fgrad[j * 2 + 0] / h;
      // interior face seen from the left cell (its face 1) / right cell (face 0)
      Fmm[ij] = sig_i * vi1 * vj1 - 0.5 * vi1 * gj1 - 0.5 * gi1 * vj1;
      op.Fmp[ij] = -sig_i * vi1 * vj0 - 0.5 * vi1 * gj0 + 0.5 * gi1 * vj0;
      op.Fpm[ij] = -sig_i * vi0 * vj1 + 0.5 * vi0 * gj1 - 0.5 * gi0 * vj1;
      Fpp[ij] = sig_i * vi0 * vj0 + 0.5 * vi0 * gj0 + 0.5 * gi0 * vj0;
      // Dirichlet faces, ghost value -u (apps/laplace_integrator.h:715-739)
      Bl[ij] = sig_b * vi0 * vj0 + vi0 * gj0 + gi0 * vj0;
      Br[ij] = sig_b * vi1 * vj1 - vi1 * gj1 - gi1 * vj1;
    }
  for(int c = 0; c < 4; ++c)
  {
    const bool has_l = c & 1, has_r = c & 2;
    for(int ij = 0; ij < np * np; ++ij)
      op.D[c][ij] = K[ij] + (has_l ? Fpp[ij] : Bl[ij]) + (has_r ? Fmm[ij] : Br[ij]);
  }
  return op;
}

void prolongation_1d(int p, double *P)
{
  const int np = p + 1;
  double nodes[MAXNP], v[MAXNP], d[MAXNP];
  gauss_lobatto_nodes(np, nodes, nullptr);
  for(int c = 0; c < 2; ++c)
    for(int i = 0; i < np; ++i)
    {
      lagrange_eval(np, nodes, 0.5 * (nodes[i] + c), v, d);
      for(int j = 0; j < np; ++j) P[(c * np + i) * np + j] = v[j];
    }
}

}  // namespace tpss
