/*
 * tpss_oracle.c -- CPU FP64 restatement of the TPSS vertex-patch Schwarz
 * smoother, the SIPG Poisson operator it smooths, and the geometric multigrid
 * around it.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path; see tpss_oracle.h).
 *
 * Reference citations are file:line in the jwitte08/TPSS tree.
 */
#include "tpss_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXK 16
#define OR_MAXNP 8

/* ========================================================================= */
/* 1D quadrature and Lagrange basis                                           */
/* ========================================================================= */

/* Legendre P_n and P_n' at x in [-1,1] by the three-term recurrence. */
static void legendre(int n, double x, double *pn, double *dpn)
{
  double p0 = 1.0, p1 = x;
  if(n == 0) { *pn = 1.0; *dpn = 0.0; return; }
  for(int k = 2; k <= n; ++k)
  {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *pn = p1;
  *dpn = n * (x * p1 - p0) / (x * x - 1.0);
}

/* Gauss-Legendre rule with nq points mapped to [0,1] (deal.II QGauss(p+1),
 * used by apps/laplace_integrator.h:233 through FDEvaluation<.., fe_degree+1>). */
void or_gauss_legendre(int nq, double *x, double *w)
{
  for(int i = 0; i < nq; ++i)
  {
    double z = cos(M_PI * (i + 0.75) / (nq + 0.5));
    for(int it = 0; it < 100; ++it)
    {
      double pn, dpn;
      legendre(nq, z, &pn, &dpn);
      double dz = pn / dpn;
      z -= dz;
      if(fabs(dz) < 1e-16) break;
    }
    double pn, dpn;
    legendre(nq, z, &pn, &dpn);
    /* ascending order on [0,1] */
    x[nq - 1 - i] = 0.5 * (1.0 + z);
    w[nq - 1 - i] = 1.0 / ((1.0 - z * z) * dpn * dpn); /* = 0.5 * 2/((1-z^2)P'^2) */
  }
}

/* Gauss-Lobatto points (np >= 2) on [0,1]: endpoints and roots of P'_{np-1}
 * (FE_DGQ support points, apps/poisson_problem.h:139-145). */
void or_gauss_lobatto(int np, double *x)
{
  const int N = np - 1;
  x[0] = 0.0;
  x[N] = 1.0;
  for(int i = 1; i < N; ++i)
  {
    double z = -cos(M_PI * i / N);
    for(int it = 0; it < 100; ++it)
    {
      /* Newton on q(z) = P_N'(z) using (1-z^2)P_N'' = 2zP_N' - N(N+1)P_N */
      double pn, dpn;
      legendre(N, z, &pn, &dpn);
      double d2 = (2.0 * z * dpn - N * (N + 1.0) * pn) / (1.0 - z * z);
      double dz = dpn / d2;
      z -= dz;
      if(fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (1.0 + z);
  }
}

/* Lagrange basis on the given nodes: values and first derivatives at x. */
void or_lagrange(int np, const double *nodes, double x, double *val, double *der)
{
  for(int j = 0; j < np; ++j)
  {
    double v = 1.0, d = 0.0;
    for(int m = 0; m < np; ++m)
    {
      if(m == j) continue;
      const double denom = nodes[j] - nodes[m];
      /* product rule: d(v * (x - x_m)/denom) */
      d = (d * (x - nodes[m]) + v) / denom;
      v = v * (x - nodes[m]) / denom;
    }
    val[j] = v;
    der[j] = d;
  }
}

/* Unit-interval shape data: values / gradients at the p+1 Gauss points and
 * at the two faces (x=0 face_no 0, x=1 face_no 1).  Same data as deal.II's
 * UnivariateShapeData consumed by FDEvaluation (fd_evaluation.templates.h:1-61). */
void or_shape_info(int p, double *nodes, double *qx, double *qw, double *val,
                   double *grad, double *fval, double *fgrad)
{
  const int np = p + 1, nq = p + 1;
  double tv[OR_MAXNP], td[OR_MAXNP];
  or_gauss_lobatto(np, nodes);
  or_gauss_legendre(nq, qx, qw);
  for(int q = 0; q < nq; ++q)
  {
    or_lagrange(np, nodes, qx[q], tv, td);
    for(int i = 0; i < np; ++i)
    {
      val[i * nq + q] = tv[i];
      grad[i * nq + q] = td[i];
    }
  }
  for(int f = 0; f < 2; ++f)
  {
    or_lagrange(np, nodes, (double)f, tv, td);
    for(int i = 0; i < np; ++i)
    {
      fval[i * 2 + f] = tv[i];
      fgrad[i * 2 + f] = td[i];
    }
  }
}

/* ========================================================================= */
/* 1D patch matrices (A7)                                                     */
/* ========================================================================= */

/* Nitsche::compute_penalty_impl, apps/common_integrator.h:30-37. */
double or_penalty(int p, double h_left, double h_right)
{
  const double one_over_h = 0.5 / h_left + 0.5 / h_right;
  const double gamma = p == 0 ? 1.0 : (double)p * (p + 1);
  return 2.0 * gamma * one_over_h;
}

/*
 * Mass M and SIPG Laplace L of one direction of a vertex patch (two 1D cells
 * with lengths h[0], h[1]); k = 2(p+1), row-major k x k.
 *  - cell terms: FD::L2::CellOperation / FD::Laplace::CellOperation,
 *    apps/common_integrator.h:161-192, 302-338 (JxW = h w_q, grad / h:
 *    fd_evaluation.templates.h:1-61);
 *  - face terms of each cell: FD::Laplace::SIPG::FaceOperation,
 *    apps/common_integrator.h:196-240 with the average factor 1 at the
 *    physical boundary and 1/2 elsewhere (fd_evaluation.h:651-688) and the
 *    normal -1 / +1 on face 0 / 1 (fd_evaluation.h:779-785);
 *  - interface cell 0 | cell 1: common_integrator.h:242-289;
 *  - assembly into the patch matrix: fd_evaluation.templates.h:192-281,
 *    346-375 with the DGQ 1D index cell_no*(p+1)+i (dof_info.h:358-369).
 */
void or_patch_matrices_1d(int p, const double *h, int bdry_left, int bdry_right,
                          double ip, double *M, double *L)
{
  or_patch_matrices_1d_coef(p, h, bdry_left, bdry_right, ip, 1.0, 1.0, M, L);
}

/*
 * Same as or_patch_matrices_1d with L = a_cons * (cell stiffness + face
 * consistency/symmetry terms) + a_pen * (face penalty terms).  The elasticity
 * local solvers use this form: for component c in direction d the strain
 * integrators scale the stiffness and consistency terms by
 * 2 mu * (1 if c == d else 1/2) but the penalty by 2 mu
 * (CellStrain / NitscheStrain, apps/linelasticity_integrator.h:753-779,
 * 869-966, basic penalty variant), and grad-div adds lambda to both in
 * direction c only (CellGradDiv / NitscheGradDiv, :839-867, 1059-1145).
 */
void or_patch_matrices_1d_coef(int p, const double *h, int bdry_left, int bdry_right,
                               double ip, double a_cons, double a_pen, double *M, double *L)
{
  const int np = p + 1, nq = p + 1, k = 2 * np;
  double nodes[OR_MAXNP], qx[OR_MAXNP], qw[OR_MAXNP];
  double val[OR_MAXNP * OR_MAXNP], grad[OR_MAXNP * OR_MAXNP];
  double fval[OR_MAXNP * 2], fgrad[OR_MAXNP * 2];
  or_shape_info(p, nodes, qx, qw, val, grad, fval, fgrad);
  memset(M, 0, sizeof(double) * k * k);
  memset(L, 0, sizeof(double) * k * k);

  for(int c = 0; c < 2; ++c)
  {
    const double hc = h[c];
    for(int i = 0; i < np; ++i)
      for(int j = 0; j < np; ++j)
      {
        double m = 0.0, l = 0.0;
        for(int q = 0; q < nq; ++q)
        {
          const double dx = hc * qw[q];
          m += val[j * nq + q] * val[i * nq + q] * dx;
          l += (grad[i * nq + q] / hc) * (grad[j * nq + q] / hc) * dx;
        }
        M[(c * np + i) * k + c * np + j] += m;
        L[(c * np + i) * k + c * np + j] += a_cons * l;
      }
  }

  /* face terms, one cell at a time */
  for(int c = 0; c < 2; ++c)
    for(int f = 0; f < 2; ++f)
    {
      const double hc = h[c];
      const int outer = (c == 0 && f == 0) || (c == 1 && f == 1);
      const int at_bdry = outer && ((c == 0) ? bdry_left : bdry_right);
      const double a = at_bdry ? 1.0 : 0.5;
      const double nrm = f == 0 ? -1.0 : 1.0;
      const double pen = ip * a * or_penalty(p, hc, hc);
      for(int i = 0; i < np; ++i)
      {
        const double vi = fval[i * 2 + f], gi = fgrad[i * 2 + f] / hc;
        for(int j = 0; j < np; ++j)
        {
          const double uj = fval[j * 2 + f], guj = fgrad[j * 2 + f] / hc;
          double v = -a_cons * a * (vi * nrm * guj + gi * uj * nrm);
          v += a_pen * pen * vi * uj * nrm * nrm;
          L[(c * np + i) * k + c * np + j] += v;
        }
      }
    }

  /* interface between cell 0 (face 1) and cell 1 (face 0) */
  {
    const double n0 = 1.0, n1 = -1.0;
    const double h0 = h[0], h1 = h[1];
    const double pen = ip * 0.5 * or_penalty(p, h0, h1);
    for(int i = 0; i < np; ++i)
    {
      const double v0 = fval[i * 2 + 1], gv0 = fgrad[i * 2 + 1] / h0;
      const double v1 = fval[i * 2 + 0], gv1 = fgrad[i * 2 + 0] / h1;
      for(int j = 0; j < np; ++j)
      {
        const double u0 = fval[j * 2 + 1], gu0 = fgrad[j * 2 + 1] / h0;
        const double u1 = fval[j * 2 + 0], gu1 = fgrad[j * 2 + 0] / h1;
        double v01 = -a_cons * 0.5 * (v0 * n0 * gu1 + gv0 * u1 * n1);
        double v10 = -a_cons * 0.5 * (v1 * n1 * gu0 + gv1 * u0 * n0);
        v01 += a_pen * pen * v0 * u1 * n0 * n1;
        v10 += a_pen * pen * v1 * u0 * n1 * n0;
        L[(0 * np + i) * k + np + j] += v01;
        L[(np + i) * k + 0 * np + j] += v10;
      }
    }
  }
}

/* ========================================================================= */
/* Symmetric-definite generalized eigenproblem (A8)                           */
/* ========================================================================= */

/* Cyclic Jacobi on a symmetric k x k matrix A (overwritten by its diagonal
 * form), accumulating the rotations into V (Golub & Van Loan, Alg. 8.5.3). */
static void jacobi_eig(int k, double *A, double *V)
{
  for(int i = 0; i < k; ++i)
    for(int j = 0; j < k; ++j)
      V[i * k + j] = i == j ? 1.0 : 0.0;
  for(int sweep = 0; sweep < 100; ++sweep)
  {
    double off = 0.0, diag = 0.0;
    for(int i = 0; i < k; ++i)
    {
      diag += A[i * k + i] * A[i * k + i];
      for(int j = i + 1; j < k; ++j)
        off += A[i * k + j] * A[i * k + j];
    }
    if(off <= 1e-32 * diag) break;
    for(int p = 0; p < k - 1; ++p)
      for(int q = p + 1; q < k; ++q)
      {
        const double apq = A[p * k + q];
        if(apq == 0.0) continue;
        const double app = A[p * k + p], aqq = A[q * k + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) /
                         (fabs(theta) + sqrt(theta * theta + 1.0));
        const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for(int r = 0; r < k; ++r)
        {
          const double arp = A[r * k + p], arq = A[r * k + q];
          A[r * k + p] = cs * arp - sn * arq;
          A[r * k + q] = sn * arp + cs * arq;
        }
        for(int r = 0; r < k; ++r)
        {
          const double apr = A[p * k + r], aqr = A[q * k + r];
          A[p * k + r] = cs * apr - sn * aqr;
          A[q * k + r] = sn * apr + cs * aqr;
        }
        for(int r = 0; r < k; ++r)
        {
          const double vrp = V[r * k + p], vrq = V[r * k + q];
          V[r * k + p] = cs * vrp - sn * vrq;
          V[r * k + q] = sn * vrp + cs * vrq;
        }
      }
  }
}

/*
 * A s = lambda B s with B s.p.d.: returns S (column i = eigenvector i,
 * S^T B S = I) and ascending lambda.  Same mathematics as deal.II's
 * TensorProductMatrixSymmetricSum::reinit(M, L) (called at
 * apps/laplace_integrator.h:248) and the reference wrapper
 * compute_generalized_eigenvalues_symm (generic_functionalities.h:80-113):
 * Cholesky B = C C^T, Jacobi on C^{-1} A C^{-T}, S = C^{-T} Q.
 * Returns 0 on success, 1 if B is not positive definite.
 */
int or_gen_eig(int k, const double *A, const double *B, double *S, double *lam)
{
  double C[OR_MAXK * OR_MAXK], W[OR_MAXK * OR_MAXK], Q[OR_MAXK * OR_MAXK];
  memset(C, 0, sizeof(C));
  for(int j = 0; j < k; ++j)
  {
    double s = B[j * k + j];
    for(int m = 0; m < j; ++m) s -= C[j * k + m] * C[j * k + m];
    if(!(s > 0.0)) return 1;
    C[j * k + j] = sqrt(s);
    for(int i = j + 1; i < k; ++i)
    {
      double t = B[i * k + j];
      for(int m = 0; m < j; ++m) t -= C[i * k + m] * C[j * k + m];
      C[i * k + j] = t / C[j * k + j];
    }
  }
  /* W = C^{-1} A (forward substitution, column by column) */
  for(int col = 0; col < k; ++col)
    for(int i = 0; i < k; ++i)
    {
      double t = A[i * k + col];
      for(int m = 0; m < i; ++m) t -= C[i * k + m] * W[m * k + col];
      W[i * k + col] = t / C[i * k + i];
    }
  /* G = W C^{-T} = (C^{-1} W^T)^T; G symmetric */
  double G[OR_MAXK * OR_MAXK];
  for(int row = 0; row < k; ++row)
    for(int i = 0; i < k; ++i)
    {
      double t = W[row * k + i];
      for(int m = 0; m < i; ++m) t -= C[i * k + m] * G[row * k + m];
      G[row * k + i] = t / C[i * k + i];
    }
  for(int i = 0; i < k; ++i)
    for(int j = i + 1; j < k; ++j)
    {
      const double s = 0.5 * (G[i * k + j] + G[j * k + i]);
      G[i * k + j] = G[j * k + i] = s;
    }
  jacobi_eig(k, G, Q);
  /* sort ascending */
  int order[OR_MAXK];
  for(int i = 0; i < k; ++i) order[i] = i;
  for(int i = 1; i < k; ++i)
  {
    int t = order[i], j = i - 1;
    while(j >= 0 && G[order[j] * k + order[j]] > G[t * k + t]) { order[j + 1] = order[j]; --j; }
    order[j + 1] = t;
  }
  /* S = C^{-T} Q (back substitution) */
  for(int c = 0; c < k; ++c)
  {
    const int oc = order[c];
    lam[c] = G[oc * k + oc];
    for(int i = k - 1; i >= 0; --i)
    {
      double t = Q[i * k + oc];
      for(int m = i + 1; m < k; ++m) t -= C[m * k + i] * S[m * k + c];
      S[i * k + c] = t / C[i * k + i];
    }
  }
  return 0;
}

/* ========================================================================= */
/* Fast-diagonalization inverse (A5)                                          */
/* ========================================================================= */

/* out = mats applied along every direction; mats[e] (m x m, row-major) acts
 * on tensor index e (index 0 fastest): out[..i_e..] = sum_j A_e(i_e, j) in[..j..].
 * trans != 0 uses A_e^T.  The mode-e product is the one of
 * internal::My::contract_general_impl (move_to_deal_ii.h:701-799). */
static void mode_product(int dim, int m, int e, const double *A, int trans,
                         const double *in, double *out)
{
  int stride = 1;
  for(int d = 0; d < e; ++d) stride *= m;
  int total = 1;
  for(int d = 0; d < dim; ++d) total *= m;
  const int n_post = total / (stride * m);
  for(int post = 0; post < n_post; ++post)
    for(int pre = 0; pre < stride; ++pre)
    {
      const int base = post * stride * m + pre;
      for(int i = 0; i < m; ++i)
      {
        double s = 0.0;
        for(int j = 0; j < m; ++j)
          s += (trans ? A[j * m + i] : A[i * m + j]) * in[base + j * stride];
        out[base + i * stride] = s;
      }
    }
}

/*
 * u = (S_{d-1} x .. x S_0) diag(1 / sum_d lambda_d) (S_{d-1}^T x .. x S_0^T) r,
 * the fast-diagonalization inverse of A = sum_d M x .. x L_d x .. x M
 * (dimension 0 rightmost, tensor_product_matrix.h:26-28, tensors.h:765-797).
 * S: dim blocks of k*k (column i = eigenvector i), lam: dim blocks of k.
 */
void or_fd_apply(int dim, int k, const double *S, const double *lam,
                 const double *in, double *out)
{
  int n = 1;
  for(int d = 0; d < dim; ++d) n *= k;
  double *a = (double *)malloc(sizeof(double) * n);
  double *b = (double *)malloc(sizeof(double) * n);
  memcpy(a, in, sizeof(double) * n);
  for(int d = 0; d < dim; ++d)
  {
    mode_product(dim, k, d, S + d * k * k, 1, a, b);
    double *t = a; a = b; b = t;
  }
  for(int i = 0; i < n; ++i)
  {
    int rem = i;
    double ev = 0.0;
    for(int d = 0; d < dim; ++d) { ev += lam[d * k + rem % k]; rem /= k; }
    a[i] /= ev;
  }
  for(int d = 0; d < dim; ++d)
  {
    mode_product(dim, k, d, S + d * k * k, 0, a, b);
    double *t = a; a = b; b = t;
  }
  memcpy(out, a, sizeof(double) * n);
  free(a);
  free(b);
}

/* ========================================================================= */
/* Smoother context                                                           */
/* ========================================================================= */

struct or_ctx
{
  int dim, p, n, np, k, ndc, ndp;
  long ncell, N;
  double ip;
  int ncomp;         /* 1: Poisson; dim: linear elasticity (component-blocked vectors) */
  double mu, lambda; /* Lame coefficients (elasticity) */
  int variant, coloring, reverse, symmetrized, n_steps, n_threads;
  double relaxation, rich_damping;
  double *vertices; /* (n+1)^dim * dim or NULL (Cartesian) */
  /* patches */
  int P, n_colors;
  int *color_ptr; /* n_colors + 1 */
  int *vtx;       /* P * dim */
  double *h;      /* P * dim * 2 */
  unsigned char *bdry; /* P * dim * 2 */
  double *S, *lam;     /* P * ncomp * dim * k * k, P * ncomp * dim * k */
  double *omega;       /* P or NULL */
  /* operator data (Cartesian) */
  double Mc[OR_MAXNP * OR_MAXNP], Kc[OR_MAXNP * OR_MAXNP];
  double Fmm[OR_MAXNP * OR_MAXNP], Fmp[OR_MAXNP * OR_MAXNP];
  double Fpm[OR_MAXNP * OR_MAXNP], Fpp[OR_MAXNP * OR_MAXNP];
  double Bl[OR_MAXNP * OR_MAXNP], Br[OR_MAXNP * OR_MAXNP];
  /* local solvers: surrogate-length FD (0) or separable fit of the exact local matrix (1) */
  int local_solver, fit_sweeps, n_fit_fallback;
  double *fit_err; /* P: relative Frobenius error of the fitted separable matrix */
  /* restricted additive Schwarz: 0 off, 1 use_ras_weights, 2 use_ras_boolean_weights
   * (schwarz_preconditioner.h:501-502) */
  int ras;
  double *ras_w;               /* N: number of patches containing each DoF */
  unsigned char *ras_flags;    /* P * ndp: 1 = restricted (not scattered) */
  int ready;
};

static int ipow(int b, int e)
{
  int r = 1;
  while(e-- > 0) r *= b;
  return r;
}

/*
 * Vertex-patch color (A13).
 * RedBlack: apps/coloring.h:286-324 -- lower-left cell x = v-1, shift_d =
 * x_d mod 2 (coloring.h:294-304), color = 2*shift_mask + (sum_d (x_d+shift_d)/2) mod 2.
 * Tiled: apps/coloring.h:359-390 -- color = shift_mask.
 */
static int patch_color(int dim, int coloring, const int *v)
{
  if(coloring == OR_COLOR_NONE) return 0;
  int mask = 0, sum = 0;
  for(int d = 0; d < dim; ++d)
  {
    const int x = v[d] - 1, s = x & 1;
    mask |= s << d;
    sum += (x + s) / 2;
  }
  if(coloring == OR_COLOR_TILED) return mask;
  return 2 * mask + (sum & 1);
}

/* A patch is a boundary patch if one of its cells touches the physical
 * boundary (patch_info.templates.h:611-645, IteratorFilters::AtBoundary). */
static int patch_at_boundary(int dim, int n, const int *v)
{
  for(int d = 0; d < dim; ++d)
    if(v[d] == 1 || v[d] == n - 1) return 1;
  return 0;
}

or_ctx *or_create(int dim, int p, int n, double ip, int variant, int coloring,
                  double relaxation, int reverse, int symmetrized,
                  double richardson_damping, int n_steps)
{
  if(dim < 1 || dim > 3 || p < 0 || p + 1 > OR_MAXNP || n < 2) return NULL;
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->dim = dim;
  c->p = p;
  c->n = n;
  c->np = p + 1;
  c->k = 2 * (p + 1);
  c->ndc = ipow(p + 1, dim);
  c->ndp = ipow(c->k, dim);
  c->ncell = ipow(n, dim);
  c->N = c->ncell * c->ndc;
  c->ncomp = 1;
  c->ip = ip;
  c->variant = variant;
  c->coloring = coloring;
  c->relaxation = relaxation;
  c->reverse = reverse;
  c->symmetrized = symmetrized;
  c->rich_damping = richardson_damping;
  c->n_steps = n_steps;
  c->n_threads = 1;
  c->n_colors = coloring == OR_COLOR_NONE ? 1 : coloring == OR_COLOR_TILED ? (1 << dim) : 2 << dim;
  return c;
}

/*
 * Linear elasticity: SIPG DG Q_p vector field with
 * dim components on the Cartesian unit cube, Lame coefficients mu, lambda
 * (apps/linelasticity_integrator.h:31-124 for the operator; block-diagonal
 * local solvers :604-647).  Vectors are component-blocked like the
 * reference's block vectors (one scalar DG block per component,
 * patch_transfer.h:285-418): x[c * ncell * (p+1)^dim + cell * (p+1)^dim + local].
 */
or_ctx *or_create_elasticity(int dim, int p, int n, double ip, double mu, double lambda,
                             int variant, int coloring, double relaxation, int reverse,
                             int symmetrized, double richardson_damping, int n_steps)
{
  if(dim < 2) return NULL;
  or_ctx *c = or_create(dim, p, n, ip, variant, coloring, relaxation, reverse, symmetrized,
                        richardson_damping, n_steps);
  if(!c) return NULL;
  c->ncomp = dim;
  c->N = c->ncell * c->ndc * dim;
  c->mu = mu;
  c->lambda = lambda;
  return c;
}

int or_n_components(const or_ctx *c) { return c->ncomp; }

void or_set_local_solver(or_ctx *c, int kind, int sweeps)
{
  c->local_solver = kind;
  c->fit_sweeps = sweeps > 0 ? sweeps : 4;
}

void or_fit_errors(const or_ctx *c, double *out)
{
  for(int j = 0; j < c->P; ++j) out[j] = c->fit_err ? c->fit_err[j] : 0.0;
}

int or_fit_fallbacks(const or_ctx *c) { return c->n_fit_fallback; }

void or_destroy(or_ctx *c)
{
  if(!c) return;
  free(c->fit_err);
  free(c->ras_w);
  free(c->ras_flags);
  free(c->vertices);
  free(c->color_ptr);
  free(c->vtx);
  free(c->h);
  free(c->bdry);
  free(c->S);
  free(c->lam);
  free(c->omega);
  free(c);
}

int or_set_vertices(or_ctx *c, const double *coords)
{
  const long nv = ipow(c->n + 1, c->dim);
  free(c->vertices);
  c->vertices = NULL;
  if(!coords) return 0;
  c->vertices = (double *)malloc(sizeof(double) * nv * c->dim);
  memcpy(c->vertices, coords, sizeof(double) * nv * c->dim);
  return 0;
}

void or_set_threads(or_ctx *c, int n_threads) { c->n_threads = n_threads > 0 ? n_threads : 1; }

/* 1D cell and face matrices of the Cartesian SIPG operator (A11 restated for
 * tensor-product cells, apps/laplace_integrator.h:613-743). */
static void build_operator_1d(or_ctx *c)
{
  const int np = c->np, nq = c->np, p = c->p;
  const double h = 1.0 / c->n;
  double nodes[OR_MAXNP], qx[OR_MAXNP], qw[OR_MAXNP];
  double val[OR_MAXNP * OR_MAXNP], grad[OR_MAXNP * OR_MAXNP];
  double fv[OR_MAXNP * 2], fg[OR_MAXNP * 2];
  or_shape_info(p, nodes, qx, qw, val, grad, fv, fg);
  const double sig_i = c->ip * 0.5 * or_penalty(p, h, h); /* interior face */
  const double sig_b = c->ip * or_penalty(p, h, h);       /* boundary face */
  for(int i = 0; i < np; ++i)
    for(int j = 0; j < np; ++j)
    {
      double m = 0, l = 0;
      for(int q = 0; q < nq; ++q)
      {
        m += val[i * nq + q] * val[j * nq + q] * h * qw[q];
        l += grad[i * nq + q] * grad[j * nq + q] / h * qw[q];
      }
      const int ij = i * np + j;
      c->Mc[ij] = m;
      c->Kc[ij] = l;
      const double vi1 = fv[i * 2 + 1], gi1 = fg[i * 2 + 1] / h;
      const double vi0 = fv[i * 2 + 0], gi0 = fg[i * 2 + 0] / h;
      const double vj1 = fv[j * 2 + 1], gj1 = fg[j * 2 + 1] / h;
      const double vj0 = fv[j * 2 + 0], gj0 = fg[j * 2 + 0] / h;
      /* interior face, "-" cell on the left (its face 1), "+" on the right */
      c->Fmm[ij] = sig_i * vi1 * vj1 - 0.5 * vi1 * gj1 - 0.5 * gi1 * vj1;
      c->Fmp[ij] = -sig_i * vi1 * vj0 - 0.5 * vi1 * gj0 + 0.5 * gi1 * vj0;
      c->Fpm[ij] = -sig_i * vi0 * vj1 + 0.5 * vi0 * gj1 - 0.5 * gi0 * vj1;
      c->Fpp[ij] = sig_i * vi0 * vj0 + 0.5 * vi0 * gj0 + 0.5 * gi0 * vj0;
      /* Dirichlet boundary faces (u_out = -u_in, laplace_integrator.h:715-739) */
      c->Bl[ij] = sig_b * vi0 * vj0 + vi0 * gj0 + gi0 * vj0;
      c->Br[ij] = sig_b * vi1 * vj1 - vi1 * gj1 - gi1 * vj1;
    }
}

/* Trilinear geometry helpers (distorted meshes). */
static const double *vertex_at(const or_ctx *c, const int *iv)
{
  long idx = 0, stride = 1;
  for(int d = 0; d < c->dim; ++d) { idx += iv[d] * stride; stride *= (c->n + 1); }
  return c->vertices + idx * c->dim;
}

/*
 * Surrogate lengths of a patch (A9): for each direction and 1D cell slot,
 * the average over the 2^{dim-1} cells in that slot of the cell's
 * Gauss-Lobatto averaged distance between its two opposite faces
 * (mapping_info.templates.h:42-240 with n_q = min(8, p+1) and straight
 * trilinear cell edges, so the arc length of a mapped line is the distance of
 * its end points for every line parallel to a reference axis).
 */
static void patch_surrogate_h(const or_ctx *c, const int *v, double *hout)
{
  const int dim = c->dim;
  if(!c->vertices)
  {
    for(int d = 0; d < dim; ++d) hout[2 * d] = hout[2 * d + 1] = 1.0 / c->n;
    return;
  }
  int nq = c->p + 1;
  if(nq > 8) nq = 8;
  if(nq < 2) nq = 2;
  double gx[OR_MAXNP + 2], gw[OR_MAXNP + 2];
  or_gauss_lobatto(nq, gx);
  /* Gauss-Lobatto weights on [0,1]: w_i = 1/(N(N+1) P_N(z_i)^2) */
  {
    const int N = nq - 1;
    for(int i = 0; i < nq; ++i)
    {
      double pn, dpn;
      const double z = 2.0 * gx[i] - 1.0;
      if(i == 0 || i == N) { pn = (i == 0 && (N & 1)) ? -1.0 : 1.0; }
      else legendre(N, z, &pn, &dpn);
      gw[i] = 1.0 / (N * (N + 1.0) * pn * pn);
    }
  }
  for(int d = 0; d < dim; ++d)
    for(int slot = 0; slot < 2; ++slot)
    {
      double acc = 0.0;
      int ncells = 0;
      for(int cb = 0; cb < (1 << dim); ++cb)
      {
        if(((cb >> d) & 1) != slot) continue;
        int cell[3];
        for(int e = 0; e < dim; ++e) cell[e] = v[e] - 1 + ((cb >> e) & 1);
        /* average over GL points of the transverse reference coords */
        double len = 0.0;
        const int nt = dim == 1 ? 1 : dim == 2 ? nq : nq * nq;
        for(int t = 0; t < nt; ++t)
        {
          double xr[3], w = 1.0;
          int tt = t, f = 0;
          for(int e = 0; e < dim; ++e)
          {
            if(e == d) continue;
            const int qi = tt % nq;
            tt /= nq;
            xr[e] = gx[qi];
            w *= gw[qi];
            ++f;
          }
          double pa[3] = {0, 0, 0}, pb[3] = {0, 0, 0};
          for(int corner = 0; corner < (1 << dim); ++corner)
          {
            double wa = 1.0, wb = 1.0;
            int iv[3];
            for(int e = 0; e < dim; ++e)
            {
              const int b = (corner >> e) & 1;
              iv[e] = cell[e] + b;
              if(e == d) { wa *= b == 0 ? 1.0 : 0.0; wb *= b == 1 ? 1.0 : 0.0; }
              else { const double s = b ? xr[e] : 1.0 - xr[e]; wa *= s; wb *= s; }
            }
            const double *X = vertex_at(c, iv);
            for(int e = 0; e < dim; ++e) { pa[e] += wa * X[e]; pb[e] += wb * X[e]; }
          }
          double dist = 0.0;
          for(int e = 0; e < dim; ++e) dist += (pb[e] - pa[e]) * (pb[e] - pa[e]);
          len += w * sqrt(dist);
        }
        acc += len;
        ++ncells;
      }
      hout[2 * d + slot] = acc / ncells;
    }
}

int or_setup(or_ctx *c)
{
  const int dim = c->dim, n = c->n, k = c->k;
  int P = ipow(n - 1, dim);
  c->P = P;
  free(c->color_ptr);
  free(c->vtx);
  free(c->h);
  free(c->bdry);
  free(c->S);
  free(c->lam);
  c->color_ptr = (int *)calloc(c->n_colors + 1, sizeof(int));
  c->vtx = (int *)malloc(sizeof(int) * P * dim);
  c->h = (double *)malloc(sizeof(double) * P * dim * 2);
  c->bdry = (unsigned char *)malloc(P * dim * 2);
  const int nb = c->ncomp;
  if(nb > 1 && c->vertices) return -2; /* elasticity: Cartesian meshes only */
  c->S = (double *)malloc(sizeof(double) * P * nb * dim * k * k);
  c->lam = (double *)malloc(sizeof(double) * P * nb * dim * k);

  /* order: color, then interior before boundary, then vertex lexicographic */
  int pos = 0;
  for(int col = 0; col < c->n_colors; ++col)
  {
    c->color_ptr[col] = pos;
    for(int pass = 0; pass < 2; ++pass)
      for(int lin = 0; lin < P; ++lin)
      {
        int v[3], t = lin;
        for(int d = 0; d < dim; ++d) { v[d] = 1 + t % (n - 1); t /= (n - 1); }
        if(patch_color(dim, c->coloring, v) != col) continue;
        if(patch_at_boundary(dim, n, v) != pass) continue;
        for(int d = 0; d < dim; ++d) c->vtx[pos * dim + d] = v[d];
        ++pos;
      }
  }
  c->color_ptr[c->n_colors] = pos;

  /* local solvers: FD::MatrixIntegrator::assemble_subspace_inverses,
   * apps/laplace_integrator.h:226-250 */
  int status = 0;
#pragma omp parallel for num_threads(c->n_threads) schedule(static) reduction(| : status)
  for(int j = 0; j < P; ++j)
  {
    const int *v = c->vtx + j * dim;
    double hh[6];
    patch_surrogate_h(c, v, hh);
    for(int d = 0; d < dim; ++d)
    {
      const int bl = v[d] - 1 == 0, br = v[d] + 1 == n;
      c->h[(j * dim + d) * 2 + 0] = hh[2 * d];
      c->h[(j * dim + d) * 2 + 1] = hh[2 * d + 1];
      c->bdry[(j * dim + d) * 2 + 0] = (unsigned char)bl;
      c->bdry[(j * dim + d) * 2 + 1] = (unsigned char)br;
      for(int b = 0; b < nb; ++b)
      {
        double M[OR_MAXK * OR_MAXK], L[OR_MAXK * OR_MAXK];
        double a_cons = 1.0, a_pen = 1.0;
        if(nb > 1) /* block b of BlockMatrixDiagonal, linelasticity_integrator.h:636-645 */
        {
          a_cons = 2.0 * c->mu * (b == d ? 1.0 : 0.5) + (b == d ? c->lambda : 0.0);
          a_pen = 2.0 * c->mu + (b == d ? c->lambda : 0.0);
        }
        or_patch_matrices_1d_coef(c->p, hh + 2 * d, bl, br, c->ip, a_cons, a_pen, M, L);
        const long off = (long)(j * nb + b) * dim + d;
        status |= or_gen_eig(k, L, M, c->S + off * k * k, c->lam + off * k);
      }
    }
  }
  build_operator_1d(c);
  if(c->local_solver == OR_LOCAL_FIT && c->ncomp == 1 && !status)
  {
    /* separable fit of the exact local matrix R_j A R_j^T, initialised with
     * the surrogate-length matrices; FD factors of the fit */
    const int ndp = c->ndp;
    double *A = (double *)malloc(sizeof(double) * ndp * ndp);
    free(c->fit_err);
    c->fit_err = (double *)calloc(P, sizeof(double));
    c->n_fit_fallback = 0;
    for(int j = 0; j < P; ++j)
    {
      const int *v = c->vtx + j * dim;
      double M[3 * OR_MAXK * OR_MAXK], L[3 * OR_MAXK * OR_MAXK];
      for(int d = 0; d < dim; ++d)
        or_patch_matrices_1d(c->p, c->h + (j * dim + d) * 2, v[d] - 1 == 0, v[d] + 1 == n, c->ip,
                             M + d * k * k, L + d * k * k);
      or_patch_local_matrix(c, j, A);
      c->fit_err[j] = or_separable_fit(dim, k, A, M, L, c->fit_sweeps);
      double S[3 * OR_MAXK * OR_MAXK], lam[3 * OR_MAXK];
      int bad = 0;
      for(int d = 0; d < dim; ++d) bad |= or_gen_eig(k, L + d * k * k, M + d * k * k, S + d * k * k, lam + d * k);
      if(!bad)
      {
        /* the fit must be s.p.d.: smallest Kronecker-sum eigenvalue positive */
        double lo = 0.0, hi = 0.0;
        for(int d = 0; d < dim; ++d) { lo += lam[d * k]; hi += fabs(lam[d * k + k - 1]); }
        bad = !(lo > 1e-12 * hi);
      }
      if(bad)
      {
        ++c->n_fit_fallback; /* keep the surrogate factors */
        c->fit_err[j] = -1.0;
        continue;
      }
      memcpy(c->S + (long)j * dim * k * k, S, sizeof(double) * dim * k * k);
      memcpy(c->lam + (long)j * dim * k, lam, sizeof(double) * dim * k);
    }
    free(A);
  }
  c->ready = 1;
  return status;
}

void or_set_patch_damping(or_ctx *c, const double *omega)
{
  free(c->omega);
  c->omega = NULL;
  if(!omega) return;
  c->omega = (double *)malloc(sizeof(double) * c->P);
  memcpy(c->omega, omega, sizeof(double) * c->P);
}

long or_n_dofs(const or_ctx *c) { return c->N; }
int or_n_patches(const or_ctx *c) { return c->P; }
int or_n_colors(const or_ctx *c) { return c->n_colors; }
void or_color_ptr(const or_ctx *c, int *out) { memcpy(out, c->color_ptr, sizeof(int) * (c->n_colors + 1)); }
void or_patch_vertices(const or_ctx *c, int *out) { memcpy(out, c->vtx, sizeof(int) * c->P * c->dim); }
void or_patch_h(const or_ctx *c, double *out) { memcpy(out, c->h, sizeof(double) * c->P * c->dim * 2); }
void or_patch_factors(const or_ctx *c, double *S, double *lam)
{
  if(S) memcpy(S, c->S, sizeof(double) * c->P * c->ncomp * c->dim * c->k * c->k);
  if(lam) memcpy(lam, c->lam, sizeof(double) * c->P * c->ncomp * c->dim * c->k);
}

/* Global DoF indices of one patch in patch-lexicographic order (A3):
 * DoFInfo patch-wise cache (dof_info.templates.h:156-182) with the DGQ 1D map
 * cell_no*(p+1)+i (dof_info.h:358-369), x-fastest uni-index (tensors.h:63-82)
 * and cells of the patch in lexicographic order (patch_info.templates.h:401-402). */
static void patch_dofs(const or_ctx *c, const int *v, uint32_t *idx)
{
  const int dim = c->dim, np = c->np, k = c->k, n = c->n;
  for(int i = 0; i < c->ndp; ++i)
  {
    int t = i;
    long cell = 0, cstride = 1, loc = 0, lstride = 1;
    for(int d = 0; d < dim; ++d)
    {
      const int id = t % k;
      t /= k;
      const int cell_no = id / np, ld = id % np;
      cell += (long)(v[d] - 1 + cell_no) * cstride;
      cstride *= n;
      loc += ld * lstride;
      lstride *= np;
    }
    idx[i] = (uint32_t)(cell * c->ndc + loc);
  }
  /* further components: the same cells in the next block (PatchTransferBlock
   * juxtaposes the blocks' index sets, patch_transfer.h:340-354) */
  const long Ns = c->ncell * c->ndc;
  for(int b = 1; b < c->ncomp; ++b)
    for(int i = 0; i < c->ndp; ++i) idx[b * c->ndp + i] = (uint32_t)(idx[i] + b * Ns);
}

void or_patch_dof_indices(const or_ctx *c, uint32_t *out)
{
  for(int j = 0; j < c->P; ++j) patch_dofs(c, c->vtx + j * c->dim, out + (long)j * c->ncomp * c->ndp);
}

/* get_color_sequence, schwarz_preconditioner.h:531-567 */
void or_color_sequence(const or_ctx *c, int transpose, int *out, int *n_out)
{
  const int nc = c->n_colors;
  int m = 0;
  for(int i = 0; i < nc; ++i) out[m++] = i;
  if(transpose || c->reverse)
    for(int i = 0; i < nc / 2; ++i) { int t = out[i]; out[i] = out[nc - 1 - i]; out[nc - 1 - i] = t; }
  if(c->symmetrized)
    for(int i = nc - 2; i >= 0; --i) out[m++] = out[i];
  *n_out = m;
}

/* ---------------- SIPG operator (A11) ------------------------------------ */

static void tensor_apply(int dim, int m, const double *const *mats, const double *in, double *out, double *tmp)
{
  int n = ipow(m, dim);
  const double *src = in;
  double *bufs[2] = {out, tmp};
  /* make the last write land in out */
  int which = (dim % 2 == 1) ? 0 : 1;
  for(int e = 0; e < dim; ++e)
  {
    double *dst = bufs[which];
    mode_product(dim, m, e, mats[e], 0, src, dst);
    src = dst;
    which ^= 1;
  }
  if(src != out) memcpy(out, src, sizeof(double) * n);
}

/* y = A x for the Cartesian unit-cube mesh: cell integrals + interior faces
 * (apps/laplace_integrator.h:613-684) + Dirichlet faces (:686-743), written per
 * cell with 1D factors (exact for tensor-product cells). */
static void operator_cartesian(const or_ctx *c, double *dst, const double *src)
{
  const int dim = c->dim, np = c->np, n = c->n, ndc = c->ndc;
#pragma omp parallel for num_threads(c->n_threads) schedule(static)
  for(long cell = 0; cell < c->ncell; ++cell)
  {
    double acc[OR_MAXNP * OR_MAXNP * OR_MAXNP], tmp[OR_MAXNP * OR_MAXNP * OR_MAXNP];
    double t2[OR_MAXNP * OR_MAXNP * OR_MAXNP], D[OR_MAXNP * OR_MAXNP];
    int cc[3];
    long r = cell;
    for(int d = 0; d < dim; ++d) { cc[d] = (int)(r % n); r /= n; }
    for(int i = 0; i < ndc; ++i) acc[i] = 0.0;
    long stride = 1;
    for(int d = 0; d < dim; ++d)
    {
      const double *mats[3];
      for(int e = 0; e < dim; ++e) mats[e] = c->Mc;
      const int has_l = cc[d] > 0, has_r = cc[d] < n - 1;
      for(int i = 0; i < np * np; ++i)
        D[i] = c->Kc[i] + (has_l ? c->Fpp[i] : c->Bl[i]) + (has_r ? c->Fmm[i] : c->Br[i]);
      mats[d] = D;
      tensor_apply(dim, np, mats, src + cell * ndc, tmp, t2);
      for(int i = 0; i < ndc; ++i) acc[i] += tmp[i];
      if(has_l)
      {
        mats[d] = c->Fpm;
        tensor_apply(dim, np, mats, src + (cell - stride) * ndc, tmp, t2);
        for(int i = 0; i < ndc; ++i) acc[i] += tmp[i];
      }
      if(has_r)
      {
        mats[d] = c->Fmp;
        tensor_apply(dim, np, mats, src + (cell + stride) * ndc, tmp, t2);
        for(int i = 0; i < ndc; ++i) acc[i] += tmp[i];
      }
      stride *= n;
    }
    for(int i = 0; i < ndc; ++i) dst[cell * ndc + i] = acc[i];
  }
}

/* Gauss-Lobatto nodes and QGauss(p+1) of the moved-vertex operator */
typedef struct
{
  double nodes[OR_MAXNP], qx[OR_MAXNP], qw[OR_MAXNP];
  int nq, nqc, nqf;
  double gamma;
} dist_quad;

static void operator_distorted(const or_ctx *c, double *dst, const double *src);

static void operator_elasticity(const or_ctx *c, double *dst, const double *src);

void or_operator_vmult(const or_ctx *c, double *dst, const double *src)
{
  if(c->ncomp > 1) operator_elasticity(c, dst, src);
  else if(c->vertices) operator_distorted(c, dst, src);
  else operator_cartesian(c, dst, src);
}

/* ---------------- exact local matrices and their separable fit ----------- */

/* A_j = R_j A R_j^T (row-major ndp x ndp), column by column from the global
 * operator applied to unit vectors: the patch matrix the reference's
 * TensorProductMatrix "basic" state and the Kronecker SVD start from
 * (tensor_product_matrix.h:136-304, kroneckersvd.h:215-296). */
static void operator_distorted_one_cell(const or_ctx *c, const dist_quad *g, long cell,
                                        double *dst, const double *src);
static void dist_quad_init(const or_ctx *c, dist_quad *g);

void or_patch_local_matrix(const or_ctx *c, int patch, double *A)
{
  const int ndp = c->ndp;
  uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * ndp * c->ncomp);
  patch_dofs(c, c->vtx + patch * c->dim, idx);
  const int one_cell = c->vertices && c->ncomp == 1;
  dist_quad g;
  if(one_cell) dist_quad_init(c, &g);
#pragma omp parallel num_threads(one_cell ? c->n_threads : 1)
  {
    double *x = (double *)calloc(c->N, sizeof(double));
    double *y = (double *)calloc(c->N, sizeof(double));
#pragma omp for schedule(static)
    for(int col = 0; col < ndp; ++col)
    {
      x[idx[col]] = 1.0;
      if(one_cell)
      {
        /* a unit vector lives in one cell: only that cell and its faces
         * contribute (and only to the rows of its face neighbours) */
        const long cell = (long)(idx[col] / (uint32_t)c->ndc);
        for(int row = 0; row < ndp; ++row) y[idx[row]] = 0.0;
        operator_distorted_one_cell(c, &g, cell, y, x);
      }
      else
        or_operator_vmult(c, y, x);
      x[idx[col]] = 0.0;
      for(int row = 0; row < ndp; ++row) A[(long)row * ndp + col] = y[idx[row]];
      if(one_cell)
      {
        /* clear what the one-cell apply wrote outside the patch rows */
        int cc[3];
        const long cell = (long)(idx[col] / (uint32_t)c->ndc);
        long r = cell, stride = 1;
        for(int d = 0; d < c->dim; ++d) { cc[d] = (int)(r % c->n); r /= c->n; }
        for(int d = 0; d < c->dim; ++d)
        {
          if(cc[d] > 0) memset(y + (cell - stride) * c->ndc, 0, sizeof(double) * c->ndc);
          if(cc[d] < c->n - 1) memset(y + (cell + stride) * c->ndc, 0, sizeof(double) * c->ndc);
          stride *= c->n;
        }
        memset(y + cell * c->ndc, 0, sizeof(double) * c->ndc);
      }
    }
    free(x);
    free(y);
  }
  free(idx);
}

/*
 * Separable fit  A ~ sum_d M_{dim-1} x .. x L_d x .. x M_0  (dimension 0
 * rightmost) by alternating least squares in the Frobenius norm: for each
 * direction d in turn the pair (L_d, M_d) solves the rank-2 Kronecker
 * least-squares problem
 *   min || A - L_d (x)_d R1 - M_d (x)_d R2 ||_F,
 *   R1 = (x)_{e != d} M_e,  R2 = sum_{e != d} (x)_{f != d} (L_f if f == e else M_f),
 * one 2x2 normal-equation system per entry (i_d, j_d).  This is the CP-ALS of
 * the reference (kroneckersvd.h:501-585) with the factors tied to the
 * separable form, so that the fit stays invertible by fast diagonalization
 * (TensorProductMatrixSymmetricSum).  M, L: dim blocks of k x k, in: initial
 * guess, out: fit.  Returns ||A - fit||_F / ||A||_F.
 */
double or_separable_fit(int dim, int k, const double *A, double *M, double *L, int sweeps)
{
  const int ndp = ipow(k, dim);
  double nM[3], nL[3], ML[3];
  for(int s = 0; s < sweeps; ++s)
    for(int d = 0; d < dim; ++d)
    {
      for(int e = 0; e < dim; ++e)
      {
        nM[e] = nL[e] = ML[e] = 0.0;
        for(int q = 0; q < k * k; ++q)
        {
          nM[e] += M[e * k * k + q] * M[e * k * k + q];
          nL[e] += L[e * k * k + q] * L[e * k * k + q];
          ML[e] += M[e * k * k + q] * L[e * k * k + q];
        }
      }
      /* Gram matrix of (R1, R2) */
      double g11 = 1.0, g12 = 0.0, g22 = 0.0;
      for(int e = 0; e < dim; ++e)
        if(e != d) g11 *= nM[e];
      for(int e = 0; e < dim; ++e)
      {
        if(e == d) continue;
        double t12 = ML[e], t22 = nL[e];
        for(int f = 0; f < dim; ++f)
          if(f != d && f != e) { t12 *= nM[f]; t22 *= nM[f]; }
        g12 += t12;
        g22 += t22;
        for(int e2 = 0; e2 < dim; ++e2)
        {
          if(e2 == d || e2 == e) continue;
          double t = ML[e] * ML[e2];
          for(int f = 0; f < dim; ++f)
            if(f != d && f != e && f != e2) t *= nM[f];
          g22 += t;
        }
      }
      double T1[OR_MAXK * OR_MAXK], T2[OR_MAXK * OR_MAXK];
      memset(T1, 0, sizeof(T1));
      memset(T2, 0, sizeof(T2));
      for(long r = 0; r < ndp; ++r)
        for(long col = 0; col < ndp; ++col)
        {
          int ri[3], ci[3];
          long tr = r, tc = col;
          for(int e = 0; e < dim; ++e) { ri[e] = (int)(tr % k); tr /= k; ci[e] = (int)(tc % k); tc /= k; }
          double w1 = 1.0, w2 = 0.0;
          for(int e = 0; e < dim; ++e)
          {
            if(e == d) continue;
            const double m = M[e * k * k + ri[e] * k + ci[e]], l = L[e * k * k + ri[e] * k + ci[e]];
            w2 = w2 * m + w1 * l;
            w1 *= m;
          }
          const double a = A[r * ndp + col];
          T1[ri[d] * k + ci[d]] += a * w1;
          T2[ri[d] * k + ci[d]] += a * w2;
        }
      const double det = g11 * g22 - g12 * g12;
      for(int q = 0; q < k * k; ++q)
      {
        L[d * k * k + q] = (g22 * T1[q] - g12 * T2[q]) / det;
        M[d * k * k + q] = (g11 * T2[q] - g12 * T1[q]) / det;
      }
    }
  /* relative Frobenius error of the fit */
  double num = 0.0, den = 0.0;
  for(long r = 0; r < ndp; ++r)
    for(long col = 0; col < ndp; ++col)
    {
      int ri[3], ci[3];
      long tr = r, tc = col;
      for(int e = 0; e < dim; ++e) { ri[e] = (int)(tr % k); tr /= k; ci[e] = (int)(tc % k); tc /= k; }
      double fit = 0.0;
      for(int d = 0; d < dim; ++d)
      {
        double t = 1.0;
        for(int e = 0; e < dim; ++e)
          t *= (e == d ? L : M)[e * k * k + ri[e] * k + ci[e]];
        fit += t;
      }
      const double a = A[r * ndp + col];
      num += (a - fit) * (a - fit);
      den += a * a;
    }
  return sqrt(num / den);
}

/* ---------------- Schwarz preconditioner (A1, A2, A12) ------------------- */

/* A_j^{-1} r: one fast-diagonalization inverse per component block
 * (BlockMatrixDiagonal::apply_inverse, block_matrix.h:1087-1130). */
void or_local_apply(const or_ctx *c, int patch, const double *r, double *u)
{
  for(int b = 0; b < c->ncomp; ++b)
  {
    const long off = (long)(patch * c->ncomp + b) * c->dim;
    or_fd_apply(c->dim, c->k, c->S + off * c->k * c->k, c->lam + off * c->k, r + b * c->ndp,
                u + b * c->ndp);
  }
}

/* apply_local_solvers, schwarz_preconditioner.templates.h:585-662: gather,
 * local FD inverse, relaxation (:631-633; per-patch omega if set), scatter_add. */
static void apply_local_solvers(const or_ctx *c, double *solution, const double *residual, int color)
{
  const int b = c->color_ptr[color], e = c->color_ptr[color + 1];
  const int ndp = c->ndp * c->ncomp;
  const int parallel = c->n_threads > 1 && c->coloring != OR_COLOR_NONE;
#pragma omp parallel num_threads(parallel ? c->n_threads : 1)
  {
    uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * ndp);
    double *rl = (double *)malloc(sizeof(double) * ndp);
    double *ul = (double *)malloc(sizeof(double) * ndp);
#pragma omp for schedule(static)
    for(int j = b; j < e; ++j)
    {
      patch_dofs(c, c->vtx + j * c->dim, idx);
      for(int i = 0; i < ndp; ++i) rl[i] = residual[idx[i]];
      or_local_apply(c, j, rl, ul);
      const double w = c->omega ? c->omega[j] : c->relaxation;
      if(w != 1.0)
        for(int i = 0; i < ndp; ++i) ul[i] *= w;
      /* weighted RAS (:636-641) or restricted scatter rscatter_add (:646-649,
       * patch_transfer.templates.h:140-158, scatter_add_impl :54-95) */
      if(c->ras == 1)
        for(int i = 0; i < ndp; ++i) ul[i] /= c->ras_w[idx[i]];
      if(c->ras == 2)
      {
        const unsigned char *fl = c->ras_flags + (long)j * ndp;
        for(int i = 0; i < ndp; ++i)
          if(!fl[i]) solution[idx[i]] += ul[i];
      }
      else
        for(int i = 0; i < ndp; ++i) solution[idx[i]] += ul[i];
    }
    free(idx);
    free(rl);
    free(ul);
  }
}

/*
 * Restricted additive Schwarz data.
 * compute_ras_weights (schwarz_preconditioner.templates.h:126-165): every patch
 * scatter_adds a local vector of ones, so w[i] = number of patches containing i.
 * compute_restricted_dofs_impl (dof_info.templates.h:192-313), DGQ layout: for
 * each DoF the pairs (patch, local index i) containing it are ordered by
 * (distance of i from the patch centre, i, patch) and only the first keeps the
 * DoF; the distance is static_cast<unsigned>(max_d |k/2 - 0.5 - m_d|) for the
 * patch multi-index m of i (comp_distance_impl_dgq, :233-244).
 * Additive, scalar (one component) only, as in the reference (:207-215).
 */
int or_set_ras(or_ctx *c, int mode)
{
  if(mode < 0 || mode > 2) return 1;
  if(mode && (c->variant != OR_ADDITIVE || c->ncomp != 1 || !c->ready)) return 1;
  c->ras = mode;
  free(c->ras_w);
  free(c->ras_flags);
  c->ras_w = NULL;
  c->ras_flags = NULL;
  if(!mode) return 0;
  const int ndp = c->ndp, dim = c->dim, k = c->k;
  uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * ndp);
  c->ras_w = (double *)calloc(c->N, sizeof(double));
  c->ras_flags = (unsigned char *)malloc((size_t)c->P * ndp);
  memset(c->ras_flags, 1, (size_t)c->P * ndp);
  /* best (distance, index, patch) per DoF, and its position */
  unsigned *best = (unsigned *)malloc(sizeof(unsigned) * 3 * c->N);
  long *pos = (long *)malloc(sizeof(long) * c->N);
  for(long i = 0; i < c->N; ++i) { best[3 * i] = 0xffffffffu; pos[i] = -1; }
  const double mid = k / 2.0 - 0.5;
  for(int j = 0; j < c->P; ++j)
  {
    patch_dofs(c, c->vtx + j * dim, idx);
    for(int i = 0; i < ndp; ++i)
    {
      const uint32_t g = idx[i];
      c->ras_w[g] += 1.0;
      double dist = 0.0;
      int t = i;
      for(int d = 0; d < dim; ++d)
      {
        const double diff = fabs(mid - (double)(t % k));
        if(diff > dist) dist = diff;
        t /= k;
      }
      const unsigned key[3] = {(unsigned)dist, (unsigned)i, (unsigned)j};
      unsigned *b = best + 3 * (long)g;
      int less = 0;
      for(int q = 0; q < 3; ++q)
        if(key[q] != b[q]) { less = key[q] < b[q]; break; }
      if(less)
      {
        b[0] = key[0]; b[1] = key[1]; b[2] = key[2];
        pos[g] = (long)j * ndp + i;
      }
    }
  }
  for(long g = 0; g < c->N; ++g)
    if(pos[g] >= 0) c->ras_flags[pos[g]] = 0;
  free(best);
  free(pos);
  free(idx);
  return 0;
}

void or_ras_weights(const or_ctx *c, double *out)
{
  for(long i = 0; i < c->N; ++i) out[i] = c->ras_w ? c->ras_w[i] : 0.0;
}

void or_restricted_flags(const or_ctx *c, unsigned char *out)
{
  if(c->ras_flags) memcpy(out, c->ras_flags, (size_t)c->P * c->ndp);
  else memset(out, 0, (size_t)c->P * c->ndp);
}

/* vmult, schwarz_preconditioner.templates.h:209-239 with
 * additive_schwarz_operation (:665-688) and multiplicative_schwarz_operation
 * (:691-732). */
void or_precond_vmult(or_ctx *c, double *dst, const double *src)
{
  memset(dst, 0, sizeof(double) * c->N);
  if(c->variant == OR_ADDITIVE)
  {
    for(int col = 0; col < c->n_colors; ++col) apply_local_solvers(c, dst, src, col);
    return;
  }
  int seq[64], ns = 0;
  or_color_sequence(c, 0, seq, &ns);
  double *res = (double *)malloc(sizeof(double) * c->N);
  apply_local_solvers(c, dst, src, seq[0]);
  for(int s = 1; s < ns; ++s)
  {
    or_operator_vmult(c, res, dst);
    for(long i = 0; i < c->N; ++i) res[i] = src[i] - res[i];
    apply_local_solvers(c, dst, res, seq[s]);
  }
  free(res);
}

/* SmootherRichardson::step_impl, smoother_richardson.h:118-145 */
void or_smoother_step(or_ctx *c, double *dst, const double *src)
{
  double *r = (double *)malloc(sizeof(double) * c->N);
  double *t = (double *)malloc(sizeof(double) * c->N);
  for(int s = 0; s < c->n_steps; ++s)
  {
    or_operator_vmult(c, r, dst);
    for(long i = 0; i < c->N; ++i) r[i] = src[i] - r[i];
    or_precond_vmult(c, t, r);
    for(long i = 0; i < c->N; ++i) dst[i] += c->rich_damping * t[i];
  }
  free(r);
  free(t);
}

/* SmootherRichardson::vmult_impl, smoother_richardson.h:147-156 */
void or_smoother_vmult(or_ctx *c, double *dst, const double *src)
{
  memset(dst, 0, sizeof(double) * c->N);
  or_smoother_step(c, dst, src);
}

/* ========================================================================= */
/* Linear elasticity SIPG operator (Cartesian)                               */
/* ========================================================================= */

/* out = (A_{dim-1} x .. x A_0)^T in (transposed 1D factors). */
static void tensor_apply_t(int dim, int m, const double *const *mats, const double *in, double *out)
{
  double a[OR_MAXNP * OR_MAXNP * OR_MAXNP], b[OR_MAXNP * OR_MAXNP * OR_MAXNP];
  const int n = ipow(m, dim);
  memcpy(a, in, sizeof(double) * n);
  for(int e = 0; e < dim; ++e)
  {
    mode_product(dim, m, e, mats[e], 1, a, b);
    memcpy(a, b, sizeof(double) * n);
  }
  memcpy(out, a, sizeof(double) * n);
}

/*
 * y = A x for the vector-valued SIPG discretization of
 *   -div(2 mu eps(u) + lambda div(u) I) = f,  u = 0 on the boundary,
 * assembled in the reference by deal.II's MeshWorker integrators
 * (apps/linelasticity_integrator.h:31-124):
 *   cell      2 mu eps(u):eps(v) + lambda div u div v
 *             (LocalIntegrators::Elasticity::cell_matrix, GradDiv::cell_matrix)
 *   interior  2 mu (sigma [u].[v] - {eps(u) n}.[v] - {eps(v) n}.[u])
 *             + lambda (sigma [u.n][v.n] - {div u}[v.n] - {div v}[u.n])   (ip_matrix)
 *   boundary  the same with a mirrored exterior state u_out = -u_in,
 *             grad u_out = grad u_in, i.e. penalty 2 sigma     (nitsche_matrix)
 * sigma = ip * gamma/h (Laplace::compute_penalty of equal cells), gamma = p(p+1).
 * Written per cell and quadrature point with the full gradient tensor; every
 * cell evaluates the traces of its face neighbours itself.
 */
static void operator_elasticity(const or_ctx *c, double *dst, const double *src)
{
  const int dim = c->dim, np = c->np, n = c->n, ndc = c->ndc, p = c->p;
  const long Ns = c->ncell * ndc;
  const double h = 1.0 / n, mu = c->mu, lam = c->lambda;
  double nodes[OR_MAXNP], qx[OR_MAXNP], qw[OR_MAXNP];
  double val[OR_MAXNP * OR_MAXNP], grad[OR_MAXNP * OR_MAXNP], fv[OR_MAXNP * 2], fg[OR_MAXNP * 2];
  or_shape_info(p, nodes, qx, qw, val, grad, fv, fg);
  /* square np x np factors, row = quadrature point, column = node */
  double V[OR_MAXNP * OR_MAXNP], G[OR_MAXNP * OR_MAXNP];
  double FV[2][OR_MAXNP * OR_MAXNP], FG[2][OR_MAXNP * OR_MAXNP];
  for(int q = 0; q < np; ++q)
    for(int i = 0; i < np; ++i)
    {
      V[q * np + i] = val[i * np + q];
      G[q * np + i] = grad[i * np + q] / h;
      for(int f = 0; f < 2; ++f)
      {
        FV[f][q * np + i] = fv[i * 2 + f]; /* every row: the trace at x = f */
        FG[f][q * np + i] = fg[i * 2 + f] / h;
      }
    }
  const double sig = c->ip * 0.5 * or_penalty(p, h, h);

#pragma omp parallel for num_threads(c->n_threads) schedule(static)
  for(long cell = 0; cell < c->ncell; ++cell)
  {
    enum { MX = OR_MAXNP * OR_MAXNP * OR_MAXNP };
    double u[3][MX], acc[3][MX], gu[3][3][MX], gw[3][3][MX], tmp[MX], t2[MX];
    int cc[3];
    long r = cell;
    for(int d = 0; d < dim; ++d) { cc[d] = (int)(r % n); r /= n; }
    for(int b = 0; b < dim; ++b)
      for(int i = 0; i < ndc; ++i)
      {
        u[b][i] = src[b * Ns + cell * ndc + i];
        acc[b][i] = 0.0;
      }

    /* cell term: stress at the Gauss points times JxW, tested with grad v */
    for(int b = 0; b < dim; ++b)
      for(int j = 0; j < dim; ++j)
      {
        const double *m[3];
        for(int e = 0; e < dim; ++e) m[e] = e == j ? G : V;
        tensor_apply(dim, np, m, u[b], gu[b][j], tmp);
      }
    for(int q = 0; q < ndc; ++q)
    {
      double w = 1.0;
      int t = q;
      for(int e = 0; e < dim; ++e) { w *= qw[t % np] * h; t /= np; }
      double div = 0.0;
      for(int b = 0; b < dim; ++b) div += gu[b][b][q];
      for(int b = 0; b < dim; ++b)
        for(int j = 0; j < dim; ++j)
          gw[b][j][q] = w * (mu * (gu[b][j][q] + gu[j][b][q]) + (b == j ? lam * div : 0.0));
    }
    for(int b = 0; b < dim; ++b)
      for(int j = 0; j < dim; ++j)
      {
        const double *m[3];
        for(int e = 0; e < dim; ++e) m[e] = e == j ? G : V;
        tensor_apply_t(dim, np, m, gw[b][j], tmp);
        for(int i = 0; i < ndc; ++i) acc[b][i] += tmp[i];
      }

    /* faces */
    long stride = 1;
    for(int d = 0; d < dim; ++d)
    {
      for(int f = 0; f < 2; ++f)
      {
        const double nrm = f ? 1.0 : -1.0;
        const int interior = f ? cc[d] < n - 1 : cc[d] > 0;
        const long nb = f ? cell + stride : cell - stride;
        /* traces of this cell (face f) and of the exterior state */
        double U[3][MX], GU[3][3][MX], Uo[3][MX], GUo[3][3][MX];
        for(int b = 0; b < dim; ++b)
        {
          const double *m[3];
          for(int e = 0; e < dim; ++e) m[e] = e == d ? FV[f] : V;
          tensor_apply(dim, np, m, u[b], U[b], tmp);
          for(int j = 0; j < dim; ++j)
          {
            for(int e = 0; e < dim; ++e) m[e] = e == d ? (j == d ? FG[f] : FV[f]) : (e == j ? G : V);
            tensor_apply(dim, np, m, u[b], GU[b][j], tmp);
          }
          if(interior)
          {
            double un[MX];
            for(int i = 0; i < ndc; ++i) un[i] = src[b * Ns + nb * ndc + i];
            for(int e = 0; e < dim; ++e) m[e] = e == d ? FV[1 - f] : V;
            tensor_apply(dim, np, m, un, Uo[b], tmp);
            for(int j = 0; j < dim; ++j)
            {
              for(int e = 0; e < dim; ++e)
                m[e] = e == d ? (j == d ? FG[1 - f] : FV[1 - f]) : (e == j ? G : V);
              tensor_apply(dim, np, m, un, GUo[b][j], tmp);
            }
          }
          else
            for(int i = 0; i < ndc; ++i)
            {
              Uo[b][i] = -U[b][i];
              for(int j = 0; j < dim; ++j) GUo[b][j][i] = GU[b][j][i];
            }
        }
        /* coefficients of v_b (A) and of d_j v_b (B) at the face points */
        double A[3][MX], B[3][3][MX];
        for(int b = 0; b < dim; ++b)
          for(int i = 0; i < ndc; ++i)
          {
            A[b][i] = 0.0;
            for(int j = 0; j < dim; ++j) B[b][j][i] = 0.0;
          }
        for(int q = 0; q < ndc; ++q)
        {
          int t = q, on_slice = 1;
          double w = 1.0;
          for(int e = 0; e < dim; ++e)
          {
            const int qe = t % np;
            t /= np;
            if(e == d) on_slice = qe == 0;
            else w *= qw[qe] * h;
          }
          if(!on_slice) continue;
          double J[3], divi = 0.0, divo = 0.0;
          for(int b = 0; b < dim; ++b)
          {
            J[b] = U[b][q] - Uo[b][q];
            divi += GU[b][b][q];
            divo += GUo[b][b][q];
          }
          for(int b = 0; b < dim; ++b)
          {
            /* {2 mu eps(u) n + lambda div(u) n}_b */
            const double fi = mu * (GU[b][d][q] + GU[d][b][q]) + (b == d ? lam * divi : 0.0);
            const double fo = mu * (GUo[b][d][q] + GUo[d][b][q]) + (b == d ? lam * divo : 0.0);
            const double F = 0.5 * (fi + fo) * nrm;
            A[b][q] = w * (-F + 2.0 * mu * sig * J[b] + (b == d ? lam * sig * J[d] : 0.0));
            /* -1/2 (2 mu eps(v) n . [u] + lambda div v [u].n) */
            B[b][d][q] += w * (-0.5 * mu * nrm * J[b]);
            if(b == d)
              for(int j = 0; j < dim; ++j) B[b][j][q] += w * (-0.5 * mu * nrm * J[j]);
            B[b][b][q] += w * (-0.5 * lam * nrm * J[d]);
          }
        }
        for(int b = 0; b < dim; ++b)
        {
          const double *m[3];
          for(int e = 0; e < dim; ++e) m[e] = e == d ? FV[f] : V;
          tensor_apply_t(dim, np, m, A[b], t2);
          for(int i = 0; i < ndc; ++i) acc[b][i] += t2[i];
          for(int j = 0; j < dim; ++j)
          {
            for(int e = 0; e < dim; ++e) m[e] = e == d ? (j == d ? FG[f] : FV[f]) : (e == j ? G : V);
            tensor_apply_t(dim, np, m, B[b][j], t2);
            for(int i = 0; i < ndc; ++i) acc[b][i] += t2[i];
          }
        }
      }
      stride *= n;
    }
    for(int b = 0; b < dim; ++b)
      for(int i = 0; i < ndc; ++i) dst[b * Ns + cell * ndc + i] = acc[b][i];
  }
}

/* ========================================================================= */
/* SIPG operator on distorted (trilinear) meshes                              */
/* ========================================================================= */

/* Reference coordinates -> physical point and Jacobian of the (bi/tri)linear
 * map of cell cc. J[a*dim+b] = d x_a / d xi_b. */
static void cell_map(const or_ctx *c, const int *cc, const double *xi, double *X, double *J)
{
  const int dim = c->dim;
  for(int a = 0; a < dim; ++a)
  {
    X[a] = 0.0;
    for(int b = 0; b < dim; ++b) J[a * dim + b] = 0.0;
  }
  for(int corner = 0; corner < (1 << dim); ++corner)
  {
    int iv[3];
    double w = 1.0, dw[3];
    for(int e = 0; e < dim; ++e)
    {
      const int bb = (corner >> e) & 1;
      iv[e] = cc[e] + bb;
      w *= bb ? xi[e] : 1.0 - xi[e];
    }
    for(int b = 0; b < dim; ++b)
    {
      double t = 1.0;
      for(int e = 0; e < dim; ++e)
      {
        const int bb = (corner >> e) & 1;
        if(e == b) t *= bb ? 1.0 : -1.0;
        else t *= bb ? xi[e] : 1.0 - xi[e];
      }
      dw[b] = t;
    }
    const double *V = vertex_at(c, iv);
    for(int a = 0; a < dim; ++a)
    {
      X[a] += w * V[a];
      for(int b = 0; b < dim; ++b) J[a * dim + b] += dw[b] * V[a];
    }
  }
}

static double inv_det(int dim, const double *J, double *Ji)
{
  double det;
  if(dim == 1) { det = J[0]; Ji[0] = 1.0 / det; return det; }
  if(dim == 2)
  {
    det = J[0] * J[3] - J[1] * J[2];
    Ji[0] = J[3] / det; Ji[1] = -J[1] / det; Ji[2] = -J[2] / det; Ji[3] = J[0] / det;
    return det;
  }
  det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
        J[2] * (J[3] * J[7] - J[4] * J[6]);
  Ji[0] = (J[4] * J[8] - J[5] * J[7]) / det;
  Ji[1] = (J[2] * J[7] - J[1] * J[8]) / det;
  Ji[2] = (J[1] * J[5] - J[2] * J[4]) / det;
  Ji[3] = (J[5] * J[6] - J[3] * J[8]) / det;
  Ji[4] = (J[0] * J[8] - J[2] * J[6]) / det;
  Ji[5] = (J[2] * J[3] - J[0] * J[5]) / det;
  Ji[6] = (J[3] * J[7] - J[4] * J[6]) / det;
  Ji[7] = (J[1] * J[6] - J[0] * J[7]) / det;
  Ji[8] = (J[0] * J[4] - J[1] * J[3]) / det;
  return det;
}

/* Tensor-product shape values / reference gradients of all cell DoFs at xi. */
static void cell_shape(const or_ctx *c, const double *nodes, const double *xi, double *phi, double *dphi)
{
  const int dim = c->dim, np = c->np;
  double v[3][OR_MAXNP], g[3][OR_MAXNP];
  for(int e = 0; e < dim; ++e) or_lagrange(np, nodes, xi[e], v[e], g[e]);
  for(int i = 0; i < c->ndc; ++i)
  {
    int id[3], t = i;
    for(int e = 0; e < dim; ++e) { id[e] = t % np; t /= np; }
    double val = 1.0;
    for(int e = 0; e < dim; ++e) val *= v[e][id[e]];
    phi[i] = val;
    for(int b = 0; b < dim; ++b)
    {
      double gg = 1.0;
      for(int e = 0; e < dim; ++e) gg *= e == b ? g[e][id[e]] : v[e][id[e]];
      dphi[i * dim + b] = gg;
    }
  }
}

/*
 * Matrix-free SIPG on the trilinear mesh with QGauss(p+1) on cells and faces
 * (apps/laplace_integrator.h:613-743): cell  int grad u . grad v;
 * interior face  sigma[[u]][[v]] - {d_n u}[[v]] - [[u]]{d_n v} with
 * sigma = ip * 0.5 * 2 gamma (1/2 h_in^-1 + 1/2 h_out^-1);
 * Dirichlet face  sigma u v - d_n u v - u d_n v, sigma = ip * 2 gamma / h.
 * h = 1 / |n^T J^{-1} e_normal| of each cell (laplace_integrator.h:651-656).
 */
static void dist_quad_init(const or_ctx *c, dist_quad *g)
{
  g->nq = c->np;
  or_gauss_lobatto(c->np, g->nodes);
  or_gauss_legendre(g->nq, g->qx, g->qw);
  g->nqc = ipow(g->nq, c->dim);
  g->nqf = ipow(g->nq, c->dim - 1);
  g->gamma = c->p == 0 ? 1.0 : (double)c->p * (c->p + 1);
}

static void cell_coords(const or_ctx *c, long cell, int *cc)
{
  long r = cell;
  for(int d = 0; d < c->dim; ++d) { cc[d] = (int)(r % c->n); r /= c->n; }
}

/* cell term  int grad u . grad v  of one cell */
static void cell_term_distorted(const or_ctx *c, const dist_quad *g, long cell, double *dst,
                                const double *src)
{
  const int dim = c->dim, nq = g->nq, ndc = c->ndc;
  int cc[3];
  cell_coords(c, cell, cc);
  double phi[512], dphi[512 * 3], X[3], J[9], Ji[9];
  const double *u = src + cell * ndc;
  double *y = dst + cell * ndc;
  for(int q = 0; q < g->nqc; ++q)
  {
    double xi[3], w = 1.0;
    int t = q;
    for(int e = 0; e < dim; ++e) { xi[e] = g->qx[t % nq]; w *= g->qw[t % nq]; t /= nq; }
    cell_map(c, cc, xi, X, J);
    const double det = inv_det(dim, J, Ji);
    cell_shape(c, g->nodes, xi, phi, dphi);
    double gu[3] = {0, 0, 0};
    for(int i = 0; i < ndc; ++i)
      for(int a = 0; a < dim; ++a)
      {
        double gg = 0.0;
        for(int b = 0; b < dim; ++b) gg += dphi[i * dim + b] * Ji[b * dim + a];
        gu[a] += u[i] * gg;
      }
    const double JxW = fabs(det) * w;
    for(int i = 0; i < ndc; ++i)
    {
      double sacc = 0.0;
      for(int a = 0; a < dim; ++a)
      {
        double gg = 0.0;
        for(int b = 0; b < dim; ++b) gg += dphi[i * dim + b] * Ji[b * dim + a];
        sacc += gg * gu[a];
      }
      y[i] += sacc * JxW;
    }
  }
}

/* face (d, side) of a cell: side 1 = upper face (interior face shared with the
 * upper neighbour, or boundary), side 0 = lower face, boundary only (an
 * interior lower face is the upper face of the lower neighbour) */
static void face_term_distorted(const or_ctx *c, const dist_quad *g, long cell, int d, int side,
                                double *dst, const double *src)
{
  const int dim = c->dim, nq = g->nq, n = c->n, ndc = c->ndc;
  int cc[3];
  cell_coords(c, cell, cc);
  long stride = 1;
  for(int e = 0; e < d; ++e) stride *= n;
  const int interior = side == 1 ? cc[d] < n - 1 : 0;
  const int boundary = side == 0 ? cc[d] == 0 : cc[d] == n - 1;
  if(!interior && !boundary) return;
  const long nb = interior ? cell + stride : -1;
  int cn[3];
  for(int e = 0; e < dim; ++e) cn[e] = cc[e];
  if(interior) cn[d] += 1;
  double phim[512], dphim[512 * 3], phip[512], dphip[512 * 3];
  double Xm[3], Jm[9], Jim[9], Xp[3], Jp[9], Jip[9];
  /* the penalty uses h = 1/|(J^{-1} n)_d| of both cells at the first face
   * quadrature point only (normal_vector(0) * inverse_jacobian(0),
   * apps/laplace_integrator.h:651-656, 692-693) */
  double hm0, hp0 = 1.0;
  {
    double xim[3], xip[3], nv[3], nn = 0.0, sq = 0.0;
    for(int e = 0; e < dim; ++e)
    {
      xim[e] = e == d ? (double)side : g->qx[0];
      xip[e] = e == d ? 0.0 : g->qx[0];
    }
    cell_map(c, cc, xim, Xm, Jm);
    inv_det(dim, Jm, Jim);
    for(int a = 0; a < dim; ++a) { nv[a] = Jim[d * dim + a] * (side == 0 ? -1.0 : 1.0); nn += nv[a] * nv[a]; }
    nn = sqrt(nn);
    for(int a = 0; a < dim; ++a) { nv[a] /= nn; sq += nv[a] * Jim[d * dim + a]; }
    hm0 = 1.0 / fabs(sq);
    if(interior)
    {
      cell_map(c, cn, xip, Xp, Jp);
      inv_det(dim, Jp, Jip);
      sq = 0.0;
      for(int a = 0; a < dim; ++a) sq += nv[a] * Jip[d * dim + a];
      hp0 = 1.0 / fabs(sq);
    }
  }
  for(int q = 0; q < g->nqf; ++q)
  {
    double xim[3], xip[3], w = 1.0;
    int t = q;
    for(int e = 0; e < dim; ++e)
    {
      if(e == d) { xim[e] = (double)side; xip[e] = 0.0; continue; }
      xim[e] = xip[e] = g->qx[t % nq];
      w *= g->qw[t % nq];
      t /= nq;
    }
    cell_map(c, cc, xim, Xm, Jm);
    const double detm = inv_det(dim, Jm, Jim);
    cell_shape(c, g->nodes, xim, phim, dphim);
    /* outward normal of the "-" cell: n ~ J^{-T} e_d * sign; surface
     * element = |det J| * |J^{-T} e_d| */
    double nv[3], nn = 0.0;
    for(int a = 0; a < dim; ++a) { nv[a] = Jim[d * dim + a] * (side == 0 ? -1.0 : 1.0); nn += nv[a] * nv[a]; }
    nn = sqrt(nn);
    for(int a = 0; a < dim; ++a) nv[a] /= nn;
    const double dS = fabs(detm) * nn * w;
    double um = 0.0, dnum = 0.0;
    double gm[512];
    for(int i = 0; i < ndc; ++i)
    {
      double gsum = 0.0;
      for(int a = 0; a < dim; ++a)
      {
        double ga = 0.0;
        for(int b = 0; b < dim; ++b) ga += dphim[i * dim + b] * Jim[b * dim + a];
        gsum += ga * nv[a];
      }
      gm[i] = gsum;
      um += phim[i] * src[cell * ndc + i];
      dnum += gsum * src[cell * ndc + i];
    }
    if(boundary)
    {
      const double sigma = c->ip * 2.0 * g->gamma / hm0;
      for(int i = 0; i < ndc; ++i)
        dst[cell * ndc + i] += (sigma * um * phim[i] - dnum * phim[i] - um * gm[i]) * dS;
      continue;
    }
    cell_map(c, cn, xip, Xp, Jp);
    inv_det(dim, Jp, Jip);
    cell_shape(c, g->nodes, xip, phip, dphip);
    double up = 0.0, dnup = 0.0;
    double gp[512];
    for(int i = 0; i < ndc; ++i)
    {
      double gsum = 0.0;
      for(int a = 0; a < dim; ++a)
      {
        double ga = 0.0;
        for(int b = 0; b < dim; ++b) ga += dphip[i * dim + b] * Jip[b * dim + a];
        gsum += ga * nv[a];
      }
      gp[i] = gsum;
      up += phip[i] * src[nb * ndc + i];
      dnup += gsum * src[nb * ndc + i];
    }
    const double sigma = c->ip * 0.5 * or_penalty(c->p, hm0, hp0);
    const double jump = um - up, avg = 0.5 * (dnum + dnup);
    const double tv = (jump * sigma - avg) * dS, tg = -0.5 * jump * dS;
    for(int i = 0; i < ndc; ++i)
    {
      dst[cell * ndc + i] += tv * phim[i] + tg * gm[i];
      dst[nb * ndc + i] += -tv * phip[i] + tg * gp[i];
    }
  }
}

/*
 * Matrix-free SIPG on the trilinear mesh with QGauss(p+1) on cells and faces
 * (apps/laplace_integrator.h:613-743): cell  int grad u . grad v;
 * interior face  sigma[[u]][[v]] - {d_n u}[[v]] - [[u]]{d_n v} with
 * sigma = ip * 0.5 * 2 gamma (1/2 h_in^-1 + 1/2 h_out^-1);
 * Dirichlet face  sigma u v - d_n u v - u d_n v, sigma = ip * 2 gamma / h.
 * h = 1 / |n^T J^{-1} e_normal| of each cell (laplace_integrator.h:651-656).
 */
static void operator_distorted(const or_ctx *c, double *dst, const double *src)
{
  dist_quad g;
  dist_quad_init(c, &g);
  memset(dst, 0, sizeof(double) * c->N);
#pragma omp parallel for num_threads(c->n_threads) schedule(static)
  for(long cell = 0; cell < c->ncell; ++cell) cell_term_distorted(c, &g, cell, dst, src);
  /* faces: every cell's upper faces plus its lower boundary faces; serial to
   * avoid write conflicts */
  for(long cell = 0; cell < c->ncell; ++cell)
    for(int d = 0; d < c->dim; ++d)
      for(int side = 0; side < 2; ++side) face_term_distorted(c, &g, cell, d, side, dst, src);
}

/* A src for src supported in one cell only: the cell term of that cell and
 * the terms of its faces (the rows of the cell and of its face neighbours) */
static void operator_distorted_one_cell(const or_ctx *c, const dist_quad *g, long cell,
                                        double *dst, const double *src)
{
  cell_term_distorted(c, g, cell, dst, src);
  int cc[3];
  cell_coords(c, cell, cc);
  long stride = 1;
  for(int d = 0; d < c->dim; ++d)
  {
    face_term_distorted(c, g, cell, d, 1, dst, src);
    if(cc[d] == 0)
      face_term_distorted(c, g, cell, d, 0, dst, src);
    else
      face_term_distorted(c, g, cell - stride, d, 1, dst, src);
    stride *= c->n;
  }
}

/* ========================================================================= */
/* Geometric multigrid                                                        */
/* ========================================================================= */

/* Embedding of the coarse Q_p space into the two children (DG level
 * transfer, what deal.II's MGTransferMatrixFree applies for FE_DGQ,
 * apps/poisson_problem.h:705-712): P[(c*np + i)*np + j] = phi_j((x_i + c)/2). */
void or_prolongation_1d(int p, double *P)
{
  const int np = p + 1;
  double nodes[OR_MAXNP], v[OR_MAXNP], g[OR_MAXNP];
  or_gauss_lobatto(np, nodes);
  for(int c = 0; c < 2; ++c)
    for(int i = 0; i < np; ++i)
    {
      or_lagrange(np, nodes, 0.5 * (nodes[i] + c), v, g);
      for(int j = 0; j < np; ++j) P[(c * np + i) * np + j] = v[j];
    }
}

void or_prolongate_add(int dim, int p, int n_coarse, const double *xc, double *xf)
{
  const int np = p + 1, ndc = ipow(np, dim), nf = 2 * n_coarse;
  double P[2 * OR_MAXNP * OR_MAXNP];
  or_prolongation_1d(p, P);
  const long ncc = ipow(n_coarse, dim);
  for(long cc = 0; cc < ncc; ++cc)
  {
    int C[3];
    long r = cc;
    for(int d = 0; d < dim; ++d) { C[d] = (int)(r % n_coarse); r /= n_coarse; }
    for(int child = 0; child < (1 << dim); ++child)
    {
      const double *mats[3];
      long fcell = 0, st = 1;
      for(int d = 0; d < dim; ++d)
      {
        const int b = (child >> d) & 1;
        mats[d] = P + b * np * np;
        fcell += (long)(2 * C[d] + b) * st;
        st *= nf;
      }
      double out[512], tmp[512];
      tensor_apply(dim, np, mats, xc + cc * ndc, out, tmp);
      for(int i = 0; i < ndc; ++i) xf[fcell * ndc + i] += out[i];
    }
  }
}

void or_restrict(int dim, int p, int n_coarse, const double *xf, double *xc)
{
  const int np = p + 1, ndc = ipow(np, dim), nf = 2 * n_coarse;
  double P[2 * OR_MAXNP * OR_MAXNP], PT[2 * OR_MAXNP * OR_MAXNP];
  or_prolongation_1d(p, P);
  for(int c = 0; c < 2; ++c)
    for(int i = 0; i < np; ++i)
      for(int j = 0; j < np; ++j) PT[c * np * np + j * np + i] = P[c * np * np + i * np + j];
  const long ncc = ipow(n_coarse, dim);
  for(long cc = 0; cc < ncc; ++cc)
  {
    int C[3];
    long r = cc;
    for(int d = 0; d < dim; ++d) { C[d] = (int)(r % n_coarse); r /= n_coarse; }
    for(int i = 0; i < ndc; ++i) xc[cc * ndc + i] = 0.0;
    for(int child = 0; child < (1 << dim); ++child)
    {
      const double *mats[3];
      long fcell = 0, st = 1;
      for(int d = 0; d < dim; ++d)
      {
        const int b = (child >> d) & 1;
        mats[d] = PT + b * np * np;
        fcell += (long)(2 * C[d] + b) * st;
        st *= nf;
      }
      double out[512], tmp[512];
      tensor_apply(dim, np, mats, xf + fcell * ndc, out, tmp);
      for(int i = 0; i < ndc; ++i) xc[cc * ndc + i] += out[i];
    }
  }
}

struct or_gmg
{
  int dim, p, n_levels, ncomp;
  double coarse_tol;
  or_ctx **pre, **post;
  double **r, **x, **b, **t;
};

or_gmg *or_gmg_create(int dim, int p, int n_coarse, int n_levels, double ip,
                      int variant, int coloring, double relaxation,
                      int n_steps, double coarse_tol)
{
  return or_gmg_create_v(dim, p, n_coarse, n_levels, ip, variant, coloring, relaxation, n_steps,
                         coarse_tol, NULL);
}

/* level vertices of a uniformly refined hierarchy: every second vertex of the
 * next finer level */
static double *coarsen_vertices(int dim, int nf, const double *vf)
{
  const int nc = nf / 2;
  const long nvc = ipow(nc + 1, dim);
  double *vc = (double *)malloc(sizeof(double) * nvc * dim);
  for(long i = 0; i < nvc; ++i)
  {
    long t = i, fi = 0, st = 1;
    for(int d = 0; d < dim; ++d) { fi += 2 * (t % (nc + 1)) * st; t /= (nc + 1); st *= (nf + 1); }
    for(int d = 0; d < dim; ++d) vc[i * dim + d] = vf[fi * dim + d];
  }
  return vc;
}

static or_gmg *gmg_alloc(int dim, int p, int n_levels, int ncomp, double coarse_tol)
{
  or_gmg *g = (or_gmg *)calloc(1, sizeof(or_gmg));
  g->dim = dim;
  g->p = p;
  g->n_levels = n_levels;
  g->ncomp = ncomp;
  g->coarse_tol = coarse_tol;
  g->pre = (or_ctx **)calloc(n_levels, sizeof(or_ctx *));
  g->post = (or_ctx **)calloc(n_levels, sizeof(or_ctx *));
  g->r = (double **)calloc(n_levels, sizeof(double *));
  g->x = (double **)calloc(n_levels, sizeof(double *));
  g->b = (double **)calloc(n_levels, sizeof(double *));
  g->t = (double **)calloc(n_levels, sizeof(double *));
  return g;
}

static void gmg_level_setup(or_gmg *g, int l)
{
  or_setup(g->pre[l]);
  or_setup(g->post[l]);
  const long N = g->pre[l]->N;
  g->r[l] = (double *)calloc(N, sizeof(double));
  g->x[l] = (double *)calloc(N, sizeof(double));
  g->b[l] = (double *)calloc(N, sizeof(double));
  g->t[l] = (double *)calloc(N, sizeof(double));
}

or_gmg *or_gmg_create_v(int dim, int p, int n_coarse, int n_levels, double ip,
                        int variant, int coloring, double relaxation,
                        int n_steps, double coarse_tol, const double *vertices)
{
  or_gmg *g = gmg_alloc(dim, p, n_levels, 1, coarse_tol);
  double **verts = (double **)calloc(n_levels, sizeof(double *));
  if(vertices)
  {
    const int nf = n_coarse << (n_levels - 1);
    const long nv = ipow(nf + 1, dim);
    verts[n_levels - 1] = (double *)malloc(sizeof(double) * nv * dim);
    memcpy(verts[n_levels - 1], vertices, sizeof(double) * nv * dim);
    for(int l = n_levels - 2; l >= 0; --l) verts[l] = coarsen_vertices(dim, n_coarse << (l + 1), verts[l + 1]);
  }
  for(int l = 0; l < n_levels; ++l)
  {
    const int n = n_coarse << l;
    /* pre-smoother forward colors, post-smoother reversed colors
     * (apps/poisson_standard.cc:122-123, apps/multigrid.h:304-353) */
    g->pre[l] = or_create(dim, p, n, ip, variant, coloring, relaxation, 0, 0, 1.0, n_steps);
    g->post[l] = or_create(dim, p, n, ip, variant, coloring, relaxation, 1, 0, 1.0, n_steps);
    if(verts[l])
    {
      or_set_vertices(g->pre[l], verts[l]);
      or_set_vertices(g->post[l], verts[l]);
      free(verts[l]);
    }
    gmg_level_setup(g, l);
  }
  free(verts);
  return g;
}

/* GMG for linear elasticity: the same hierarchy, level smoothers with the
 * block-diagonal local solvers, transfers applied per component block. */
or_gmg *or_gmg_create_elasticity(int dim, int p, int n_coarse, int n_levels, double ip,
                                 double mu, double lambda, int variant, int coloring,
                                 double relaxation, int n_steps, double coarse_tol)
{
  or_gmg *g = gmg_alloc(dim, p, n_levels, dim, coarse_tol);
  for(int l = 0; l < n_levels; ++l)
  {
    const int n = n_coarse << l;
    g->pre[l] = or_create_elasticity(dim, p, n, ip, mu, lambda, variant, coloring, relaxation, 0,
                                     0, 1.0, n_steps);
    g->post[l] = or_create_elasticity(dim, p, n, ip, mu, lambda, variant, coloring, relaxation, 1,
                                      0, 1.0, n_steps);
    gmg_level_setup(g, l);
  }
  return g;
}

void or_gmg_destroy(or_gmg *g)
{
  if(!g) return;
  for(int l = 0; l < g->n_levels; ++l)
  {
    or_destroy(g->pre[l]);
    or_destroy(g->post[l]);
    free(g->r[l]);
    free(g->x[l]);
    free(g->b[l]);
    free(g->t[l]);
  }
  free(g->pre); free(g->post); free(g->r); free(g->x); free(g->b); free(g->t);
  free(g);
}

or_ctx *or_gmg_level(or_gmg *g, int level) { return g->pre[level]; }
or_ctx *or_gmg_smoother(or_gmg *g, int level, int post) { return post ? g->post[level] : g->pre[level]; }

/* OpenMP threads of every level's operator and patch loops (the CG's vector
 * operations stay serial) */
void or_gmg_set_threads(or_gmg *g, int n_threads)
{
  for(int l = 0; l < g->n_levels; ++l)
  {
    or_set_threads(g->pre[l], n_threads);
    or_set_threads(g->post[l], n_threads);
  }
}

static double dot(long n, const double *a, const double *b)
{
  double s = 0.0;
  for(long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* Unpreconditioned CG to relative accuracy (CoarseGridSolver IterativeAcc,
 * apps/multigrid.h:141-161). */
static int cg_plain(const or_ctx *c, double *x, const double *b, double tol, int maxit)
{
  const long N = c->N;
  double *r = (double *)malloc(sizeof(double) * N), *p = (double *)malloc(sizeof(double) * N);
  double *Ap = (double *)malloc(sizeof(double) * N);
  memset(x, 0, sizeof(double) * N);
  memcpy(r, b, sizeof(double) * N);
  memcpy(p, b, sizeof(double) * N);
  double rr = dot(N, r, r);
  const double stop = tol * tol * rr;
  int it = 0;
  while(rr > stop && it < maxit && rr > 0.0)
  {
    or_operator_vmult(c, Ap, p);
    const double alpha = rr / dot(N, p, Ap);
    for(long i = 0; i < N; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
    const double rr1 = dot(N, r, r);
    for(long i = 0; i < N; ++i) p[i] = r[i] + (rr1 / rr) * p[i];
    rr = rr1;
    ++it;
  }
  free(r); free(p); free(Ap);
  return it;
}

static void vcycle_level(or_gmg *g, int l, double *x, const double *b)
{
  if(l == 0)
  {
    cg_plain(g->pre[0], x, b, g->coarse_tol, 100000);
    return;
  }
  or_ctx *c = g->pre[l];
  const long N = c->N;
  or_smoother_vmult(g->pre[l], x, b);
  or_operator_vmult(c, g->r[l], x);
  for(long i = 0; i < N; ++i) g->r[l][i] = b[i] - g->r[l][i];
  const int nc = g->pre[l - 1]->n;
  const long Nsf = N / g->ncomp, Nsc = g->pre[l - 1]->N / g->ncomp;
  for(int b = 0; b < g->ncomp; ++b)
    or_restrict(g->dim, g->p, nc, g->r[l] + b * Nsf, g->b[l - 1] + b * Nsc);
  vcycle_level(g, l - 1, g->x[l - 1], g->b[l - 1]);
  for(int b = 0; b < g->ncomp; ++b)
    or_prolongate_add(g->dim, g->p, nc, g->x[l - 1] + b * Nsc, x + b * Nsf);
  or_smoother_step(g->post[l], x, b);
}

void or_gmg_vcycle(or_gmg *g, double *dst, const double *src)
{
  vcycle_level(g, g->n_levels - 1, dst, src);
}

/* Preconditioned CG (deal.II SolverCG as called at apps/poisson_problem.h:726-755),
 * x0 = 0, stop at ||r|| <= rel_tol ||b||.  Returns iterations (-1 if not converged). */
int or_gmg_cg(or_gmg *g, double *x, const double *b, double rel_tol, int max_it, double *res_hist)
{
  or_ctx *c = g->pre[g->n_levels - 1];
  const long N = c->N;
  double *r = (double *)malloc(sizeof(double) * N), *z = (double *)malloc(sizeof(double) * N);
  double *p = (double *)malloc(sizeof(double) * N), *Ap = (double *)malloc(sizeof(double) * N);
  memset(x, 0, sizeof(double) * N);
  memcpy(r, b, sizeof(double) * N);
  const double r0 = sqrt(dot(N, r, r));
  if(res_hist) res_hist[0] = r0;
  or_gmg_vcycle(g, z, r);
  memcpy(p, z, sizeof(double) * N);
  double rz = dot(N, r, z);
  int it = 0, conv = r0 == 0.0;
  while(!conv && it < max_it)
  {
    or_operator_vmult(c, Ap, p);
    const double alpha = rz / dot(N, p, Ap);
    for(long i = 0; i < N; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
    ++it;
    const double rn = sqrt(dot(N, r, r));
    if(res_hist) res_hist[it] = rn;
    if(rn <= rel_tol * r0) { conv = 1; break; }
    or_gmg_vcycle(g, z, r);
    const double rz1 = dot(N, r, z);
    for(long i = 0; i < N; ++i) p[i] = z[i] + (rz1 / rz) * p[i];
    rz = rz1;
  }
  free(r); free(z); free(p); free(Ap);
  return conv ? it : -1;
}

/* ========================================================================= */
/* Kronecker SVD (A10)                                                        */
/* ========================================================================= */

/*
 * Nearest Kronecker-rank-r approximation A ~ sum_i B1_i (x) B0_i of an
 * (m1 m0) x (n1 n0) matrix (Van Loan & Pitsianis), the problem solved by
 * compute_ksvd (kroneckersvd.h:215-296): SVD of the rearrangement
 * R(A)[(i1*n1+j1), (i0*n0+j0)] = A[i1*m0+i0, j1*n0+j0] (reshuffle_full,
 * kroneckersvd.h:92-117), factors scaled by sqrt(sigma) (:286-295).
 * The SVD is computed here by one-sided Jacobi (Hestenes) in FP64.
 * B1: rank * m1*n1, B0: rank * m0*n0, sigma: rank.  Returns 0 on success.
 */
int or_ksvd(int m1, int n1, int m0, int n0, const double *A, int rank,
            double *B1, double *B0, double *sigma)
{
  const int R = m1 * n1, C = m0 * n0;
  const int ncol = A ? C : 0;
  if(ncol == 0) return 1;
  /* work on W = R(A) (R x C), rotate columns (C of them) */
  double *W = (double *)malloc(sizeof(double) * R * C);
  double *V = (double *)calloc((size_t)C * C, sizeof(double));
  for(int i1 = 0; i1 < m1; ++i1)
    for(int j1 = 0; j1 < n1; ++j1)
      for(int i0 = 0; i0 < m0; ++i0)
        for(int j0 = 0; j0 < n0; ++j0)
          W[(i1 * n1 + j1) * C + (i0 * n0 + j0)] = A[(i1 * m0 + i0) * (n1 * n0) + j1 * n0 + j0];
  for(int i = 0; i < C; ++i) V[i * C + i] = 1.0;
  for(int sweep = 0; sweep < 60; ++sweep)
  {
    double off = 0.0;
    for(int p = 0; p < C - 1; ++p)
      for(int q = p + 1; q < C; ++q)
      {
        double a = 0, b = 0, g = 0;
        for(int i = 0; i < R; ++i)
        {
          a += W[i * C + p] * W[i * C + p];
          b += W[i * C + q] * W[i * C + q];
          g += W[i * C + p] * W[i * C + q];
        }
        if(fabs(g) <= 1e-300 || fabs(g) <= 1e-15 * sqrt(a * b)) continue;
        off = fmax(off, fabs(g) / sqrt(a * b));
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        for(int i = 0; i < R; ++i)
        {
          const double wp = W[i * C + p], wq = W[i * C + q];
          W[i * C + p] = cs * wp - sn * wq;
          W[i * C + q] = sn * wp + cs * wq;
        }
        for(int i = 0; i < C; ++i)
        {
          const double vp = V[i * C + p], vq = V[i * C + q];
          V[i * C + p] = cs * vp - sn * vq;
          V[i * C + q] = sn * vp + cs * vq;
        }
      }
    if(off < 1e-15) break;
  }
  /* singular values = column norms; pick the largest `rank` */
  double *sv = (double *)malloc(sizeof(double) * C);
  int *ord = (int *)malloc(sizeof(int) * C);
  for(int j = 0; j < C; ++j)
  {
    double s = 0;
    for(int i = 0; i < R; ++i) s += W[i * C + j] * W[i * C + j];
    sv[j] = sqrt(s);
    ord[j] = j;
  }
  for(int i = 1; i < C; ++i)
  {
    int t = ord[i], j = i - 1;
    while(j >= 0 && sv[ord[j]] < sv[t]) { ord[j + 1] = ord[j]; --j; }
    ord[j + 1] = t;
  }
  for(int r = 0; r < rank; ++r)
  {
    const int j = r < C ? ord[r] : ord[C - 1];
    const double s = r < C ? sv[j] : 0.0;
    sigma[r] = s;
    const double sq = sqrt(s);
    /* u = W[:,j]/s, v = V[:,j]; B1 = sqrt(s) u, B0 = sqrt(s) v */
    for(int i = 0; i < R; ++i) B1[r * R + i] = s > 0 ? W[i * C + j] / s * sq : 0.0;
    for(int i = 0; i < C; ++i) B0[r * C + i] = V[i * C + j] * sq;
    /* sign convention: the first non-negligible entry of B0 is positive (the
     * golden values of tests/kroneckersvd.cc:242-250, 275-283 follow it) */
    double bmax = 0.0;
    for(int i = 0; i < C; ++i) bmax = fmax(bmax, fabs(B0[r * C + i]));
    int ifirst = 0;
    while(ifirst < C - 1 && fabs(B0[r * C + ifirst]) <= 1e-12 * bmax) ++ifirst;
    if(B0[r * C + ifirst] < 0)
    {
      for(int i = 0; i < R; ++i) B1[r * R + i] = -B1[r * R + i];
      for(int i = 0; i < C; ++i) B0[r * C + i] = -B0[r * C + i];
    }
  }
  free(W); free(V); free(sv); free(ord);
  return 0;
}

/* KSVD of a diagonal matrix diag(d) of size (m1 m0): compute_ksvd for
 * diagonal input (kroneckersvd.h:412-490); factors are diagonal, returned as
 * their diagonals (B1: rank*m1, B0: rank*m0). */
void or_ksvd_diag(int m1, int m0, const double *d, int rank, double *B1, double *B0, double *sigma)
{
  /* rearranged matrix of a diagonal is R(i1, i0) = d[i1*m0+i0] restricted to
   * the diagonal entries of each factor */
  double *A = (double *)calloc((size_t)m1 * m0 * m1 * m0, sizeof(double));
  const int N = m1 * m0;
  for(int i = 0; i < N; ++i) A[(long)i * N + i] = d[i];
  double *F1 = (double *)malloc(sizeof(double) * rank * m1 * m1);
  double *F0 = (double *)malloc(sizeof(double) * rank * m0 * m0);
  or_ksvd(m1, m1, m0, m0, A, rank, F1, F0, sigma);
  for(int r = 0; r < rank; ++r)
  {
    for(int i = 0; i < m1; ++i) B1[r * m1 + i] = F1[r * m1 * m1 + i * m1 + i];
    for(int i = 0; i < m0; ++i) B0[r * m0 + i] = F0[r * m0 * m0 + i * m0 + i];
  }
  free(A); free(F1); free(F0);
}
