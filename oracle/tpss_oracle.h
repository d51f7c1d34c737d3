/*
 * tpss_oracle.h -- CPU FP64 restatement of the TPSS vertex-patch Schwarz smoother.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only the test-suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  Nothing in the shipped library links or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates
 * (paths relative to the jwitte08/TPSS source tree).
 *
 * Conventions (shared with the HIP library):
 *   - mesh: n^dim cells on [0,1]^dim, cell index c = sum_d c_d n^d (x fastest);
 *     an optional vertex array moves interior vertices (distorted meshes).
 *   - DG Q_p, Lagrange basis on Gauss-Lobatto nodes, DoFs lexicographic inside
 *     a cell, global DoF = cell * (p+1)^dim + local.
 *   - vertex patch = the 2^dim cells around an interior vertex v (1<=v_d<=n-1),
 *     patch-local DoF index lexicographic with x fastest, k = 2(p+1) per dir.
 */
#ifndef TPSS_ORACLE_H
#define TPSS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_ADDITIVE = 0, OR_MULTIPLICATIVE = 1 };
enum { OR_COLOR_NONE = 0, OR_COLOR_TILED = 1, OR_COLOR_REDBLACK = 2 };
enum { OR_LOCAL_SURROGATE = 0, OR_LOCAL_FIT = 1 };

/* ---- 1D building blocks ------------------------------------------------ */
void or_gauss_legendre(int nq, double *x, double *w);
void or_gauss_lobatto(int np, double *x);
void or_lagrange(int np, const double *nodes, double x, double *val, double *der);
void or_shape_info(int p, double *nodes, double *qx, double *qw, double *val,
                   double *grad, double *fval, double *fgrad);

double or_penalty(int p, double h_left, double h_right);
void or_patch_matrices_1d(int p, const double *h, int bdry_left, int bdry_right,
                          double ip, double *M, double *L);
void or_patch_matrices_1d_coef(int p, const double *h, int bdry_left, int bdry_right,
                               double ip, double a_cons, double a_pen, double *M, double *L);
int or_gen_eig(int k, const double *A, const double *B, double *S, double *lam);
void or_fd_apply(int dim, int k, const double *S, const double *lam,
                 const double *in, double *out);

/* ---- smoother context ---------------------------------------------------- */
typedef struct or_ctx or_ctx;

or_ctx *or_create(int dim, int p, int n, double ip, int variant, int coloring,
                  double relaxation, int reverse, int symmetrized,
                  double richardson_damping, int n_steps);
/* linear elasticity: dim components, component-blocked vectors, Cartesian */
or_ctx *or_create_elasticity(int dim, int p, int n, double ip, double mu, double lambda,
                             int variant, int coloring, double relaxation, int reverse,
                             int symmetrized, double richardson_damping, int n_steps);
int or_n_components(const or_ctx *c);
/* local solvers: OR_LOCAL_SURROGATE (surrogate patch lengths, the reference's
 * distorted-mesh path) or OR_LOCAL_FIT (separable ALS fit of the exact local
 * matrix, `sweeps` sweeps); call before or_setup */
void or_set_local_solver(or_ctx *c, int kind, int sweeps);
void or_fit_errors(const or_ctx *c, double *out);
/* restricted additive Schwarz: 0 off, 1 weights, 2 boolean (after or_setup) */
int or_set_ras(or_ctx *c, int mode);
void or_ras_weights(const or_ctx *c, double *out);
void or_restricted_flags(const or_ctx *c, unsigned char *out);
int or_fit_fallbacks(const or_ctx *c);
void or_patch_local_matrix(const or_ctx *c, int patch, double *A);
double or_separable_fit(int dim, int k, const double *A, double *M, double *L, int sweeps);
void or_destroy(or_ctx *c);
int or_set_vertices(or_ctx *c, const double *coords);
int or_setup(or_ctx *c);
void or_set_patch_damping(or_ctx *c, const double *omega);
void or_set_threads(or_ctx *c, int n_threads);

long or_n_dofs(const or_ctx *c);
int or_n_patches(const or_ctx *c);
int or_n_colors(const or_ctx *c);
void or_color_ptr(const or_ctx *c, int *out);
void or_patch_vertices(const or_ctx *c, int *out);
void or_patch_dof_indices(const or_ctx *c, uint32_t *out);
void or_patch_h(const or_ctx *c, double *out);
void or_patch_factors(const or_ctx *c, double *S, double *lam);
void or_color_sequence(const or_ctx *c, int transpose, int *out, int *n_out);

void or_operator_vmult(const or_ctx *c, double *dst, const double *src);
void or_local_apply(const or_ctx *c, int patch, const double *r, double *u);
void or_precond_vmult(or_ctx *c, double *dst, const double *src);
void or_smoother_step(or_ctx *c, double *dst, const double *src);
void or_smoother_vmult(or_ctx *c, double *dst, const double *src);

/* ---- geometric multigrid ------------------------------------------------- */
void or_prolongation_1d(int p, double *P);
void or_prolongate_add(int dim, int p, int n_coarse, const double *xc, double *xf);
void or_restrict(int dim, int p, int n_coarse, const double *xf, double *xc);

typedef struct or_gmg or_gmg;
or_gmg *or_gmg_create(int dim, int p, int n_coarse, int n_levels, double ip,
                      int variant, int coloring, double relaxation,
                      int n_steps, double coarse_tol);
or_gmg *or_gmg_create_v(int dim, int p, int n_coarse, int n_levels, double ip,
                        int variant, int coloring, double relaxation,
                        int n_steps, double coarse_tol, const double *vertices);
or_gmg *or_gmg_create_elasticity(int dim, int p, int n_coarse, int n_levels, double ip,
                                 double mu, double lambda, int variant, int coloring,
                                 double relaxation, int n_steps, double coarse_tol);
void or_gmg_destroy(or_gmg *g);
or_ctx *or_gmg_level(or_gmg *g, int level);
or_ctx *or_gmg_smoother(or_gmg *g, int level, int post);
void or_gmg_set_threads(or_gmg *g, int n_threads);
void or_gmg_vcycle(or_gmg *g, double *dst, const double *src);
int or_gmg_cg(or_gmg *g, double *x, const double *b, double rel_tol, int max_it,
              double *res_hist);

/* ---- Kronecker SVD ------------------------------------------------------- */
int or_ksvd(int m1, int n1, int m0, int n0, const double *A, int rank,
            double *B1, double *B0, double *sigma);
void or_ksvd_diag(int m1, int m0, const double *d, int rank, double *B1,
                  double *B0, double *sigma);

#ifdef __cplusplus
}
#endif
#endif
