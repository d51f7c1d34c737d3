/*
 * tpss.h -- C ABI of the MI355X vertex-patch Schwarz smoother library
 * (libtpss_amd.so).
 *
 * Plain C types only: device buffers are passed as `void *` / `const void *`
 * (HIP device pointers whose element type is the level's vector precision),
 * streams as `void *` (a hipStream_t, NULL = the null stream), host arrays as
 * typed pointers.  Every entry point returns a tpss_status; no C++ exception
 * crosses this boundary.  The message of the last failure on the calling
 * thread is available from tpss_last_error().
 *
 * Each entry point names the reference interface it replaces
 * (jwitte08/TPSS, file:line).
 */
#ifndef TPSS_H
#define TPSS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
  TPSS_OK = 0,
  TPSS_E_INVALID_ARG = 1, /* AssertThrow on parameters (e.g. relaxation <= 0,
                             schwarz_preconditioner.templates.h:32) */
  TPSS_E_NONFINITE = 2,   /* AssertIsFinite(dst.l2_norm()), :238 */
  TPSS_E_SINGULAR = 3,    /* zero eigenvalues / indefinite mass,
                             tensor_product_matrix.h:369-371 */
  TPSS_E_HIP = 4,         /* a HIP runtime call failed */
  TPSS_E_NOT_CONVERGED = 5, /* SolverControl::NoConvergence */
  TPSS_E_UNSUPPORTED = 6  /* configuration outside this build (ExcNotImplemented) */
} tpss_status;

typedef enum { TPSS_ADDITIVE = 0, TPSS_MULTIPLICATIVE = 1 } tpss_smoother_variant;
typedef enum { TPSS_COLOR_NONE = 0, TPSS_COLOR_TILED = 1, TPSS_COLOR_REDBLACK = 2 } tpss_coloring;
typedef enum { TPSS_F64 = 0, TPSS_F32 = 1, TPSS_BF16 = 2 } tpss_precision;
/* TPSS_ELASTICITY: SIPG linear elasticity with dim components
 * (LinElasticity::MW::MatrixIntegrator, apps/linelasticity_integrator.h:31-124),
 * component-blocked vectors like the reference's block vectors, Cartesian
 * meshes; the Schwarz smoother uses the block-diagonal local solvers
 * (FD::MatrixIntegrator::assemble_subspace_inverses, :604-647). */
typedef enum { TPSS_POISSON = 0, TPSS_ELASTICITY = 1 } tpss_equation;
/* Local solvers of distorted patches: TPSS_LOCAL_SURROGATE = fast diagonalization
 * of the Cartesian patch with surrogate lengths (MappingInfo::compute_average_scaling,
 * mapping_info.templates.h:42-240, the reference's path); TPSS_LOCAL_FIT = fast
 * diagonalization of the separable Kronecker fit (alternating least squares,
 * rank-2 Kronecker projection per direction) of the exact local matrix
 * R_j A R_j^T, initialised with the surrogate matrices (kroneckersvd.h:215-585). */
typedef enum { TPSS_LOCAL_SURROGATE = 0, TPSS_LOCAL_FIT = 1 } tpss_local_solver;
/* Restricted additive Schwarz.  TPSS_RAS_WEIGHTS: each local solution is
 * divided by the number of patches containing the DoF before scatter_add
 * (compute_ras_weights, schwarz_preconditioner.templates.h:126-165, :636-641).
 * TPSS_RAS_BOOLEAN: restricted prolongation rscatter_add, each DoF taken from
 * one patch only (patch_transfer.templates.h:140-158, restricted flags
 * dof_info.templates.h:192-313). */
typedef enum { TPSS_RAS_NONE = 0, TPSS_RAS_WEIGHTS = 1, TPSS_RAS_BOOLEAN = 2 } tpss_ras;

typedef struct tpss_level tpss_level;
typedef struct tpss_smoother tpss_smoother;
typedef struct tpss_gmg tpss_gmg;

/* One multigrid level: mesh, DG Q_p space, SIPG Poisson operator.
 * Replaces Poisson::ModelProblem mesh/DoF setup (apps/poisson_problem.h:262-306)
 * + Laplace::DG::MF::Operator (apps/laplace_integrator.h:578-743).
 * Limits: at most 2^24 cells (3D: n_cells_1d <= 255) and 2^31 DoFs per level
 * (TPSS_E_INVALID_ARG beyond); with vertices, every cell must have a positive
 * Jacobian determinant at each corner (TPSS_E_SINGULAR otherwise). */
typedef struct
{
  int dim;           /* 2 or 3 */
  int degree;        /* p, 1..4 */
  int n_cells_1d;    /* n: n^dim cells on [0,1]^dim */
  double ip_factor;  /* EquationData::ip_factor (apps/equation_data.h:35-39) */
  const double *vertices; /* host, (n+1)^dim x dim coordinates, x fastest;
                             NULL = uniform Cartesian mesh */
  int vector_precision;   /* TPSS_F64 or TPSS_F32: element type of level vectors */
  int equation;           /* tpss_equation (0 = Poisson) */
  double mu, lambda;      /* Lame coefficients (TPSS_ELASTICITY, EquationData
                             apps/equation_data.h:382-385) */
} tpss_level_desc;

/* Smoother settings: SchwarzPreconditioner::AdditionalData
 * (schwarz_preconditioner.h:495-515) + SmootherRichardson::AdditionalData
 * (smoother_richardson.h:41-48) + the patch-solve arithmetic. */
typedef struct
{
  int variant;           /* tpss_smoother_variant */
  int coloring;          /* tpss_coloring (AVP: NONE = one color like the
                            reference without TBB, patch_info.templates.h:450-459) */
  double relaxation;     /* local damping omega (relaxation), > 0 */
  int reverse;           /* reverse color order */
  int symmetrized;       /* colors 0..C-1..0 */
  double damping_factor; /* Richardson omega_R */
  int n_steps;           /* Richardson steps per step() */
  int local_precision;   /* tpss_precision of the patch solves; TPSS_BF16 (BF16-MFMA
                          * products, FP32 sums): multiplicative, red-black, Cartesian 3D
                          * Q3 Poisson or 3D Q2 elasticity */
  int local_solver;      /* tpss_local_solver (0 = surrogate lengths) */
  int fit_sweeps;        /* ALS sweeps of TPSS_LOCAL_FIT (0 = 4) */
  int ras;               /* restricted additive Schwarz (additive variant, Poisson):
                            0 off, TPSS_RAS_WEIGHTS = use_ras_weights,
                            TPSS_RAS_BOOLEAN = use_ras_boolean_weights
                            (schwarz_preconditioner.h:501-502) */
} tpss_smoother_desc;

typedef struct
{
  double time;       /* seconds accumulated (device time, HIP events) */
  const char *description;
  long n_calls;
} tpss_time_info;

/* ---- library ------------------------------------------------------------ */
const char *tpss_last_error(void);
const char *tpss_version(void);

/* ---- level / operator ---------------------------------------------------- */
int tpss_level_create(const tpss_level_desc *desc, tpss_level **out);
int tpss_level_destroy(tpss_level *level);
int tpss_level_n_dofs(const tpss_level *level, int64_t *n_dofs);
/* dst = A src; Laplace::DG::MF::Operator::vmult (apps/laplace_integrator.h:578-594);
 * dst must not alias src */
int tpss_operator_vmult(tpss_level *level, void *dst, const void *src, void *stream);
/* dst = b - A x (the residual update sadd(-1, 1, rhs) after vmult,
 * schwarz_preconditioner.templates.h:721-725); dst aliases neither b nor x */
int tpss_operator_residual(tpss_level *level, void *dst, const void *b, const void *x,
                           void *stream);

/* ---- smoother / preconditioner ------------------------------------------ */
/* SchwarzPreconditioner::initialize(subdomain_handler, op, data)
 * (schwarz_preconditioner.templates.h:3-60) incl. SubdomainHandler::reinit
 * (subdomain_handler.templates.h:34-154) and compute_inverses (:101-123);
 * local solvers are set up in FP64 on the device. */
int tpss_smoother_create(tpss_level *level, const tpss_smoother_desc *desc, tpss_smoother **out);
/* SchwarzPreconditioner::initialize(other, data): shares the local solvers
 * (schwarz_preconditioner.templates.h:63-86), e.g. post-smoother with reversed colors */
int tpss_smoother_clone_shallow(const tpss_smoother *other, const tpss_smoother_desc *desc,
                                tpss_smoother **out);
int tpss_smoother_destroy(tpss_smoother *sm);
/* SchwarzPreconditioner::vmult (schwarz_preconditioner.templates.h:209-239); dst is overwritten.
 * All apply entry points of a smoother require dst != src and first report
 * TPSS_E_NONFINITE (doing nothing else) if the non-finite check has flagged
 * the output of an earlier call (see tpss_smoother_set_check_finite). */
int tpss_precond_vmult(tpss_smoother *sm, void *dst, const void *src, void *stream);
/* SchwarzPreconditioner::vmult_add (schwarz_preconditioner.templates.h:187-194): dst += P src */
int tpss_precond_vmult_add(tpss_smoother *sm, void *dst, const void *src, void *stream);
/* SchwarzPreconditioner::Tvmult (schwarz_preconditioner.h:187, templates.h:243-268):
 * dst = P^T src.  The reference declares it but throws ("transpose operation is
 * not implemented", templates.h:672, 698) while its get_color_sequence(transpose)
 * reverses the colors (schwarz_preconditioner.h:531-567).  Here: the
 * multiplicative sweep over the reversed color sequence of vmult (the exact
 * transpose, local solvers being symmetric); the additive operator is symmetric. */
int tpss_precond_Tvmult(tpss_smoother *sm, void *dst, const void *src, void *stream);
/* SchwarzPreconditioner::Tvmult_add (templates.h:197-206): dst += P^T src */
int tpss_precond_Tvmult_add(tpss_smoother *sm, void *dst, const void *src, void *stream);
/* Allocation-free variant of tpss_precond_vmult.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The scratch
 * the preconditioner needs (the additive slot buffer, or the residual of
 * multiplicative sweeps that update it) comes from one file-scope device array
 * of TPSS_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends;
 * smoothers that need more are rejected with TPSS_E_INVALID_ARG.  The call
 * allocates nothing (no hipMalloc / hipFree), records no section timers, and
 * shares the array with every other such call on the GPU: calls that overlap
 * on one GPU race.  tpss_noalloc_init looks up the array's device address for
 * the current device (otherwise the first call on a device does). */
#define TPSS_NOALLOC_SCRATCH_BYTES (64u << 20)
int tpss_noalloc_init(void);
int tpss_precond_vmult_noalloc(tpss_smoother *sm, void *dst, const void *src, void *stream);
/* SmootherRichardson::step (smoother_richardson.h:118-145) */
int tpss_smoother_step(tpss_smoother *sm, void *dst, const void *src, void *stream);
/* SmootherRichardson::vmult (smoother_richardson.h:147-156) */
int tpss_smoother_vmult(tpss_smoother *sm, void *dst, const void *src, void *stream);
/* Per-patch damping omega_j replacing the scalar relaxation at
 * schwarz_preconditioner.templates.h:631-633 (host array of n_patches in the
 * library's patch order; NULL restores the scalar relaxation). */
int tpss_smoother_set_patch_damping(tpss_smoother *sm, const double *omega);
int tpss_smoother_info(const tpss_smoother *sm, int *n_patches, int *n_colors, int *n_dofs_per_patch);
/* PartitionData::get_patch_range (patch_info.h:763-770): host int[n_colors+1] */
int tpss_smoother_color_ptr(const tpss_smoother *sm, int *out);
/* vertex integer coordinates of each patch, host int[n_patches*dim] */
int tpss_smoother_patch_vertices(const tpss_smoother *sm, int *out);
/* DoFInfo::dof_indices_patchwise (dof_info.templates.h:156-182): the gather /
 * scatter index table as the device kernels compute it, host
 * uint32[n_patches * n_dofs_per_patch] */
int tpss_smoother_patch_dof_indices(const tpss_smoother *sm, uint32_t *out);
/* DoFInfo::restricted_dof_flags_patchwise (dof_info.templates.h:192-313) as the
 * device computes them: host uint8[n_patches * n_dofs_per_patch], 1 where the
 * restricted prolongation of the patch skips the DoF (scalar levels) */
int tpss_smoother_restricted_dof_flags(const tpss_smoother *sm, uint8_t *out);
/* surrogate lengths h[patch][direction][cell_no_1d] (mapping_info.h:126-130), host */
int tpss_smoother_patch_h(const tpss_smoother *sm, double *out);
/* relative Frobenius error ||A_j - fit_j|| / ||A_j|| of every patch's separable
 * fit (TPSS_LOCAL_FIT; host double[n_patches], patches whose fitted mass matrix
 * is not s.p.d. keep the surrogate factors and report -1) */
int tpss_smoother_fit_errors(const tpss_smoother *sm, double *out);
/* SchwarzPreconditioner::get_time_data (schwarz_preconditioner.h:87-199):
 * sections "Update residuals", "Apply inverses", "Compute inverses",
 * "Copy Locally Owned", "Step"; *n in: capacity, out: count. */
int tpss_smoother_get_timings(const tpss_smoother *sm, tpss_time_info *out, int *n);
int tpss_smoother_set_timing(tpss_smoother *sm, int enabled);
/* AssertIsFinite(dst.l2_norm()) (schwarz_preconditioner.templates.h:238, debug-only
 * in the reference), on by default: the values written to dst are checked and a
 * flag in pinned host memory is set when one is NaN/Inf -- by the residual-free
 * color passes themselves where they write all of dst (Cartesian multiplicative
 * vmult / Tvmult / step), otherwise by one pass over dst at the end of the call.
 * mode 1 (default): no host synchronisation; the flag is reported as
 * TPSS_E_NONFINITE by the next apply entry point of this smoother or by
 * tpss_smoother_finite_status.  mode 2: the call synchronises its stream and
 * reports at once.  mode 0: off.  Smoothers inside tpss_gmg_solve are not
 * checked per call: the CG's dot products reduce every entry, and a non-finite
 * residual ends the solve with TPSS_E_NONFINITE. */
int tpss_smoother_set_check_finite(tpss_smoother *sm, int mode);
/* synchronises stream, then returns TPSS_E_NONFINITE (and clears the flag) if
 * an output since the last report was non-finite, else TPSS_OK */
int tpss_smoother_finite_status(tpss_smoother *sm, void *stream);
/* Multiplicative variant on Cartesian meshes with red-black colors: each color
 * pass computes b_j - B_j x_out inside the patch kernel instead of a separate
 * residual update (default on; 0 selects the residual-update path that the
 * reference follows, schwarz_preconditioner.templates.h:719-731). */
int tpss_smoother_set_fused_residual(tpss_smoother *sm, int enabled);

/* ---- geometric multigrid + CG (caller of the smoother) ------------------- */
/* ModelProblem::prepare_multigrid + solve (apps/poisson_problem.h:642-755),
 * MGSmootherSchwarz (apps/multigrid.h:201-399), CoarseGridSolver (:137-197). */
typedef struct
{
  int dim, degree;
  int n_cells_coarse;      /* cells per direction on the coarsest level */
  int n_levels;            /* fine level has n_cells_coarse * 2^(n_levels-1) cells */
  double ip_factor;
  const double *vertices;  /* fine-level vertices (host) or NULL; coarser
                              levels take every second vertex */
  int mg_precision;        /* TPSS_F64 or TPSS_F32: V-cycle vector precision */
  tpss_smoother_desc smoother; /* pre-smoother; post-smoother = reversed colors */
  double coarse_tol;       /* relative accuracy of the FP64 coarse CG */
  int equation;            /* tpss_equation (0 = Poisson) */
  double mu, lambda;       /* Lame coefficients (TPSS_ELASTICITY) */
} tpss_gmg_desc;

typedef struct
{
  int iterations;
  double initial_residual, final_residual;
  double reduction_rate;   /* rho = (r_n/r_0)^(1/n), apps/postprocess.h:40-60 */
  double n_frac;           /* log(tol)/log(rho) */
  double seconds;          /* wall time of the solve (device synchronised) */
} tpss_solve_info;

int tpss_gmg_create(const tpss_gmg_desc *desc, tpss_gmg **out);
int tpss_gmg_destroy(tpss_gmg *g);
int tpss_gmg_level(tpss_gmg *g, int level, tpss_level **out);
int tpss_gmg_smoother(tpss_gmg *g, int level, int post, tpss_smoother **out);
/* one V-cycle dst = B src on the fine level (FP64 vectors; dst may equal src) */
int tpss_gmg_vcycle(tpss_gmg *g, void *dst, const void *src, void *stream);
/* GMG-preconditioned CG in FP64 (x, b: FP64 device vectors of the fine level) */
int tpss_gmg_solve(tpss_gmg *g, void *x, const void *b, double rel_tol, int max_iter,
                   void *stream, tpss_solve_info *info, double *residual_history);

/* ---- learned per-patch damping ------------------------------------------- */
/* No counterpart in the reference: the reference applies one scalar
 * relaxation to every patch (schwarz_preconditioner.templates.h:631-633).
 * An MLP maps the singular values of the 2^dim cell Jacobians of a patch to
 * omega_j in [omega_min, omega_max]; BF16 weights, FP32 accumulation. */
typedef struct tpss_mlp tpss_mlp;
int tpss_mlp_create(int n_in, int n_hidden, double omega_min, double omega_max, uint64_t seed,
                    tpss_mlp **out);
int tpss_mlp_destroy(tpss_mlp *m);
int tpss_mlp_n_params(const tpss_mlp *m, size_t *n);
int tpss_mlp_set_params(tpss_mlp *m, const float *host, size_t n);
int tpss_mlp_get_params(const tpss_mlp *m, float *host, size_t n);
/* out[b] = omega(X[b, :]) for device X (batch x n_in) */
int tpss_mlp_forward(tpss_mlp *m, const float *X, float *out, int64_t batch, void *stream);
/* mean-squared loss 1/(2B) sum (omega - target)^2 and its gradient (device grad_out) */
int tpss_mlp_loss_grad(tpss_mlp *m, const float *X, const float *target, int64_t batch,
                       float *grad_out, double *loss, void *stream);
/* Adam over (X, target) on the device, minibatches in a seeded order */
int tpss_mlp_train(tpss_mlp *m, const float *X, const float *target, int64_t n, int epochs,
                   int batch, double lr, uint64_t seed, double *final_loss);
/* per-patch features (device float[n_patches * dim * 2^dim]) */
int tpss_smoother_patch_features(const tpss_smoother *sm, float *dev_out, void *stream);
/* per-patch optimal local damping omega_j = sum_e d_j^T r / sum_e d_j^T A d_j
 * (d_j = R_j^T A_j^{-1} R_j r, r = A e, e random), clamped (device float[n_patches]);
 * FP64 levels; on elasticity levels A_j^{-1} is the block-diagonal local solver and
 * the products run over all component blocks of the patch */
int tpss_smoother_optimal_damping(tpss_smoother *sm, int n_samples, uint64_t seed,
                                  double omega_min, double omega_max, float *dev_out);
/* per-patch largest eigenvalue of M_j^{-1} A_j (M_j: the patch's local solver,
 * A_j = R_j A R_j^T the exact local matrix; 1 for exact local solvers) by
 * n_iter power iterations in the A_j inner product, the patches of one
 * red-black color at a time (device double[n_patches]; FP64 Poisson levels) */
int tpss_smoother_local_lambda_max(tpss_smoother *sm, int n_iter, uint64_t seed, double *dev_lambda);
/* omega_j = MLP(features_j) for every patch (NULL restores the scalar relaxation) */
int tpss_smoother_set_damping_mlp(tpss_smoother *sm, tpss_mlp *m);

/* ---- Kronecker SVD --------------------------------------------------------- */
/* compute_ksvd (kroneckersvd.h:215-296) for a batch of matrices, FP64 on the
 * device: A (device, batch x (m1 m0) x (n1 n0), row-major) ~ sum_{i<rank}
 * B1_i (x) B0_i; B1: batch x rank x m1 x n1, B0: batch x rank x m0 x n0,
 * sigma: batch x rank (device). */
int tpss_ksvd_batched(const double *A, int batch, int m1, int n1, int m0, int n0, int rank,
                      double *B1, double *B0, double *sigma, void *stream);
/* compute_ksvd of a sum of Kronecker products (kroneckersvd.h:306-402):
 * A = sum_{i < in_rank} Abig_i (x) Bsmall_i, Abig: batch x in_rank x m1 x n1,
 * Bsmall: batch x in_rank x m0 x n0 (device, row-major); the Lanczos products use
 * the terms (A is never formed); outputs as tpss_ksvd_batched. in_rank <= 16. */
int tpss_ksvd_sum_batched(const double *Abig, const double *Bsmall, int batch, int in_rank, int m1,
                          int n1, int m0, int n0, int rank, double *B1, double *B0, double *sigma,
                          void *stream);
/* compute_kcp (kroneckersvd.h:501-585): rank-`rank` CP decomposition by
 * alternating least squares of sum_{i < in_rank} C_i (x) B_i (x) A_i, initialised
 * with the first `rank` input terms, `iterations` sweeps (0: rank^2 + 10).
 * in: batch x in_rank x [C_i | B_i | A_i] (row-major, sizes m[k] x n[k] of the
 * k-th matrix C, B, A); out: batch x rank x [C | B | A].  rank <= in_rank <= 8,
 * matrices of at most 144 entries. */
int tpss_kcp_batched(const double *in, int batch, int in_rank, const int *m, const int *n,
                     int rank, int iterations, double *out, void *stream);

/* ---- generic Kronecker-sum matrices --------------------------------------- */
/* Tensors::TensorProductMatrix<order> in the basic state (tensor_product_matrix.h:
 * 136-304): A = sum_{r < rank} A_{r,order-1} (x) ... (x) A_{r,0} (dimension 0
 * rightmost; vector index sum_d i_d prod_{e<d} size_e), A_{r,d} of size
 * m[d] x n[d], order 1..3.  factors (device): batch x rank x order matrices,
 * row-major, d = 0 first.  vmult / Tvmult / vmult_add
 * (tensor_product_matrix.templates.h:123-291): dst = (add ? dst : 0) + A src
 * (transpose: A^T src), src and dst batch x (prod n | prod m) (swapped for
 * transpose); intermediate tensors of at most 4096 entries. */
int tpss_kron_vmult(int order, int rank, const int *m, const int *n, int batch,
                    const double *factors, const double *src, double *dst, int transpose, int add,
                    void *stream);
/* dense matrices (as_table / as_transpose_table): batch x M x N (or N x M) */
int tpss_kron_table(int order, int rank, const int *m, const int *n, int batch,
                    const double *factors, int transpose, double *out, void *stream);
/* InverseTable (matrix_helper.h:21-68): SVD pseudo-inverse of a batch of n x n
 * matrices (device, row-major), singular values s_i < threshold * s_max dropped
 * (LAPACKFullMatrix::compute_inverse_svd), or with kernel_size >= 0 the
 * kernel_size smallest ones (compute_inverse_svd_with_kernel); n <= 256 */
int tpss_kron_pinv(const double *A, int n, int batch, double threshold, int kernel_size,
                   double *out, void *stream);

/* ---- grid transfer (MGTransferMatrixFree for FE_DGQ) --------------------- */
int tpss_transfer_prolongate_add(tpss_level *coarse, tpss_level *fine, void *dst_fine,
                                 const void *src_coarse, void *stream);
int tpss_transfer_restrict(tpss_level *coarse, tpss_level *fine, void *dst_coarse,
                           const void *src_fine, void *stream);

/* ---- device memory (plumbing for hosts without their own HIP allocator) --- */
int tpss_device_count(int *count);
int tpss_device_malloc(void **ptr, size_t bytes);
int tpss_device_free(void *ptr);
int tpss_memcpy_h2d(void *dst, const void *src, size_t bytes);
int tpss_memcpy_d2h(void *dst, const void *src, size_t bytes);
int tpss_device_synchronize(void);

/* ---- vectors -------------------------------------------------------------- */
/* x[i] = U[-1,1] keyed by (seed, i) with SplitMix64 (no reference counterpart:
 * the reference fills with rand(), generic_functionalities.h:245-278) */
int tpss_vector_fill_random(void *x, int64_t n, int precision, uint64_t seed, void *stream);
int tpss_vector_convert(void *dst, int dst_precision, const void *src, int src_precision,
                        int64_t n, void *stream);

#ifdef __cplusplus
}
#endif
#endif
