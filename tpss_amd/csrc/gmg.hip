// gmg.hip -- geometric multigrid V-cycle and the FP64 preconditioned CG around
// it (the callers of the Schwarz smoother): ModelProblem::prepare_multigrid /
// solve (apps/poisson_problem.h:642-755), MGSmootherSchwarz
// (apps/multigrid.h:201-399), CoarseGridSolver (apps/multigrid.h:137-197).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <functional>
#include <memory>
#include <utility>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "ops.hpp"
#include "smoother.hpp"

struct tpss_gmg
{
  tpss_gmg_desc desc{};
  int L = 0;
  int mg_prec = TPSS_F64;
  std::vector<tpss_level *> levels;        // vectors in mg precision
  std::vector<tpss_smoother *> pre, post;  // index 0 unused (coarse level)
  tpss_level *fine64 = nullptr;            // FP64 fine level of the outer CG
  tpss_level *coarse64 = nullptr;          // FP64 coarse level of the coarse solve
  bool own_fine64 = false, own_coarse64 = false;
  std::vector<tpss::DeviceArray<char>> vx, vb, vr;  // per level
  // coarse solve
  int Nc = 0;
  bool dense_coarse = false;
  tpss::DeviceArray<double> coarse_inv, cb64, cx64;
  // CG (FP64)
  tpss::DeviceArray<double> r, z, p, Ap, scal, cr, cz, cp, cAp;
  tpss::DotWorkspace dots;
  // per-block parts of p.Ap formed by the operator pass of the outer CG (FP64
  // 3D Q3 Cartesian fine levels; empty elsewhere)
  tpss::DeviceArray<double> op_partial;
  std::vector<double> fine_vertices;
};

namespace tpss
{
namespace
{
// x_c from x_f: every second vertex (the coarse mesh of a uniform refinement)
std::vector<double> coarsen_vertices(const std::vector<double> &vf, int dim, int nf)
{
  const int nc = nf / 2;
  std::vector<double> vc(size_t(ipow(nc + 1, dim)) * dim);
  for(int64_t i = 0; i < ipow(nc + 1, dim); ++i)
  {
    int64_t t = i, fi = 0, st = 1;
    for(int d = 0; d < dim; ++d)
    {
      fi += 2 * (t % (nc + 1)) * st;
      t /= (nc + 1);
      st *= (nf + 1);
    }
    for(int d = 0; d < dim; ++d) vc[size_t(i) * dim + d] = vf[size_t(fi) * dim + d];
  }
  return vc;
}

tpss_level *make_level(const tpss_gmg_desc &d, int n, const std::vector<double> *verts, int prec)
{
  tpss_level_desc ld{};
  ld.dim = d.dim;
  ld.degree = d.degree;
  ld.n_cells_1d = n;
  ld.ip_factor = d.ip_factor;
  ld.vertices = verts ? verts->data() : nullptr;
  ld.vector_precision = prec;
  ld.equation = d.equation;
  ld.mu = d.mu;
  ld.lambda = d.lambda;
  tpss_level *lv = nullptr;
  const int st = tpss_level_create(&ld, &lv);
  if(st != TPSS_OK) throw Error(st, std::string("level: ") + tpss_last_error());
  return lv;
}

void dot64(tpss_gmg *g, const double *a, const double *b, int64_t n, double *out, hipStream_t s)
{
  g->dots.dot<double>(a, b, n, out, s);
}

double read_scalar(const double *dptr, hipStream_t s)
{
  double h = 0.0;
  TPSS_HIP(hipMemcpyAsync(&h, dptr, sizeof(double), hipMemcpyDeviceToHost, s));
  TPSS_HIP(hipStreamSynchronize(s));
  return h;
}

// Preconditioned CG in FP64 (deal.II SolverCG semantics, x0 = 0,
// stop when ||r_k|| <= tol ||r_0||).  Scalars stay on the device (sc: 5
// values of this solve's own, the coarse CG runs inside the outer CG's
// preconditioner); one host synchronisation per iteration for the
// convergence test.  precond(z, r, rz, r_low_ready) sets z = B r and the
// device scalar rz = r.z.  The update x += alpha p, r -= alpha Ap forms r.r in
// the same pass and, if r_low is given, writes the updated r in FP32 there
// (the preconditioner is then told that its FP32 input is ready).
using Precond = std::function<void(double *, const double *, double *, bool)>;

int pcg(tpss_gmg *g, tpss_level *A, double *x, const double *b, double tol, int maxit,
        const Precond &precond, hipStream_t s, std::vector<double> *hist, double *r, double *z,
        double *p, double *Ap, double *sc, float *r_low)
{
  const int64_t N = A->N;
  // sc: [rz, pAp, rr, rz'] -- r.z of the current and of the next iteration in
  // sc[0] and sc[3] by turns (rz / rz_next swap on the host, no copies)
  fill_zero(x, TPSS_F64, N, s);
  TPSS_HIP(hipMemcpyAsync(r, b, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  dot64(g, r, r, N, sc + 2, s);
  const double r0 = std::sqrt(read_scalar(sc + 2, s));
  if(hist) hist->assign(1, r0);
  if(r0 == 0.0) return 0;
  auto apply_precond = [&](double *rz, bool r_low_ready) {
    if(precond)
      precond(z, r, rz, r_low_ready);
    else
    {
      TPSS_HIP(hipMemcpyAsync(z, r, N * sizeof(double), hipMemcpyDeviceToDevice, s));
      dot64(g, r, z, N, rz, s);
    }
  };
  double *rz = sc + 0, *rz_next = sc + 3;
  apply_precond(rz, false);
  TPSS_HIP(hipMemcpyAsync(p, z, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  const unsigned grid = blocks_for(N, 256);
  const int nb = DotWorkspace::blocks(N);
  for(int it = 1; it <= maxit; ++it)
  {
    // Ap and p.Ap in one pass where the level has the fused form
    const int nbo = operator_apply_dot_f64(A, Ap, p, g->op_partial.ptr,
                                           int64_t(g->op_partial.n), s);
    if(nbo > 0)
    {
      hipLaunchKernelGGL(k_dot_final_wide, dim3(1), dim3(1024), 0, s, g->op_partial.ptr, nbo,
                         sc + 1);
      TPSS_HIP(hipGetLastError());
    }
    else
    {
      operator_apply_any(A, Ap, p, nullptr, nullptr, 0, s);
      dot64(g, p, Ap, N, sc + 1, s);
    }
    if(r_low)
      hipLaunchKernelGGL(k_cg_update_dot<true>, dim3(nb), dim3(256), 0, s, x, r, p, Ap, N, rz,
                         sc + 1, g->dots.partial.ptr, r_low);
    else
      hipLaunchKernelGGL(k_cg_update_dot<false>, dim3(nb), dim3(256), 0, s, x, r, p, Ap, N, rz,
                         sc + 1, g->dots.partial.ptr, nullptr);
    TPSS_HIP(hipGetLastError());
    g->dots.finish(nb, sc + 2, s);
    const double rn = std::sqrt(read_scalar(sc + 2, s));
    if(hist) hist->push_back(rn);
    if(!std::isfinite(rn)) throw Error(TPSS_E_NONFINITE, "non-finite residual in CG");
    if(rn <= tol * r0) return it;
    apply_precond(rz_next, r_low != nullptr);
    hipLaunchKernelGGL(k_cg_direction, dim3(grid), dim3(256), 0, s, p, z, N, rz_next, rz);
    TPSS_HIP(hipGetLastError());
    std::swap(rz, rz_next);
  }
  return -1;
}

// Cholesky inverse of an s.p.d. matrix (host, FP64), row-major in place
void spd_inverse(std::vector<double> &A, int n)
{
  std::vector<double> Lm(size_t(n) * n, 0.0);
  for(int j = 0; j < n; ++j)
  {
    double s = A[size_t(j) * n + j];
    for(int k = 0; k < j; ++k) s -= Lm[size_t(j) * n + k] * Lm[size_t(j) * n + k];
    if(!(s > 0.0)) throw Error(TPSS_E_SINGULAR, "coarse matrix is not positive definite");
    const double ljj = std::sqrt(s);
    Lm[size_t(j) * n + j] = ljj;
    for(int i = j + 1; i < n; ++i)
    {
      double t = A[size_t(i) * n + j];
      for(int k = 0; k < j; ++k) t -= Lm[size_t(i) * n + k] * Lm[size_t(j) * n + k];
      Lm[size_t(i) * n + j] = t / ljj;
    }
  }
  // Linv (lower)
  std::vector<double> Li(size_t(n) * n, 0.0);
  for(int j = 0; j < n; ++j)
  {
    Li[size_t(j) * n + j] = 1.0 / Lm[size_t(j) * n + j];
    for(int i = j + 1; i < n; ++i)
    {
      double t = 0.0;
      for(int k = j; k < i; ++k) t -= Lm[size_t(i) * n + k] * Li[size_t(k) * n + j];
      Li[size_t(i) * n + j] = t / Lm[size_t(i) * n + i];
    }
  }
  // A^{-1} = Li^T Li
  for(int i = 0; i < n; ++i)
    for(int j = 0; j <= i; ++j)
    {
      double t = 0.0;
      for(int k = i; k < n; ++k) t += Li[size_t(k) * n + i] * Li[size_t(k) * n + j];
      A[size_t(i) * n + j] = t;
      A[size_t(j) * n + i] = t;
    }
}

void setup_coarse(tpss_gmg *g, hipStream_t s)
{
  tpss_level *c = g->coarse64;
  g->Nc = int(c->N);
  g->cb64.resize(c->N);
  g->cx64.resize(c->N);
  g->dense_coarse = c->N <= 2048;
  if(!g->dense_coarse)
  {
    g->cr.resize(c->N);
    g->cz.resize(c->N);
    g->cp.resize(c->N);
    g->cAp.resize(c->N);
    return;
  }
  const int n = g->Nc;
  std::vector<double> A(size_t(n) * n);
  DeviceArray<double> e(n), col(n);
  for(int j = 0; j < n; ++j)
  {
    fill_zero(e.ptr, TPSS_F64, n, s);
    const double one = 1.0;
    TPSS_HIP(hipMemcpyAsync(e.ptr + j, &one, sizeof(double), hipMemcpyHostToDevice, s));
    operator_apply_any(c, col.ptr, e.ptr, nullptr, nullptr, 0, s);
    TPSS_HIP(hipMemcpyAsync(A.data() + size_t(j) * n, col.ptr, n * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    TPSS_HIP(hipStreamSynchronize(s));
  }
  // A is symmetric; stored column-wise above == row-wise
  spd_inverse(A, n);
  g->coarse_inv.upload(A);
}

void coarse_solve(tpss_gmg *g, double *x, const double *b, hipStream_t s)
{
  if(g->dense_coarse)
  {
    hipLaunchKernelGGL(k_dense_matvec, dim3((g->Nc + 3) / 4), dim3(256), 0, s, g->coarse_inv.ptr, b,
                       x, g->Nc);
    TPSS_HIP(hipGetLastError());
    return;
  }
  const int it = pcg(g, g->coarse64, x, b, g->desc.coarse_tol, 100000, nullptr, s, nullptr,
                     g->cr.ptr, g->cz.ptr, g->cp.ptr, g->cAp.ptr, g->scal.ptr + 8, nullptr);
  if(it < 0) throw Error(TPSS_E_NOT_CONVERGED, "coarse CG did not converge");
}

// one V-cycle x_l = B_l b_l; xo / bo (optional) stand in for the level's own
// x_l / b_l (an FP64 hierarchy runs its finest level on the CG's vectors)
void vcycle(tpss_gmg *g, int l, hipStream_t s, void *xo = nullptr, const void *bo = nullptr)
{
  tpss_level *lv = g->levels[l];
  void *x = xo ? xo : g->vx[l].ptr, *r = g->vr[l].ptr;
  const void *b = bo ? bo : g->vb[l].ptr;
  if(l == 0)
  {
    convert(g->cb64.ptr, TPSS_F64, b, g->mg_prec, lv->N, s);
    coarse_solve(g, g->cx64.ptr, g->cb64.ptr, s);
    convert(x, g->mg_prec, g->cx64.ptr, TPSS_F64, lv->N, s);
    return;
  }
  // pre-smoothing from a zero initial guess: SmootherRichardson::vmult
  fill_zero(x, g->mg_prec, lv->N, s);
  smoother_step_any(g->pre[l], x, b, s, true);
  operator_apply_any(lv, r, x, b, nullptr, 0, s);
  tpss_level *lc = g->levels[l - 1];
  const int st1 = tpss_transfer_restrict(lc, lv, g->vb[l - 1].ptr, r, s);
  if(st1 != TPSS_OK) throw Error(st1, tpss_last_error());
  vcycle(g, l - 1, s);
  const int st2 = tpss_transfer_prolongate_add(lc, lv, x, g->vx[l - 1].ptr, s);
  if(st2 != TPSS_OK) throw Error(st2, tpss_last_error());
  smoother_step_any(g->post[l], x, b, s);
}

// z = B r on the fine level; r, z FP64.  rz (optional device scalar)
// receives r.z, formed in the pass that brings z to FP64; r_low_ready: the
// finest level's right-hand side already holds r in the multigrid precision.
void apply_mg(tpss_gmg *g, double *z, const double *r, hipStream_t s, double *rz = nullptr,
              bool r_low_ready = false)
{
  const int l = g->L - 1;
  const int64_t N = g->levels[l]->N;
  if(g->mg_prec == TPSS_F64 && z != r)
  {
    vcycle(g, l, s, z, r);
    if(rz) dot64(g, r, z, N, rz, s);
    return;
  }
  if(!r_low_ready) convert(g->vb[l].ptr, g->mg_prec, r, TPSS_F64, N, s);
  vcycle(g, l, s);
  if(!rz)
  {
    convert(z, TPSS_F64, g->vx[l].ptr, g->mg_prec, N, s);
    return;
  }
  const int nb = DotWorkspace::blocks(N);
  if(g->mg_prec == TPSS_F32)
    hipLaunchKernelGGL(k_convert_dot<float>, dim3(nb), dim3(256), 0, s, z,
                       (const float *)g->vx[l].ptr, r, N, g->dots.partial.ptr);
  else
    hipLaunchKernelGGL(k_convert_dot<double>, dim3(nb), dim3(256), 0, s, z,
                       (const double *)g->vx[l].ptr, r, N, g->dots.partial.ptr);
  TPSS_HIP(hipGetLastError());
  g->dots.finish(nb, rz, s);
}
}  // namespace
}  // namespace tpss

using namespace tpss;

extern "C" {

int tpss_gmg_create(const tpss_gmg_desc *desc, tpss_gmg **out)
{
  return guarded([&] {
    TPSS_REQUIRE(desc && out, "null argument");
    TPSS_REQUIRE(desc->n_levels >= 1 && desc->n_levels <= 12, "n_levels must be in 1..12");
    TPSS_REQUIRE(desc->n_cells_coarse >= 1, "n_cells_coarse must be >= 1");
    TPSS_REQUIRE(desc->mg_precision == TPSS_F64 || desc->mg_precision == TPSS_F32,
                 "mg precision must be F64 or F32");
    TPSS_REQUIRE(desc->coarse_tol > 0.0, "coarse_tol must be positive");
    auto g = std::make_unique<tpss_gmg>();
    g->desc = *desc;
    g->L = desc->n_levels;
    g->mg_prec = desc->mg_precision;
    const int nf = desc->n_cells_coarse << (desc->n_levels - 1);
    std::vector<std::vector<double>> verts(g->L);
    if(desc->vertices)
    {
      verts[g->L - 1].assign(desc->vertices,
                             desc->vertices + size_t(ipow(nf + 1, desc->dim)) * desc->dim);
      for(int l = g->L - 2; l >= 0; --l)
        verts[l] = coarsen_vertices(verts[l + 1], desc->dim, desc->n_cells_coarse << (l + 1));
    }
    hipStream_t s = nullptr;
    g->levels.resize(g->L, nullptr);
    g->pre.resize(g->L, nullptr);
    g->post.resize(g->L, nullptr);
    g->vx.resize(g->L);
    g->vb.resize(g->L);
    g->vr.resize(g->L);
    tpss_gmg *gp = g.get();
    auto cleanup = [gp] { tpss_gmg_destroy(gp); };
    try
    {
      for(int l = 0; l < g->L; ++l)
      {
        const int n = desc->n_cells_coarse << l;
        g->levels[l] = make_level(*desc, n, desc->vertices ? &verts[l] : nullptr, g->mg_prec);
        const size_t bytes = size_t(g->levels[l]->N) * elem_size(g->mg_prec);
        g->vx[l].resize(bytes);
        g->vb[l].resize(bytes);
        g->vr[l].resize(bytes);
        if(l > 0)
        {
          tpss_smoother_desc pd = desc->smoother;
          pd.reverse = 0;
          int st = tpss_smoother_create(g->levels[l], &pd, &g->pre[l]);
          if(st != TPSS_OK) throw Error(st, tpss_last_error());
          // post-smoother: reversed colors (apps/poisson_standard.cc:122-123), shared solvers
          tpss_smoother_desc qd = desc->smoother;
          qd.reverse = 1;
          st = tpss_smoother_clone_shallow(g->pre[l], &qd, &g->post[l]);
          if(st != TPSS_OK) throw Error(st, tpss_last_error());
        }
      }
      if(g->mg_prec == TPSS_F64)
      {
        g->fine64 = g->levels[g->L - 1];
        g->coarse64 = g->levels[0];
      }
      else
      {
        g->fine64 = make_level(*desc, nf, desc->vertices ? &verts[g->L - 1] : nullptr, TPSS_F64);
        g->own_fine64 = true;
        g->coarse64 = make_level(*desc, desc->n_cells_coarse, desc->vertices ? &verts[0] : nullptr,
                                 TPSS_F64);
        g->own_coarse64 = true;
      }
      setup_coarse(g.get(), s);
      const int64_t N = g->fine64->N;
      g->r.resize(N);
      g->z.resize(N);
      g->p.resize(N);
      g->Ap.resize(N);
      g->scal.resize(16);  // outer CG [0, 4), coarse CG [8, 12)
      {
        const tpss_level *f = g->fine64;
        if(f->ncomp == 1 && !f->distorted && f->dim == 3 && f->np == 4)
          g->op_partial.resize(blocks_for(f->ncell, 16));
      }
    }
    catch(...)
    {
      g.release();
      cleanup();
      throw;
    }
    *out = g.release();
  });
}

int tpss_gmg_destroy(tpss_gmg *g)
{
  if(!g) return TPSS_OK;
  for(auto *s : g->pre) tpss_smoother_destroy(s);
  for(auto *s : g->post) tpss_smoother_destroy(s);
  if(g->own_fine64) tpss_level_destroy(g->fine64);
  if(g->own_coarse64) tpss_level_destroy(g->coarse64);
  for(auto *l : g->levels) tpss_level_destroy(l);
  delete g;
  return TPSS_OK;
}

int tpss_gmg_level(tpss_gmg *g, int level, tpss_level **out)
{
  return guarded([&] {
    TPSS_REQUIRE(g && out && level >= 0 && level < g->L, "invalid level");
    *out = g->levels[level];
  });
}

int tpss_gmg_smoother(tpss_gmg *g, int level, int post, tpss_smoother **out)
{
  return guarded([&] {
    TPSS_REQUIRE(g && out && level >= 1 && level < g->L, "invalid level (level 0 is solved directly)");
    *out = post ? g->post[level] : g->pre[level];
  });
}

int tpss_gmg_vcycle(tpss_gmg *g, void *dst, const void *src, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(g && dst && src, "null argument");
    apply_mg(g, (double *)dst, (const double *)src, as_stream(stream));
  });
}

int tpss_gmg_solve(tpss_gmg *g, void *x, const void *b, double rel_tol, int max_iter, void *stream,
                   tpss_solve_info *info, double *residual_history)
{
  return guarded([&] {
    TPSS_REQUIRE(g && x && b, "null argument");
    TPSS_REQUIRE(rel_tol > 0.0 && max_iter >= 1, "invalid tolerance / iteration limit");
    hipStream_t s = as_stream(stream);
    TPSS_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<double> hist;
    auto pre = [&](double *z, const double *r, double *rz, bool r_low_ready) {
      apply_mg(g, z, r, s, rz, r_low_ready);
    };
    float *r_low = g->mg_prec == TPSS_F32 ? (float *)g->vb[g->L - 1].ptr : nullptr;
    const int it = pcg(g, g->fine64, (double *)x, (const double *)b, rel_tol, max_iter, pre, s,
                       &hist, g->r.ptr, g->z.ptr, g->p.ptr, g->Ap.ptr, g->scal.ptr, r_low);
    TPSS_HIP(hipStreamSynchronize(s));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if(residual_history)
      for(size_t i = 0; i < hist.size(); ++i) residual_history[i] = hist[i];
    if(info)
    {
      const int n = int(hist.size()) - 1;
      info->iterations = it;
      info->initial_residual = hist.front();
      info->final_residual = hist.back();
      info->reduction_rate = n > 0 ? std::pow(hist.back() / hist.front(), 1.0 / n) : 0.0;
      info->n_frac = (n > 0 && info->reduction_rate > 0.0 && info->reduction_rate < 1.0)
                       ? std::log(rel_tol) / std::log(info->reduction_rate)
                       : double(n);
      info->seconds = secs;
    }
    if(it < 0) throw Error(TPSS_E_NOT_CONVERGED, "GMG-CG did not converge");
  });
}

}  // extern "C"
