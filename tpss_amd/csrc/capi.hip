// capi.hip -- the C ABI of libtpss_amd: levels (SIPG operator), Schwarz
// smoothers, grid transfer, vectors, geometric multigrid + CG.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"
#include "setup_kernels.hpp"
#include "ops.hpp"
#include "kernels_general.hpp"
#include "kernels_elasticity.hpp"

namespace tpss
{
const char *last_error_cstr();


// ---------------------------------------------------------------------------- levels

template<typename T, int NP>
OpTab<T, NP> make_optab(const Operator1D &op)
{
  OpTab<T, NP> t;
  for(int e = 0; e < NP * NP; ++e)
  {
    t.M[e] = T(op.M[e]);
    t.Fpm[e] = T(op.Fpm[e]);
    t.Fmp[e] = T(op.Fmp[e]);
    for(int c = 0; c < 4; ++c) t.D[c][e] = T(op.D[c][e]);
  }
  return t;
}

template<int DIM, int NP>
constexpr int cells_per_block()
{
  constexpr int ndc = cpow(NP, DIM);
  return ndc >= 256 ? 1 : 256 / ndc;
}

// y = A x (or y = b - A x), optionally only on a list of cells; vectors of type VT
template<int NP, typename T>
ElastTab<NP, T> make_elast_tab(const tpss_level *lv)
{
  ElastTab<NP, T> tab;
  const Basis1D bs = make_basis(NP - 1);
  const double h = lv->h;
  for(int q = 0; q < NP; ++q)
  {
    tab.w[q] = T(h * bs.qw[q]);
    for(int i = 0; i < NP; ++i)
    {
      tab.V[q * NP + i] = T(bs.val[i * NP + q]);
      tab.G[q * NP + i] = T(bs.grad[i * NP + q] / h);
    }
  }
  for(int f = 0; f < 2; ++f)
    for(int i = 0; i < NP; ++i)
    {
      tab.f[f][i] = T(bs.fval[i * 2 + f]);
      tab.fd[f][i] = T(bs.fgrad[i * 2 + f] / h);
    }
  const int p = NP - 1;
  const double gamma = p == 0 ? 1.0 : double(p) * (p + 1);
  tab.mu = T(lv->mu);
  tab.lambda = T(lv->lambda);
  tab.sigma = T(lv->ip * gamma / h);  // Laplace::compute_penalty of two cells of size h
  return tab;
}

template<typename VT>
void operator_apply(const tpss_level *lv, VT *y, const VT *x, const VT *b, const uint32_t *cells,
                    int64_t n_list, hipStream_t s)
{
  if(lv->ncomp > 1)
  {
    dispatch_dim_np(lv->dim, lv->np, [&](auto D, auto NPc) {
      constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value;
      constexpr int NDC = cpow(NP, DIM);
      const auto tab = make_elast_tab<NP, VT>(lv);
      const int64_t count = cells ? n_list : lv->ncell;
      if(count == 0) return;
      (void)NDC;
      constexpr int CPB = elasticity_cells_per_block<DIM, NP>();
      const unsigned threads = unsigned(elasticity_threads<DIM, NP>());
      const unsigned grid = blocks_for(count, CPB);
      if(b)
        hipLaunchKernelGGL((k_elasticity<DIM, NP, VT, true>), dim3(grid), dim3(threads), 0, s, x,
                           y, b, tab, lv->n, lv->Ns, cells, count);
      else
        hipLaunchKernelGGL((k_elasticity<DIM, NP, VT, false>), dim3(grid), dim3(threads), 0, s, x,
                           y, b, tab, lv->n, lv->Ns, cells, count);
      TPSS_HIP(hipGetLastError());
    });
    return;
  }
  if(lv->distorted)
  {
    dispatch_dim_np(lv->dim, lv->np, [&](auto D, auto NPc) {
      constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value;
      constexpr int NDC = cpow(NP, DIM);
      GeneralTab<NP, VT> tab;
      const Basis1D bs = make_basis(NP - 1);
      for(int q = 0; q < NP; ++q)
      {
        tab.qx[q] = VT(bs.qx[q]);
        tab.qw[q] = VT(bs.qw[q]);
        for(int i = 0; i < NP; ++i)
        {
          tab.V[q * NP + i] = VT(bs.val[i * NP + q]);
          tab.Dv[q * NP + i] = VT(bs.grad[i * NP + q]);
        }
      }
      for(int f = 0; f < 2; ++f)
        for(int i = 0; i < NP; ++i)
        {
          tab.f[f][i] = VT(bs.fval[i * 2 + f]);
          tab.fd[f][i] = VT(bs.fgrad[i * 2 + f]);
        }
      const int64_t count = cells ? n_list : lv->ncell;
      if(count == 0) return;
      const unsigned threads = NDC <= 64 ? 64u : 128u;
      if(b)
        hipLaunchKernelGGL((k_sipg_general<DIM, NP, VT, true>), dim3(unsigned(count)),
                           dim3(threads), 0, s, x, y, b, lv->d_vertices.ptr, tab, lv->n, lv->ip,
                           NP - 1, cells);
      else
        hipLaunchKernelGGL((k_sipg_general<DIM, NP, VT, false>), dim3(unsigned(count)),
                           dim3(threads), 0, s, x, y, b, lv->d_vertices.ptr, tab, lv->n, lv->ip,
                           NP - 1, cells);
      TPSS_HIP(hipGetLastError());
    });
    return;
  }
  dispatch_dim_np(lv->dim, lv->np, [&](auto D, auto NPc) {
    constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value;
    using CT = VT;
    const auto tab = make_optab<CT, NP>(lv->op1d);
    if(cells && n_list == 0) return;
    if(cells)
    {
      // only the listed cells (the cells a multiplicative color reads)
      constexpr int CPB = cells_per_block<DIM, NP>();
      constexpr int NDC = cpow(NP, DIM);
      const unsigned grid = blocks_for(n_list, CPB);
      const unsigned threads = unsigned(((CPB * NDC + 63) / 64) * 64);
      if(b)
        hipLaunchKernelGGL((k_sipg_cartesian<DIM, NP, VT, CT, CPB, true>), dim3(grid),
                           dim3(threads), 0, s, x, y, b, tab, lv->n, lv->ncell, cells, n_list);
      else
        hipLaunchKernelGGL((k_sipg_cartesian<DIM, NP, VT, CT, CPB, false>), dim3(grid),
                           dim3(threads), 0, s, x, y, b, tab, lv->n, lv->ncell, cells, n_list);
      TPSS_HIP(hipGetLastError());
      return;
    }
    if constexpr(DIM == 3 && NP == 4)
    {
      const unsigned grid = blocks_for(lv->ncell, 16);
      if(b)
        hipLaunchKernelGGL((k_sipg_lines_q3<VT, CT, true>), dim3(grid), dim3(256), 0, s, x, y, b,
                           tab, lv->n, lv->ncell);
      else
        hipLaunchKernelGGL((k_sipg_lines_q3<VT, CT, false>), dim3(grid), dim3(256), 0, s, x, y,
                           b, tab, lv->n, lv->ncell);
      TPSS_HIP(hipGetLastError());
      return;
    }
    if constexpr(DIM == 3)
    {
      constexpr int CPB = lines_cells_per_block<NP>();
      const unsigned grid = blocks_for(lv->ncell, CPB);
      if(b)
        hipLaunchKernelGGL((k_sipg_lines<NP, VT, CT, true>), dim3(grid), dim3(256), 0, s, x, y, b,
                           tab, lv->n, lv->ncell);
      else
        hipLaunchKernelGGL((k_sipg_lines<NP, VT, CT, false>), dim3(grid), dim3(256), 0, s, x, y,
                           b, tab, lv->n, lv->ncell);
      TPSS_HIP(hipGetLastError());
      return;
    }
    constexpr int TX = row_tile<DIM, NP>();
    const int64_t rows = DIM == 3 ? int64_t(lv->n) * lv->n : lv->n;
    const unsigned grid = unsigned(rows * ((lv->n + TX - 1) / TX));
    if(b)
      hipLaunchKernelGGL((k_sipg_rows<DIM, NP, VT, CT, true>), dim3(grid), dim3(256), 0, s, x, y, b,
                         tab, lv->n);
    else
      hipLaunchKernelGGL((k_sipg_rows<DIM, NP, VT, CT, false>), dim3(grid), dim3(256), 0, s, x, y,
                         b, tab, lv->n);
    TPSS_HIP(hipGetLastError());
  });
}

void operator_apply_any(const tpss_level *lv, void *y, const void *x, const void *b,
                        const uint32_t *cells, int64_t n_list, hipStream_t s)
{
  if(lv->vec_prec == TPSS_F64)
    operator_apply<double>(lv, (double *)y, (const double *)x, (const double *)b, cells, n_list, s);
  else
    operator_apply<float>(lv, (float *)y, (const float *)x, (const float *)b, cells, n_list, s);
}

void operator_apply_f64(const tpss_level *lv, double *y, const double *x, const uint32_t *cells,
                        int64_t n_list, hipStream_t s)
{
  operator_apply<double>(lv, y, x, nullptr, cells, n_list, s);
}

int operator_apply_dot_f64(const tpss_level *lv, double *y, const double *x, double *partial,
                           int64_t cap, hipStream_t s)
{
  if(lv->vec_prec != TPSS_F64 || lv->ncomp != 1 || lv->distorted || lv->dim != 3 || lv->np != 4)
    return 0;
  const unsigned grid = blocks_for(lv->ncell, 16);
  if(int64_t(grid) > cap) return 0;
  const auto tab = make_optab<double, 4>(lv->op1d);
  hipLaunchKernelGGL((k_sipg_lines_q3<double, double, false, true>), dim3(grid), dim3(256), 0, s,
                     x, y, nullptr, tab, lv->n, lv->ncell, partial);
  TPSS_HIP(hipGetLastError());
  return int(grid);
}

size_t elem_size(int prec) { return prec == TPSS_F64 ? 8 : 4; }

template<typename T>
void fill(T *x, int64_t n, T v, hipStream_t s)
{
  if(n == 0) return;
  hipLaunchKernelGGL((k_fill<T>), dim3(blocks_for(n, 256)), dim3(256), 0, s, x, n, v);
  TPSS_HIP(hipGetLastError());
}

void fill_zero(void *x, int prec, int64_t n, hipStream_t s)
{
  if(prec == TPSS_F64)
    fill<double>((double *)x, n, 0.0, s);
  else
    fill<float>((float *)x, n, 0.0f, s);
}

// y = a x + b y
void axpby(void *y, const void *x, int prec, int64_t n, double a, double b, hipStream_t s)
{
  if(n == 0) return;
  if(prec == TPSS_F64)
    hipLaunchKernelGGL((k_axpby<double>), dim3(blocks_for(n, 256)), dim3(256), 0, s, (double *)y,
                       (const double *)x, n, a, b);
  else
    hipLaunchKernelGGL((k_axpby<float>), dim3(blocks_for(n, 256)), dim3(256), 0, s, (float *)y,
                       (const float *)x, n, a, b);
  TPSS_HIP(hipGetLastError());
}

void convert(void *dst, int dp, const void *src, int sp, int64_t n, hipStream_t s)
{
  if(n == 0) return;
  const unsigned g = blocks_for(n, 256);
  if(dp == TPSS_F64 && sp == TPSS_F64)
    TPSS_HIP(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToDevice, s));
  else if(dp == TPSS_F32 && sp == TPSS_F32)
    TPSS_HIP(hipMemcpyAsync(dst, src, n * 4, hipMemcpyDeviceToDevice, s));
  else if(dp == TPSS_F64 && sp == TPSS_F32)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(256), 0, s, (double *)dst,
                       (const float *)src, n);
  else if(dp == TPSS_F32 && sp == TPSS_F64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(256), 0, s, (float *)dst,
                       (const double *)src, n);
  else
    throw Error(TPSS_E_INVALID_ARG, "unsupported vector precision");
  TPSS_HIP(hipGetLastError());
}

// Rejects collapsed or inverted cells of a moved-vertex mesh: the Jacobian
// determinant of the multilinear map must be positive at every corner of
// every cell (host, before anything is allocated on the device).
void check_cells(int dim, int n, const double *V)
{
  const int64_t ncell = ipow(n, dim);
  for(int64_t c = 0; c < ncell; ++c)
  {
    int cc[3] = {0, 0, 0};
    int64_t r = c;
    for(int d = 0; d < dim; ++d)
    {
      cc[d] = int(r % n);
      r /= n;
    }
    auto vtx = [&](int corner, int a) {
      int64_t idx = 0, st = 1;
      for(int d = 0; d < dim; ++d)
      {
        idx += int64_t(cc[d] + ((corner >> d) & 1)) * st;
        st *= n + 1;
      }
      return V[idx * dim + a];
    };
    for(int corner = 0; corner < (1 << dim); ++corner)
    {
      // J[a][b] = dx_a / dxhat_b at the corner: edge from the corner along b
      double J[3][3];
      for(int b = 0; b < dim; ++b)
      {
        const int other = corner ^ (1 << b);
        const double sgn = ((corner >> b) & 1) ? -1.0 : 1.0;
        for(int a = 0; a < dim; ++a) J[a][b] = sgn * (vtx(other, a) - vtx(corner, a));
      }
      const double det =
        dim == 2 ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                 : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      if(!(det > 0.0))
        throw Error(TPSS_E_SINGULAR, "cell " + std::to_string(c) +
                                       " is collapsed or inverted (Jacobian determinant " +
                                       std::to_string(det) + " at a corner)");
    }
  }
}

}  // namespace tpss

using namespace tpss;

extern "C" {

const char *tpss_last_error(void) { return last_error_cstr(); }
const char *tpss_version(void) { return "tpss_amd 0.1 (gfx950)"; }

int tpss_level_create(const tpss_level_desc *desc, tpss_level **out)
{
  return guarded([&] {
    TPSS_REQUIRE(desc && out, "null argument");
    TPSS_REQUIRE(desc->dim == 2 || desc->dim == 3, "dim must be 2 or 3");
    TPSS_REQUIRE(desc->degree >= 1 && desc->degree <= 4, "degree must be in 1..4");
    TPSS_REQUIRE(desc->n_cells_1d >= 2 && desc->n_cells_1d <= 1000, "n_cells_1d must be in 2..1000");
    TPSS_REQUIRE(desc->ip_factor > 0, "ip_factor must be positive");
    TPSS_REQUIRE(desc->vector_precision == TPSS_F64 || desc->vector_precision == TPSS_F32,
                 "vector precision must be F64 or F32");
    TPSS_REQUIRE(desc->equation == TPSS_POISSON || desc->equation == TPSS_ELASTICITY,
                 "invalid equation");
    const bool elasticity = desc->equation == TPSS_ELASTICITY;
    if(elasticity)
    {
      TPSS_REQUIRE(!desc->vertices, "linear elasticity: Cartesian meshes only (vertices = NULL)");
      TPSS_REQUIRE(desc->mu > 0.0 && desc->lambda >= 0.0, "Lame coefficients: mu > 0, lambda >= 0");
    }
    auto lv = std::make_unique<tpss_level>();
    lv->dim = desc->dim;
    lv->p = desc->degree;
    lv->np = desc->degree + 1;
    lv->n = desc->n_cells_1d;
    lv->ndc = ipow(lv->np, lv->dim);
    {
      // every launch of this library keeps gridDim.x * blockDim.x below 2^32
      // while a level has fewer than 2^32 / 256 cells and patches (one block of
      // at most 256 threads per cell or patch) and fewer than 2^31 DoFs
      // (32-bit DoF and cell indices in the index tables and cell lists)
      const int64_t cells = int64_t(ipow(lv->n, lv->dim));
      const int64_t dofs = cells * lv->ndc * (elasticity ? lv->dim : 1);
      TPSS_REQUIRE(cells * 256 < (int64_t(1) << 32) && dofs < (int64_t(1) << 31),
                   "level too large: at most 2^24 cells (3D: n_cells_1d <= 255) and 2^31 DoFs");
    }
    if(desc->vertices) check_cells(lv->dim, lv->n, desc->vertices);
    lv->ncell = ipow(lv->n, lv->dim);
    lv->Ns = lv->ncell * lv->ndc;
    lv->ncomp = elasticity ? lv->dim : 1;
    lv->N = lv->Ns * lv->ncomp;
    lv->mu = elasticity ? desc->mu : 0.0;
    lv->lambda = elasticity ? desc->lambda : 0.0;
    lv->ip = desc->ip_factor;
    lv->h = 1.0 / lv->n;
    lv->vec_prec = desc->vector_precision;
    if(desc->vertices)
    {
      const size_t nv = size_t(ipow(lv->n + 1, lv->dim)) * lv->dim;
      lv->vertices.assign(desc->vertices, desc->vertices + nv);
      lv->d_vertices.upload(lv->vertices);
      lv->distorted = true;
    }
    lv->op1d = make_operator_1d(lv->p, lv->h, lv->ip);
    *out = lv.release();
  });
}

int tpss_level_destroy(tpss_level *level)
{
  delete level;
  return TPSS_OK;
}

int tpss_level_n_dofs(const tpss_level *level, int64_t *n_dofs)
{
  return guarded([&] {
    TPSS_REQUIRE(level && n_dofs, "null argument");
    *n_dofs = level->N;
  });
}

int tpss_operator_vmult(tpss_level *level, void *dst, const void *src, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(level && dst && src, "null argument");
    // the kernels read face-neighbour cells of src while writing dst
    TPSS_REQUIRE(dst != src, "dst and src must not alias");
    operator_apply_any(level, dst, src, nullptr, nullptr, 0, as_stream(stream));
  });
}

int tpss_operator_residual(tpss_level *level, void *dst, const void *b, const void *x, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(level && dst && b && x, "null argument");
    TPSS_REQUIRE(dst != x && dst != b, "dst must alias neither b nor x");
    operator_apply_any(level, dst, x, b, nullptr, 0, as_stream(stream));
  });
}

int tpss_device_count(int *count)
{
  return guarded([&] {
    TPSS_REQUIRE(count, "null argument");
    TPSS_HIP(hipGetDeviceCount(count));
  });
}

int tpss_device_malloc(void **ptr, size_t bytes)
{
  return guarded([&] {
    TPSS_REQUIRE(ptr, "null argument");
    *ptr = nullptr;
    if(bytes) TPSS_HIP(hipMalloc(ptr, bytes));
  });
}

int tpss_device_free(void *ptr)
{
  return guarded([&] {
    if(ptr) TPSS_HIP(hipFree(ptr));
  });
}

int tpss_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
  return guarded([&] {
    TPSS_REQUIRE((dst && src) || bytes == 0, "null argument");
    if(bytes) TPSS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}

int tpss_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
  return guarded([&] {
    TPSS_REQUIRE((dst && src) || bytes == 0, "null argument");
    if(bytes) TPSS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  });
}

int tpss_device_synchronize(void)
{
  return guarded([&] { TPSS_HIP(hipDeviceSynchronize()); });
}

int tpss_vector_fill_random(void *x, int64_t n, int precision, uint64_t seed, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(x || n == 0, "null argument");
    if(n == 0) return;
    hipStream_t s = as_stream(stream);
    if(precision == TPSS_F64)
      hipLaunchKernelGGL((k_random<double>), dim3(blocks_for(n, 256)), dim3(256), 0, s,
                         (double *)x, n, seed);
    else if(precision == TPSS_F32)
      hipLaunchKernelGGL((k_random<float>), dim3(blocks_for(n, 256)), dim3(256), 0, s,
                         (float *)x, n, seed);
    else
      throw Error(TPSS_E_INVALID_ARG, "precision must be F64 or F32");
    TPSS_HIP(hipGetLastError());
  });
}

int tpss_vector_convert(void *dst, int dst_precision, const void *src, int src_precision,
                        int64_t n, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE((dst && src) || n == 0, "null argument");
    convert(dst, dst_precision, src, src_precision, n, as_stream(stream));
  });
}

int tpss_transfer_prolongate_add(tpss_level *coarse, tpss_level *fine, void *dst_fine,
                                 const void *src_coarse, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(coarse && fine && dst_fine && src_coarse, "null argument");
    TPSS_REQUIRE(fine->n == 2 * coarse->n && fine->dim == coarse->dim && fine->p == coarse->p &&
                   fine->vec_prec == coarse->vec_prec && fine->ncomp == coarse->ncomp,
                 "levels are not a refinement pair");
    hipStream_t s = as_stream(stream);
    dispatch_dim_np(fine->dim, fine->np, [&](auto D, auto NPc) {
      constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value;
      TransferTab<NP> tab;
      double P[2 * MAXNP * MAXNP];
      prolongation_1d(NP - 1, P);
      for(int c = 0; c < 2; ++c)
        for(int e = 0; e < NP * NP; ++e) tab.P[c][e] = P[c * NP * NP + e];
      // one scalar DG transfer per component block (MGTransferMatrixFree per block)
      const unsigned g = unsigned(coarse->ncell);
      const unsigned nt = unsigned(TransferShape<DIM, NP>::THREADS);
      for(int b = 0; b < fine->ncomp; ++b)
      {
        if(fine->vec_prec == TPSS_F64)
          hipLaunchKernelGGL((k_prolongate_add<DIM, NP, double>), dim3(g), dim3(nt), 0, s,
                             (double *)dst_fine + b * fine->Ns,
                             (const double *)src_coarse + b * coarse->Ns, tab, coarse->n);
        else
          hipLaunchKernelGGL((k_prolongate_add<DIM, NP, float>), dim3(g), dim3(nt), 0, s,
                             (float *)dst_fine + b * fine->Ns,
                             (const float *)src_coarse + b * coarse->Ns, tab, coarse->n);
        TPSS_HIP(hipGetLastError());
      }
    });
  });
}

int tpss_transfer_restrict(tpss_level *coarse, tpss_level *fine, void *dst_coarse,
                           const void *src_fine, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(coarse && fine && dst_coarse && src_fine, "null argument");
    TPSS_REQUIRE(fine->n == 2 * coarse->n && fine->dim == coarse->dim && fine->p == coarse->p &&
                   fine->vec_prec == coarse->vec_prec && fine->ncomp == coarse->ncomp,
                 "levels are not a refinement pair");
    hipStream_t s = as_stream(stream);
    dispatch_dim_np(fine->dim, fine->np, [&](auto D, auto NPc) {
      constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value;
      TransferTab<NP> tab;
      double P[2 * MAXNP * MAXNP];
      prolongation_1d(NP - 1, P);
      for(int c = 0; c < 2; ++c)
        for(int e = 0; e < NP * NP; ++e) tab.P[c][e] = P[c * NP * NP + e];
      const unsigned g = unsigned(coarse->ncell);
      const unsigned nt = unsigned(TransferShape<DIM, NP>::THREADS);
      for(int b = 0; b < fine->ncomp; ++b)
      {
        if(fine->vec_prec == TPSS_F64)
          hipLaunchKernelGGL((k_restrict<DIM, NP, double>), dim3(g), dim3(nt), 0, s,
                             (double *)dst_coarse + b * coarse->Ns,
                             (const double *)src_fine + b * fine->Ns, tab, coarse->n);
        else
          hipLaunchKernelGGL((k_restrict<DIM, NP, float>), dim3(g), dim3(nt), 0, s,
                             (float *)dst_coarse + b * coarse->Ns,
                             (const float *)src_fine + b * fine->Ns, tab, coarse->n);
        TPSS_HIP(hipGetLastError());
      }
    });
  });
}

}  // extern "C"
