// smoother.hip -- SchwarzPreconditioner / SchwarzSmoother on the device.
//
// Setup (SubdomainHandler::reinit + compute_inverses): patch gathering and
// coloring on the host (index arithmetic only), surrogate lengths and FP64
// local-solver factors (1D generalized eigenproblems) on the device.
// Apply: fused gather -> fast-diagonalization -> damping -> scatter kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>

#include "common.hpp"
#include "kernels.hpp"
#include "ops.hpp"
#include "setup_kernels.hpp"
#include "kernels_mvp.hpp"
#include "kernels_mvp_split.hpp"
#include "fit_kernels.hpp"
#include "smoother.hpp"

namespace tpss
{
namespace
{
// RedBlack / Tiled vertex-patch colors (apps/coloring.h:286-324, 359-390)
int patch_color(int dim, int coloring, const int *v)
{
  if(coloring == TPSS_COLOR_NONE) return 0;
  int mask = 0, sum = 0;
  for(int d = 0; d < dim; ++d)
  {
    const int x = v[d] - 1, s = x & 1;
    mask |= s << d;
    sum += (x + s) / 2;
  }
  if(coloring == TPSS_COLOR_TILED) return mask;
  return 2 * mask + (sum & 1);
}

bool patch_at_boundary(int dim, int n, const int *v)
{
  for(int d = 0; d < dim; ++d)
    if(v[d] == 1 || v[d] == n - 1) return true;
  return false;
}

int n_colors_of(int dim, int coloring)
{
  return coloring == TPSS_COLOR_NONE ? 1 : coloring == TPSS_COLOR_TILED ? (1 << dim) : (2 << dim);
}

template<int NP>
Shape1D<NP> make_shape(int p)
{
  const Basis1D b = make_basis(p);
  Shape1D<NP> s;
  for(int e = 0; e < NP * NP; ++e)
  {
    s.val[e] = b.val[e];
    s.grad[e] = b.grad[e];
  }
  for(int e = 0; e < NP * 2; ++e)
  {
    s.fval[e] = b.fval[e];
    s.fgrad[e] = b.fgrad[e];
  }
  for(int q = 0; q < NP; ++q) s.qw[q] = b.qw[q];
  return s;
}
}  // namespace

// ---------------------------------------------------------------------------- setup

// Separable fit of the exact local matrices (TPSS_LOCAL_FIT): the local
// matrices of batches of same-red-black-color patches column by column from
// the FP64 operator, alternating least squares from the surrogate matrices,
// FD factors of the fit replacing the surrogate ones where the fit is s.p.d.
// (fitted masses s.p.d., smallest Kronecker-sum eigenvalue positive).  Requires
// one factor row per (patch, direction).
void fit_local_solvers(const tpss_level *lv, LocalSolvers *ls, int sweeps, const double *d_h,
                       const uint8_t *d_b, const double *d_coef, hipStream_t s)
{
  const int dim = ls->dim, K = ls->K, P = ls->P, n = ls->n;
  const int npd = ipow(K, dim), nprob = P * dim;
  DeviceArray<double> Mf{size_t(nprob) * K * K}, Lf{size_t(nprob) * K * K};
  DeviceArray<double> d_err{size_t(P)};
  dispatch_dim_np(dim, lv->np, [&](auto, auto NPc) {
    constexpr int NP = decltype(NPc)::value;
    const auto sh = make_shape<NP>(NP - 1);
    hipLaunchKernelGGL((k_assemble_1d<NP>), dim3(blocks_for(nprob, 64)), dim3(64), 0, s, d_h, d_b,
                       d_coef, sh, NP - 1, lv->ip, nprob, Mf.ptr, Lf.ptr);
    TPSS_HIP(hipGetLastError());
  });
  // batches: patches of one red-black color (no shared cells or faces)
  std::vector<std::vector<int>> by_color(2 << dim);
  for(int jp = 0; jp < P; ++jp)
    by_color[patch_color(dim, TPSS_COLOR_REDBLACK, &ls->vtx_unpacked[size_t(jp) * dim])].push_back(jp);
  // local matrices of a batch: the more patches one operator launch covers, the
  // fewer launches the npd columns take (a third of the free device memory, at
  // most 48 GiB: one batch per color at 48^3 Q4 on a 288 GB device)
  size_t free_bytes = 0, total_bytes = 0;
  TPSS_HIP(hipMemGetInfo(&free_bytes, &total_bytes));
  const size_t cap_bytes =
    std::max<size_t>(size_t(1) << 30, std::min<size_t>(size_t(48) << 30, free_bytes / 3));
  const int batch = int(std::max<size_t>(1, std::min<size_t>(
    size_t(P), cap_bytes / (size_t(npd) * npd * sizeof(double)))));
  DeviceArray<double> x{size_t(lv->Ns)}, y{size_t(lv->Ns)};
  fill_zero(x.ptr, TPSS_F64, lv->Ns, s);
  DeviceArray<double> Abuf{size_t(std::min(batch, P)) * npd * npd};
  DeviceArray<int> d_plist;
  // per position cb of the unit vector's cell in the patch: that cell and its
  // face neighbours inside the patch (the only patch cells where A e is not zero)
  std::vector<DeviceArray<uint32_t>> d_cells(size_t(1) << dim);
  const int ncb = 1 + dim;  // cells per patch in each list
  for(const auto &members : by_color)
    for(size_t b0 = 0; b0 < members.size(); b0 += size_t(batch))
    {
      const int nb = int(std::min(members.size() - b0, size_t(batch)));
      std::vector<int> plist(members.begin() + b0, members.begin() + b0 + nb);
      auto cell_at = [&](int jp, int cb) {
        int64_t c = 0, st = 1;
        for(int d = 0; d < dim; ++d)
        {
          c += int64_t(ls->vtx_unpacked[size_t(jp) * dim + d] - 1 + ((cb >> d) & 1)) * st;
          st *= n;
        }
        return uint32_t(c);
      };
      TPSS_HIP(hipStreamSynchronize(s));  // the previous batch is done with the lists
      for(int cb = 0; cb < (1 << dim); ++cb)
      {
        std::vector<uint32_t> cells;
        cells.reserve(size_t(nb) * ncb);
        for(int jp : plist)
        {
          cells.push_back(cell_at(jp, cb));
          for(int d = 0; d < dim; ++d) cells.push_back(cell_at(jp, cb ^ (1 << d)));
        }
        d_cells[cb].upload(cells);
      }
      d_plist.upload(plist);
      dispatch_dim_np(dim, lv->np, [&](auto D, auto NPc) {
        constexpr int DIM = decltype(D)::value, KC = 2 * decltype(NPc)::value;
        constexpr int NP = KC / 2, NDC = cpow(NP, DIM);
        for(int cb = 0; cb < (1 << DIM); ++cb)
        {
          // the patch cells outside the list keep A e = 0 for every column of this cb
          fill_zero(y.ptr, TPSS_F64, lv->Ns, s);
          for(int i = 0; i < NDC; ++i)
          {
            // patch-local index of local node i of the cell at position cb
            int col = 0, st = 1, r = i;
            for(int d = 0; d < DIM; ++d)
            {
              col += (((cb >> d) & 1) * NP + r % NP) * st;
              r /= NP;
              st *= KC;
            }
            hipLaunchKernelGGL((k_fit_unit<DIM, KC>), dim3(blocks_for(nb, 256)), dim3(256), 0, s,
                               x.ptr, ls->d_vtx.ptr, d_plist.ptr, nb, n, col, 1.0);
            operator_apply_f64(lv, y.ptr, x.ptr, d_cells[cb].ptr, int64_t(nb) * ncb, s);
            hipLaunchKernelGGL((k_fit_gather<DIM, KC>), dim3(blocks_for(int64_t(nb) * npd, 256)),
                               dim3(256), 0, s, y.ptr, ls->d_vtx.ptr, d_plist.ptr, nb, n, col,
                               Abuf.ptr);
            hipLaunchKernelGGL((k_fit_unit<DIM, KC>), dim3(blocks_for(nb, 256)), dim3(256), 0, s,
                               x.ptr, ls->d_vtx.ptr, d_plist.ptr, nb, n, col, 0.0);
          }
        }
        hipLaunchKernelGGL((k_separable_fit<DIM, KC>), dim3(nb), dim3(256), 0, s, Abuf.ptr,
                           d_plist.ptr, nb, Mf.ptr, Lf.ptr, sweeps, d_err.ptr);
        TPSS_HIP(hipGetLastError());
      });
    }
  // FD factors of the fits; patches with an indefinite fitted mass keep the surrogate
  DeviceArray<double> Sf{size_t(nprob) * K * K}, lamf{size_t(nprob) * K};
  DeviceArray<int> ok{size_t(P)};
  {
    std::vector<int> ones(size_t(P), 1);
    ok.upload(ones);
  }
  dispatch_dim_np(dim, lv->np, [&](auto, auto NPc) {
    constexpr int KC = 2 * decltype(NPc)::value;
    hipLaunchKernelGGL((k_fit_eig<KC>), dim3(blocks_for(nprob, 64)), dim3(64), 0, s, Mf.ptr, Lf.ptr,
                       nprob, dim, Sf.ptr, lamf.ptr, ok.ptr);
    hipLaunchKernelGGL((k_fit_validate<KC>), dim3(blocks_for(P, 64)), dim3(64), 0, s, lamf.ptr, P,
                       dim, ok.ptr);
    hipLaunchKernelGGL((k_fit_select<KC>), dim3(blocks_for(nprob, 64)), dim3(64), 0, s, Sf.ptr,
                       lamf.ptr, ok.ptr, nprob, dim, ls->S64.ptr, ls->lam64.ptr, d_err.ptr);
    TPSS_HIP(hipGetLastError());
  });
  TPSS_HIP(hipStreamSynchronize(s));
  ls->fit_err = d_err.download();
}

// Launch order of one color's patches for the residual-free pass (3D, red-black):
// every outer neighbour cell of a color is read by three of its patches, the
// ones across its x-, y- and z-face.  In vertex-lexicographic order the z-reader
// comes one plane of patches (about 500 blocks, 10 MB of traffic) after the
// other two, too far for the 4 MB L2 of an XCD, and the hardware deals
// consecutive blocks to the 8 XCDs in turn.  Here the (x, y) extent of the color
// is cut into kTx x (8 / kTx) column tiles, one per XCD; a tile's patches go
// z-slab by z-slab (lexicographic inside a slab), and the tiles' lists are
// interleaved patch by patch (one patch per block), so that block i runs on XCD
// i mod 8 and the three readers of a cell run on one XCD a slab apart.  Tiles
// measured per 64^3 Q3 step: FP64 passes 0.914 ms with 1 x 8 (full-width y
// bands), 0.920 with 2 x 4, 0.927 with 4 x 2, 0.969 with 8 x 1; FP32 passes
// 0.453 ms with 4 x 2 or 2 x 4, 0.459 with 1 x 8.
// Returns the positions [begin, end) in launch order.
std::vector<int> xcd_launch_order(const LocalSolvers &ls, int begin, int end, int kTx)
{
  constexpr int kXcd = 8;
  const int kTy = kXcd / kTx;
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
  auto block = [&](int j, int d) {
    const int x = ls.vtx_unpacked[size_t(j) * 3 + d] - 1;
    return (x + (x & 1)) / 2;
  };
  for(int j = begin; j < end; ++j)
    for(int d = 0; d < 3; ++d)
    {
      lo[d] = std::min(lo[d], block(j, d));
      hi[d] = std::max(hi[d], block(j, d));
    }
  std::vector<std::vector<int>> tiles(kXcd);
  for(int j = begin; j < end; ++j)
  {
    const int tx = (block(j, 0) - lo[0]) * kTx / (hi[0] - lo[0] + 1);
    const int ty = (block(j, 1) - lo[1]) * kTy / (hi[1] - lo[1] + 1);
    tiles[size_t(tx + kTx * ty)].push_back(j);
  }
  for(auto &t : tiles)
    std::stable_sort(t.begin(), t.end(), [&](int a, int b) {
      for(int d = 2; d >= 0; --d)
        if(block(a, d) != block(b, d)) return block(a, d) < block(b, d);
      return false;
    });
  std::vector<int> order;
  order.reserve(size_t(end - begin));
  for(size_t at = 0; order.size() < size_t(end - begin); ++at)
    for(const auto &t : tiles)
      if(at < t.size()) order.push_back(t[at]);
  return order;
}

std::shared_ptr<LocalSolvers> build_local_solvers(const tpss_level *lv,
                                                  const tpss_smoother_desc &desc, hipStream_t s)
{
  const int coloring = desc.coloring;
  const bool fit = desc.local_solver == TPSS_LOCAL_FIT;
  auto t0 = std::chrono::steady_clock::now();
  auto ls = std::make_shared<LocalSolvers>();
  const int dim = lv->dim, n = lv->n;
  ls->dim = dim;
  ls->p = lv->p;
  ls->np = lv->np;
  ls->K = 2 * lv->np;
  ls->n = n;
  ls->coloring = coloring;
  ls->ncomp = lv->ncomp;
  ls->n_colors = n_colors_of(dim, coloring);
  const int P = ipow(n - 1, dim);
  ls->P = P;

  // patch order: color, interior before boundary (patch_info.templates.h:611-645),
  // vertex lexicographic inside each group
  ls->color_ptr.assign(ls->n_colors + 1, 0);
  ls->vtx.resize(P);
  ls->vtx_unpacked.resize(size_t(P) * dim);
  int pos = 0;
  for(int col = 0; col < ls->n_colors; ++col)
  {
    ls->color_ptr[col] = pos;
    for(int pass = 0; pass < 2; ++pass)
      for(int lin = 0; lin < P; ++lin)
      {
        int v[3] = {1, 1, 1}, t = lin;
        for(int d = 0; d < dim; ++d)
        {
          v[d] = 1 + t % (n - 1);
          t /= (n - 1);
        }
        if(patch_color(dim, coloring, v) != col) continue;
        if(int(patch_at_boundary(dim, n, v)) != pass) continue;
        ls->vtx[pos] = uint32_t(v[0]) | (uint32_t(v[1]) << 10) | (uint32_t(v[2]) << 20);
        for(int d = 0; d < dim; ++d) ls->vtx_unpacked[size_t(pos) * dim + d] = v[d];
        ++pos;
      }
  }
  ls->color_ptr[ls->n_colors] = pos;
  ls->d_vtx.upload(ls->vtx);

  // 1D problems: (h_left, h_right, boundary bits) per patch and direction
  std::vector<double> hs(size_t(P) * dim * 2);
  std::vector<uint8_t> bd(size_t(P) * dim * 2);
  if(lv->distorted)
  {
    DeviceArray<double> d_h(hs.size());
    DeviceArray<uint8_t> d_b(bd.size());
    const int nq = std::min(8, std::max(2, lv->p + 1));  // apps/poisson_distorted.cc:88
    double gx[8], gw[8];
    gauss_lobatto_nodes(nq, gx, gw);
    DeviceArray<double> d_gx, d_gw;
    d_gx.upload(std::vector<double>(gx, gx + nq));
    d_gw.upload(std::vector<double>(gw, gw + nq));
    auto launch = [&](auto D, auto NQ) {
      constexpr int DIM = decltype(D)::value, Q = decltype(NQ)::value;
      hipLaunchKernelGGL((k_surrogate_h<DIM, Q>), dim3(blocks_for(P, 128)), dim3(128), 0, s,
                         lv->d_vertices.ptr, ls->d_vtx.ptr, P, n, d_gx.ptr, d_gw.ptr, d_h.ptr,
                         d_b.ptr);
    };
    if(dim == 2)
    {
      if(nq == 2) launch(ic<2>{}, ic<2>{});
      else if(nq == 3) launch(ic<2>{}, ic<3>{});
      else if(nq == 4) launch(ic<2>{}, ic<4>{});
      else launch(ic<2>{}, ic<5>{});
    }
    else
    {
      if(nq == 2) launch(ic<3>{}, ic<2>{});
      else if(nq == 3) launch(ic<3>{}, ic<3>{});
      else if(nq == 4) launch(ic<3>{}, ic<4>{});
      else launch(ic<3>{}, ic<5>{});
    }
    TPSS_HIP(hipGetLastError());
    TPSS_HIP(hipStreamSynchronize(s));
    hs = d_h.download();
    bd = d_b.download();
    ls->nfac = P * dim;
    ls->fac.resize(size_t(P) * dim);
    for(size_t i = 0; i < ls->fac.size(); ++i) ls->fac[i] = uint32_t(i);
    ls->prob_h = hs;
    ls->prob_b = bd;
    ls->prob_coef.assign(size_t(ls->nfac) * 2, 1.0);
  }
  else
  {
    for(int jp = 0; jp < P; ++jp)
      for(int d = 0; d < dim; ++d)
      {
        const int vd = ls->vtx_unpacked[size_t(jp) * dim + d];
        hs[(size_t(jp) * dim + d) * 2 + 0] = lv->h;
        hs[(size_t(jp) * dim + d) * 2 + 1] = lv->h;
        bd[(size_t(jp) * dim + d) * 2 + 0] = vd == 1;
        bd[(size_t(jp) * dim + d) * 2 + 1] = vd == n - 1;
      }
    // deduplicate identical 1D problems; with component blocks (elasticity) the
    // 1D matrix of block c in direction d also depends on whether c == d
    // (the fitted local solvers need one factor row per patch and direction)
    std::map<std::tuple<double, double, int, int, double, double>, uint32_t> uniq;
    const size_t rows = size_t(P) * dim;
    ls->fac.resize(rows * ls->ncomp);
    for(int c = 0; c < ls->ncomp; ++c)
      for(size_t i = 0; i < rows; ++i)
      {
        if(fit)
        {
          ls->prob_h.push_back(hs[2 * i]);
          ls->prob_h.push_back(hs[2 * i + 1]);
          ls->prob_b.push_back(bd[2 * i]);
          ls->prob_b.push_back(bd[2 * i + 1]);
          ls->prob_coef.push_back(1.0);
          ls->prob_coef.push_back(1.0);
          ls->fac[i] = uint32_t(i);
          continue;
        }
        const int d = int(i % dim);
        double a_cons = 1.0, a_pen = 1.0;
        if(ls->ncomp > 1)
        {
          a_cons = 2.0 * lv->mu * (c == d ? 1.0 : 0.5) + (c == d ? lv->lambda : 0.0);
          a_pen = 2.0 * lv->mu + (c == d ? lv->lambda : 0.0);
        }
        auto key = std::make_tuple(hs[2 * i], hs[2 * i + 1], int(bd[2 * i]), int(bd[2 * i + 1]),
                                   a_cons, a_pen);
        auto it = uniq.find(key);
        if(it == uniq.end())
        {
          const uint32_t id = uint32_t(uniq.size());
          uniq.emplace(key, id);
          ls->prob_h.push_back(hs[2 * i]);
          ls->prob_h.push_back(hs[2 * i + 1]);
          ls->prob_b.push_back(bd[2 * i]);
          ls->prob_b.push_back(bd[2 * i + 1]);
          ls->prob_coef.push_back(a_cons);
          ls->prob_coef.push_back(a_pen);
          ls->fac[c * rows + i] = id;
        }
        else
          ls->fac[c * rows + i] = it->second;
      }
    ls->nfac = fit ? int(rows) : int(uniq.size());
  }
  ls->h_host = hs;
  ls->d_fac.upload(ls->fac);
  if(dim == 3 && ls->K == 8 && ls->ncomp == 1 && coloring == TPSS_COLOR_REDBLACK && !lv->distorted)
  {
    std::vector<uint32_t> vtx_l(static_cast<size_t>(P)), fac_l(static_cast<size_t>(P) * 3);
    for(int c = 0; c < ls->n_colors; ++c)
    {
      const std::vector<int> order = xcd_launch_order(
        *ls, ls->color_ptr[c], ls->color_ptr[c + 1], desc.local_precision == TPSS_F64 ? 1 : 4);
      for(size_t q = 0; q < order.size(); ++q)
      {
        const size_t at = size_t(ls->color_ptr[c]) + q, j = size_t(order[q]);
        vtx_l[at] = ls->vtx[j];
        for(int d = 0; d < 3; ++d) fac_l[at * 3 + d] = ls->fac[j * 3 + d];
      }
    }
    ls->d_vtx_launch.upload(vtx_l);
    ls->d_fac_launch.upload(fac_l);
  }

  // FP64 factors on the device
  const int K = ls->K, nf = ls->nfac;
  ls->S64.resize(size_t(nf) * K * K);
  ls->lam64.resize(size_t(nf) * K);
  {
    DeviceArray<double> d_h, d_coef;
    DeviceArray<uint8_t> d_b;
    d_h.upload(ls->prob_h);
    d_b.upload(ls->prob_b);
    d_coef.upload(ls->prob_coef);
    DeviceArray<int> d_status(2);
    TPSS_HIP(hipMemsetAsync(d_status.ptr, 0, 2 * sizeof(int), s));
    const int nrows = P * ls->ncomp;
    dispatch_dim_np(dim, lv->np, [&](auto D, auto NPc) {
      constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value;
      const auto sh = make_shape<NP>(NP - 1);
      hipLaunchKernelGGL((k_setup_fd<NP>), dim3(blocks_for(nf, 64)), dim3(64), 0, s, d_h.ptr,
                         d_b.ptr, d_coef.ptr, sh, NP - 1, lv->ip, nf, ls->S64.ptr, ls->lam64.ptr,
                         d_status.ptr);
      TPSS_HIP(hipGetLastError());
      hipLaunchKernelGGL((k_check_eigenvalue_sums<DIM, 2 * NP>), dim3(blocks_for(nrows, 64)),
                         dim3(64), 0, s, ls->d_fac.ptr, ls->lam64.ptr, nrows, d_status.ptr + 1);
      TPSS_HIP(hipGetLastError());
    });
    int status[2] = {0, 0};
    TPSS_HIP(hipMemcpyAsync(status, d_status.ptr, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    TPSS_HIP(hipStreamSynchronize(s));
    if(status[0]) throw Error(TPSS_E_SINGULAR, "local solver setup: mass matrix not s.p.d.");
    if(status[1]) throw Error(TPSS_E_SINGULAR, "local solver setup: has zero eigenvalues");
    if(fit)
    {
      fit_local_solvers(lv, ls.get(), desc.fit_sweeps > 0 ? desc.fit_sweeps : 4, d_h.ptr, d_b.ptr,
                        d_coef.ptr, s);
      TPSS_HIP(hipMemsetAsync(d_status.ptr + 1, 0, sizeof(int), s));
      dispatch_dim_np(dim, lv->np, [&](auto D, auto NPc) {
        constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value;
        hipLaunchKernelGGL((k_check_eigenvalue_sums<DIM, 2 * NP>), dim3(blocks_for(nrows, 64)),
                           dim3(64), 0, s, ls->d_fac.ptr, ls->lam64.ptr, nrows, d_status.ptr + 1);
        TPSS_HIP(hipGetLastError());
      });
      TPSS_HIP(hipMemcpyAsync(status, d_status.ptr, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
      TPSS_HIP(hipStreamSynchronize(s));
      if(status[1]) throw Error(TPSS_E_SINGULAR, "local solver setup: fitted solver has zero eigenvalues");
    }
  }
  // symmetric-first order for mirror-symmetric problems (even-odd solves)
  ls->d_eo.resize(size_t(nf));
  dispatch_dim_np(dim, lv->np, [&](auto, auto NPc) {
    constexpr int NP = decltype(NPc)::value;
    hipLaunchKernelGGL((k_even_odd_order<2 * NP>), dim3(blocks_for(nf, 64)), dim3(64), 0, s,
                       ls->S64.ptr, ls->lam64.ptr, nf, ls->d_eo.ptr);
    TPSS_HIP(hipGetLastError());
  });
  ls->S32.resize(ls->S64.n);
  ls->lam32.resize(ls->lam64.n);
  convert(ls->S32.ptr, TPSS_F32, ls->S64.ptr, TPSS_F64, int64_t(ls->S64.n), s);
  convert(ls->lam32.ptr, TPSS_F32, ls->lam64.ptr, TPSS_F64, int64_t(ls->lam64.n), s);

  // cells covered by each color (targets of the localized residual update)
  ls->color_cells.resize(ls->n_colors);
  ls->color_ncells.assign(ls->n_colors, 0);
  for(int col = 0; col < ls->n_colors; ++col)
  {
    std::vector<uint32_t> cells;
    for(int jp = ls->color_ptr[col]; jp < ls->color_ptr[col + 1]; ++jp)
      for(int cb = 0; cb < (1 << dim); ++cb)
      {
        int64_t c = 0, st = 1;
        for(int d = 0; d < dim; ++d)
        {
          c += int64_t(ls->vtx_unpacked[size_t(jp) * dim + d] - 1 + ((cb >> d) & 1)) * st;
          st *= n;
        }
        cells.push_back(uint32_t(c));
      }
    std::sort(cells.begin(), cells.end());
    cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
    ls->color_ncells[col] = int64_t(cells.size());
    ls->color_cells[col] = std::make_shared<DeviceArray<uint32_t>>();
    ls->color_cells[col]->upload(cells);
  }
  TPSS_HIP(hipStreamSynchronize(s));
  ls->setup_seconds =
    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return ls;
}

// ---------------------------------------------------------------------------- apply

// one launch over the patches [begin, end) of component block comp: gather,
// fast-diagonalization inverse, damping, scatter (mode 0: scatter-add into dst;
// mode 1: write the additive slot buffer)
template<typename VT, typename CT>
void patch_pass(tpss_smoother *sm, const VT *src, VT *dst, CT *ubuf, int begin, int end, int mode,
                hipStream_t s, int comp = 0)
{
  const LocalSolvers &ls = *sm->ls;
  if(end <= begin) return;
  const int64_t off = comp * sm->level->Ns;
  src += off;
  if(dst) dst += off;
  const uint32_t *fac = ls.d_fac.ptr + size_t(comp) * ls.P * ls.dim;
  const CT *S = std::is_same<CT, double>::value ? (const CT *)ls.S64.ptr : (const CT *)ls.S32.ptr;
  const CT *lam =
    std::is_same<CT, double>::value ? (const CT *)ls.lam64.ptr : (const CT *)ls.lam32.ptr;
  const float *om = sm->use_omega ? sm->d_omega.ptr : nullptr;
  const CT relax = CT(sm->desc.relaxation);
  dispatch_dim_np(ls.dim, ls.np, [&](auto D, auto NPc) {
    constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value, K = 2 * NP;
    using PS = PatchShape<DIM, K>;
    const unsigned grid = unsigned((end - begin + PS::PPB - 1) / PS::PPB);
    const unsigned threads = unsigned(PS::THREADS);
    if(mode == 0)
      hipLaunchKernelGGL((k_patch_fd<DIM, K, VT, CT, 0>), dim3(grid), dim3(threads), 0, s, src,
                         dst, ubuf, ls.d_vtx.ptr, fac, S, lam, om, relax, begin, end, ls.n);
    else
      hipLaunchKernelGGL((k_patch_fd<DIM, K, VT, CT, 1>), dim3(grid), dim3(threads), 0, s, src,
                         dst, ubuf, ls.d_vtx.ptr, fac, S, lam, om, relax, begin, end, ls.n);
    TPSS_HIP(hipGetLastError());
  });
}

// residual-free multiplicative color pass (Cartesian mesh, red-black colors);
// FD = 1 selects the BF16-MFMA patch solve (3D, Q3)
// colors of at most this many patches (3D, K = 8) run k_patch_mvp_split: its
// four waves per patch shorten the pass where a color holds too few patches to
// fill the GPU (measured per pass, FP32 / FP64: 16^3 cells, about 220 patches
// per color, 5.1 / 6.1 us against 5.8 / 7.8 us with one wave per patch), and
// lengthen it where the patches fill the GPU (32^3, about 1860 patches per
// color: 12.3 / 17.2 against 7.6 / 12.1 us)
constexpr int kSplitMaxPatches = 512;

template<typename VT, typename CT, int FD = 0>
void mvp_fused_pass(tpss_smoother *sm, const VT *b, VT *x, int begin, int end, bool x_is_zero,
                    hipStream_t s)
{
  const LocalSolvers &ls = *sm->ls;
  if(end <= begin) return;
  const CT *S = std::is_same<CT, double>::value ? (const CT *)ls.S64.ptr : (const CT *)ls.S32.ptr;
  const CT *lam =
    std::is_same<CT, double>::value ? (const CT *)ls.lam64.ptr : (const CT *)ls.lam32.ptr;
  const float *om = sm->use_omega ? sm->d_omega.ptr : nullptr;
  const CT relax = CT(sm->desc.relaxation);
  dispatch_dim_np(ls.dim, ls.np, [&](auto D, auto NPc) {
    constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value, K = 2 * NP;
    using PS = PatchShape<DIM, K>;
    OpTab<CT, NP> tab;
    const Operator1D &op = sm->level->op1d;
    for(int e = 0; e < NP * NP; ++e)
    {
      tab.M[e] = CT(op.M[e]);
      tab.Fpm[e] = CT(op.Fpm[e]);
      tab.Fmp[e] = CT(op.Fmp[e]);
      for(int c = 0; c < 4; ++c) tab.D[c][e] = CT(op.D[c][e]);
    }
    using MS = MvpShape<DIM, K, FD>;
    const unsigned grid = unsigned((end - begin + MS::PPB - 1) / MS::PPB);
    if constexpr(FD == 1)
    {
      if constexpr(DIM == 3 && K == 8 && std::is_same<CT, float>::value)
        hipLaunchKernelGGL((k_patch_mvp_fused<DIM, K, VT, CT, 1>), dim3(grid), dim3(MS::THREADS), 0,
                           s, b, x, tab, ls.d_vtx.ptr, ls.d_fac.ptr, S, lam, nullptr, om, relax,
                           begin, end, ls.n, x_is_zero ? 1 : 0, sm->pass_flag);
      else
        throw Error(TPSS_E_UNSUPPORTED, "BF16-MFMA patch solves are built for 3D Q3 (k = 8)");
    }
    else
    {
      if constexpr(DIM == 3 && K == 8)
        if(end - begin <= kSplitMaxPatches)
        {
          // few patches (coarse levels): one patch over four waves
          hipLaunchKernelGGL((k_patch_mvp_split<VT, CT>), dim3(unsigned(end - begin)), dim3(256),
                             0, s, b, x, tab, ls.d_vtx.ptr, ls.d_fac.ptr, S, lam, om, relax,
                             begin, ls.n, x_is_zero ? 1 : 0, sm->pass_flag);
          TPSS_HIP(hipGetLastError());
          return;
        }
      // the XCD launch order where no per-patch damping is indexed by position
      const bool xcd = om == nullptr && ls.d_vtx_launch.ptr != nullptr && MS::PPB == 1;
      hipLaunchKernelGGL((k_patch_mvp_fused<DIM, K, VT, CT>), dim3(grid), dim3(MS::THREADS), 0, s,
                         b, x, tab, xcd ? ls.d_vtx_launch.ptr : ls.d_vtx.ptr,
                         xcd ? ls.d_fac_launch.ptr : ls.d_fac.ptr, S, lam, ls.d_eo.ptr, om, relax,
                         begin, end, ls.n, x_is_zero ? 1 : 0, sm->pass_flag);
    }
    TPSS_HIP(hipGetLastError());
  });
}

template<typename VT, typename CT>
void avp_sum(tpss_smoother *sm, const CT *ubuf, VT *dst, bool accumulate, double alpha,
             hipStream_t s, int comp = 0)
{
  const tpss_level *lv = sm->level;
  dst += comp * lv->Ns;
  dispatch_dim_np(lv->dim, lv->np, [&](auto D, auto NPc) {
    constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value;
    const unsigned g = blocks_for(lv->Ns, 256);
    auto launch = [&](auto ACC, auto RASc) {
      hipLaunchKernelGGL((k_avp_sum<DIM, NP, VT, CT, decltype(ACC)::value != 0,
                                    decltype(RASc)::value>),
                         dim3(g), dim3(256), 0, s, ubuf, dst, lv->n, lv->Ns, alpha);
    };
    const int ras = sm->desc.ras;
    if(accumulate)
    {
      if(ras == 1) launch(ic<1>{}, ic<1>{});
      else if(ras == 2) launch(ic<1>{}, ic<2>{});
      else launch(ic<1>{}, ic<0>{});
    }
    else
    {
      if(ras == 1) launch(ic<0>{}, ic<1>{});
      else if(ras == 2) launch(ic<0>{}, ic<2>{});
      else launch(ic<0>{}, ic<0>{});
    }
    TPSS_HIP(hipGetLastError());
  });
}

// Device-time sections: a pair of HIP events around the section's launches on
// the caller's stream; no synchronisation here, the pending pairs are resolved
// by tpss_smoother_get_timings.
struct SectionTimer
{
  tpss_smoother *sm;
  int idx;
  hipStream_t s;
  hipEvent_t a = nullptr;
  SectionTimer(tpss_smoother *sm_, int idx_, hipStream_t s_) : sm(sm_), idx(idx_), s(s_)
  {
    if(!sm->timing) return;
    TPSS_HIP(hipEventCreate(&a));
    TPSS_HIP(hipEventRecord(a, s));
  }
  ~SectionTimer()
  {
    if(!sm->timing || !a) return;
    hipEvent_t b = nullptr;
    if(hipEventCreate(&b) != hipSuccess || hipEventRecord(b, s) != hipSuccess)
    {
      (void)hipEventDestroy(a);
      if(b) (void)hipEventDestroy(b);
      return;
    }
    sm->pending.push_back({idx, a, b});
  }
};

void resolve_timings(tpss_smoother *sm)
{
  for(auto &p : sm->pending)
  {
    float ms = 0.f;
    if(hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess)
    {
      sm->timers[p.idx].time += 1e-3 * ms;
      sm->timers[p.idx].n_calls += 1;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  sm->pending.clear();
}

// dst = P src  (or, if accumulate: dst += alpha * P src for the additive variant)
template<typename VT, typename CT>
void precond_vmult_t(tpss_smoother *sm, VT *dst, const VT *src, hipStream_t s, bool accumulate,
                     double alpha, const std::vector<int> &seq)
{
  LocalSolvers &ls = *sm->ls;
  const tpss_level *lv = sm->level;
  SectionTimer step_timer(sm, 4, s);
  if(sm->desc.variant == TPSS_ADDITIVE)
  {
    CT *ub = (CT *)sm->scratch;
    if(!ub)
    {
      sm->ubuf.resize(size_t(lv->Ns) * (1 << lv->dim) * sizeof(CT));
      ub = (CT *)sm->ubuf.ptr;
    }
    {
      SectionTimer t(sm, 1, s);
      for(int c = 0; c < ls.ncomp; ++c)
      {
        patch_pass<VT, CT>(sm, src, nullptr, ub, 0, ls.P, 1, s, c);
        avp_sum<VT, CT>(sm, ub, dst, accumulate, accumulate ? alpha : 1.0, s, c);
      }
    }
    return;
  }
  // multiplicative: colors in sequence, localized residual before each color
  VT *out = dst;
  if(accumulate)
  {
    sm->tmp2.resize(size_t(lv->N) * sizeof(VT));
    out = (VT *)sm->tmp2.ptr;
  }
  if(!lv->distorted && lv->ncomp == 1 && ls.coloring == TPSS_COLOR_REDBLACK && sm->fused_mvp)
  {
    // cells not yet reached by a color are read as neighbours: start from zero
    fill_zero(out, lv->vec_prec, lv->N, s);
    for(size_t k = 0; k < seq.size(); ++k)
    {
      const int c = seq[k];
      SectionTimer t(sm, 1, s);
      mvp_fused_pass<VT, CT>(sm, src, out, ls.color_ptr[c], ls.color_ptr[c + 1], k == 0, s);
    }
    if(accumulate)
      axpby(dst, out, lv->vec_prec, lv->N, alpha, 1.0, s);
    else if(sm->pass_flag)
      sm->pass_checked = true;  // every cell belongs to a patch: dst is covered
    return;
  }
  VT *res = (VT *)sm->scratch;
  if(!res)
  {
    sm->res.resize(size_t(lv->N) * sizeof(VT));
    res = (VT *)sm->res.ptr;
  }
  fill_zero(out, lv->vec_prec, lv->N, s);
  for(size_t k = 0; k < seq.size(); ++k)
  {
    const int c = seq[k];
    const VT *r = src;
    if(k > 0)
    {
      SectionTimer t(sm, 0, s);
      operator_apply_any(lv, res, out, src, ls.color_cells[c]->ptr, ls.color_ncells[c], s);
      r = res;
    }
    SectionTimer t(sm, 1, s);
    for(int comp = 0; comp < ls.ncomp; ++comp)
      patch_pass<VT, CT>(sm, r, out, nullptr, ls.color_ptr[c], ls.color_ptr[c + 1], 0, s, comp);
  }
  if(accumulate) axpby(dst, out, lv->vec_prec, lv->N, alpha, 1.0, s);
}

// BF16-MFMA patch solves (FP32 accumulation) for the multiplicative variant on
// Cartesian meshes with red-black colors, 3D Q3; vectors FP32 or FP64
// BF16-MFMA patch solves of the elasticity smoother (3D Q2, block diagonal):
// the residual-update color sweep, the residual in the vector precision, the
// component blocks' solves on the matrix cores
template<typename VT>
void precond_vmult_bf16_elasticity(tpss_smoother *sm, VT *dst, const VT *src, hipStream_t s,
                                   bool accumulate, double alpha, const std::vector<int> &seq)
{
  LocalSolvers &ls = *sm->ls;
  tpss_level *lv = sm->level;
  SectionTimer step_timer(sm, 4, s);
  VT *out = dst;
  if(accumulate)
  {
    sm->tmp2.resize(size_t(lv->N) * sizeof(VT));
    out = (VT *)sm->tmp2.ptr;
  }
  VT *res = (VT *)sm->scratch;
  if(!res)
  {
    sm->res.resize(size_t(lv->N) * sizeof(VT));
    res = (VT *)sm->res.ptr;
  }
  const float *om = sm->use_omega ? sm->d_omega.ptr : nullptr;
  fill_zero(out, lv->vec_prec, lv->N, s);
  for(size_t k = 0; k < seq.size(); ++k)
  {
    const int c = seq[k];
    const int begin = ls.color_ptr[c], end = ls.color_ptr[c + 1];
    const VT *r = src;
    if(k > 0)
    {
      SectionTimer t(sm, 0, s);
      operator_apply_any(lv, res, out, src, ls.color_cells[c]->ptr, ls.color_ncells[c], s);
      r = res;
    }
    if(end <= begin) continue;
    SectionTimer t(sm, 1, s);
    for(int comp = 0; comp < ls.ncomp; ++comp)
      hipLaunchKernelGGL((k_patch_fd_bf16_q2<VT>), dim3(unsigned((end - begin + 1) / 2)), dim3(128),
                         0, s, r + comp * lv->Ns, out + comp * lv->Ns, ls.d_vtx.ptr,
                         ls.d_fac.ptr + size_t(comp) * ls.P * 3, ls.S32.ptr, ls.lam32.ptr, om,
                         float(sm->desc.relaxation), begin, end, ls.n);
    TPSS_HIP(hipGetLastError());
  }
  if(accumulate) axpby(dst, out, lv->vec_prec, lv->N, alpha, 1.0, s);
}

template<typename VT>
void precond_vmult_bf16(tpss_smoother *sm, VT *dst, const VT *src, hipStream_t s, bool accumulate,
                        double alpha, const std::vector<int> &seq)
{
  LocalSolvers &ls = *sm->ls;
  const tpss_level *lv = sm->level;
  if(sm->desc.variant == TPSS_MULTIPLICATIVE && lv->ncomp == 3 && ls.coloring == TPSS_COLOR_REDBLACK &&
     ls.dim == 3 && ls.K == 6)
  {
    precond_vmult_bf16_elasticity(sm, dst, src, s, accumulate, alpha, seq);
    return;
  }
  if(sm->desc.variant != TPSS_MULTIPLICATIVE || lv->distorted || lv->ncomp > 1 ||
     ls.coloring != TPSS_COLOR_REDBLACK || ls.dim != 3 || ls.K != 8)
    throw Error(TPSS_E_UNSUPPORTED, "BF16-MFMA patch solves: multiplicative, red-black, Cartesian "
                                    "3D Q3 (Poisson) or 3D Q2 (elasticity) only");
  SectionTimer step_timer(sm, 4, s);
  VT *out = dst;
  if(accumulate)
  {
    sm->tmp2.resize(size_t(lv->N) * sizeof(VT));
    out = (VT *)sm->tmp2.ptr;
  }
  fill_zero(out, lv->vec_prec, lv->N, s);
  for(size_t k = 0; k < seq.size(); ++k)
  {
    const int c = seq[k];
    SectionTimer t(sm, 1, s);
    mvp_fused_pass<VT, float, 1>(sm, src, out, ls.color_ptr[c], ls.color_ptr[c + 1], k == 0, s);
  }
  if(accumulate) axpby(dst, out, lv->vec_prec, lv->N, alpha, 1.0, s);
}
template void precond_vmult_bf16<double>(tpss_smoother *, double *, const double *, hipStream_t,
                                         bool, double, const std::vector<int> &);
template void precond_vmult_bf16<float>(tpss_smoother *, float *, const float *, hipStream_t, bool,
                                        double, const std::vector<int> &);

template<typename VT>
void precond_dispatch(tpss_smoother *sm, VT *dst, const VT *src, hipStream_t s, bool accumulate,
                      double alpha, const std::vector<int> &seq)
{
  const int lp = sm->desc.local_precision;
  if(std::is_same<VT, double>::value && lp == TPSS_F64)
    precond_vmult_t<VT, double>(sm, dst, src, s, accumulate, alpha, seq);
  else if(lp == TPSS_F32)
    precond_vmult_t<VT, float>(sm, dst, src, s, accumulate, alpha, seq);
  else if(lp == TPSS_BF16)
    precond_vmult_bf16(sm, dst, src, s, accumulate, alpha, seq);
  else
    throw Error(TPSS_E_UNSUPPORTED, "local precision F64 needs F64 level vectors");
}

// transpose: the additive operator is symmetric (symmetric local solvers,
// scalar weights); the transpose of a multiplicative sweep is the sweep over
// the reversed color sequence
void precond_vmult_any(tpss_smoother *sm, void *dst, const void *src, hipStream_t s,
                       bool accumulate, double alpha, bool transpose)
{
  const std::vector<int> &seq = transpose ? sm->seq_t : sm->seq;
  if(sm->level->vec_prec == TPSS_F64)
    precond_dispatch<double>(sm, (double *)dst, (const double *)src, s, accumulate, alpha, seq);
  else
    precond_dispatch<float>(sm, (float *)dst, (const float *)src, s, accumulate, alpha, seq);
}

// True if a Richardson step with omega_R = 1 can run as residual-free color
// passes on (x, b) directly: for the multiplicative variant
//   x + P (b - A x)  ==  the color sweep started from x,
// because each patch update x_j <- (1 - w) x_j + w A_j^{-1} (b_j - B_j x_out)
// equals y_j <- y_j + w A_j^{-1} (r - A y)_j on y = x - x_0, r = b - A x_0.
bool fused_step_applies(const tpss_smoother *sm)
{
  const tpss_level *lv = sm->level;
  const LocalSolvers &ls = *sm->ls;
  if(sm->desc.variant != TPSS_MULTIPLICATIVE || lv->distorted || lv->ncomp != 1 ||
     ls.coloring != TPSS_COLOR_REDBLACK || !sm->fused_mvp || sm->desc.damping_factor != 1.0)
    return false;
  // BF16 patch solves stay on the correction (precond on the residual): in the
  // residual-free step they would produce the solution values themselves and
  // cap the smoother's accuracy at BF16 resolution
  if(sm->desc.local_precision == TPSS_BF16) return false;
  return !(lv->vec_prec == TPSS_F32 && sm->desc.local_precision == TPSS_F64);
}

template<typename VT>
void fused_sweep(tpss_smoother *sm, VT *x, const VT *b, bool x_is_zero, hipStream_t s)
{
  const LocalSolvers &ls = *sm->ls;
  const int lp = sm->desc.local_precision;
  for(size_t k = 0; k < sm->seq.size(); ++k)
  {
    const int c = sm->seq[k];
    const bool zero = x_is_zero && k == 0;
    SectionTimer t(sm, 1, s);
    if(lp == TPSS_F64)
      mvp_fused_pass<VT, double>(sm, b, x, ls.color_ptr[c], ls.color_ptr[c + 1], zero, s);
    else
      mvp_fused_pass<VT, float>(sm, b, x, ls.color_ptr[c], ls.color_ptr[c + 1], zero, s);
  }
}

// SmootherRichardson::step_impl: x <- x + omega_R P (b - A x), n_steps times.
// x_is_zero: the caller guarantees dst == 0 (SmootherRichardson::vmult).
void smoother_step_any(tpss_smoother *sm, void *dst, const void *src, hipStream_t s,
                       bool x_is_zero)
{
  const tpss_level *lv = sm->level;
  if(fused_step_applies(sm))
  {
    SectionTimer step_timer(sm, 4, s);
    for(int k = 0; k < sm->desc.n_steps; ++k)
    {
      if(lv->vec_prec == TPSS_F64)
        fused_sweep<double>(sm, (double *)dst, (const double *)src, x_is_zero && k == 0, s);
      else
        fused_sweep<float>(sm, (float *)dst, (const float *)src, x_is_zero && k == 0, s);
    }
    if(sm->pass_flag) sm->pass_checked = true;
    return;
  }
  const size_t bytes = size_t(lv->N) * elem_size(lv->vec_prec);
  sm->rich_r.resize(bytes);
  for(int k = 0; k < sm->desc.n_steps; ++k)
  {
    {
      // the Richardson residual b - A x is timed with the residual updates
      SectionTimer t(sm, 0, s);
      operator_apply_any(lv, sm->rich_r.ptr, dst, src, nullptr, 0, s);
    }
    precond_vmult_any(sm, dst, sm->rich_r.ptr, s, true, sm->desc.damping_factor);
  }
}

// device scratch (bytes) one precond vmult (dst overwritten) needs: the additive
// slot buffer, or the residual of the multiplicative sweeps that update it
size_t precond_scratch_bytes(const tpss_smoother *sm)
{
  const tpss_level *lv = sm->level;
  const LocalSolvers &ls = *sm->ls;
  const int lp = sm->desc.local_precision;
  const size_t ct = lp == TPSS_F64 ? 8 : 4;
  if(sm->desc.variant == TPSS_ADDITIVE) return size_t(lv->Ns) * (size_t(1) << lv->dim) * ct;
  if(lp == TPSS_BF16) return lv->ncomp > 1 ? size_t(lv->N) * elem_size(lv->vec_prec) : 0;
  if(!lv->distorted && lv->ncomp == 1 && ls.coloring == TPSS_COLOR_REDBLACK && sm->fused_mvp)
    return 0;
  return size_t(lv->N) * elem_size(lv->vec_prec);
}

// throws (once) if a check kernel of an earlier call has flagged its output
void report_nonfinite(tpss_smoother *sm)
{
  if(sm->nonfinite.host && __atomic_load_n(sm->nonfinite.host, __ATOMIC_ACQUIRE))
  {
    __atomic_store_n(sm->nonfinite.host, 0, __ATOMIC_RELEASE);
    throw Error(TPSS_E_NONFINITE, "non-finite entries in the output of this smoother");
  }
}

// AssertIsFinite(dst.l2_norm()) (schwarz_preconditioner.templates.h:238), on
// always by default: one pass over the output that sets a pinned host flag
// when it meets a NaN or Inf.  Mode 1 reports the flag at the next entry
// point of this smoother (or tpss_smoother_finite_status) without a host
// synchronisation here; mode 2 synchronises and reports at once.
void check_finite(tpss_smoother *sm, const void *x, hipStream_t s)
{
  if(!sm->check_finite) return;
  const tpss_level *lv = sm->level;
  sm->nonfinite.ensure();
  // a residual-free sweep that wrote the whole output has checked every value
  // it stored (PassCheck); otherwise one pass over the output
  const bool checked = sm->pass_checked;
  sm->pass_checked = false;
  if(!checked)
  {
    const unsigned grid = unsigned(std::min<int64_t>(2048, (lv->N + 255) / 256));
    if(lv->vec_prec == TPSS_F64)
      hipLaunchKernelGGL((k_nonfinite<double>), dim3(grid), dim3(256), 0, s, (const double *)x,
                         lv->N, sm->nonfinite.dev);
    else
      hipLaunchKernelGGL((k_nonfinite<float>), dim3(grid), dim3(256), 0, s, (const float *)x,
                         lv->N, sm->nonfinite.dev);
    TPSS_HIP(hipGetLastError());
  }
  if(sm->check_finite == 2)
  {
    TPSS_HIP(hipStreamSynchronize(s));
    report_nonfinite(sm);
  }
}

std::vector<int> color_sequence(int n_colors, bool reverse, bool symmetrized)
{
  // get_color_sequence, schwarz_preconditioner.h:531-567
  std::vector<int> seq(n_colors);
  for(int i = 0; i < n_colors; ++i) seq[i] = i;
  if(reverse) std::reverse(seq.begin(), seq.end());
  if(symmetrized)
    for(int i = n_colors - 2; i >= 0; --i) seq.push_back(seq[i]);
  return seq;
}

void validate_desc(const tpss_smoother_desc *d)
{
  TPSS_REQUIRE(d, "null smoother descriptor");
  TPSS_REQUIRE(d->variant == TPSS_ADDITIVE || d->variant == TPSS_MULTIPLICATIVE,
               "invalid smoother variant");
  TPSS_REQUIRE(d->coloring >= TPSS_COLOR_NONE && d->coloring <= TPSS_COLOR_REDBLACK,
               "invalid coloring");
  TPSS_REQUIRE(d->relaxation > 0.0, "relaxation must be positive");
  TPSS_REQUIRE(d->n_steps >= 1, "n_steps must be >= 1");
  TPSS_REQUIRE(d->local_precision >= TPSS_F64 && d->local_precision <= TPSS_BF16,
               "invalid local precision");
  TPSS_REQUIRE(d->local_solver == TPSS_LOCAL_SURROGATE || d->local_solver == TPSS_LOCAL_FIT,
               "invalid local solver");
  TPSS_REQUIRE(d->fit_sweeps >= 0 && d->fit_sweeps <= 100, "fit_sweeps must be in 0..100");
  TPSS_REQUIRE(!(d->variant == TPSS_MULTIPLICATIVE && d->coloring == TPSS_COLOR_NONE),
               "the multiplicative variant needs a coloring (RedBlack for DG)");
  TPSS_REQUIRE(d->ras >= 0 && d->ras <= 2, "ras must be 0, 1 or 2");
  // dof_info.templates.h:213-214: "Additive scheme is only supported."
  TPSS_REQUIRE(d->ras == 0 || d->variant == TPSS_ADDITIVE,
               "restricted additive Schwarz needs the additive variant");
}

}  // namespace tpss

using namespace tpss;

extern "C" {

int tpss_smoother_create(tpss_level *level, const tpss_smoother_desc *desc, tpss_smoother **out)
{
  return guarded([&] {
    TPSS_REQUIRE(level && out, "null argument");
    validate_desc(desc);
    if(level->vec_prec == TPSS_F32 && desc->local_precision == TPSS_F64)
      throw Error(TPSS_E_UNSUPPORTED, "local precision F64 needs F64 level vectors");
    if(level->ncomp > 1 && desc->local_precision == TPSS_BF16 &&
       !(level->dim == 3 && level->p == 2 && desc->variant == TPSS_MULTIPLICATIVE &&
         desc->coloring == TPSS_COLOR_REDBLACK))
      throw Error(TPSS_E_UNSUPPORTED, "linear elasticity with BF16-MFMA patch solves: 3D Q2, "
                                      "multiplicative, red-black colors only");
    if(level->ncomp > 1 && desc->local_solver == TPSS_LOCAL_FIT)
      throw Error(TPSS_E_UNSUPPORTED, "fitted local solvers: Poisson levels only");
    // dof_info.templates.h:208: "TODO more than scalar elements"
    if(level->ncomp > 1 && desc->ras != 0)
      throw Error(TPSS_E_UNSUPPORTED, "restricted additive Schwarz: scalar (Poisson) levels only");
    auto sm = std::make_unique<tpss_smoother>();
    sm->level = level;
    sm->desc = *desc;
    sm->ls = build_local_solvers(level, *desc, nullptr);
    sm->seq = color_sequence(sm->ls->n_colors, desc->reverse != 0, desc->symmetrized != 0);
    sm->seq_t.assign(sm->seq.rbegin(), sm->seq.rend());
    sm->nonfinite.ensure();
    sm->timers[2].time = sm->ls->setup_seconds;
    sm->timers[2].n_calls = 1;
    *out = sm.release();
  });
}

int tpss_smoother_clone_shallow(const tpss_smoother *other, const tpss_smoother_desc *desc,
                                tpss_smoother **out)
{
  return guarded([&] {
    TPSS_REQUIRE(other && out, "null argument");
    validate_desc(desc);
    TPSS_REQUIRE(desc->coloring == other->desc.coloring,
                 "a shallow copy shares the patches and must keep the coloring");
    auto sm = std::make_unique<tpss_smoother>();
    sm->level = other->level;
    sm->desc = *desc;
    sm->ls = other->ls;
    sm->seq = color_sequence(sm->ls->n_colors, desc->reverse != 0, desc->symmetrized != 0);
    sm->seq_t.assign(sm->seq.rbegin(), sm->seq.rend());
    sm->check_finite = other->check_finite;
    sm->nonfinite.ensure();
    if(other->use_omega)
    {
      sm->d_omega.resize(other->d_omega.n);
      TPSS_HIP(hipMemcpy(sm->d_omega.ptr, other->d_omega.ptr, other->d_omega.n * sizeof(float),
                         hipMemcpyDeviceToDevice));
      sm->use_omega = true;
    }
    *out = sm.release();
  });
}

int tpss_smoother_destroy(tpss_smoother *sm)
{
  if(sm) resolve_timings(sm);
  delete sm;
  return TPSS_OK;
}

int tpss_precond_vmult(tpss_smoother *sm, void *dst, const void *src, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && dst && src, "null argument");
    TPSS_REQUIRE(dst != src, "dst and src must not alias");
    report_nonfinite(sm);
    PassCheck armed(sm);
    precond_vmult_any(sm, dst, src, as_stream(stream), false, 1.0);
    check_finite(sm, dst, as_stream(stream));
  });
}

int tpss_precond_vmult_add(tpss_smoother *sm, void *dst, const void *src, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && dst && src, "null argument");
    TPSS_REQUIRE(dst != src, "dst and src must not alias");
    report_nonfinite(sm);
    PassCheck armed(sm);
    precond_vmult_any(sm, dst, src, as_stream(stream), true, 1.0);
    check_finite(sm, dst, as_stream(stream));
  });
}

int tpss_precond_Tvmult(tpss_smoother *sm, void *dst, const void *src, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && dst && src, "null argument");
    TPSS_REQUIRE(dst != src, "dst and src must not alias");
    report_nonfinite(sm);
    PassCheck armed(sm);
    precond_vmult_any(sm, dst, src, as_stream(stream), false, 1.0, true);
    check_finite(sm, dst, as_stream(stream));
  });
}

int tpss_precond_Tvmult_add(tpss_smoother *sm, void *dst, const void *src, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && dst && src, "null argument");
    TPSS_REQUIRE(dst != src, "dst and src must not alias");
    report_nonfinite(sm);
    PassCheck armed(sm);
    precond_vmult_any(sm, dst, src, as_stream(stream), true, 1.0, true);
    check_finite(sm, dst, as_stream(stream));
  });
}

int tpss_smoother_step(tpss_smoother *sm, void *dst, const void *src, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && dst && src, "null argument");
    TPSS_REQUIRE(dst != src, "dst and src must not alias");
    report_nonfinite(sm);
    PassCheck armed(sm);
    smoother_step_any(sm, dst, src, as_stream(stream));
    check_finite(sm, dst, as_stream(stream));
  });
}

int tpss_smoother_vmult(tpss_smoother *sm, void *dst, const void *src, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && dst && src, "null argument");
    TPSS_REQUIRE(dst != src, "dst and src must not alias");
    report_nonfinite(sm);
    PassCheck armed(sm);
    fill_zero(dst, sm->level->vec_prec, sm->level->N, as_stream(stream));
    smoother_step_any(sm, dst, src, as_stream(stream), true);
    check_finite(sm, dst, as_stream(stream));
  });
}

int tpss_smoother_set_patch_damping(tpss_smoother *sm, const double *omega)
{
  return guarded([&] {
    TPSS_REQUIRE(sm, "null argument");
    if(!omega)
    {
      sm->use_omega = false;
      return;
    }
    std::vector<float> h(sm->ls->P);
    for(int i = 0; i < sm->ls->P; ++i)
    {
      TPSS_REQUIRE(omega[i] > 0.0 && std::isfinite(omega[i]), "patch damping must be positive");
      h[i] = float(omega[i]);
    }
    sm->d_omega.upload(h);
    sm->use_omega = true;
  });
}

int tpss_smoother_info(const tpss_smoother *sm, int *n_patches, int *n_colors, int *n_dofs_per_patch)
{
  return guarded([&] {
    TPSS_REQUIRE(sm, "null argument");
    if(n_patches) *n_patches = sm->ls->P;
    if(n_colors) *n_colors = sm->ls->n_colors;
    if(n_dofs_per_patch) *n_dofs_per_patch = sm->ls->ncomp * ipow(sm->ls->K, sm->ls->dim);
  });
}

int tpss_smoother_color_ptr(const tpss_smoother *sm, int *out)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && out, "null argument");
    std::copy(sm->ls->color_ptr.begin(), sm->ls->color_ptr.end(), out);
  });
}

int tpss_smoother_patch_vertices(const tpss_smoother *sm, int *out)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && out, "null argument");
    std::copy(sm->ls->vtx_unpacked.begin(), sm->ls->vtx_unpacked.end(), out);
  });
}

int tpss_smoother_patch_dof_indices(const tpss_smoother *sm, uint32_t *out)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && out, "null argument");
    patch_dof_indices_device(*sm->ls, out);
  });
}

int tpss_smoother_restricted_dof_flags(const tpss_smoother *sm, uint8_t *out)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && out, "null argument");
    const LocalSolvers &ls = *sm->ls;
    TPSS_REQUIRE(ls.ncomp == 1, "restricted DoF flags: scalar levels only");
    const int64_t total = int64_t(ls.P) * ipow(ls.K, ls.dim);
    DeviceArray<uint8_t> d_out{size_t(total)};
    dispatch_dim_np(ls.dim, ls.np, [&](auto D, auto NPc) {
      constexpr int DIM = decltype(D)::value, NP = decltype(NPc)::value;
      hipLaunchKernelGGL((k_restricted_flags<DIM, NP>), dim3(blocks_for(total, 256)), dim3(256), 0,
                         nullptr, ls.d_vtx.ptr, ls.P, ls.n, d_out.ptr);
      TPSS_HIP(hipGetLastError());
    });
    TPSS_HIP(hipMemcpy(out, d_out.ptr, size_t(total), hipMemcpyDeviceToHost));
  });
}

int tpss_smoother_patch_h(const tpss_smoother *sm, double *out)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && out, "null argument");
    std::copy(sm->ls->h_host.begin(), sm->ls->h_host.end(), out);
  });
}

int tpss_smoother_fit_errors(const tpss_smoother *sm, double *out)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && out, "null argument");
    TPSS_REQUIRE(!sm->ls->fit_err.empty(), "the smoother has no fitted local solvers");
    std::copy(sm->ls->fit_err.begin(), sm->ls->fit_err.end(), out);
  });
}

int tpss_smoother_get_timings(const tpss_smoother *sm, tpss_time_info *out, int *n)
{
  return guarded([&] {
    TPSS_REQUIRE(sm && n, "null argument");
    static const char *names[5] = {"[SchwarzPrecond] Update residuals:",
                                   "[SchwarzPrecond] Apply inverses:",
                                   "[SchwarzPrecond] Compute inverses:",
                                   "[SchwarzPrecond] Copy Locally Owned:",
                                   "[SchwarzPrecond] Step:"};
    const int cap = *n;
    *n = 5;
    resolve_timings(const_cast<tpss_smoother *>(sm));
    if(!out) return;
    for(int i = 0; i < std::min(cap, 5); ++i)
    {
      out[i].time = sm->timers[i].time;
      out[i].description = names[i];
      out[i].n_calls = sm->timers[i].n_calls;
    }
  });
}

int tpss_smoother_set_timing(tpss_smoother *sm, int enabled)
{
  return guarded([&] {
    TPSS_REQUIRE(sm, "null argument");
    resolve_timings(sm);
    if(enabled && !sm->timing)
      for(int i = 0; i < 5; ++i)
        if(i != 2) sm->timers[i] = Timer{};
    sm->timing = enabled != 0;
  });
}

int tpss_smoother_set_fused_residual(tpss_smoother *sm, int enabled)
{
  return guarded([&] {
    TPSS_REQUIRE(sm, "null argument");
    sm->fused_mvp = enabled != 0;
  });
}

int tpss_smoother_set_check_finite(tpss_smoother *sm, int mode)
{
  return guarded([&] {
    TPSS_REQUIRE(sm, "null argument");
    TPSS_REQUIRE(mode >= 0 && mode <= 2, "check mode must be 0, 1 or 2");
    sm->check_finite = mode;
  });
}

int tpss_smoother_finite_status(tpss_smoother *sm, void *stream)
{
  return guarded([&] {
    TPSS_REQUIRE(sm, "null argument");
    TPSS_HIP(hipStreamSynchronize(as_stream(stream)));
    report_nonfinite(sm);
  });
}

}  // extern "C"

namespace tpss
{
void patch_dof_indices_device(const LocalSolvers &ls, uint32_t *host_out)
{
  const int64_t total = int64_t(ls.P) * ipow(ls.K, ls.dim);
  DeviceArray<uint32_t> d_out{size_t(total)};
  dispatch_dim_np(ls.dim, ls.np, [&](auto D, auto NPc) {
    constexpr int DIM = decltype(D)::value, K = 2 * decltype(NPc)::value;
    hipLaunchKernelGGL((k_patch_dof_indices<DIM, K>), dim3(blocks_for(total, 256)), dim3(256), 0,
                       nullptr, ls.d_vtx.ptr, ls.P, ls.n, d_out.ptr);
    TPSS_HIP(hipGetLastError());
  });
  if(ls.ncomp == 1)
  {
    TPSS_HIP(hipMemcpy(host_out, d_out.ptr, size_t(total) * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return;
  }
  // component blocks juxtaposed per patch (PatchTransferBlock, patch_transfer.h:340-354)
  std::vector<uint32_t> base(static_cast<size_t>(total));
  TPSS_HIP(hipMemcpy(base.data(), d_out.ptr, size_t(total) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  const int ndp = ipow(ls.K, ls.dim);
  const int64_t Ns = int64_t(ipow(ls.n, ls.dim)) * ipow(ls.np, ls.dim);
  for(int64_t j = 0; j < ls.P; ++j)
    for(int c = 0; c < ls.ncomp; ++c)
      for(int i = 0; i < ndp; ++i)
        host_out[(j * ls.ncomp + c) * ndp + i] = uint32_t(base[j * ndp + i] + c * Ns);
}
}  // namespace tpss
