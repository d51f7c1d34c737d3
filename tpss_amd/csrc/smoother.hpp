// smoother.hpp -- internal state of the Schwarz smoother shared by the
// translation units of libtpss_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "common.hpp"

namespace tpss
{
// Patches, colors and FP64/FP32 local-solver factors of one level; shared
// between shallow copies (SchwarzPreconditioner::initialize(other, data)).
struct LocalSolvers
{
  int dim = 0, p = 0, np = 0, K = 0, n = 0, P = 0, n_colors = 0, coloring = 0;
  int ncomp = 1;                   // component blocks (block-diagonal local solvers)
  std::vector<int> color_ptr;
  std::vector<uint32_t> vtx;       // packed vertex per patch
  std::vector<int> vtx_unpacked;   // P * dim
  DeviceArray<uint32_t> d_vtx;
  std::vector<uint32_t> fac;       // ncomp * P * dim: row (c * P + patch) * dim + direction
  DeviceArray<uint32_t> d_fac;
  int nfac = 0;
  std::vector<double> prob_h;      // nfac * 2
  std::vector<uint8_t> prob_b;     // nfac * 2
  std::vector<double> prob_coef;   // nfac * 2: (a_cons, a_pen) of the 1D matrix
  DeviceArray<double> S64, lam64;  // nfac * K * K, nfac * K
  DeviceArray<float> S32, lam32;
  DeviceArray<uint32_t> d_eo;      // nfac: 1 = eigenpairs in even-odd order
  // vtx and fac of every color in the launch order of the residual-free pass
  // (3D, K = 8, Cartesian): each color's patches dealt to the 8 XCDs in column
  // tiles (y bands for FP64 patch solves, 4 x 2 tiles for FP32), z-slab by
  // z-slab inside a tile (empty where that pass does not run)
  DeviceArray<uint32_t> d_vtx_launch, d_fac_launch;
  std::vector<double> h_host;      // P * dim * 2 surrogate lengths
  std::vector<double> fit_err;     // P: relative error of the separable fit (-1: surrogate kept)
  std::vector<std::shared_ptr<DeviceArray<uint32_t>>> color_cells;
  std::vector<int64_t> color_ncells;
  double setup_seconds = 0.0;
};

void patch_dof_indices_device(const LocalSolvers &ls, uint32_t *host_out);

// One int in pinned host memory that kernels set through its device address:
// the non-finite check reports without a host synchronisation.
struct PinnedFlag
{
  int *host = nullptr;
  int *dev = nullptr;
  PinnedFlag() = default;
  PinnedFlag(const PinnedFlag &) = delete;
  PinnedFlag &operator=(const PinnedFlag &) = delete;
  ~PinnedFlag()
  {
    if(host) (void)hipHostFree(host);
  }
  void ensure()
  {
    if(host) return;
    void *h = nullptr;
    TPSS_HIP(hipHostMalloc(&h, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    host = static_cast<int *>(h);
    *host = 0;
    void *d = nullptr;
    TPSS_HIP(hipHostGetDevicePointer(&d, h, 0));
    dev = static_cast<int *>(d);
  }
};

}  // namespace tpss

struct tpss_smoother
{
  tpss_level *level = nullptr;
  tpss_smoother_desc desc{};
  std::shared_ptr<tpss::LocalSolvers> ls;
  std::vector<int> seq;    // color sequence of vmult (schwarz_preconditioner.h:531-567)
  std::vector<int> seq_t;  // of Tvmult: seq reversed
  tpss::DeviceArray<float> d_omega;
  bool use_omega = false;
  bool timing = false;
  // non-finite check of every output: 0 off, 1 on and reported at the next
  // entry point or by tpss_smoother_finite_status (default), 2 on with a
  // synchronisation in the call itself
  int check_finite = 1;
  tpss::PinnedFlag nonfinite;
  // set by an entry point for its own call: the residual-free color passes
  // check the values they store and set this flag (the device address of
  // `nonfinite`); pass_checked records that they covered the whole output, so
  // that check_finite needs no separate pass over it
  int *pass_flag = nullptr;
  bool pass_checked = false;
  bool fused_mvp = true;  // residual-free color passes where the mesh allows them
  tpss::Timer timers[5];
  struct PendingSection
  {
    int idx;
    hipEvent_t a, b;
  };
  std::vector<PendingSection> pending;
  // scratch (bytes)
  tpss::DeviceArray<char> res, tmp2, rich_r, ubuf;
  // caller-provided device scratch for the current call, used instead of res /
  // ubuf when set (the allocation-free entry point)
  char *scratch = nullptr;
};

namespace tpss
{
void precond_vmult_any(tpss_smoother *sm, void *dst, const void *src, hipStream_t s,
                       bool accumulate, double alpha, bool transpose = false);
size_t precond_scratch_bytes(const tpss_smoother *sm);
void report_nonfinite(tpss_smoother *sm);
void check_finite(tpss_smoother *sm, const void *x, hipStream_t s);

// arms the in-pass non-finite check of the residual-free color passes for the
// duration of one entry-point call
struct PassCheck
{
  tpss_smoother *sm;
  explicit PassCheck(tpss_smoother *s) : sm(s)
  {
    sm->pass_checked = false;
    sm->pass_flag = sm->check_finite ? sm->nonfinite.dev : nullptr;
  }
  PassCheck(const PassCheck &) = delete;
  PassCheck &operator=(const PassCheck &) = delete;
  ~PassCheck() { sm->pass_flag = nullptr; }
};
void smoother_step_any(tpss_smoother *sm, void *dst, const void *src, hipStream_t s,
                       bool x_is_zero = false);
template<typename VT>
void precond_vmult_bf16(tpss_smoother *sm, VT *dst, const VT *src, hipStream_t s, bool accumulate,
                        double alpha, const std::vector<int> &seq);
}  // namespace tpss
