// tpss_schwarz.hpp -- header-only C++ host mirror of the reference interface
// on top of the C ABI in tpss.h.
//
//   tpss::Level                    ~ Laplace::DG::MF::Operator (vmult)
//                                    apps/laplace_integrator.h:578-594
//   tpss::SchwarzPreconditioner    ~ SchwarzPreconditioner<dim,Op,Vec,Mat>::vmult
//                                    preconditioner/schwarz_preconditioner.h:87-199
//   tpss::SchwarzSmoother          ~ SmootherRichardson / SchwarzSmoother::{vmult,step}
//                                    smoother/smoother_richardson.h:73-98
//   tpss::GMGSolver                ~ ModelProblem::solve with the GMG preconditioner
//                                    apps/poisson_problem.h:726-755
//
// Vectors are device buffers (tpss::DeviceVector<Number> owns one); every call
// throws tpss::Error with the library's status and message on failure, like
// the reference's AssertThrow.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "tpss.h"

namespace tpss
{
struct Error : std::runtime_error
{
  int status;
  Error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void check(int status)
{
  if(status != TPSS_OK) throw Error(status, tpss_last_error());
}

template<typename Number>
struct precision_of;
template<>
struct precision_of<double>
{
  static constexpr int value = TPSS_F64;
};
template<>
struct precision_of<float>
{
  static constexpr int value = TPSS_F32;
};

template<typename Number>
class DeviceVector
{
public:
  explicit DeviceVector(std::size_t n = 0) { reinit(n); }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  DeviceVector(DeviceVector &&o) noexcept : ptr_(o.ptr_), n_(o.n_) { o.ptr_ = nullptr, o.n_ = 0; }
  ~DeviceVector() { tpss_device_free(ptr_); }
  void reinit(std::size_t n)
  {
    tpss_device_free(ptr_);
    ptr_ = nullptr;
    n_ = n;
    check(tpss_device_malloc(&ptr_, n * sizeof(Number)));
  }
  std::size_t size() const { return n_; }
  Number *data() { return static_cast<Number *>(ptr_); }
  const Number *data() const { return static_cast<const Number *>(ptr_); }
  void upload(const std::vector<Number> &h)
  {
    if(h.size() != n_) throw Error(TPSS_E_INVALID_ARG, "size mismatch");
    check(tpss_memcpy_h2d(ptr_, h.data(), n_ * sizeof(Number)));
  }
  std::vector<Number> download() const
  {
    std::vector<Number> h(n_);
    check(tpss_memcpy_d2h(h.data(), ptr_, n_ * sizeof(Number)));
    return h;
  }

private:
  void *ptr_ = nullptr;
  std::size_t n_ = 0;
};

template<typename Number>
class Level
{
public:
  Level(int dim, int degree, int n_cells_1d, double ip_factor = 1.0,
        const double *vertices = nullptr)
  {
    tpss_level_desc d{dim, degree, n_cells_1d, ip_factor, vertices, precision_of<Number>::value,
                      TPSS_POISSON, 0.0, 0.0};
    check(tpss_level_create(&d, &h_));
  }
  // linear elasticity with Lame coefficients mu, lambda (LinElasticity::ModelProblem)
  Level(int dim, int degree, int n_cells_1d, double ip_factor, double mu, double lambda)
  {
    tpss_level_desc d{dim, degree, n_cells_1d, ip_factor, nullptr, precision_of<Number>::value,
                      TPSS_ELASTICITY, mu, lambda};
    check(tpss_level_create(&d, &h_));
  }
  Level(const Level &) = delete;
  ~Level() { tpss_level_destroy(h_); }
  std::int64_t n_dofs() const
  {
    std::int64_t n = 0;
    check(tpss_level_n_dofs(h_, &n));
    return n;
  }
  void vmult(DeviceVector<Number> &dst, const DeviceVector<Number> &src) const
  {
    check(tpss_operator_vmult(h_, dst.data(), src.data(), nullptr));
  }
  tpss_level *handle() const { return h_; }

private:
  tpss_level *h_ = nullptr;
};

// AdditionalData of the reference (schwarz_preconditioner.h:495-515) plus the
// Richardson data (smoother_richardson.h:41-48) and the patch-solve precision
struct AdditionalData
{
  int variant = TPSS_MULTIPLICATIVE;
  int coloring = TPSS_COLOR_REDBLACK;
  double relaxation = 1.0;
  bool reverse = false;
  bool symmetrized = false;
  double damping_factor = 1.0;
  int n_steps = 1;
  int local_precision = -1;  // -1: the level's vector precision
  int local_solver = TPSS_LOCAL_SURROGATE;
  int fit_sweeps = 0;
  // use_ras_weights / use_ras_boolean_weights (schwarz_preconditioner.h:501-502)
  bool use_ras_weights = false;
  bool use_ras_boolean_weights = false;

  tpss_smoother_desc desc(int vec_prec) const
  {
    return tpss_smoother_desc{variant,       coloring,       relaxation,
                              int(reverse),  int(symmetrized), damping_factor,
                              n_steps,       local_precision < 0 ? vec_prec : local_precision,
                              local_solver,  fit_sweeps,
                              use_ras_weights ? TPSS_RAS_WEIGHTS
                                              : use_ras_boolean_weights ? TPSS_RAS_BOOLEAN : 0};
  }
};

template<typename Number>
class SchwarzPreconditioner
{
public:
  using value_type = Number;
  using VectorType = DeviceVector<Number>;

  SchwarzPreconditioner() = default;
  SchwarzPreconditioner(const SchwarzPreconditioner &) = delete;
  ~SchwarzPreconditioner() { tpss_smoother_destroy(h_); }

  // initialize(subdomain_handler, operator, data) (schwarz_preconditioner.templates.h:3-60)
  void initialize(const Level<Number> &level, const AdditionalData &data)
  {
    clear();
    auto d = data.desc(precision_of<Number>::value);
    check(tpss_smoother_create(level.handle(), &d, &h_));
  }
  // initialize(other, data): shallow copy sharing the local solvers (:63-86)
  void initialize(const SchwarzPreconditioner &other, const AdditionalData &data)
  {
    clear();
    auto d = data.desc(precision_of<Number>::value);
    check(tpss_smoother_clone_shallow(other.h_, &d, &h_));
  }
  void clear()
  {
    tpss_smoother_destroy(h_);
    h_ = nullptr;
  }
  // vmult (schwarz_preconditioner.templates.h:209-239); dst is overwritten
  void vmult(VectorType &dst, const VectorType &src) const
  {
    check(tpss_precond_vmult(h_, dst.data(), src.data(), nullptr));
  }
  // vmult_add / Tvmult / Tvmult_add (schwarz_preconditioner.h:185-194)
  void vmult_add(VectorType &dst, const VectorType &src) const
  {
    check(tpss_precond_vmult_add(h_, dst.data(), src.data(), nullptr));
  }
  void Tvmult(VectorType &dst, const VectorType &src) const
  {
    check(tpss_precond_Tvmult(h_, dst.data(), src.data(), nullptr));
  }
  void Tvmult_add(VectorType &dst, const VectorType &src) const
  {
    check(tpss_precond_Tvmult_add(h_, dst.data(), src.data(), nullptr));
  }
  std::vector<tpss_time_info> get_time_data() const
  {
    std::vector<tpss_time_info> t(8);
    int n = int(t.size());
    check(tpss_smoother_get_timings(h_, t.data(), &n));
    t.resize(n);
    return t;
  }
  tpss_smoother *handle() const { return h_; }

protected:
  tpss_smoother *h_ = nullptr;
};

template<typename Number>
class SchwarzSmoother : public SchwarzPreconditioner<Number>
{
public:
  using VectorType = DeviceVector<Number>;
  // SmootherRichardson::step (smoother_richardson.h:118-145)
  void step(VectorType &dst, const VectorType &src) const
  {
    check(tpss_smoother_step(this->h_, dst.data(), src.data(), nullptr));
  }
  // SmootherRichardson::vmult (smoother_richardson.h:147-156)
  void vmult(VectorType &dst, const VectorType &src) const
  {
    check(tpss_smoother_vmult(this->h_, dst.data(), src.data(), nullptr));
  }
};

class GMGSolver
{
public:
  explicit GMGSolver(const tpss_gmg_desc &d) { check(tpss_gmg_create(&d, &h_)); }
  GMGSolver(const GMGSolver &) = delete;
  ~GMGSolver() { tpss_gmg_destroy(h_); }
  tpss_solve_info solve(DeviceVector<double> &x, const DeviceVector<double> &b, double rel_tol,
                        int max_iter = 200) const
  {
    tpss_solve_info info{};
    check(tpss_gmg_solve(h_, x.data(), b.data(), rel_tol, max_iter, nullptr, &info, nullptr));
    return info;
  }

private:
  tpss_gmg *h_ = nullptr;
};

}  // namespace tpss
