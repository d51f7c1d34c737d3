// common.hpp -- shared host/device definitions of libtpss_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tpss.h"

namespace tpss
{
// ------------------------------------------------------------------ errors

struct Error : std::runtime_error
{
  int code;
  Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string &msg);

#define TPSS_HIP(call)                                                                     \
  do                                                                                       \
  {                                                                                        \
    hipError_t e_ = (call);                                                                \
    if(e_ != hipSuccess)                                                                   \
      throw ::tpss::Error(TPSS_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while(0)

#define TPSS_REQUIRE(cond, msg)                                                            \
  do                                                                                       \
  {                                                                                        \
    if(!(cond))                                                                            \
      throw ::tpss::Error(TPSS_E_INVALID_ARG, msg);                                        \
  } while(0)

// Runs f, converting exceptions into a status code plus last-error message.
template<typename F>
int guarded(F &&f)
{
  try
  {
    f();
    return TPSS_OK;
  }
  catch(const Error &e)
  {
    set_last_error(e.what());
    return e.code;
  }
  catch(const std::exception &e)
  {
    set_last_error(e.what());
    return TPSS_E_INVALID_ARG;
  }
}

// ------------------------------------------------------------------ device memory

template<typename T>
struct DeviceArray
{
  T *ptr = nullptr;
  size_t n = 0;
  DeviceArray() = default;
  explicit DeviceArray(size_t count) { resize(count); }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  DeviceArray(DeviceArray &&o) noexcept : ptr(o.ptr), n(o.n)
  {
    o.ptr = nullptr;
    o.n = 0;
  }
  DeviceArray &operator=(DeviceArray &&o) noexcept
  {
    if(this != &o)
    {
      release();
      ptr = o.ptr;
      n = o.n;
      o.ptr = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DeviceArray() { release(); }
  void resize(size_t count)
  {
    if(count == n) return;
    release();
    if(count)
    {
      void *p = nullptr;
      TPSS_HIP(hipMalloc(&p, count * sizeof(T)));
      ptr = static_cast<T *>(p);
    }
    n = count;
  }
  void release()
  {
    if(ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  void upload(const std::vector<T> &h)
  {
    resize(h.size());
    if(!h.empty()) TPSS_HIP(hipMemcpy(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  std::vector<T> download() const
  {
    std::vector<T> h(n);
    if(n) TPSS_HIP(hipMemcpy(h.data(), ptr, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
};

// ------------------------------------------------------------------ 1D basis (host, FP64)

constexpr int MAXNP = 6;  // degree <= 5
constexpr int MAXK = 2 * MAXNP;

struct Basis1D
{
  int p = 0, np = 0, nq = 0;
  double nodes[MAXNP], qx[MAXNP], qw[MAXNP];
  double val[MAXNP * MAXNP], grad[MAXNP * MAXNP];  // [i*nq+q]
  double fval[MAXNP * 2], fgrad[MAXNP * 2];        // [i*2+face]
};

Basis1D make_basis(int p);
void lagrange_eval(int np, const double *nodes, double x, double *val, double *der);
void gauss_lobatto_nodes(int np, double *x, double *w);
double nitsche_penalty(int p, double h_left, double h_right);

// 1D matrices of the Cartesian SIPG operator on one cell (row-major np x np)
struct Operator1D
{
  double M[MAXNP * MAXNP];      // cell mass
  double D[4][MAXNP * MAXNP];   // K + left-face + right-face; index = has_left | has_right<<1
  double Fpm[MAXNP * MAXNP];    // coupling to the left neighbour
  double Fmp[MAXNP * MAXNP];    // coupling to the right neighbour
};
Operator1D make_operator_1d(int p, double h, double ip);

// prolongation coarse -> two children: P[(c*np+i)*np+j] = phi_j((x_i+c)/2)
void prolongation_1d(int p, double *P);

// ------------------------------------------------------------------ timing

struct Timer
{
  double time = 0.0;
  long n_calls = 0;
  const char *description = "";
};

inline int ipow(int b, int e)
{
  int r = 1;
  while(e-- > 0) r *= b;
  return r;
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace tpss

// ------------------------------------------------------------------ opaque handles

struct tpss_level
{
  int dim = 0, p = 0, np = 0, n = 0, ndc = 0;
  int64_t ncell = 0, N = 0;
  int ncomp = 1;          // 1: Poisson; dim: linear elasticity (component blocks)
  int64_t Ns = 0;         // DoFs of one component block (ncell * ndc)
  double mu = 0.0, lambda = 0.0;
  double ip = 1.0, h = 0.0;
  int vec_prec = TPSS_F64;
  bool distorted = false;
  std::vector<double> vertices;            // host copy ((n+1)^dim * dim)
  tpss::DeviceArray<double> d_vertices;
  tpss::Operator1D op1d;
};
