// Host-side use of the C++ mirror (include/tpss_schwarz.hpp): one smoother
// vmult and one preconditioner vmult on a seeded vector, written to argv[1]
// for comparison with the oracle.
#include <cstdio>
#include <vector>

#include "tpss_schwarz.hpp"

int main(int argc, char **argv)
{
  if(argc < 2) return 2;
  const int dim = 3, degree = 2, n = 6;
  tpss::Level<double> level(dim, degree, n);
  tpss::AdditionalData data;  // multiplicative, red-black colors, relaxation 1
  tpss::SchwarzSmoother<double> smoother;
  smoother.initialize(level, data);
  tpss::SchwarzPreconditioner<double> post;
  tpss::AdditionalData rev = data;
  rev.reverse = true;
  post.initialize(smoother, rev);

  const auto N = std::size_t(level.n_dofs());
  tpss::DeviceVector<double> src(N), dst(N), dst2(N);
  tpss::check(tpss_vector_fill_random(src.data(), std::int64_t(N), TPSS_F64, 21, nullptr));
  smoother.vmult(dst, src);
  post.vmult(dst2, src);
  const auto a = dst.download(), b = dst2.download();
  FILE *f = std::fopen(argv[1], "wb");
  if(!f) return 3;
  std::fwrite(a.data(), sizeof(double), N, f);
  std::fwrite(b.data(), sizeof(double), N, f);
  std::fclose(f);
  std::printf("ok %zu\n", N);
  return 0;
}
