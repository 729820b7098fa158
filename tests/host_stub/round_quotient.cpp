// RoundQuotient (fvs_physics.hpp) on the host: every quotient perturbed by
// one ulp either way must come back as the correctly rounded a / b.
#include "fvs_physics.hpp"

#include <cmath>
#include <cstdio>
#include <random>

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> m(1.0, 2.0);
  std::uniform_int_distribution<int> e(-60, 60), pert(-1, 1), sg(0, 1);
  long bad = 0;
  const long n = 2000000;
  for (long i = 0; i < n; ++i) {
    const double a = std::ldexp(m(rng), e(rng)) * (sg(rng) ? -1 : 1);
    const double b = std::ldexp(m(rng), e(rng)) * (sg(rng) ? -1 : 1);
    const double q = a / b;
    const int p = pert(rng);
    const double qp = p == 0 ? q : std::nextafter(q, p > 0 ? INFINITY : -INFINITY);
    if (fvs::RoundQuotient(a, b, qp) != q) ++bad;
  }
  // the quotient the gfx950 division sequence rounds the wrong way
  const double a = 0x1.6666666666663p+1, b = 0x1.ffffffffffffbp+0;
  if (fvs::RoundQuotient(a, b, 0x1.6666666666666p+0) != 0x1.6666666666667p+0) ++bad;
  // zero, infinite and NaN quotients pass through
  if (fvs::RoundQuotient(0.0, 3.0, 0.0) != 0.0) ++bad;
  if (!std::isinf(fvs::RoundQuotient(1.0, 0.0, INFINITY))) ++bad;
  if (!std::isnan(fvs::RoundQuotient(0.0, 0.0, NAN))) ++bad;
  if (fvs::RoundQuotient(2.0, INFINITY, 0.0) != 0.0) ++bad;
  std::printf("%ld\n", bad);
  return bad != 0;
}
