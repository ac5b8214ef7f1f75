// Checks the select-form helpers of kah_internal.h against the formulas they replace, on
// special values and random operands (run by tests/test_point_math.py, host build only).
#define __host__
#define __device__
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include "../../kitamr.jl_amd/csrc/kah_internal.h"

static bool same(double a, double b) {   // equal values, or both NaN (zeros of any sign)
  if (a != a || b != b) return a != a && b != b;
  return a == b;
}

int main() {
  using kah::minmod;
  using kah::minmod_sel;
  const double inf = std::numeric_limits<double>::infinity();
  const double nan = std::numeric_limits<double>::quiet_NaN();
  const double den = std::numeric_limits<double>::denorm_min();
  const double sp[] = {0.0, -0.0, 1.0, -1.0, 0.5, -2.0, 1e-300, -1e-300, den, -den,
                       1e300, -1e300, inf, -inf, nan, 3.0, -3.0};
  long bad = 0, n = 0;
  for (double a : sp)
    for (double b : sp) {
      ++n;
      if (!same(minmod(a, b), minmod_sel(a, b))) {
        ++bad;
        std::printf("minmod(%g, %g): %g vs %g\n", a, b, minmod(a, b), minmod_sel(a, b));
      }
    }
  std::mt19937_64 rng(20261202);
  std::uniform_int_distribution<uint64_t> bits;
  std::normal_distribution<double> nd;
  for (int t = 0; t < 2000000; ++t) {
    double a, b;
    if (t & 1) {
      uint64_t x = bits(rng), y = bits(rng);
      std::memcpy(&a, &x, 8);
      std::memcpy(&b, &y, 8);
    } else {
      a = nd(rng);
      b = (t & 2) ? -nd(rng) * 1e-3 : nd(rng);
    }
    ++n;
    if (!same(minmod(a, b), minmod_sel(a, b))) ++bad;
  }
  std::printf("minmod_sel: %ld of %ld differ\n", bad, n);
  // div_by (the reciprocal, a product and one FMA step on the exact residual) against IEEE
  // division: quotients of operands whose exponents stay clear of overflow and underflow, as
  // the kernels' divisions by cell sizes, 2 lambda and the shave's d are
  long dbad = 0, dn = 0;
  std::uniform_int_distribution<int> ea(-120, 40), eb(-40, 10);
  std::uniform_int_distribution<uint64_t> mant(0, (uint64_t(1) << 52) - 1);
  auto draw = [&](int e) {
    double x = std::ldexp(1.0 + std::ldexp(double(mant(rng)), -52), e);
    return (bits(rng) & 1) ? -x : x;
  };
  for (int t = 0; t < 20000000; ++t) {
    double a = (t % 7 == 0) ? double(bits(rng) % 1000) * 0.001 - 0.5 : draw(ea(rng));
    double b = draw(eb(rng));
    ++dn;
    if (!same(kah::div_by(a, b, 1.0 / b), a / b)) {
      if (dbad < 5) std::printf("div_by(%a, %a) = %a, a/b = %a\n", a, b,
                                kah::div_by(a, b, 1.0 / b), a / b);
      ++dbad;
    }
  }
  std::printf("div_by: %ld of %ld differ from a/b\n", dbad, dn);
  return (bad || dbad) ? 1 : 0;
}
