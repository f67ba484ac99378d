// Host check of bes::RcpDiv (kernels_common.hpp): every quotient must be the
// bitwise double of a / b, over random operands of every exponent, operands
// one ulp off an exact quotient, zeros of both signs, subnormals, infinities
// and NaN.  Prints the sample and mismatch counts.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../bes_amd/csrc/kernels_common.hpp"

static uint64_t state = 0x9E3779B97F4A7C15ull;
static uint64_t next() {
  state ^= state << 13;
  state ^= state >> 7;
  state ^= state << 17;
  return state;
}
// random double with a random mantissa and a binary exponent in [emin, emax]
static double rnd(int emin, int emax) {
  const uint64_t m = next() & ((1ull << 52) - 1);
  const int e = emin + (int)(next() % (uint64_t)(emax - emin + 1));
  uint64_t bits = ((uint64_t)(e + 1023) << 52) | m;
  if (next() & 1) bits |= 1ull << 63;
  double d;
  std::memcpy(&d, &bits, 8);
  return d;
}
static bool same(double x, double y) {
  uint64_t a, b;
  std::memcpy(&a, &x, 8);
  std::memcpy(&b, &y, 8);
  return a == b || (x != x && y != y);
}

int main(int argc, char **argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 10000000L;
  long bad = 0, cnt = 0;
  auto check = [&](double a, double b) {
    const bes::RcpDiv d(b);
    const double q = a / b, r = d(a);
    cnt++;
    if (!same(q, r)) {
      if (bad < 5) std::printf("a=%.17g b=%.17g a/b=%.17g got=%.17g\n", a, b, q, r);
      bad++;
    }
  };
  const double special[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 5e-324, -2.5e-320, 1e-310, 1e308, -1.7e308};
  for (double a : special)
    for (int k = 0; k < 1000; k++) check(a, rnd(-60, 60));
  for (double b : special)
    for (int k = 0; k < 1000; k++) check(rnd(-60, 60), b);
  for (long k = 0; k < n; k++) {   // every exponent, and the usual ranges more often
    const double a = (k & 7) ? rnd(-60, 60) : rnd(-1022, 1023);
    const double b = (k & 15) ? rnd(-60, 60) : rnd(-1022, 1023);
    check(a, b);
  }
  for (long k = 0; k < n / 4; k++) {   // a one ulp off an exact multiple of b
    const double b = rnd(-30, 30), t = rnd(-5, 5);
    check(std::nextafter(b * t, (next() & 1) ? INFINITY : -INFINITY), b);
  }
  std::printf("samples %ld mismatches %ld\n", cnt, bad);
  return bad != 0;
}
