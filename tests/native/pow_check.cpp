// Accuracy of pow_pos (acme-ecp_amd/csrc/crm_pow.hpp) against the C library's pow, which
// glibc rounds correctly: prints the largest error in ulp over seeded samples of the
// microphysics' exponents and of x in [1e-300, 1e3], y in (0, 2], then the special cases.
#include <cmath>
#include <cstdio>
#include <random>

#include "crm_pow.hpp"

static double ulps(double a, double b) {
  if (a == b) return 0;
  double u = std::nextafter(std::fabs(b), INFINITY) - std::fabs(b);
  return std::fabs(a - b) / u;
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 2000000;
  std::mt19937_64 g(1);
  std::uniform_real_distribution<double> U(0, 1);
  // terminal velocities (b/4), accretion ((3+b)/4), evaporation ((5+b)/8), ice fall
  const double ys[] = {0.2, 0.0275, 0.125, 0.95, 0.7775, 0.875, 0.725, 0.63875, 0.6875, 0.24, 1.5, 1.94};
  double worst = 0;
  for (long t = 0; t < n; t++) {
    double x = (t % 3 == 0) ? std::pow(10.0, -14 + 14 * U(g)) : std::pow(10.0, -300 + 303 * U(g));
    double y = (t % 2) ? ys[t % 12] : 2 * U(g) + 1e-300;
    worst = std::fmax(worst, ulps(crm::pow_pos(x, y), std::pow(x, y)));
  }
  for (int k = -1074; k <= 10; k++) {  // every power of two, subnormals included
    double x = std::ldexp(1.0, k);
    for (double y : ys) worst = std::fmax(worst, ulps(crm::pow_pos(x, y), std::pow(x, y)));
  }
  printf("worst_ulp %.4f\n", worst);
  printf("zero %a\ninf %a\nneg %a\nnan %a\none %a\n", crm::pow_pos(0.0, 0.2), crm::pow_pos(INFINITY, 0.2),
         crm::pow_pos(-1.0, 0.2), crm::pow_pos(NAN, 0.2), crm::pow_pos(1.0, 0.7));
  return 0;
}
