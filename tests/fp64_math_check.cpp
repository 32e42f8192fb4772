// Host check of the FP64 routines of the exact-range kernel
// (srg_amd/csrc/tdbp_geom.h): sincos_2pi against long-double sin/cos of
// 2*pi*f, goldschmidt_sqrt against the correctly rounded sqrt from
// reciprocal-root estimates 1e-7 off.  Prints one line of figures; exits 1
// when a bound fails.
#include <cmath>
#include <cstdio>
#include <random>

#include "../srg_amd/csrc/tdbp_geom.h"

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> uf(-0.5, 0.5);
  const long double pi = 3.141592653589793238462643383279502884L;
  double es = 0.0, ec = 0.0;
  for (int i = 0; i < 4000000; ++i) {
    // The quadrant boundaries and the ends first, then random phases.
    const double f = i < 17 ? -0.5 + 0.0625 * i : uf(rng);
    double s, c;
    tdbp::sincos_2pi(f, &s, &c);
    const long double a = 2.0L * pi * (long double)f;
    es = std::fmax(es, std::fabs((double)((long double)s - sinl(a))));
    ec = std::fmax(ec, std::fabs((double)((long double)c - cosl(a))));
  }
  std::uniform_real_distribution<double> ur(1.0e5, 4.0e7);
  std::uniform_real_distribution<double> ue(-1e-7, 1e-7);
  double er = 0.0;
  long exact = 0;
  const long n = 2000000;
  for (long i = 0; i < n; ++i) {
    const double R = ur(rng), x = R * R;
    const double y0 = (1.0 / std::sqrt(x)) * (1.0 + ue(rng));
    const double g = tdbp::goldschmidt_sqrt(x, y0), t = std::sqrt(x);
    er = std::fmax(er, std::fabs(g - t) / t);
    exact += g == t;
  }
  const double frac = (double)exact / (double)n;
  std::printf("sincos_2pi max err sin %.3e cos %.3e; goldschmidt_sqrt max rel err %.3e, "
              "correctly rounded %.6f\n", es, ec, er, frac);
  return (es <= 1e-11 && ec <= 1e-11 && er <= 2.3e-16 && frac >= 0.999) ? 0 : 1;
}
