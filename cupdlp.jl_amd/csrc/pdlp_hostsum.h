// pdlp_hostsum.h — compensated sums on the host.  Same arithmetic as the
// device's: error-free products (TwoProd with fma) and TwoSum, errors added in
// FP64, rounded once (Dot2: error <= eps |sum| + gamma_n^2 sum |terms|), so
// host and device sums of the same terms agree in any order unless the sum
// cancels to ~n^2 eps of its terms' magnitude.
#pragma once
#include <math.h>
#include <stddef.h>

#include <vector>

namespace pdlp {

struct HostDD {
  double s = 0.0, e = 0.0;
  void add(double v) {
    const double t = s + v;
    const double bb = t - s;
    const double err = (s - (t - bb)) + (v - bb);
    s = t;
    e += err;
  }
  void add_prod(double x, double y) {
    const double p = x * y;
    const double q = fma(x, y, -p);
    add(p);
    e += q;
  }
  double value() const { return s + e; }
};

inline double dd_norm2(const std::vector<double> &v) {
  HostDD a;
  for (double x : v) a.add_prod(x, x);
  return sqrt(a.value());
}

}  // namespace pdlp
