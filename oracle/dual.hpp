// dual.hpp -- TEST INFRASTRUCTURE ONLY: forward-mode dual numbers (value and
// one tangent) for instantiating the oracle as the tangent-linear model of the
// reference routines. Branches and MAX/MIN/SIGN follow the primal values (the
// oracle compares pv(x)), so the tangent is the derivative along the branch the
// primal took; sqrt at 0 has a zero tangent (the convention the adjoint uses).
#pragma once
#include <cmath>

namespace ora {

struct Dual {
  double v = 0.0, t = 0.0;
  Dual() = default;
  Dual(double x) : v(x), t(0.0) {}  // NOLINT: implicit, like a double literal
  Dual(double x, double dx) : v(x), t(dx) {}
};

inline double pv(const Dual& a) { return a.v; }
inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.t); }
inline Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.t + b.t); }
inline Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.t - b.t); }
inline Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.t * b.v + a.v * b.t); }
inline Dual operator/(const Dual& a, const Dual& b) {
  const double q = a.v / b.v;
  return Dual(q, (a.t - q * b.t) / b.v);
}
inline Dual operator+(double a, const Dual& b) { return Dual(a) + b; }
inline Dual operator-(double a, const Dual& b) { return Dual(a) - b; }
inline Dual operator*(double a, const Dual& b) { return Dual(a * b.v, a * b.t); }
inline Dual operator/(double a, const Dual& b) { return Dual(a) / b; }
inline Dual operator+(const Dual& a, double b) { return a + Dual(b); }
inline Dual operator-(const Dual& a, double b) { return a - Dual(b); }
inline Dual operator*(const Dual& a, double b) { return Dual(a.v * b, a.t * b); }
inline Dual operator/(const Dual& a, double b) { return Dual(a.v / b, a.t / b); }
inline Dual sqrt(const Dual& a) {
  const double r = std::sqrt(a.v);
  return Dual(r, r > 0.0 ? 0.5 * a.t / r : 0.0);
}
inline Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }

}  // namespace ora
