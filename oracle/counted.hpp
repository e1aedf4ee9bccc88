// counted.hpp -- TEST INFRASTRUCTURE ONLY: a double that counts the FP64
// operations applied to it (add/sub, mul, div, sqrt, compare-select), for
// instantiating the oracle to measure the FP64 work of one cell-update of
// the reference algorithm. Counters are per thread; run single-threaded.
#pragma once
#include <cmath>

namespace ora {

struct OpCount {
  unsigned long long add = 0, mul = 0, div = 0, sqrt = 0;
};
inline OpCount& op_count() {
  static thread_local OpCount c;
  return c;
}

struct Counted {
  double v = 0.0;
  Counted() = default;
  Counted(double x) : v(x) {}  // NOLINT: implicit, like a double literal
};

inline double pv(const Counted& a) { return a.v; }
inline Counted operator-(const Counted& a) { return Counted(-a.v); }
inline Counted operator+(const Counted& a, const Counted& b) { op_count().add++; return Counted(a.v + b.v); }
inline Counted operator-(const Counted& a, const Counted& b) { op_count().add++; return Counted(a.v - b.v); }
inline Counted operator*(const Counted& a, const Counted& b) { op_count().mul++; return Counted(a.v * b.v); }
inline Counted operator/(const Counted& a, const Counted& b) { op_count().div++; return Counted(a.v / b.v); }
inline Counted operator+(double a, const Counted& b) { return Counted(a) + b; }
inline Counted operator-(double a, const Counted& b) { return Counted(a) - b; }
inline Counted operator*(double a, const Counted& b) { return Counted(a) * b; }
inline Counted operator/(double a, const Counted& b) { return Counted(a) / b; }
inline Counted operator+(const Counted& a, double b) { return a + Counted(b); }
inline Counted operator-(const Counted& a, double b) { return a - Counted(b); }
inline Counted operator*(const Counted& a, double b) { return a * Counted(b); }
inline Counted operator/(const Counted& a, double b) { return a / Counted(b); }
inline Counted sqrt(const Counted& a) { op_count().sqrt++; return Counted(std::sqrt(a.v)); }
inline Counted fabs(const Counted& a) { return Counted(std::fabs(a.v)); }

}  // namespace ora
