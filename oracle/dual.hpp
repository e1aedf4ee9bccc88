// dual.hpp -- TEST INFRASTRUCTURE ONLY: forward-mode dual numbers (value and
// one tangent) for instantiating the oracle as the tangent-linear model of the
// reference routines. Branches and MAX/MIN/SIGN follow the primal values (the
// oracle compares pv(x)), so the tangent is the derivative along the branch the
// primal took; sqrt at 0 has a zero tangent (the convention the adjoint uses).
#pragma once
#include <omp.h>

#include <cmath>

namespace ora {

// Branch log of a rollout: with `on`, every comparison the oracle makes on
// primal values (it compares pv(x), below) and every |x| folds its outcome
// into a digest of the thread making it. With the OpenMP loops' static
// schedule and one thread count, each thread sees the same decisions in the
// same order on the same input, so the sum of the thread digests is a
// fingerprint of the branches the whole rollout took.
struct BranchLog {
  static constexpr int kThreads = 256;
  struct alignas(64) Slot {
    unsigned long long h, n;
  };
  bool on = false;
  Slot slot[kThreads];
  void reset() {
    for (Slot& s : slot) s.h = 1469598103934665603ull, s.n = 0;
  }
};
BranchLog& branch_log();
// A decision between operands within kTieUlps of each other is logged as a
// tie, not by its outcome: roundoff decides it, and the two branches agree
// to roundoff in value (the reference's clamps meet their bounds exactly
// wherever the limiter's 2x-slope branch is active, lr_states_plm.c:165-170).
// Crossing a genuine kink between x - eps v and x + eps v separates the
// operands by about |slope| eps at the two ends, far more than kTieUlps.
constexpr double kTieUlps = 64.0;
inline bool note(bool c, double a, double b) {
  BranchLog& L = branch_log();
  if (L.on) {
    const double m = std::fmax(std::fabs(a), std::fabs(b));
    const bool tie = std::fabs(a - b) <= kTieUlps * 2.220446049250313e-16 * m;
    const unsigned long long key =
        tie ? 0x94D049BB133111EBull : (c ? 0x9E3779B97F4A7C15ull : 0xC2B2AE3D27D4EB4Full);
    BranchLog::Slot& s = L.slot[omp_get_thread_num() % BranchLog::kThreads];
    s.h = (s.h ^ key) * 1099511628211ull;
    s.n++;
  }
  return c;
}
// The primal value as the oracle's comparisons see it: the comparison
// operators record their outcome; arithmetic uses the plain double.
struct PvVal {
  double v;
  operator double() const { return v; }  // NOLINT: implicit, for arithmetic on pv(x)
};
inline bool operator<(PvVal a, double b) { return note(a.v < b, a.v, b); }
inline bool operator>(PvVal a, double b) { return note(a.v > b, a.v, b); }
inline bool operator<=(PvVal a, double b) { return note(a.v <= b, a.v, b); }
inline bool operator>=(PvVal a, double b) { return note(a.v >= b, a.v, b); }
inline bool operator==(PvVal a, double b) { return note(a.v == b, a.v, b); }
inline bool operator!=(PvVal a, double b) { return note(a.v != b, a.v, b); }
inline bool operator<(PvVal a, PvVal b) { return note(a.v < b.v, a.v, b.v); }
inline bool operator>(PvVal a, PvVal b) { return note(a.v > b.v, a.v, b.v); }
inline bool operator<=(PvVal a, PvVal b) { return note(a.v <= b.v, a.v, b.v); }
inline bool operator>=(PvVal a, PvVal b) { return note(a.v >= b.v, a.v, b.v); }

struct Dual {
  double v = 0.0, t = 0.0;
  Dual() = default;
  Dual(double x) : v(x), t(0.0) {}  // NOLINT: implicit, like a double literal
  Dual(double x, double dx) : v(x), t(dx) {}
};

inline PvVal pv(const Dual& a) { return PvVal{a.v}; }
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
inline Dual fabs(const Dual& a) { return note(a.v < 0.0, a.v, 0.0) ? -a : a; }

}  // namespace ora
