// flop_count.hpp -- an operation-counting scalar for cloud_oracle.hpp.
//
// TEST INFRASTRUCTURE ONLY (see cloud_oracle.hpp).  Instantiating the
// oracle's templates with orc_flop::Flop runs the FP64 march (its values are
// doubles, so every branch and trip count is the FP64 twin's) and counts the
// arithmetic the GLSL source states, by work region (FlopRegion): add, sub,
// mul, div, sqrt and each transcendental (exp, pow, acos, cos) count 1, an
// fma counts 2; comparisons, min / max / clamp, floor / fract's floor, abs,
// rounding and conversions count 0.  These are the per-unit FLOP weights of
// the roofline (bench.py's FLOPS table is held to them by a CPU test).
#pragma once
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "cloud_oracle.hpp"

namespace orc_flop {

struct Tally {
  uint64_t ops[orc::FR_COUNT] = {};
  int region = orc::FR_RAY;
};
inline thread_local Tally tally;
inline void count(int n) { tally.ops[tally.region] += (uint64_t)n; }

struct Flop {
  double v = 0.0;
  Flop() = default;
  template <class A, class = std::enable_if_t<std::is_arithmetic<A>::value>>
  Flop(A a) : v((double)a) {}
  template <class A, class = std::enable_if_t<std::is_arithmetic<A>::value>>
  explicit operator A() const { return (A)v; }
};

inline Flop mk(double v) {
  Flop f;
  f.v = v;
  return f;
}
inline Flop operator+(Flop a, Flop b) { count(1); return mk(a.v + b.v); }
inline Flop operator-(Flop a, Flop b) { count(1); return mk(a.v - b.v); }
inline Flop operator*(Flop a, Flop b) { count(1); return mk(a.v * b.v); }
inline Flop operator/(Flop a, Flop b) { count(1); return mk(a.v / b.v); }
inline Flop operator-(Flop a) { return mk(-a.v); }
inline Flop& operator+=(Flop& a, Flop b) { a = a + b; return a; }
inline Flop& operator-=(Flop& a, Flop b) { a = a - b; return a; }
inline Flop& operator*=(Flop& a, Flop b) { a = a * b; return a; }
inline bool operator<(Flop a, Flop b) { return a.v < b.v; }
inline bool operator>(Flop a, Flop b) { return a.v > b.v; }
inline bool operator<=(Flop a, Flop b) { return a.v <= b.v; }
inline bool operator>=(Flop a, Flop b) { return a.v >= b.v; }
inline bool operator==(Flop a, Flop b) { return a.v == b.v; }
inline bool operator!=(Flop a, Flop b) { return a.v != b.v; }

// the oracle's math (found by argument-dependent lookup)
inline Flop m_floor(Flop x) { return mk(std::floor(x.v)); }
inline Flop m_sqrt(Flop x) { count(1); return mk(std::sqrt(x.v)); }
inline Flop m_exp(Flop x) { count(1); return mk(std::exp(x.v)); }
inline Flop m_pow(Flop x, Flop y) { count(1); return mk(std::pow(x.v, y.v)); }
inline Flop m_acos(Flop x) { count(1); return mk(std::acos(x.v)); }
inline Flop m_cos(Flop x) { count(1); return mk(std::cos(x.v)); }
inline Flop m_fabs(Flop x) { return mk(std::fabs(x.v)); }
inline Flop m_fma(Flop a, Flop b, Flop c) { count(2); return mk(std::fma(a.v, b.v, c.v)); }
inline Flop m_fmax(Flop a, Flop b) { return mk(std::fmax(a.v, b.v)); }
inline Flop m_fmin(Flop a, Flop b) { return mk(std::fmin(a.v, b.v)); }
inline Flop m_nearbyint(Flop x) { return mk(std::nearbyint(x.v)); }

}  // namespace orc_flop

namespace orc {
template <> struct Region<orc_flop::Flop> {
  static void set(int r) { orc_flop::tally.region = r; }
};
}  // namespace orc
