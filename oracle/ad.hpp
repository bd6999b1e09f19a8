// Oracle: forward-mode dual numbers for the CPU restatement of the FIM step.
//
// TEST INFRASTRUCTURE ONLY. This directory is the correctness oracle for the
// HIP product path; only tests, the smoke check and the bench's CPU-baseline
// leg may load it. Nothing in the product links it.
//
// Restates rs/fim/ad.rs:21-319 (Ad<N>) and the Scalar trait
// rs/fim/ad.rs:325-417: branch-selecting max/min act on the value, a clamp
// keeps the live derivative when the value sits exactly on the bound
// (value >= floor keeps `self`).
#pragma once
#include <array>
#include <cmath>

namespace orc {

template <int N>
struct Ad {
  double v = 0.0;
  std::array<double, N> d{};
  Ad() = default;
  Ad(double value) : v(value) { d.fill(0.0); }  // constant
  static Ad var(double value, int slot) {
    Ad a(value);
    a.d[slot] = 1.0;
    return a;
  }
};

// ---- scalar helpers, double ------------------------------------------------
inline double val(double x) { return x; }
inline double cst(double, double c) { return c; }
inline double max_floor(double x, double f) { return x >= f ? x : f; }  // f64::max
inline double min_ceil(double x, double c) { return x <= c ? x : c; }
inline double max_of(double a, double b) { return a >= b ? a : b; }
inline double min_of(double a, double b) { return a <= b ? a : b; }
inline double recip(double x) { return 1.0 / x; }
inline double powf_(double x, double p) { return std::pow(x, p); }
inline double exp_(double x) { return std::exp(x); }
inline double sqrt_(double x) { return std::sqrt(x); }

// ---- Ad<N> -----------------------------------------------------------------
template <int N> inline double val(const Ad<N>& a) { return a.v; }
template <int N> inline Ad<N> operator+(Ad<N> a, const Ad<N>& b) {
  a.v += b.v;
  for (int i = 0; i < N; ++i) a.d[i] += b.d[i];
  return a;
}
template <int N> inline Ad<N> operator-(Ad<N> a, const Ad<N>& b) {
  a.v -= b.v;
  for (int i = 0; i < N; ++i) a.d[i] -= b.d[i];
  return a;
}
template <int N> inline Ad<N> operator-(Ad<N> a) {
  a.v = -a.v;
  for (int i = 0; i < N; ++i) a.d[i] = -a.d[i];
  return a;
}
template <int N> inline Ad<N> operator*(const Ad<N>& a, const Ad<N>& b) {
  Ad<N> r;
  r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N> inline Ad<N> operator/(const Ad<N>& a, const Ad<N>& b) {
  Ad<N> r;
  r.v = a.v / b.v;
  double inv2 = 1.0 / (b.v * b.v);
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) * inv2;
  return r;
}
template <int N> inline Ad<N> operator+(Ad<N> a, double c) { a.v += c; return a; }
template <int N> inline Ad<N> operator-(Ad<N> a, double c) { a.v -= c; return a; }
template <int N> inline Ad<N> operator*(Ad<N> a, double c) {
  a.v *= c;
  for (int i = 0; i < N; ++i) a.d[i] *= c;
  return a;
}
template <int N> inline Ad<N> operator/(Ad<N> a, double c) {
  a.v /= c;
  for (int i = 0; i < N; ++i) a.d[i] /= c;
  return a;
}
template <int N> inline Ad<N> operator+(double c, const Ad<N>& a) { return a + c; }
template <int N> inline Ad<N> operator-(double c, const Ad<N>& a) { return Ad<N>(c) - a; }
template <int N> inline Ad<N> operator*(double c, const Ad<N>& a) { return a * c; }
template <int N> inline Ad<N> operator/(double c, const Ad<N>& a) { return Ad<N>(c) / a; }

template <int N> inline Ad<N> cst(const Ad<N>&, double c) { return Ad<N>(c); }
template <int N> inline Ad<N> max_floor(const Ad<N>& a, double f) { return a.v >= f ? a : Ad<N>(f); }
template <int N> inline Ad<N> min_ceil(const Ad<N>& a, double c) { return a.v <= c ? a : Ad<N>(c); }
template <int N> inline Ad<N> max_of(const Ad<N>& a, const Ad<N>& b) { return a.v >= b.v ? a : b; }
template <int N> inline Ad<N> min_of(const Ad<N>& a, const Ad<N>& b) { return a.v <= b.v ? a : b; }
template <int N> inline Ad<N> recip(const Ad<N>& a) {
  Ad<N> r;
  r.v = 1.0 / a.v;
  double c = -r.v * r.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * c;
  return r;
}
template <int N> inline Ad<N> powf_(const Ad<N>& a, double p) {
  Ad<N> r;
  r.v = std::pow(a.v, p);
  double c = p * std::pow(a.v, p - 1.0);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * c;
  return r;
}
template <int N> inline Ad<N> exp_(const Ad<N>& a) {
  Ad<N> r;
  r.v = std::exp(a.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * r.v;
  return r;
}
template <int N> inline Ad<N> sqrt_(const Ad<N>& a) {
  Ad<N> r;
  r.v = std::sqrt(a.v);
  double c = 0.5 / r.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * c;
  return r;
}

}  // namespace orc
