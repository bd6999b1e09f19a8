// Per-cell forward-mode dual number for the HIP FIM path.
//
// Cell properties are evaluated once per cell with partials against that
// cell's own primaries [p, sw, hc]; face terms are then composed from two
// cells' duals with the chain rule (one face owns six partials). Branch
// selection follows the value: a clamp keeps the live derivative when the
// value sits exactly on its bound, the convention of rs/fim/ad.rs:116-130.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#define RSF_HD __host__ __device__ __forceinline__

namespace rsf {

template <int N>
struct Dual {
  double v;
  double d[N];
};

template <int N> RSF_HD Dual<N> dconst(double c) {
  Dual<N> r;
  r.v = c;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = 0.0;
  return r;
}
template <int N> RSF_HD Dual<N> dvar(double c, int slot) {
  Dual<N> r = dconst<N>(c);
  r.d[slot] = 1.0;
  return r;
}

// Scalar-generic helpers: T is double or Dual<N>.
RSF_HD double value(double x) { return x; }
template <int N> RSF_HD double value(const Dual<N>& x) { return x.v; }
RSF_HD double lift(double, double c) { return c; }
template <int N> RSF_HD Dual<N> lift(const Dual<N>&, double c) { return dconst<N>(c); }

template <int N> RSF_HD Dual<N> operator+(Dual<N> a, const Dual<N>& b) {
  a.v += b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) a.d[i] += b.d[i];
  return a;
}
template <int N> RSF_HD Dual<N> operator-(Dual<N> a, const Dual<N>& b) {
  a.v -= b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) a.d[i] -= b.d[i];
  return a;
}
template <int N> RSF_HD Dual<N> operator-(Dual<N> a) {
  a.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) a.d[i] = -a.d[i];
  return a;
}
template <int N> RSF_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N> RSF_HD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  double ib = 1.0 / b.v;
  r.v = a.v / b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N> RSF_HD Dual<N> operator+(Dual<N> a, double c) { a.v += c; return a; }
template <int N> RSF_HD Dual<N> operator+(double c, Dual<N> a) { a.v += c; return a; }
template <int N> RSF_HD Dual<N> operator-(Dual<N> a, double c) { a.v -= c; return a; }
template <int N> RSF_HD Dual<N> operator-(double c, const Dual<N>& a) { return dconst<N>(c) - a; }
template <int N> RSF_HD Dual<N> operator*(Dual<N> a, double c) {
  a.v *= c;
#pragma unroll
  for (int i = 0; i < N; ++i) a.d[i] *= c;
  return a;
}
template <int N> RSF_HD Dual<N> operator*(double c, const Dual<N>& a) { return a * c; }
template <int N> RSF_HD Dual<N> operator/(Dual<N> a, double c) {
  double ic = 1.0 / c;
  a.v = a.v / c;
#pragma unroll
  for (int i = 0; i < N; ++i) a.d[i] *= ic;
  return a;
}
template <int N> RSF_HD Dual<N> operator/(double c, const Dual<N>& a) { return dconst<N>(c) / a; }

RSF_HD double floor_at(double x, double f) { return x >= f ? x : f; }
RSF_HD double ceil_at(double x, double c) { return x <= c ? x : c; }
RSF_HD double pick_max(double a, double b) { return a >= b ? a : b; }
RSF_HD double pick_min(double a, double b) { return a <= b ? a : b; }
RSF_HD double inv(double x) { return 1.0 / x; }
RSF_HD double pw(double x, double p) { return pow(x, p); }
RSF_HD double ex(double x) { return exp(x); }

template <int N> RSF_HD Dual<N> floor_at(const Dual<N>& x, double f) { return x.v >= f ? x : dconst<N>(f); }
template <int N> RSF_HD Dual<N> ceil_at(const Dual<N>& x, double c) { return x.v <= c ? x : dconst<N>(c); }
template <int N> RSF_HD Dual<N> pick_max(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v ? a : b; }
template <int N> RSF_HD Dual<N> pick_min(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v ? a : b; }
template <int N> RSF_HD Dual<N> inv(const Dual<N>& a) {
  Dual<N> r;
  r.v = 1.0 / a.v;
  double s = -r.v * r.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s;
  return r;
}
template <int N> RSF_HD Dual<N> pw(const Dual<N>& a, double p) {
  Dual<N> r;
  r.v = pow(a.v, p);
  double s = p * pow(a.v, p - 1.0);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s;
  return r;
}
template <int N> RSF_HD Dual<N> ex(const Dual<N>& a) {
  Dual<N> r;
  r.v = exp(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * r.v;
  return r;
}

}  // namespace rsf
