// oracle/dual.hpp -- forward-mode dual numbers for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY.  The oracle differentiates the restated forward
// models with these duals, which is an independent derivation from the
// hand-written adjoints of the HIP kernels (it plays the role MATLAB's
// dlgradient plays in utils/askadam.m:105).
//
// Subgradient conventions at kinks (dlgradient's own are unpinned offline):
//   abs'(0) = 0, max/min pass the derivative of the first argument on ties.
#pragma once
#include <cmath>

namespace orc {

template <class S, int N>
struct Dual {
  S v;
  S d[N];
  Dual() : v(0) { for (int i = 0; i < N; ++i) d[i] = 0; }
  Dual(S x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0; }  // NOLINT: implicit by design
  static Dual var(S x, int k) { Dual r(x); r.d[k] = 1; return r; }
};

template <class S, int N> inline Dual<S, N> operator+(const Dual<S, N>& a, const Dual<S, N>& b) {
  Dual<S, N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <class S, int N> inline Dual<S, N> operator-(const Dual<S, N>& a, const Dual<S, N>& b) {
  Dual<S, N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <class S, int N> inline Dual<S, N> operator-(const Dual<S, N>& a) {
  Dual<S, N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <class S, int N> inline Dual<S, N> operator*(const Dual<S, N>& a, const Dual<S, N>& b) {
  Dual<S, N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <class S, int N> inline Dual<S, N> operator/(const Dual<S, N>& a, const Dual<S, N>& b) {
  Dual<S, N> r; r.v = a.v / b.v; S ib = S(1) / b.v;
  for (int i = 0; i < N; ++i) { r.d[i] = (a.d[i] - r.v * b.d[i]) * ib; }
  return r; }

#define ORC_DUAL_SCALAR_OPS(OP)                                                                   \
  template <class S, int N> inline Dual<S, N> operator OP(const Dual<S, N>& a, S b) { return a OP Dual<S, N>(b); } \
  template <class S, int N> inline Dual<S, N> operator OP(S a, const Dual<S, N>& b) { return Dual<S, N>(a) OP b; }
ORC_DUAL_SCALAR_OPS(+)
ORC_DUAL_SCALAR_OPS(-)
ORC_DUAL_SCALAR_OPS(*)
ORC_DUAL_SCALAR_OPS(/)
#undef ORC_DUAL_SCALAR_OPS

template <class S, int N> inline Dual<S, N> exp(const Dual<S, N>& a) {
  Dual<S, N> r; r.v = std::exp(a.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * r.v; return r; }
template <class S, int N> inline Dual<S, N> log(const Dual<S, N>& a) {
  Dual<S, N> r; r.v = std::log(a.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / a.v; return r; }
template <class S, int N> inline Dual<S, N> sqrt(const Dual<S, N>& a) {
  Dual<S, N> r; r.v = std::sqrt(a.v); S h = S(0.5) / r.v;
  for (int i = 0; i < N; ++i) { r.d[i] = a.d[i] * h; }
  return r; }
template <class S, int N> inline Dual<S, N> sin(const Dual<S, N>& a) {
  Dual<S, N> r; r.v = std::sin(a.v); S c = std::cos(a.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * c; return r; }
template <class S, int N> inline Dual<S, N> cos(const Dual<S, N>& a) {
  Dual<S, N> r; r.v = std::cos(a.v); S s = -std::sin(a.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s; return r; }
template <class S, int N> inline Dual<S, N> fabs(const Dual<S, N>& a) {
  S sg = a.v > 0 ? S(1) : (a.v < 0 ? S(-1) : S(0));
  Dual<S, N> r; r.v = std::fabs(a.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * sg; return r; }
template <class S, int N> inline Dual<S, N> vmax(const Dual<S, N>& a, const Dual<S, N>& b) { return (a.v >= b.v) ? a : b; }
template <class S, int N> inline Dual<S, N> vmin(const Dual<S, N>& a, const Dual<S, N>& b) { return (a.v <= b.v) ? a : b; }
template <class S, int N> inline Dual<S, N> vmax(const Dual<S, N>& a, S b) { return (a.v >= b) ? a : Dual<S, N>(b); }
template <class S, int N> inline Dual<S, N> vmin(const Dual<S, N>& a, S b) { return (a.v <= b) ? a : Dual<S, N>(b); }
template <class S, int N> inline S value(const Dual<S, N>& a) { return a.v; }
template <class S, int N> inline bool finite(const Dual<S, N>& a) { return std::isfinite(a.v); }

template <class S, int N> inline S vval(const Dual<S, N>& a) { return a.v; }
inline double vval(double a) { return a; }
inline float vval(float a) { return a; }
inline double vmax(double a, double b) { return a >= b ? a : b; }
inline double vmin(double a, double b) { return a <= b ? a : b; }
inline float vmax(float a, float b) { return a >= b ? a : b; }
inline float vmin(float a, float b) { return a <= b ? a : b; }
inline double value(double a) { return a; }
inline float value(float a) { return a; }
inline bool finite(double a) { return std::isfinite(a); }
inline bool finite(float a) { return std::isfinite(a); }

// scalar type underlying a value type
template <class V> struct scalar_of { using type = V; };
template <class S, int N> struct scalar_of<Dual<S, N>> { using type = S; };

}  // namespace orc
