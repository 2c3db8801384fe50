// Forward-mode dual numbers for the oracle's tangent-linear build
// (oracle_fwd.cpp).  TEST INFRASTRUCTURE ONLY, like the rest of oracle/.
//
// A Dual carries a value and one directional derivative, x + x' eps with
// eps^2 = 0.  The value part of every operation is the same IEEE double
// operation the plain-C oracle performs, so a dual run reproduces the oracle's
// iterates and control flow exactly (branches and convergence tests read the
// value only) and carries d(.)/d(theta) along one parameter direction.
// CDual is the complex counterpart standing in for C99 `double complex` in the
// FFT and Gamma code: the componentwise formulas of C99 complex addition and
// multiplication for finite operands.
#pragma once

#include <cmath>

namespace orc_fwd {

struct Dual {
  double v, d;
  Dual() = default;
  constexpr Dual(double x) : v(x), d(0.0) {}
  constexpr Dual(double x, double dx) : v(x), d(dx) {}
  explicit operator double() const { return v; }
  explicit operator float() const { return (float)v; }
  explicit operator int() const { return (int)v; }
  explicit operator long() const { return (long)v; }
  explicit operator unsigned long() const { return (unsigned long)v; }
  explicit operator long long() const { return (long long)v; }
  explicit operator bool() const { return v != 0.0; }
  Dual& operator+=(Dual o) { v += o.v; d += o.d; return *this; }
  Dual& operator-=(Dual o) { v -= o.v; d -= o.d; return *this; }
  Dual& operator*=(Dual o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
  Dual& operator/=(Dual o) { v /= o.v; d = (d - v * o.d) / o.v; return *this; }
};

inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
// mixed operands: exact matches, so a double never has to pick between
// becoming a Dual or a CDual
inline Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
inline Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
inline Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
inline Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator+(Dual a) { return a; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
inline bool operator==(Dual a, Dual b) { return a.v == b.v; }
inline bool operator!=(Dual a, Dual b) { return a.v != b.v; }
inline bool operator!(Dual a) { return a.v == 0.0; }

using std::acos;
using std::cbrt;
using std::copysign;
using std::cos;
using std::exp;
using std::fabs;
using std::isnan;
using std::log;
using std::pow;
using std::sin;
using std::sqrt;
using std::tanh;

// chain rule d f(x) = f'(x) x'; a zero tangent stays zero where f' is
// infinite (sqrt(0), acos(+-1)), as the derivative of a constant.  An
// overflowed value carries no tangent (the Jacobi rotation angle of the
// NGMRES least squares overflows theta^2 when a pivot is tiny: there the
// rotation is the identity, in value and derivative).
inline Dual sqrt(Dual a) {
  const double r = std::sqrt(a.v);
  return Dual(r, a.d == 0.0 || std::isinf(r) ? 0.0 : a.d / (2.0 * r));
}
inline Dual cbrt(Dual a) {
  const double r = std::cbrt(a.v);
  return Dual(r, a.d == 0.0 ? 0.0 : a.d / (3.0 * r * r));
}
inline Dual fabs(Dual a) { return a.v < 0.0 ? Dual(-a.v, -a.d) : Dual(a.v, a.d); }
inline Dual exp(Dual a) {
  const double e = std::exp(a.v);
  return Dual(e, e * a.d);
}
inline Dual log(Dual a) { return Dual(std::log(a.v), a.d == 0.0 ? 0.0 : a.d / a.v); }
inline Dual sin(Dual a) { return Dual(std::sin(a.v), a.d == 0.0 ? 0.0 : std::cos(a.v) * a.d); }
inline Dual cos(Dual a) { return Dual(std::cos(a.v), a.d == 0.0 ? 0.0 : -std::sin(a.v) * a.d); }
inline Dual acos(Dual a) {
  return Dual(std::acos(a.v), a.d == 0.0 ? 0.0 : -a.d / std::sqrt(1.0 - a.v * a.v));
}
inline Dual copysign(Dual a, Dual b) {
  const double r = std::copysign(a.v, b.v);
  return Dual(r, std::signbit(r) == std::signbit(a.v) ? a.d : -a.d);
}
inline Dual tanh(Dual a) {
  const double t = std::tanh(a.v);
  return Dual(t, a.d == 0.0 ? 0.0 : (1.0 - t * t) * a.d);
}
inline bool isnan(Dual a) { return std::isnan(a.v); }
// x^y: d = y x^(y-1) x' + x^y ln(x) y'
inline Dual pow(Dual a, Dual b) {
  const double r = std::pow(a.v, b.v);
  double d = 0.0;
  if (a.d != 0.0) d += b.v * std::pow(a.v, b.v - 1.0) * a.d;
  if (b.d != 0.0 && a.v > 0.0) d += r * std::log(a.v) * b.d;
  return Dual(r, d);
}
inline Dual pow(Dual a, double b) { return pow(a, Dual(b)); }
inline Dual pow(double a, Dual b) { return pow(Dual(a), b); }
inline Dual pow(Dual a, int b) { return pow(a, Dual((double)b)); }

struct CDual {
  Dual re, im;
  CDual() = default;
  constexpr CDual(double r) : re(r), im(0.0) {}
  constexpr CDual(Dual r) : re(r), im(0.0) {}
  constexpr CDual(Dual r, Dual i) : re(r), im(i) {}
  CDual& operator+=(CDual o) { re += o.re; im += o.im; return *this; }
  CDual& operator-=(CDual o) { re -= o.re; im -= o.im; return *this; }
  CDual& operator*=(CDual o) { *this = CDual(re * o.re - im * o.im, re * o.im + im * o.re); return *this; }
};
inline CDual operator+(CDual a, CDual b) { return CDual(a.re + b.re, a.im + b.im); }
inline CDual operator-(CDual a, CDual b) { return CDual(a.re - b.re, a.im - b.im); }
inline CDual operator-(CDual a) { return CDual(-a.re, -a.im); }
// C99 complex multiply (a + ib)(c + id) = (ac - bd) + i(ad + bc); a real
// operand multiplies componentwise, as C does after the usual conversions
inline CDual operator*(CDual a, CDual b) { return CDual(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
inline CDual operator*(CDual a, Dual b) { return CDual(a.re * b, a.im * b); }
inline CDual operator*(Dual a, CDual b) { return CDual(a * b.re, a * b.im); }
inline CDual operator*(CDual a, double b) { return CDual(a.re * b, a.im * b); }
inline CDual operator*(double a, CDual b) { return CDual(a * b.re, a * b.im); }
inline CDual operator+(CDual a, Dual b) { return CDual(a.re + b, a.im); }
inline CDual operator+(Dual a, CDual b) { return CDual(a + b.re, b.im); }
inline CDual operator+(double a, CDual b) { return CDual(a + b.re, b.im); }
// textbook quotient (only the coordinate utility divides complex numbers)
inline CDual operator/(CDual a, CDual b) {
  const Dual den = b.re * b.re + b.im * b.im;
  return CDual((a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den);
}
inline Dual creal(CDual a) { return a.re; }
inline Dual cimag(CDual a) { return a.im; }
inline CDual conj(CDual a) { return CDual(a.re, -a.im); }

}  // namespace orc_fwd
