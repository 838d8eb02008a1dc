/*
 * cstep.h -- complex-step scalar for the C++ build of the oracle
 * (TEST INFRASTRUCTURE ONLY).
 *
 * z = re + i im carries a value and, in its imaginary part, h times its
 * derivative along one input direction.  The complex-step method evaluates
 * f(x + i h) and takes Im f / h; for h far below FP64 resolution the terms of
 * order h^2 vanish, so every operation here keeps exactly the first-order
 * imaginary part and computes the real part with the same double operation as
 * the plain oracle.  The real trajectory (and every branch it takes: the C+-
 * selection, the Newton stopping test, pivoting, upwinding) is therefore the
 * plain oracle's bit for bit, and Im J / h is the derivative of exactly that
 * computation.
 *
 * Non-analytic functions follow the usual complex-step conventions: |z| =
 * z sign(Re z); comparisons, min, max and isnan look at the real part.
 */
#ifndef ORACLE_CSTEP_H
#define ORACLE_CSTEP_H
#include <cmath>

struct cs_t {
  double re, im;
  cs_t() : re(0.0), im(0.0) {}
  cs_t(double r) : re(r), im(0.0) {}
  cs_t(double r, double i) : re(r), im(i) {}
  cs_t& operator+=(const cs_t& b) { re += b.re; im += b.im; return *this; }
  cs_t& operator-=(const cs_t& b) { re -= b.re; im -= b.im; return *this; }
  cs_t& operator*=(const cs_t& b) { *this = cs_t(re * b.re, re * b.im + im * b.re); return *this; }
  cs_t& operator/=(const cs_t& b) {
    const double q = re / b.re;
    *this = cs_t(q, (im - q * b.im) / b.re);
    return *this;
  }
};

inline cs_t operator+(const cs_t& a) { return a; }
inline cs_t operator-(const cs_t& a) { return cs_t(-a.re, -a.im); }
inline cs_t operator+(const cs_t& a, const cs_t& b) { return cs_t(a.re + b.re, a.im + b.im); }
inline cs_t operator-(const cs_t& a, const cs_t& b) { return cs_t(a.re - b.re, a.im - b.im); }
inline cs_t operator*(const cs_t& a, const cs_t& b) { return cs_t(a.re * b.re, a.re * b.im + a.im * b.re); }
inline cs_t operator/(const cs_t& a, const cs_t& b) {
  const double q = a.re / b.re;
  return cs_t(q, (a.im - q * b.im) / b.re);
}
inline cs_t operator+(const cs_t& a, double b) { return cs_t(a.re + b, a.im); }
inline cs_t operator+(double a, const cs_t& b) { return cs_t(a + b.re, b.im); }
inline cs_t operator-(const cs_t& a, double b) { return cs_t(a.re - b, a.im); }
inline cs_t operator-(double a, const cs_t& b) { return cs_t(a - b.re, -b.im); }
inline cs_t operator*(const cs_t& a, double b) { return cs_t(a.re * b, a.im * b); }
inline cs_t operator*(double a, const cs_t& b) { return cs_t(a * b.re, a * b.im); }
inline cs_t operator/(const cs_t& a, double b) { return cs_t(a.re / b, a.im / b); }
inline cs_t operator/(double a, const cs_t& b) {
  const double q = a / b.re;
  return cs_t(q, -q * b.im / b.re);
}

#define CS_CMP(op)                                                                    \
  inline bool operator op(const cs_t& a, const cs_t& b) { return a.re op b.re; }     \
  inline bool operator op(const cs_t& a, double b) { return a.re op b; }             \
  inline bool operator op(double a, const cs_t& b) { return a op b.re; }
CS_CMP(<)
CS_CMP(>)
CS_CMP(<=)
CS_CMP(>=)
CS_CMP(==)
CS_CMP(!=)
#undef CS_CMP

inline cs_t fabs(const cs_t& a) { return a.re < 0.0 ? -a : a; }
inline cs_t sqrt(const cs_t& a) {
  const double r = std::sqrt(a.re);
  return cs_t(r, a.im / (2.0 * r));
}
inline cs_t acos(const cs_t& a) { return cs_t(std::acos(a.re), -a.im / std::sqrt(1.0 - a.re * a.re)); }
inline cs_t cos(const cs_t& a) { return cs_t(std::cos(a.re), -std::sin(a.re) * a.im); }
inline cs_t log(const cs_t& a) { return cs_t(std::log(a.re), a.im / a.re); }
inline cs_t fmin(const cs_t& a, const cs_t& b) { return b.re < a.re ? b : a; }
inline cs_t fmax(const cs_t& a, const cs_t& b) { return b.re > a.re ? b : a; }
inline bool isnan(const cs_t& a) { return std::isnan(a.re); }
inline double rv(const cs_t& a) { return a.re; }

#endif
