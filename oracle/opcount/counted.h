/*
 * Operation-counting scalar for the oracle (TEST AND MEASUREMENT INFRASTRUCTURE).
 *
 * Each translation unit of oracle/opcount/ includes this header, defines `double` as `cd` and includes one of
 * the oracle's C sources, so the reference algorithm as the oracle restates it runs on a scalar that tallies
 * every floating-point operation it performs at run time.  Constant expressions of literals stay plain double
 * (the compiler folds them in the reference too); library functions count as one operation of their kind.
 */
#ifndef GWO_COUNTED_H
#define GWO_COUNTED_H
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>

struct OpTally {
    unsigned long long add, mul, div, sqrt, exp, log, pow, sincos, atan, cmp, misc;
};
extern OpTally g_ops;

struct cd {
    double v;
    cd() = default;
    constexpr cd(double x) : v(x) {}
    constexpr cd(int x) : v(x) {}
    constexpr cd(long x) : v((double)x) {}
    constexpr cd(unsigned x) : v(x) {}
    constexpr cd(unsigned long x) : v((double)x) {}
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
    explicit operator unsigned long() const { return (unsigned long)v; }
    explicit operator bool() const { return v != 0; }
    cd &operator+=(cd o) { g_ops.add++; v += o.v; return *this; }
    cd &operator-=(cd o) { g_ops.add++; v -= o.v; return *this; }
    cd &operator*=(cd o) { g_ops.mul++; v *= o.v; return *this; }
    cd &operator/=(cd o) { g_ops.div++; v /= o.v; return *this; }
};

inline cd operator+(cd a, cd b) { g_ops.add++; return cd(a.v + b.v); }
inline cd operator-(cd a, cd b) { g_ops.add++; return cd(a.v - b.v); }
inline cd operator*(cd a, cd b) { g_ops.mul++; return cd(a.v * b.v); }
inline cd operator/(cd a, cd b) { g_ops.div++; return cd(a.v / b.v); }
inline cd operator-(cd a) { return cd(-a.v); } /* a sign flip folds into the consumer */
inline cd operator+(cd a) { return a; }
inline bool operator<(cd a, cd b) { g_ops.cmp++; return a.v < b.v; }
inline bool operator>(cd a, cd b) { g_ops.cmp++; return a.v > b.v; }
inline bool operator<=(cd a, cd b) { g_ops.cmp++; return a.v <= b.v; }
inline bool operator>=(cd a, cd b) { g_ops.cmp++; return a.v >= b.v; }
inline bool operator==(cd a, cd b) { g_ops.cmp++; return a.v == b.v; }
inline bool operator!=(cd a, cd b) { g_ops.cmp++; return a.v != b.v; }
inline bool operator!(cd a) { return a.v == 0; }

#define GWO_CD_UNARY(fn, field)                                                                                    \
    inline cd fn(cd a) { g_ops.field++; return cd(std::fn(a.v)); }
GWO_CD_UNARY(sqrt, sqrt)
GWO_CD_UNARY(exp, exp)
GWO_CD_UNARY(log, log)
GWO_CD_UNARY(log10, log)
GWO_CD_UNARY(cbrt, pow)
GWO_CD_UNARY(atan, atan)
GWO_CD_UNARY(acos, atan)
GWO_CD_UNARY(asin, atan)
GWO_CD_UNARY(tan, sincos)
GWO_CD_UNARY(floor, misc)
GWO_CD_UNARY(ceil, misc)
GWO_CD_UNARY(round, misc)
GWO_CD_UNARY(rint, misc)
/* sin and cos of one argument are one sincos pair in the frozen convention: each counts half */
inline cd sin(cd a) { g_ops.sincos++; return cd(std::sin(a.v)); }
inline cd cos(cd a) { g_ops.sincos++; return cd(std::cos(a.v)); }
inline cd fabs(cd a) { return cd(std::fabs(a.v)); }
inline cd pow(cd a, cd b) { g_ops.pow++; return cd(std::pow(a.v, b.v)); }
inline cd pow(cd a, double b) { g_ops.pow++; return cd(std::pow(a.v, b)); }
inline cd pow(double a, cd b) { g_ops.pow++; return cd(std::pow(a, b.v)); }
inline cd pow(cd a, int b) { g_ops.pow++; return cd(std::pow(a.v, b)); }
inline cd atan2(cd a, cd b) { g_ops.atan++; return cd(std::atan2(a.v, b.v)); }
inline cd fmin(cd a, cd b) { g_ops.cmp++; return cd(std::fmin(a.v, b.v)); }
inline cd fmax(cd a, cd b) { g_ops.cmp++; return cd(std::fmax(a.v, b.v)); }
inline cd copysign(cd a, cd b) { return cd(std::copysign(a.v, b.v)); }
inline bool isnan(cd a) { return std::isnan(a.v); }
inline bool isinf(cd a) { return std::isinf(a.v); }
inline bool isfinite(cd a) { return std::isfinite(a.v); }
#endif
