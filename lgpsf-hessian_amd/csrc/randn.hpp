// randn.hpp — the hashed Gaussian as a pure function of (seed, i, j), for
// host and device: glr_common_impl.h:35-73 (lgh_randn / lgh_randn_at).
//
// The integer part (splitmix64 mix, xorshift64* pair, 53-bit uniforms) and
// the IEEE operations (*, /, sqrt; correctly rounded on both the host and
// gfx950, with FMA contraction disabled) are the reference's.  The one
// libm call, log(q), is evaluated on the device in double-double:
//     q = 2^e m, m in [sqrt(1/2), sqrt(2)),  c = k/64 nearest to m,
//     t = (m - c) / (m + c),  |t| <= 2^-7.5,
//     log q = e ln2 + log c + 2 atanh(t) = ... + 2 t sum_k t^(2k) / (2k + 1),
// with log c from a double-double table (log_table.hpp) and the series to
// t^14 (t^16 < 2^-120): the terms from t^8 on in double, the first four by
// double-double Horner.  hi is the correctly rounded log and lo the rest
// (error ~2^-100 relative).
//
// The reference's draw uses the platform log (glibc), which is not
// correctly rounded: its documented error is at most 0.519 ulp (0.507 ulp
// for arguments within 1/16 of 1), so it can only differ from hi where the
// exact log lies within 0.019 ulp of the midpoint between hi and its
// neighbour.  log_mid_window flags exactly those arguments (with margin);
// a draw is exact unless flagged, and a flagged draw whose value (in the
// output precision) is the same with either neighbour is exact too.  The
// remaining draws are finished with the platform log itself (k_randn), so
// the device sketch equals the reference's bit for bit.
#pragma once
#include <stdint.h>
#include <math.h>

#include "log_table.hpp"

#if defined(__HIPCC__)
#define LGH_HD __host__ __device__
#define LGH_H __host__
#else
#define LGH_HD
#define LGH_H
#endif

namespace lgh {

struct dd { double hi, lo; };

LGH_HD inline dd dd_two_sum (double a, double b)
{
#pragma clang fp contract(off)
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return { s, e };
}

LGH_HD inline dd dd_quick (double a, double b)
{
#pragma clang fp contract(off)
  double s = a + b;
  double e = b - (s - a);
  return { s, e };
}

LGH_HD inline dd dd_add (dd a, dd b)
{
#pragma clang fp contract(off)
  dd s = dd_two_sum (a.hi, b.hi);
  dd t = dd_two_sum (a.lo, b.lo);
  s.lo += t.hi;
  s = dd_quick (s.hi, s.lo);
  s.lo += t.lo;
  return dd_quick (s.hi, s.lo);
}

LGH_HD inline dd dd_mul (dd a, dd b)
{
#pragma clang fp contract(off)
  double p = a.hi * b.hi;
  double e = fma (a.hi, b.hi, -p);
  e += a.hi * b.lo + a.lo * b.hi;
  return dd_quick (p, e);
}

LGH_HD inline dd dd_mul_d (dd a, double b)
{
#pragma clang fp contract(off)
  double p = a.hi * b;
  double e = fma (a.hi, b, -p);
  e += a.lo * b;
  return dd_quick (p, e);
}

// 1/d as a double-double (d a small odd integer)
LGH_HD inline dd dd_recip (double d)
{
#pragma clang fp contract(off)
  double h = 1.0 / d;
  double r = fma (-h, d, 1.0);         // exact remainder 1 - h d
  return { h, r / d };
}

LGH_HD inline dd log_dd2 (double q)
{
#pragma clang fp contract(off)
  // q is a positive normal double (q >= 2^-104 for the polar draws)
  union { double d; uint64_t u; } cv;
  cv.d = q;
  int e = (int) ((cv.u >> 52) & 0x7ff) - 1023;
  cv.u = (cv.u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;   // [1,2)
  double m = cv.d;
  if (m > 1.4142135623730951) { m = 0.5 * m; e += 1; }              // exact
  // c = k/64 nearest to m (k in 45..91); m - c is exact (Sterbenz)
  const int k = (int) (m * 64.0 + 0.5);
  const double c = (double) k * 0.015625;
  const double num = m - c;
  dd den = dd_two_sum (m, c);
  double th = num / den.hi;
  // residual num - th * den, exact enough for the low part of t
  double p = th * den.hi;
  double pe = fma (th, den.hi, -p);
  double r = ((num - p) - pe) - th * den.lo;
  dd t = dd_quick (th, r / den.hi);
  dd t2 = dd_mul (t, t);
  // sum_{j=0..7} t^(2j) / (2j + 1): j = 4..7 in double, j = 3..0 in dd
  const double h = t2.hi;
  double sd = 1.0 / 15.0;
  sd = sd * h + 1.0 / 13.0;
  sd = sd * h + 1.0 / 11.0;
  sd = sd * h + 1.0 / 9.0;
  dd S = { sd, 0.0 };
  for (int j = 3; j >= 0; j--) {
    S = dd_mul (S, t2);
    S = dd_add (S, dd_recip (2.0 * j + 1.0));
  }
  dd L = dd_mul (t, S);
  L.hi *= 2.0; L.lo *= 2.0;                                         // exact
  const double *tab = log_table ();
  const dd Lc = { tab[2 * (k - LOG_TAB_K0)], tab[2 * (k - LOG_TAB_K0) + 1] };
  // e ln2, ln2 = 0x1.62e42fefa39efp-1 + 0x1.abc9e3b39803fp-56
  const double ln2_hi = 6.93147180559945286227e-01;
  const double ln2_lo = 2.31904681384629955842e-17;
  double de = (double) e;
  double a = de * ln2_hi;
  double ae = fma (de, ln2_hi, -a);
  dd E = dd_quick (a, ae + de * ln2_lo);
  return dd_add (dd_add (E, Lc), L);
}

LGH_HD inline double log_dd (double q) { return log_dd2 (q).hi; }

// Half-widths (in ulps of the result) of the window around a rounding
// midpoint inside which glibc's log may round to either side: its error
// bounds 0.5 + 2.06/128 (+ 0.001 without FMA) for |log q| > 1/16 and 0.507
// near 1 (q >= 1 - 1/16, its separate polynomial path), each with margin.
static constexpr double LOG_WIN_FAR = 0.025, LOG_WIN_NEAR1 = 0.012;

// L = log_dd2(q) for q in (0, 1) (so L.hi < 0).  True when the platform log
// may return *alt, the neighbour of L.hi on the side of the exact value,
// instead of L.hi.
LGH_HD inline bool log_mid_window (dd L, double q, double *alt)
{
#pragma clang fp contract(off)
  if (L.lo == 0.0) return false;
  union { double d; uint64_t u; } h, a, o;
  h.d = L.hi;
  // L.hi < 0: a larger magnitude is the next integer encoding
  const bool toward_zero = L.lo > 0.0;
  a.u = toward_zero ? h.u - 1 : h.u + 1;
  o.u = toward_zero ? h.u + 1 : h.u - 1;
  const double gap = fabs (a.d - L.hi);
  const double ulp = fmax (gap, fabs (o.d - L.hi));
  const double w = q >= 0.9375 ? LOG_WIN_NEAR1 : LOG_WIN_FAR;
  *alt = a.d;
  return 0.5 * gap - fabs (L.lo) <= w * ulp;
}

// the accepted polar pair of lgh_randn_at: u1 and q = u1^2 + u2^2 in (0, 1)
LGH_HD inline void randn_polar (uint64_t seed, int64_t i, int64_t j,
                                double *u1_out, double *q_out)
{
#pragma clang fp contract(off)
  uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (uint64_t) (i + 1)
               + 0xC2B2AE3D27D4EB4FULL * (uint64_t) (j + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  uint64_t s = z ? z : 0x853C49E6748FEA9BULL;
  for (;;) {
    uint64_t x = s;
    x ^= x >> 12; x ^= x << 25; x ^= x >> 27; s = x;
    double u1 = (double) ((x * 2685821657736338717ULL) >> 11) / 9007199254740992.0;
    x = s;
    x ^= x >> 12; x ^= x << 25; x ^= x >> 27; s = x;
    double u2 = (double) ((x * 2685821657736338717ULL) >> 11) / 9007199254740992.0;
    u1 = 2. * u1 - 1.;
    u2 = 2. * u2 - 1.;
    double a = u1 * u1;
    double b = u2 * u2;
    double q = a + b;
    if (q > 0. && q < 1.) { *u1_out = u1; *q_out = q; return; }
  }
}

// the draw from its polar pair and a value of log(q)
LGH_HD inline double randn_from (double u1, double q, double logq)
{
#pragma clang fp contract(off)
  double f = sqrt (-2. * logq / q);
  return u1 * f;
}

// the draw with the correctly rounded log
LGH_HD inline double randn_at (uint64_t seed, int64_t i, int64_t j)
{
  double u1, q;
  randn_polar (seed, i, j, &u1, &q);
  return randn_from (u1, q, log_dd (q));
}

// the draw in output type T and whether it may differ from the platform-log
// draw (see log_mid_window): then *u1, *q are the pair to finish it with
template <typename T>
LGH_HD inline T randn_at_flag (uint64_t seed, int64_t i, int64_t j,
                               bool *flag, double *u1_out, double *q_out)
{
  double u1, q;
  randn_polar (seed, i, j, &u1, &q);
  const dd L = log_dd2 (q);
  const T v = (T) randn_from (u1, q, L.hi);
  double alt;
  *flag = log_mid_window (L, q, &alt) && (T) randn_from (u1, q, alt) != v;
  *u1_out = u1;
  *q_out = q;
  return v;
}

// the reference's draw on the host, with the platform log
LGH_H inline double randn_at_libm (uint64_t seed, int64_t i, int64_t j)
{
  double u1, q;
  randn_polar (seed, i, j, &u1, &q);
  return randn_from (u1, q, ::log (q));
}

}  // namespace lgh
