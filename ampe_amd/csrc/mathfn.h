// Elementary functions the hot kernels call per cell, written for the
// argument ranges they see.  Plain C++ (host and device), so the CPU tests
// compile the same source.
#pragma once

#include <cmath>

#ifndef AMPE_HD
#define AMPE_HD inline
#endif

namespace ampe {

// Natural logarithm within 1 ulp (log 0 = -inf, log inf = inf, NaN for
// negative or NaN arguments).  x = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m - 1 and
// s = f / (2 + f), so log m = 2 atanh s = f - s f + s R(s^2) with the Taylor
// series R(z) = sum_{k>=1} 2 z^k / (2k + 1); |s| <= 0.1716 and ten terms leave
// a remainder below 1e-18 relative.  Assembled as f - (h - s (h + R)) with
// h = f^2 / 2 (the arrangement of FDLIBM's e_log.c, which keeps the leading
// terms exact), plus e ln 2 with ln 2 split into a 32-bit head (e ln2_hi is
// exact) and a tail.  About a third of the instructions of the general
// library routine.
AMPE_HD double log_pos(double x)
{
   if (!(x > 0.0)) return (x == 0.0) ? -HUGE_VAL : __builtin_nan("");
   if (x == HUGE_VAL) return x;
   const double ln2_hi = 0.6931471803691238;     // 0x3fe62e42fee00000
   const double ln2_lo = 1.9082149292705877e-10;  // ln 2 - ln2_hi
   int e;
   double m = std::frexp(x, &e);  // [0.5, 1), subnormal x included
   if (m < 0.7071067811865476) {
      m = m + m;
      e -= 1;
   }
   double f = m - 1.0;
   double s = f / (2.0 + f);
   double z = s * s;
   double t = 2.0 / 21.0;
   t = std::fma(t, z, 2.0 / 19.0);
   t = std::fma(t, z, 2.0 / 17.0);
   t = std::fma(t, z, 2.0 / 15.0);
   t = std::fma(t, z, 2.0 / 13.0);
   t = std::fma(t, z, 2.0 / 11.0);
   t = std::fma(t, z, 2.0 / 9.0);
   t = std::fma(t, z, 2.0 / 7.0);
   t = std::fma(t, z, 2.0 / 5.0);
   t = std::fma(t, z, 2.0 / 3.0);
   double R = t * z;
   double h = 0.5 * f * f;
   double dk = (double)e;
   return dk * ln2_hi - ((h - (s * (h + R) + dk * ln2_lo)) - f);
}

}  // namespace ampe
