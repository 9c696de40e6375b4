// pq_fixed.h -- exact fixed-point accumulation for the deterministic J/K
// build.
//
// A contribution x (FP64) is split exactly into three integer limbs
//     x ~= l2 + l1 * 2^-32 + l0 * 2^-64,
// l2 = trunc(x), |l1| < 2^32, |l0| <= 2^32, all three with the sign of x.
// The only rounding is l0 to the nearest integer, so every contribution is
// represented to within 2^-65 absolute.  An accumulator is three int64
// sums, one per limb; integer addition is associative, so the accumulated
// value does not depend on the order in which the contributions arrive
// (atomics from concurrent waves, blocks and streams, LDS or global, any
// tiling).  The limb sums hold at least 2^31 contributions each (|l1|, |l0|
// <= 2^32) before they could overflow, and |x| must stay below 2^62.
//
// fx_value() turns the three sums back into FP64: carry-normalise to
// s2 * 2^64 + r1 * 2^32 + r0 with r1, r0 in [0, 2^32), take the magnitude as
// a 128-bit integer and convert it with at most two roundings.  Same sums,
// same bits.
//
// The functions are host-callable too, so the arithmetic can be tested on the
// CPU with a host compiler (PQ_HD is empty there).
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define PQ_HD __host__ __device__
#else
#define PQ_HD
#endif

namespace pq {

struct FxLimbs {
  long long l0, l1, l2;
};

PQ_HD inline FxLimbs fx_split(double x) {
  const double a = fabs(x);
  const double hi = floor(a);           // integer part, exact
  const double f1 = (a - hi) * 0x1p32;  // fraction * 2^32, exact
  const double mid = floor(f1);
  const double lo = rint((f1 - mid) * 0x1p32);
  FxLimbs r;
  r.l2 = (long long)hi;
  r.l1 = (long long)mid;
  r.l0 = (long long)lo;
  if (x < 0.0) {
    r.l2 = -r.l2;
    r.l1 = -r.l1;
    r.l0 = -r.l0;
  }
  return r;
}

PQ_HD inline double fx_value(long long s0, long long s1, long long s2) {
  // carry-normalise: s0 = q0 2^32 + r0 with r0 in [0, 2^32), and so on
  // (the arithmetic right shift rounds towards minus infinity)
  const long long q0 = s0 >> 32;
  const uint64_t r0 = (uint64_t)(s0 - (q0 << 32));
  s1 += q0;
  const long long q1 = s1 >> 32;
  const uint64_t r1 = (uint64_t)(s1 - (q1 << 32));
  s2 += q1;
  // value * 2^64 = hi * 2^64 + lo as a two's-complement 128-bit integer
  uint64_t lo = (r1 << 32) | r0;
  uint64_t hi = (uint64_t)s2;
  const bool neg = s2 < 0;
  if (neg) {  // magnitude: negate the 128-bit integer
    lo = ~lo + 1;
    hi = ~hi + (lo == 0 ? 1 : 0);
  }
  const double mag = (double)hi + (double)lo * 0x1p-64;
  return neg ? -mag : mag;
}

}  // namespace pq
