// Host harness over pyquante_amd/csrc/pq_fixed.h for the CPU tests: the same
// split / accumulate / convert steps the deterministic J/K kernels run, with
// the limb sums kept in unsigned 64-bit arithmetic as the device's integer
// atomics keep them.
#include "../../pyquante_amd/csrc/pq_fixed.h"

extern "C" {

void fx_split3(double x, long long* out) {
  const pq::FxLimbs l = pq::fx_split(x);
  out[0] = l.l0;
  out[1] = l.l1;
  out[2] = l.l2;
}

// sum of x[order[0]], x[order[1]], ... through the limbs
double fx_sum(const double* x, const int* order, int n) {
  unsigned long long s[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const pq::FxLimbs l = pq::fx_split(x[order[i]]);
    s[0] += (unsigned long long)l.l0;
    s[1] += (unsigned long long)l.l1;
    s[2] += (unsigned long long)l.l2;
  }
  return pq::fx_value((long long)s[0], (long long)s[1], (long long)s[2]);
}

double fx_value3(long long s0, long long s1, long long s2) { return pq::fx_value(s0, s1, s2); }
}
