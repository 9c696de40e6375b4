// pq_eri_cls.hip -- one angular-momentum class and one mode of the Rys ERI
// kernels per translation unit (PQ_LA..PQ_LD and PQ_MODE given on the
// compiler command line: 0 J/K FP64, 1 J/K FP32, 2 store, 3 Schwarz
// diagonal, 4 nuclear gradient, 5 J/K of two densities FP64, 6 deterministic
// J/K FP64 with fixed-point accumulators, 7 the same for FP32 quartets, 8
// deterministic J/K of two densities), so the instantiations compile in
// parallel.
#include "pq_eri_impl.h"
#include "pq_grad_impl.h"

#define PQ_CAT2(a, b, c, d, m) launch_cls_##a##b##c##d##_m##m
#define PQ_CAT(a, b, c, d, m) PQ_CAT2(a, b, c, d, m)

namespace pq {
#if PQ_MODE == 4
hipError_t PQ_CAT(PQ_LA, PQ_LB, PQ_LC, PQ_LD, PQ_MODE)(const EngineView& ev, const Tile* tiles, int ntiles,
                                                       const int* kets, const double* D, double* gacc, int natom,
                                                       hipStream_t st) {
  return launch_grad_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD>(ev, tiles, ntiles, kets, D, gacc, natom, st);
}
#else
hipError_t PQ_CAT(PQ_LA, PQ_LB, PQ_LC, PQ_LD, PQ_MODE)(const EngineView& ev, const Tile* tiles, int ntiles,
                                                       const int* kets, const double* D, double* Jacc,
                                                       double* Kacc, double* ints, double* qout, hipStream_t st,
                                                       const int2* groups, int ngroups, int gflags) {
#if PQ_MODE == 0
  return launch_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD, double, MODE_JK>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints,
                                                                 qout, st, groups, ngroups, gflags);
#elif PQ_MODE == 1
  return launch_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD, float, MODE_JK>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints,
                                                                qout, st, groups, ngroups, gflags);
#elif PQ_MODE == 2
  return launch_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD, double, MODE_STORE>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints,
                                                                    qout, st, nullptr, 0, 0);
#elif PQ_MODE == 5
  return launch_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD, double, MODE_JK2>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints,
                                                                  qout, st, groups, ngroups, gflags);
#elif PQ_MODE == 6
  return launch_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD, double, MODE_JKD>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints,
                                                                  qout, st, groups, ngroups, gflags);
#elif PQ_MODE == 7
  return launch_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD, float, MODE_JKD>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints,
                                                                 qout, st, groups, ngroups, gflags);
#elif PQ_MODE == 8
  return launch_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD, double, MODE_JK2D>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints,
                                                                   qout, st, groups, ngroups, gflags);
#else
  return launch_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD, double, MODE_DIAG>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints,
                                                                   qout, st, nullptr, 0, 0);
#endif
}
#endif
}  // namespace pq
