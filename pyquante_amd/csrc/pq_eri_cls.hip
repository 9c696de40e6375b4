// pq_eri_cls.hip -- one angular-momentum class of the Rys ERI kernels per
// translation unit (PQ_LA..PQ_LD given on the compiler command line), so the
// classes compile in parallel.
#include "pq_eri_impl.h"

#define PQ_CAT2(a, b, c, d) launch_cls_##a##b##c##d
#define PQ_CAT(a, b, c, d) PQ_CAT2(a, b, c, d)

namespace pq {
hipError_t PQ_CAT(PQ_LA, PQ_LB, PQ_LC, PQ_LD)(int mode, int fp32, const EngineView& ev, const Tile* tiles,
                                              int ntiles, const int* kets, const double* D, double* Jacc,
                                              double* Kacc, double* ints, double* qout, hipStream_t st) {
  return launch_cls<PQ_LA, PQ_LB, PQ_LC, PQ_LD>(mode, fp32, ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout,
                                                st);
}
}  // namespace pq
