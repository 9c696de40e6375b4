// pq_hgp_cls.hip -- the shell-quartet Head-Gordon-Pople store kernel of one
// angular-momentum class per translation unit (PQ_LA..PQ_LD given on the
// compiler command line), so the instantiations compile in parallel.
#include "pq_hgp_impl.h"

#define PQ_HCAT2(a, b, c, d) launch_hgp_##a##b##c##d
#define PQ_HCAT(a, b, c, d) PQ_HCAT2(a, b, c, d)

namespace pq {
hipError_t PQ_HCAT(PQ_LA, PQ_LB, PQ_LC, PQ_LD)(const EngineView& ev, const Tile* tiles, int ntiles,
                                               const int* kets, double* ints, hipStream_t st) {
  return launch_hgp_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD>(ev, tiles, ntiles, kets, ints, st);
}
}  // namespace pq
