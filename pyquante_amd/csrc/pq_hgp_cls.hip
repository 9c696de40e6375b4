// pq_hgp_cls.hip -- the shell-quartet Head-Gordon-Pople store kernel of one
// angular-momentum class per translation unit (PQ_LA..PQ_LD given on the
// compiler command line), so the instantiations compile in parallel.
#include <cstdint>

#include "pq_hgp_impl.h"

#define PQ_HCAT2(a, b, c, d) launch_hgp_##a##b##c##d
#define PQ_HCAT(a, b, c, d) PQ_HCAT2(a, b, c, d)
#define PQ_PCAT2(a, b, c, d) hgp_private_##a##b##c##d
#define PQ_PCAT(a, b, c, d) PQ_PCAT2(a, b, c, d)

namespace pq {
hipError_t PQ_HCAT(PQ_LA, PQ_LB, PQ_LC, PQ_LD)(const EngineView& ev, const Tile* tiles, int ntiles,
                                               const int* kets, double* ints, hipStream_t st) {
  return launch_hgp_one<PQ_LA, PQ_LB, PQ_LC, PQ_LD>(ev, tiles, ntiles, kets, ints, st);
}

// Per-lane private memory of this class's kernel (its recurrence table or
// register spills), or SIZE_MAX when the runtime cannot tell.
size_t PQ_PCAT(PQ_LA, PQ_LB, PQ_LC, PQ_LD)() {
  hipFuncAttributes a{};
  if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_hgp_shell<PQ_LA, PQ_LB, PQ_LC, PQ_LD>)) !=
      hipSuccess)
    return SIZE_MAX;
  return a.localSizeBytes;
}
}  // namespace pq
