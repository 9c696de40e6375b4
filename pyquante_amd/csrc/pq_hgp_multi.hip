// pq_hgp_multi.hip -- every shell-quartet HGP class in one launch.  The
// grids of single classes are small (144 to 15,000 waves, the d classes
// below 1,024, against 1,024 SIMDs) and the waves of the step-table classes
// run long serial recurrences, so one class at a time leaves most of the chip
// idle; one launch runs them together on one queue (side streams would each
// need their own private-memory backing).  Block b belongs to the segment
// with the largest first block <= b; within it, block (tile * split + part)
// runs output part `part` of tile `tile`.
#include "pq_hgp_impl.h"

namespace pq {

#define PQ_HGP_CLASSES(X) \
  X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(2, 0, 0, 0) X(2, 1, 0, 0) \
  X(1, 1, 1, 0) X(1, 1, 1, 1) X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 1, 0) X(2, 1, 1, 1) X(2, 1, 2, 0) \
  X(2, 1, 2, 1) X(2, 2, 0, 0) X(2, 2, 1, 0) X(2, 2, 1, 1) X(2, 2, 2, 0) X(2, 2, 2, 1) X(2, 2, 2, 2)


__global__ __launch_bounds__(64) void k_hgp_multi(EngineView ev, HgpMulti m, double* __restrict__ ints) {
  const int b = (int)blockIdx.x;
  int s = 0;
  while (s + 1 < m.nseg && b >= m.seg[s + 1].b0) ++s;
  const HgpSeg& g = m.seg[s];
  const int local = b - g.b0, tile = local / g.split, part = local % g.split;
  const Tile tl = g.tiles[tile];
  switch (g.code) {
#define PQ_HGP_CASE(A, B, C_, D_)                                   \
  case hgp_code(A, B, C_, D_):                                      \
    if constexpr (!kHgpAlone<A, B, C_, D_>)                         \
      hgp_tile<A, B, C_, D_>(ev, tl, g.kets, ints, part);           \
    break;
    PQ_HGP_CLASSES(PQ_HGP_CASE)
#undef PQ_HGP_CASE
    default:
      break;
  }
}

bool hgp_multi_class(int la, int lb, int lc, int ld, int* split) {
#define PQ_HGP_Q(A, B, C_, D_)                             \
  if (la == A && lb == B && lc == C_ && ld == D_) {        \
    *split = kHgpSplit<A, B, C_, D_>;                      \
    return !kHgpAlone<A, B, C_, D_>;                       \
  }
  PQ_HGP_CLASSES(PQ_HGP_Q)
#undef PQ_HGP_Q
  return false;
}

hipError_t launch_hgp_multi(const EngineView& ev, const HgpMulti& m, int nblocks, double* ints, hipStream_t st) {
  if (m.nseg <= 0 || nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_hgp_multi, dim3((unsigned)nblocks), dim3(64), 0, st, ev, m, ints);
  return hipGetLastError();
}

}  // namespace pq
