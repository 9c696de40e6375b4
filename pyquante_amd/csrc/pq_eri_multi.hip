// pq_eri_multi.hip -- one launch over the bra runs (FP64 tiles) of the light
// J/K classes (ss|ss), (ps|ss), (ps|ps), (pp|ss) and (ds|ss): the flat
// (class, bra run) work list of a Fock build in a single grid, each block
// mapped to its class by a segment table in the kernel arguments (heaviest
// class first), running the same block body as the per-class k_eri_group.
// The fused kernel carries the registers and LDS of its largest member, which
// is the price of the single launch; pq_engine_set_fused selects it.
#include "pq_eri_impl.h"

namespace pq {

namespace {
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void seg_body(const EngineView& ev, const JkSeg& g, int gidx, const double* D,
                                         double* Jacc, double* Kacc, double* lds) {
  group_body<LA, LB, LC, LD, double, MODE_JK>(ev, g.tiles, g.groups, gidx, g.kets, D, Jacc, Kacc, g.stage_d, lds);
}
}  // namespace

__global__ __launch_bounds__(256) void k_eri_group_multi(EngineView ev, JkMulti m, const double* __restrict__ D,
                                                         double* Jacc, double* Kacc) {
  extern __shared__ double pq_lds[];
  const int b = blockIdx.x;
  int s = 0;
  while (s + 1 < m.nseg && b >= m.seg[s + 1].b0) ++s;
  const JkSeg& g = m.seg[s];
  const int gidx = b - g.b0;
  switch (g.code) {
    case hgp_code(0, 0, 0, 0): seg_body<0, 0, 0, 0>(ev, g, gidx, D, Jacc, Kacc, pq_lds); break;
    case hgp_code(1, 0, 0, 0): seg_body<1, 0, 0, 0>(ev, g, gidx, D, Jacc, Kacc, pq_lds); break;
    case hgp_code(1, 0, 1, 0): seg_body<1, 0, 1, 0>(ev, g, gidx, D, Jacc, Kacc, pq_lds); break;
    case hgp_code(1, 1, 0, 0): seg_body<1, 1, 0, 0>(ev, g, gidx, D, Jacc, Kacc, pq_lds); break;
    case hgp_code(2, 0, 0, 0): seg_body<2, 0, 0, 0>(ev, g, gidx, D, Jacc, Kacc, pq_lds); break;
    default: break;
  }
}

bool jk_multi_class(int la, int lb, int lc, int ld) {
  const int c = hgp_code(la, lb, lc, ld);
  return c == hgp_code(0, 0, 0, 0) || c == hgp_code(1, 0, 0, 0) || c == hgp_code(1, 0, 1, 0) ||
         c == hgp_code(1, 1, 0, 0) || c == hgp_code(2, 0, 0, 0);
}

size_t jk_multi_lds(int la, int lb, int lc, int ld, int nbf, int stage_d) {
  switch (hgp_code(la, lb, lc, ld)) {
    case hgp_code(0, 0, 0, 0): return group_lds_bytes<0, 0, 0, 0, double, MODE_JK>(nbf, stage_d);
    case hgp_code(1, 0, 0, 0): return group_lds_bytes<1, 0, 0, 0, double, MODE_JK>(nbf, stage_d);
    case hgp_code(1, 0, 1, 0): return group_lds_bytes<1, 0, 1, 0, double, MODE_JK>(nbf, stage_d);
    case hgp_code(1, 1, 0, 0): return group_lds_bytes<1, 1, 0, 0, double, MODE_JK>(nbf, stage_d);
    case hgp_code(2, 0, 0, 0): return group_lds_bytes<2, 0, 0, 0, double, MODE_JK>(nbf, stage_d);
    default: return 0;
  }
}

hipError_t launch_jk_multi(const EngineView& ev, const JkMulti& m, int nblocks, size_t lds, const double* D,
                           double* Jacc, double* Kacc, hipStream_t st) {
  if (nblocks <= 0 || m.nseg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_eri_group_multi, dim3((unsigned)nblocks), dim3(256), lds, st, ev, m, D, Jacc, Kacc);
  return hipGetLastError();
}

}  // namespace pq
