// pq_hgp_impl.h -- Head-Gordon-Pople ERIs per shell quartet (chgp.c:78-433,
// the reference's HGP backend, evaluated there one function quartet and one
// primitive quartet at a time).
//
// One lane computes one contracted shell quartet (ab|cd) of class
// (LA LB|LC LD), or part of its primitive sum, from tiles of one bra pair and
// up to 64 kets (split over bra primitives and ket-primitive lanes as the J/K
// tiles are; a partial block is added into the zeroed array with atomics):
//   1. per primitive quartet, the Obara-Saika vertical recurrence builds
//      [e0|f0]^(m) for |e| <= LA+LB, |f| <= LC+LD from the Boys values
//      F_m(X), m <= L (chgp.c:244-419, vrr);
//   2. the m = 0 values with |e| >= LA, |f| >= LC are summed over the
//      contraction;
//   3. the horizontal recurrence (a, b+1_i| = (a+1_i, b| + (A-B)_i (a, b|
//      (chgp.c:78-170, contr_hrr) runs once per contracted block, in closed
//      form:  (ab|cd) = sum_{k<=b, l<=d} prod_i C(b_i,k_i) AB_i^(b_i-k_i)
//      C(d_i,l_i) CD_i^(d_i-l_i) [a+k,0|c+l,0].
// The Boys values come from the Rys rule of order L/2+1 at the same X:
// F_m(X) = sum_r w_r t_r^(2m) holds exactly for m < 2 (L/2+1), which covers
// every m <= L, and every term is positive.
//
// Up to kHgpAloneMax table entries (every class but (dp|dp), (dd|pp),
// (dd|ds), (dd|dp) and (dd|dd)) every index of the recurrences is a
// compile-time constant (static_for over the Cartesian triples), and the
// tables stay in registers; the classes above kHgpUnrollMax among them run in
// launches of their own.  The five larger classes walk step tables in
// constant memory instead (uniform loads), with the recurrence table in
// per-lane private memory (at most 3,366 doubles): unrolled, (dp|pp) at 285
// entries already spills 310 registers, and (dd|dd) took 12 minutes to
// compile.
#pragma once
#include <type_traits>

#include "pq_eri_impl.h"
#include "pq_hgp_core.h"

namespace pq {

// Output split of (dd|dd): blockIdx.y picks the bra function a, and each of
// the 6 waves of a tile redoes the vertical recurrence and contraction but
// runs the horizontal one (1,296 outputs, most of the lane's work with
// single-primitive d shells) for its own a only.  Measured: (dd|dd) 2.4x
// faster; the same split slowed (dp|dp), (dd|pp) and (dd|ds), whose vertical
// recurrence over several primitives outweighs their horizontal one.
template <int LA, int LB, int LC, int LD>
constexpr int kHgpSplit =
    (HgpCls<LA, LB, LC, LD>::SIZE > kHgpUnrollMax && ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD) >= 1296)
        ? ncart(LA)
        : 1;

// One tile (and output part `split` of kHgpSplit) of class (LA LB|LC LD).
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void hgp_tile(const EngineView& ev, const Tile& tl, const int* __restrict__ kets,
                                         double* __restrict__ ints, int split) {
  using H = HgpCls<LA, LB, LC, LD>;
  const TileLane tn = tile_lane(tl.nket, threadIdx.x);
  if (!tn.active) return;
  const int bra = tl.bra, ket = kets[tl.ket0 + tn.slot];
  const PairInfo bp = ev.pairs[bra], kp = ev.pairs[ket];
  const int i0 = tl.prim & 0xffff, ni = tl.prim >> 16;
  const int i1 = ni ? i0 + ni : bp.pp_n;
  const int k0 = ket_chunk(tn.kc, kp.pp_n, tn.ks), k1 = ket_chunk(tn.kc + 1, kp.pp_n, tn.ks);
  // a lane holding part of the primitive sum adds its block into the zeroed
  // array; a whole quartet is stored
  const bool part = tl.prim != 0 || tn.ks > 1;
  // RS: the step-table classes whose targets live in private memory; a
  // quartet of one primitive quartet (single) leaves them in the recurrence
  // table V, and the horizontal recurrence reads them there
  constexpr bool RS = !kHgpUnrolled<LA, LB, LC, LD>;
  constexpr int NACC = RS ? H::NTP : H::NT;
  const bool single = RS && i1 - i0 == 1 && k1 - k0 == 1;
  double acc[NACC];
  double V[kHgpUnrolled<LA, LB, LC, LD> ? 1 : H::SIZE + H::L + 1];
  if (!single) {
#pragma unroll
    for (int c = 0; c < NACC; ++c) acc[c] = 0.0;
  }
#pragma unroll 1
  for (int k = k0; k < k1; ++k) {
    const PrimPair q = ev.prims[kp.pp_off + k];
#pragma unroll 1
    for (int i = i0; i < i1; ++i) {
      if constexpr (kHgpUnrolled<LA, LB, LC, LD>)
        hgp_vrr<LA, LB, LC, LD>(bp, ev.prims[bp.pp_off + i], kp, q, acc);
      else
        hgp_vrr_rolled<LA, LB, LC, LD>(bp, ev.prims[bp.pp_off + i], kp, q, acc, V, !single);
    }
  }
  // powers of A - B and C - D for the horizontal recurrence
  double ABp[3][3], CDp[3][3];
  const double AB[3] = {bp.ABx, bp.ABy, bp.ABz}, CD[3] = {kp.ABx, kp.ABy, kp.ABz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ABp[d][0] = CDp[d][0] = 1.0;
    ABp[d][1] = AB[d];
    ABp[d][2] = AB[d] * AB[d];
    CDp[d][1] = CD[d];
    CDp[d][2] = CD[d] * CD[d];
  }
  const bool same = bra == ket;
  if constexpr (RS) {
    using S = HgpSteps<LA, LB, LC, LD>;
    const S& T = kHgpSteps<LA, LB, LC, LD>;
    constexpr int NABS = S::NAB / kHgpSplit<LA, LB, LC, LD>;
    const int ab0 = split * NABS;
    // the contracted targets: acc, or V through the target table
    auto hrr = [&](auto from_v) {
      constexpr bool FV = decltype(from_v)::value;
      auto tgt = [&](int t) -> double { return FV ? V[T.tgt[t]] : acc[t]; };
#pragma unroll 1
    for (int ab = ab0; ab < ab0 + NABS; ++ab) {
      const int a = ab / H::NB, bb = ab % H::NB;
#pragma unroll 1
      for (int cd = 0; cd < S::NCD; ++cd) {
        const int c = cd / H::ND, d = cd % H::ND;
        const int fi = bp.bfA + a, fj = bp.bfB + bb, fk = kp.bfA + c, fl = kp.bfB + d;
        if (!store_canonical(bp, kp, same, fi, fj, fk, fl)) continue;
        // every term slot (unused ones have cb = 0 and read acc[idx * NFC]),
        // so that the acc loads of one output are in flight together
        constexpr int MTB = LB == 0 ? 1 : (LB == 1 ? 2 : S::MAXT), MTK = LD == 0 ? 1 : (LD == 1 ? 2 : S::MAXT);
        double v = 0.0;
#pragma unroll
        for (int x = 0; x < MTB; ++x) {
          const HgpTerm tb = T.bt[ab][x];
          double r = 0.0;
#pragma unroll
          for (int y = 0; y < MTK; ++y) {
            const HgpTerm tk = T.kt[cd][y];
            r += tk.cb * CDp[0][tk.p0] * CDp[1][tk.p1] * CDp[2][tk.p2] * tgt(tb.idx * H::NFC + tk.idx);
          }
          v += tb.cb * ABp[0][tb.p0] * ABp[1][tb.p1] * ABp[2][tb.p2] * r;
        }
        v *= cart_norm(LA, a) * cart_norm(LB, bb) * cart_norm(LC, c) * cart_norm(LD, d);
        double* dst = ints + ijkl2intindex(fi, fj, fk, fl);
        if (part) unsafeAtomicAdd(dst, v);
        else *dst = v;
      }
    }
    };
    if (single) hrr(std::true_type{});
    else hrr(std::false_type{});
    return;
  }
  double out[H::NA * H::NB * H::NC * H::ND];
  hgp_hrr_block<LA, LB, LC, LD>(acc, AB, CD, out);
  static_for<0, H::NA>([&](auto A_) {
    static_for<0, H::NB>([&](auto B_) {
      static_for<0, H::NC>([&](auto C_) {
        static_for<0, H::ND>([&](auto D_) {
          constexpr int a = decltype(A_)::value, bb = decltype(B_)::value;
          constexpr int c = decltype(C_)::value, d = decltype(D_)::value;
          const int fi = bp.bfA + a, fj = bp.bfB + bb, fk = kp.bfA + c, fl = kp.bfB + d;
          if (!store_canonical(bp, kp, same, fi, fj, fk, fl)) return;
          constexpr double nf = cart_norm(LA, a) * cart_norm(LB, bb) * cart_norm(LC, c) * cart_norm(LD, d);
          const double v = out[((a * H::NB + bb) * H::NC + c) * H::ND + d];
          double* dst = ints + ijkl2intindex(fi, fj, fk, fl);
          if (part) unsafeAtomicAdd(dst, v * nf);
          else *dst = v * nf;
        });
      });
    });
  });
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_hgp_shell(EngineView ev, const Tile* __restrict__ tiles,
                                                  const int* __restrict__ kets, double* __restrict__ ints) {
  hgp_tile<LA, LB, LC, LD>(ev, tiles[blockIdx.x], kets, ints, (int)blockIdx.y);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_hgp_one(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets, double* ints,
                          hipStream_t st) {
  if (ntiles <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_hgp_shell<LA, LB, LC, LD>), dim3((unsigned)ntiles, kHgpSplit<LA, LB, LC, LD>), dim3(64), 0,
                     st, ev, tiles, kets, ints);
  return hipGetLastError();
}

}  // namespace pq
