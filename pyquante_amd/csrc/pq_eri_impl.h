// pq_eri_impl.h -- Rys-quadrature two-electron-integral kernels for gfx950.
//
// One template per angular-momentum class (la lb|lc ld), l <= 2, la >= lb,
// lc >= ld, pair class (la,lb) >= (lc,ld).  Work arrives as tiles: one bra
// shell pair and up to 64 ket shell pairs of equal contraction length taken
// from the ket class's pair list (sorted by Schwarz bound), so every lane of
// a wavefront shares the bra pair and the primitive loop trip count.
//
// One lane computes one shell quartet.  For the large classes the Cartesian
// block is split over blockIdx.y: each lane then owns the components of one
// bra function a (SPLIT 1) or one bra function pair (a, b) (SPLIT 2), keeps
// only those accumulators in registers, and recomputes the (cheap) roots and
// 1-D recurrences; the compiler drops the transfer rows it does not need.
//
// Modes: MODE_JK digests the block into J and K (the reference's
// getJ/getK/get2JmK, Ints.py:195-238, done directly from the integrals),
// MODE_JK2 digests one pass of integrals into J and K of two densities (the
// four getJ/getK calls of an unrestricted iteration, hartree_fock.py:233-236),
// MODE_STORE writes the reference's flat unique-ERI array (Ints.py:146-167,
// index cints.c:413-425), MODE_DIAG produces Schwarz bounds from (ab|ab).
#pragma once
#include <type_traits>

#include "pq_device.h"
#include "pq_fixed.h"
#include "pq_hgp_core.h"
#include "pq_kernels.h"

namespace pq {

template <int LA, int LB, int LC, int LD>
struct Cls {
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int NCOMP = NA * NB * NC * ND;
  static constexpr int L = LA + LB + LC + LD;
  static constexpr int NR = L / 2 + 1;
  // component split: 0 none, 1 over a, 2 over (a, b)
  static constexpr int SPLIT = NCOMP <= kLaneMaxComp ? 0 : (NCOMP / NA <= kLaneMaxComp ? 1 : 2);
  static constexpr int NSPLIT = SPLIT == 0 ? 1 : (SPLIT == 1 ? NA : NA * NB);
};

// Contracted block of one shell quartet for bra components a in [A0, A0+NAS),
// b in [B0, B0+NBS) and every ket component; PyQuante component order, without
// the per-component normalisation factor.
// UB: the bra pair is the same for the whole wavefront (scalar loads of its
// primitive pairs).
template <int LA, int LB, int LC, int LD, int A0, int NAS, int B0, int NBS, typename T, bool UB = true,
          bool LT = false>
__device__ __forceinline__ void eri_block(const EngineView& ev, const PairInfo& bp, const PairInfo& kp, int i0,
                                          int i1, int k0, int k1, T (&out)[NAS * NBS * ncart(LC) * ncart(LD)],
                                          const void* rtab = nullptr) {
  using C = Cls<LA, LB, LC, LD>;
  constexpr int NR = C::NR;
#pragma unroll
  for (int c = 0; c < NAS * NBS * C::NC * C::ND; ++c) out[c] = (T)0;
  const T AB[3] = {(T)bp.ABx, (T)bp.ABy, (T)bp.ABz};
  const T CD[3] = {(T)kp.ABx, (T)kp.ABy, (T)kp.ABz};
  // one primitive quartet of the contraction, accumulated into out
  auto quartet = [&](const PrimPair& b, const PrimPair& q) __attribute__((always_inline)) {
    const PrimQ s = prim_setup(bp, b, kp, q);
    const typename RysRule<NR, T>::Point rp = RysRule<NR, T>::template prepare<LT>(s.X, rtab);
    const T pref = (T)s.pref, qfac = (T)s.qfac, pfac = (T)s.pfac;
    const T half_rpq = (T)(0.5 * s.rpq), i2p = (T)s.i2p, i2q = (T)s.i2q;
    // roots unrolled for the small orders; rolled above to bound code size
    constexpr int UNR = NR <= 2 ? NR : 1;
#pragma unroll UNR
    for (int r = 0; r < NR; ++r) {
      T ur, wr;
      RysRule<NR, T>::root(rp, r, ur, wr);
      const T B00 = half_rpq * ur;
      const T B10 = i2p * ((T)1 - qfac * ur);
      const T B01 = i2q * ((T)1 - pfac * ur);
      T Ix[LA + 1][LB + 1][LC + 1][LD + 1], Iy[LA + 1][LB + 1][LC + 1][LD + 1],
          Iz[LA + 1][LB + 1][LC + 1][LD + 1];
      rys_1d<LA, LB, LC, LD, T>((T)1, (T)s.PA[0] - qfac * (T)s.PQ[0] * ur,
                                (T)s.QC[0] + pfac * (T)s.PQ[0] * ur, B10, B01, B00, AB[0], CD[0], Ix);
      rys_1d<LA, LB, LC, LD, T>((T)1, (T)s.PA[1] - qfac * (T)s.PQ[1] * ur,
                                (T)s.QC[1] + pfac * (T)s.PQ[1] * ur, B10, B01, B00, AB[1], CD[1], Iy);
      rys_1d<LA, LB, LC, LD, T>(wr * pref, (T)s.PA[2] - qfac * (T)s.PQ[2] * ur,
                                (T)s.QC[2] + pfac * (T)s.PQ[2] * ur, B10, B01, B00, AB[2], CD[2], Iz);
#pragma unroll
      for (int a = 0; a < NAS; ++a)
#pragma unroll
        for (int bb = 0; bb < NBS; ++bb)
#pragma unroll
          for (int c = 0; c < C::NC; ++c)
#pragma unroll
            for (int d = 0; d < C::ND; ++d) {
              const int o = ((a * NBS + bb) * C::NC + c) * C::ND + d;
              out[o] += Ix[cart_pow(LA, A0 + a, 0)][cart_pow(LB, B0 + bb, 0)][cart_pow(LC, c, 0)][cart_pow(LD, d, 0)] *
                        Iy[cart_pow(LA, A0 + a, 1)][cart_pow(LB, B0 + bb, 1)][cart_pow(LC, c, 1)][cart_pow(LD, d, 1)] *
                        Iz[cart_pow(LA, A0 + a, 2)][cart_pow(LB, B0 + bb, 2)][cart_pow(LC, c, 2)][cart_pow(LD, d, 2)];
            }
    }
  };
  auto bra_prim = [&](int i) __attribute__((always_inline)) {
    return UB ? ld_uniform(ev.prims + bp.pp_off + i) : ev.prims[bp.pp_off + i];
  };
  // ket primitives (per lane) outside, bra primitives (shared by the
  // wavefront, scalar loads) inside.  Light classes load the next ket
  // primitive pair while the current one is processed (its latency is then
  // hidden behind the bra loop); heavy ones keep the registers.
  constexpr bool PF = NAS * NBS * C::NC * C::ND <= 27;
  if constexpr (PF) {
    if (k0 >= k1) return;
    PrimPair q = ev.prims[kp.pp_off + k0];
#pragma unroll 1
    for (int k = k0; k < k1; ++k) {
      const PrimPair qn = ev.prims[kp.pp_off + (k + 1 < k1 ? k + 1 : k)];
#pragma unroll 1
      for (int i = i0; i < i1; ++i) quartet(bra_prim(i), q);
      q = qn;
    }
  } else {
#pragma unroll 1
    for (int k = k0; k < k1; ++k) {
      const PrimPair q = ev.prims[kp.pp_off + k];
#pragma unroll 1
      for (int i = i0; i < i1; ++i) quartet(bra_prim(i), q);
    }
  }
}

// ---------------------------------------------------------------------
// The J/K classes of total angular momentum 2 and 3 ((ps|ps), (pp|ss),
// (ds|ss); (pp|ps), (ds|ps), (dp|ss)): Head-Gordon-Pople recurrences instead
// of Rys quadrature.  Per primitive quartet the Boys values F_0..F_L come
// from the grid table (staged in LDS by the bra-run kernels; L + 11 doubles
// and a Taylor sum per order) and the Obara-Saika vertical recurrence builds
// the [e0|f0] targets (tables of 15-52 entries, every index a compile-time
// constant, so they stay in registers); those are summed over the
// contraction and the horizontal recurrence runs once per contracted block.
// Two-root Rys quadrature reads 48 table coefficients and spends 44 FMAs on
// the roots alone.  Measured on (H2O)_8 6-31G**, kernel time inside the
// concurrent build: (pp|ss) 0.206 -> 0.148 ms, (ds|ss) 0.102 -> 0.093,
// (ps|ps) 0.309 -> 0.303; one class at a time: (pp|ps) 0.214 -> 0.166 ms,
// (ds|ps) 0.083 -> 0.064, (dp|ss) 0.099 -> 0.077.  With one root (L <= 1)
// the two cost the same per primitive quartet and the Rys kernels hold fewer
// registers ((ss|ss) 0.117 against 0.128 ms), so they stay Rys; from L = 4
// the tables outgrow the register file.  Same block layout as eri_block;
// FP64 only (the FP32 tiles of mixed builds keep the Rys path).
// ---------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename T, int MODE>
constexpr bool kHgpJK = mode_is_jk(MODE) && std::is_same<T, double>::value &&
                        (LA + LB + LC + LD == 2 || LA + LB + LC + LD == 3);

template <int LA, int LB, int LC, int LD, bool LT>
__device__ __forceinline__ void hgp_jk_block(const EngineView& ev, const PairInfo& bp, const PairInfo& kp, int i0,
                                             int i1, int k0, int k1,
                                             double (&out)[ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD)],
                                             const double* btab) {
  using H = HgpCls<LA, LB, LC, LD>;
  double acc[H::NEA * H::NFC];
#pragma unroll
  for (int c = 0; c < H::NEA * H::NFC; ++c) acc[c] = 0.0;
  if (k0 < k1) {
    // the next ket primitive pair loads while the bra loop runs
    PrimPair q = ev.prims[kp.pp_off + k0];
#pragma unroll 1
    for (int k = k0; k < k1; ++k) {
      const PrimPair qn = ev.prims[kp.pp_off + (k + 1 < k1 ? k + 1 : k)];
#pragma unroll 1
      for (int i = i0; i < i1; ++i)
        hgp_vrr<LA, LB, LC, LD, true, LT>(bp, ld_uniform(ev.prims + bp.pp_off + i), kp, q, acc, btab);
      q = qn;
    }
  }
  const double AB[3] = {bp.ABx, bp.ABy, bp.ABz}, CD[3] = {kp.ABx, kp.ABy, kp.ABz};
  hgp_hrr_block<LA, LB, LC, LD>(acc, AB, CD, out);
}

// ---------------------------------------------------------------------
// Digestion of one normalised integral block into J/K accumulators.
// With f the permutational degeneracy factor of the quartet (1/2 for each of
// A==B, C==D, AB==CD) every unique (ab|cd) contributes
//   Jacc[a,b] += 2 f v D[c,d],  Jacc[c,d] += 2 f v D[a,b],
//   Kacc[a,c] += f v D[b,d],    Kacc[a,d] += f v D[b,c],
//   Kacc[b,c] += f v D[a,d],    Kacc[b,d] += f v D[a,c],
// and J = Jacc + Jacc^T, K = Kacc + Kacc^T (symmetric D).
// ---------------------------------------------------------------------
// krow, when non-null, is the block's LDS copy of the bra rows of K (rows
// 0..NA-1 for the functions of shell A, NA..NA+NB-1 for shell B, nbf columns
// each): the four K blocks accumulate there with LDS atomics and leave the
// block as coalesced global atomics (flush_block), one cache line per 16
// columns instead of one per lane.  jsum, when non-null, is the block's LDS
// accumulator of the ket-side J block per lane (shared by the component
// splits of one tile), flushed once per block.  drow, when non-null, is the
// block's LDS copy of the bra rows of D (same layout as krow): the D elements
// of the bra rows then come from LDS and only D[c,d] is gathered from memory.
// jrow, when non-null, is the block's LDS accumulator of the bra J block
// (NA x NB), flushed once per block instead of one global atomic per tile.
//
// The digestion runs in three phases so that few D elements and accumulators
// are live at once (the register peak of the kernel, and so its occupancy):
// the J blocks; then the K rows of each bra function a (from the D rows of
// b), each row flushed as soon as it is complete; then the K rows of each b
// (from the D rows of a).
// Adds x into accumulator element p: an FP64 atomic (global memory, or LDS
// when GLB is false), or in the fixed-point modes the three limbs of x into
// the int64 sums at p, p + stride and p + 2 stride.
//
// The atomics go through pointers of their own address space (LDS or global
// memory).  Through generic pointers the compiler merges an LDS branch and a
// global branch of the same add into one FLAT atomic on a selected address,
// which then takes the vector-memory path (and its latency) for the LDS
// accumulators as well.
template <typename T>
__device__ __forceinline__ void atomic_add_as(T* p, T x, bool glb) {
  using lds_t = __attribute__((address_space(3))) T;
  using glb_t = __attribute__((address_space(1))) T;
  if (glb) __hip_atomic_fetch_add((glb_t*)p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else __hip_atomic_fetch_add((lds_t*)p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int MODE, bool GLB>
__device__ __forceinline__ void acc_add(double* p, int64_t stride, double x) {
  if constexpr (mode_limbs(MODE) == 3) {
    const FxLimbs l = fx_split(x);
    unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
    if (l.l0) atomic_add_as(q, (unsigned long long)l.l0, GLB);
    if (l.l1) atomic_add_as(q + stride, (unsigned long long)l.l1, GLB);
    if (l.l2) atomic_add_as(q + 2 * stride, (unsigned long long)l.l2, GLB);
  } else {
    atomic_add_as(p, x, GLB);
  }
}

// Flushes accumulator element i of a block's LDS array (limbs stride lstride
// apart in the fixed-point modes) into global element p (limbs gstride apart); zero
// elements are skipped.
template <int MODE>
__device__ __forceinline__ void acc_flush(const double* lds, int64_t i, int64_t lstride, double* p,
                                          int64_t gstride) {
  if constexpr (mode_limbs(MODE) == 3) {
    const long long* l = reinterpret_cast<const long long*>(lds);
    const long long s0 = l[i], s1 = l[i + lstride], s2 = l[i + 2 * lstride];
    unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
    if (s0) atomic_add_as(q, (unsigned long long)s0, true);
    if (s1) atomic_add_as(q + gstride, (unsigned long long)s1, true);
    if (s2) atomic_add_as(q + 2 * gstride, (unsigned long long)s2, true);
  } else {
    const double v = lds[i];
    if (v != 0.0) atomic_add_as(p, v, true);
  }
}

// Bra runs (k_eri_group) of classes whose bra block has at most 3 elements
// per density keep each lane's J_ab partial sums in registers over the whole
// run, with one wave reduction per run instead of one per tile.
template <int LA, int LB, int MODE>
constexpr bool kRunJab = ncart(LA) * ncart(LB) * mode_ndens(MODE) <= 3;

// RUN: called from a bra run of k_eri_group, where krow and jrow are always
// the block's LDS arrays and jsum is null (compile-time, so that no global
// fallback path is generated); jab_acc (RUN only, kRunJab classes): the
// lane's J_ab accumulators of the run, added to instead of reduced per tile.
template <int LA, int LB, int LC, int LD, int A0, int NAS, int B0, int NBS, typename T, int MODE, bool RUN = false>
__device__ __forceinline__ void digest_lane(const EngineView& ev, const PairInfo& bp, const PairInfo& kp,
                                            const T (&out)[NAS * NBS * ncart(LC) * ncart(LD)], double f,
                                            bool active, const double* __restrict__ D, double* Jacc,
                                            double* Kacc, double* krow, double* jsum,
                                            const double* drow, double* jrow, double* jpair, int jn, int kpos,
                                            double* jab_acc = nullptr) {
  using C = Cls<LA, LB, LC, LD>;
  constexpr int NC = C::NC, ND = C::ND, NO = NAS * NBS * NC * ND;
  // the quartet's molecule block (bra and ket share it)
  const int n = bp.mstride;
  D += bp.moff;
  Jacc += bp.moff;
  Kacc += bp.moff;
  const int a0 = bp.bfA + A0, b0 = bp.bfB + B0, c0 = kp.bfA, d0 = kp.bfB;
  // explicit address spaces keep the two paths apart (a merged load through
  // a generic pointer would be a FLAT load on both)
  using lds_cd = const __attribute__((address_space(3))) double;
  using glb_cd = const __attribute__((address_space(1))) double;
  glb_cd* Dg = (glb_cd*)D;
  lds_cd* Dl = (lds_cd*)drow;
  // D[row a of shell A, col] and D[row b of shell B, col]
  auto d_a = [&](int a, int col) -> double { return drow ? Dl[(A0 + a) * n + col] : Dg[(a0 + a) * n + col]; };
  auto d_b = [&](int b, int col) -> double { return drow ? Dl[(C::NA + B0 + b) * n + col] : Dg[(b0 + b) * n + col]; };
  // normalised integrals times the degeneracy factor
  double w[NO];
#pragma unroll
  for (int a = 0; a < NAS; ++a)
#pragma unroll
    for (int b = 0; b < NBS; ++b)
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          const int o = ((a * NBS + b) * NC + c) * ND + d;
          w[o] = (double)out[o] * (cart_norm(LA, A0 + a) * cart_norm(LB, B0 + b) * cart_norm(LC, c) *
                                   cart_norm(LD, d) * f);
        }
  const int lane = threadIdx.x & 63;
  // ---- J: J_ab (shared by the wavefront: reduced, one atomic) and J_cd
  {
    double Dcd[NC][ND], Dab[NAS][NBS], jcd[NC][ND];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        Dcd[c][d] = Dg[(c0 + c) * n + d0 + d];
        jcd[c][d] = 0.0;
      }
#pragma unroll
    for (int a = 0; a < NAS; ++a)
#pragma unroll
      for (int b = 0; b < NBS; ++b) Dab[a][b] = d_a(a, b0 + b);
#pragma unroll
    for (int a = 0; a < NAS; ++a)
#pragma unroll
      for (int b = 0; b < NBS; ++b) {
        double jab = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            const double v = w[((a * NBS + b) * NC + c) * ND + d];
            jab += v * Dcd[c][d];
            jcd[c][d] += v * Dab[a][b];
          }
        if constexpr (RUN && kRunJab<LA, LB, MODE>) {
          jab_acc[a * NBS + b] += jab;
        } else {
          const double s = wave_sum(jab);
          if (lane == 0) {
            if (RUN || jrow) acc_add<MODE, false>(&jrow[(A0 + a) * C::NB + B0 + b], C::NA * C::NB, 2.0 * s);
            else acc_add<MODE, true>(&Jacc[(a0 + a) * n + b0 + b], ev.msize, 2.0 * s);
          }
        }
      }
    if (!active) return;
    if (!RUN && jsum) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d)
          acc_add<MODE, false>(&jsum[lane * NC * ND + c * ND + d], 64 * NC * ND, 2.0 * jcd[c][d]);
    } else {
      // ket-side J by ket position: the lanes of a tile hit consecutive
      // addresses (coalesced atomics), scattered into J after the build
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d)
          acc_add<MODE, true>(&jpair[(c * ND + d) * jn + kpos], ev.jp_size, 2.0 * jcd[c][d]);
    }
  }
  // one K row (bra function row r of the block layout, matrix row grow):
  // columns c0.. (NC values) and d0.. (ND values)
  // (LDS rows: NA + NB rows of the largest molecule's nbf columns per limb)
  const int64_t kstride = (int64_t)(C::NA + C::NB) * ev.nbf;
  auto flush_row = [&](int r, int grow, const double (&kc)[NC], const double (&kd)[ND]) {
    if (RUN || krow) {
#pragma unroll
      for (int c = 0; c < NC; ++c) acc_add<MODE, false>(&krow[r * n + c0 + c], kstride, kc[c]);
#pragma unroll
      for (int d = 0; d < ND; ++d) acc_add<MODE, false>(&krow[r * n + d0 + d], kstride, kd[d]);
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) acc_add<MODE, true>(&Kacc[grow * n + c0 + c], ev.msize, kc[c]);
#pragma unroll
      for (int d = 0; d < ND; ++d) acc_add<MODE, true>(&Kacc[grow * n + d0 + d], ev.msize, kd[d]);
    }
  };
  // ---- K rows of a: K_ac += v D_bd, K_ad += v D_bc
  {
    double Dbc[NBS][NC], Dbd[NBS][ND];
#pragma unroll
    for (int b = 0; b < NBS; ++b) {
#pragma unroll
      for (int c = 0; c < NC; ++c) Dbc[b][c] = d_b(b, c0 + c);
#pragma unroll
      for (int d = 0; d < ND; ++d) Dbd[b][d] = d_b(b, d0 + d);
    }
#pragma unroll
    for (int a = 0; a < NAS; ++a) {
      double kc[NC], kd[ND];
#pragma unroll
      for (int c = 0; c < NC; ++c) kc[c] = 0.0;
#pragma unroll
      for (int d = 0; d < ND; ++d) kd[d] = 0.0;
#pragma unroll
      for (int b = 0; b < NBS; ++b)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            const double v = w[((a * NBS + b) * NC + c) * ND + d];
            kc[c] += v * Dbd[b][d];
            kd[d] += v * Dbc[b][c];
          }
      flush_row(A0 + a, a0 + a, kc, kd);
    }
  }
  // ---- K rows of b: K_bc += v D_ad, K_bd += v D_ac
  {
    double Dac[NAS][NC], Dad[NAS][ND];
#pragma unroll
    for (int a = 0; a < NAS; ++a) {
#pragma unroll
      for (int c = 0; c < NC; ++c) Dac[a][c] = d_a(a, c0 + c);
#pragma unroll
      for (int d = 0; d < ND; ++d) Dad[a][d] = d_a(a, d0 + d);
    }
#pragma unroll
    for (int b = 0; b < NBS; ++b) {
      double kc[NC], kd[ND];
#pragma unroll
      for (int c = 0; c < NC; ++c) kc[c] = 0.0;
#pragma unroll
      for (int d = 0; d < ND; ++d) kd[d] = 0.0;
#pragma unroll
      for (int a = 0; a < NAS; ++a)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            const double v = w[((a * NBS + b) * NC + c) * ND + d];
            kc[c] += v * Dad[a][d];
            kd[d] += v * Dac[a][c];
          }
      flush_row(C::NA + B0 + b, b0 + b, kc, kd);
    }
  }
}

__device__ __forceinline__ double quartet_degeneracy(const PairInfo& bp, const PairInfo& kp, int bra,
                                                     int ket) {
  double f = 1.0;
  if (bp.shA == bp.shB) f *= 0.5;
  if (kp.shA == kp.shB) f *= 0.5;
  if (bra == ket) f *= 0.5;
  return f;
}

// Is (i j|k l) the copy of its unique integral that the reference stores?
__device__ __forceinline__ bool store_canonical(const PairInfo& bp, const PairInfo& kp, bool same_pair,
                                                int i, int j, int k, int l) {
  if (bp.shA == bp.shB && i < j) return false;
  if (kp.shA == kp.shB && k < l) return false;
  if (same_pair) {
    const int64_t ij = i >= j ? (int64_t)i * (i + 1) / 2 + j : (int64_t)j * (j + 1) / 2 + i;
    const int64_t kl = k >= l ? (int64_t)k * (k + 1) / 2 + l : (int64_t)l * (l + 1) / 2 + k;
    if (ij < kl) return false;
  }
  return true;
}

// Schwarz bounds are non-negative: their IEEE bit patterns order like the
// values, so a 64-bit integer max combines the component splits.
__device__ __forceinline__ void atomic_max_nonneg(double* p, double v) {
  atomicMax((unsigned long long*)p, (unsigned long long)__double_as_longlong(v));
}

// One lane, one shell quartet, bra components [A0, A0+NAS) x [B0, B0+NBS).
template <int LA, int LB, int LC, int LD, typename T, int MODE, int A0, int NAS, int B0, int NBS, bool LT = false,
          bool RUN = false>
__device__ __forceinline__ void lane_body(const EngineView& ev, const Tile& tl, bool active, int bra, int ket,
                                          const double* __restrict__ D, double* Jacc, double* Kacc,
                                          double* ints, double* qout, double* krow, double* jsum,
                                          const void* rtab = nullptr, int kc = 0, int ks = 1, int nk = 64,
                                          const double* drow = nullptr, double* jrow = nullptr, int kpos = 0,
                                          double* jab_acc = nullptr) {
  using C = Cls<LA, LB, LC, LD>;
  constexpr bool UB = MODE != MODE_DIAG;  // Schwarz diagonal: bra = ket per lane
  const PairInfo bp = UB ? ld_uniform(ev.pairs + bra) : ev.pairs[bra];
  const PairInfo kp = ev.pairs[ket];
  constexpr int NO = NAS * NBS * C::NC * C::ND;
  T out[NO];
  if (active) {
    const int i0 = tl.prim & 0xffff, ni = tl.prim >> 16;
    const int k0 = ket_chunk(kc, kp.pp_n, ks), k1 = ket_chunk(kc + 1, kp.pp_n, ks);
    if constexpr (kHgpJK<LA, LB, LC, LD, T, MODE>) {
      static_assert(NAS == C::NA && NBS == C::NB, "the HGP J/K classes are unsplit");
      hgp_jk_block<LA, LB, LC, LD, LT>(ev, bp, kp, i0, ni ? i0 + ni : bp.pp_n, k0, k1, out,
                                       static_cast<const double*>(rtab));
    } else {
      eri_block<LA, LB, LC, LD, A0, NAS, B0, NBS, T, UB, LT>(ev, bp, kp, i0, ni ? i0 + ni : bp.pp_n, k0, k1, out,
                                                             rtab);
    }
  } else {
#pragma unroll
    for (int c = 0; c < NO; ++c) out[c] = (T)0;
  }
  if constexpr (mode_is_jk(MODE)) {
    // ket-primitive chunks of one quartet (lanes slot + j*nk) are summed into
    // the chunk-0 lane, which alone digests
    if (ks > 1) {
      for (int step = ks >> 1; step >= 1; step >>= 1) {
#pragma unroll
        for (int c = 0; c < NO; ++c) {
          const T v = __shfl_down(out[c], step * nk);
          if (kc < step) out[c] += v;
        }
      }
      active = active && kc == 0;
    }
    const double f = active ? quartet_degeneracy(bp, kp, bra, ket) : 0.0;
    // one block of integrals, digested once per density (stacked n x n
    // matrices; LDS K rows and J sums stacked the same way)
    // (accumulators: NL limb arrays per density)
    const size_t nn = (size_t)ev.msize;
    constexpr int Y = pair_class_of(LC, LD);
    constexpr int NL = mode_limbs(MODE);
#pragma unroll
    for (int d = 0; d < mode_ndens(MODE); ++d)
      digest_lane<LA, LB, LC, LD, A0, NAS, B0, NBS, T, MODE, RUN>(
          ev, bp, kp, out, f, active, D + d * nn, Jacc + d * NL * nn, Kacc + d * NL * nn,
          krow ? krow + d * NL * (C::NA + C::NB) * ev.nbf : nullptr,
          jsum ? jsum + d * NL * 64 * C::NC * C::ND : nullptr,
          drow ? drow + d * (C::NA + C::NB) * ev.nbf : nullptr, jrow ? jrow + d * NL * C::NA * C::NB : nullptr,
          ev.jpair + d * NL * ev.jp_size + ev.jp_off[Y], ev.jp_n[Y], kpos,
          jab_acc ? jab_acc + d * C::NA * C::NB : nullptr);
  } else if constexpr (MODE == MODE_STORE) {
    if (!active) return;
    if (ev.blk) {
      // one requested shell quartet (s1 s2|s3 s4) in the libint layout,
      // row-major ((i n2 + j) n3 + k) n4 + l: flags say whether the request's
      // first pair is this lane's ket pair (the kernels need bra class >=
      // ket class) and whether either requested pair lists its shells in the
      // reverse of the stored order
      const uint64_t code = ev.blk[kpos];
      const int64_t off = (int64_t)(code & ((1ull << 61) - 1));
      const int fl = (int)(code >> 61);
#pragma unroll
      for (int a = 0; a < NAS; ++a)
#pragma unroll
        for (int b = 0; b < NBS; ++b)
#pragma unroll
          for (int c = 0; c < C::NC; ++c)
#pragma unroll
            for (int d = 0; d < C::ND; ++d) {
              int x1 = A0 + a, x2 = B0 + b, y1 = c, y2 = d;
              int n1 = C::NA, n2 = C::NB, m1 = C::NC, m2 = C::ND;
              if (fl & 1) {
                int t = x1; x1 = y1; y1 = t;
                t = x2; x2 = y2; y2 = t;
                t = n1; n1 = m1; m1 = t;
                t = n2; n2 = m2; m2 = t;
              }
              if (fl & 2) {
                int t = x1; x1 = x2; x2 = t;
                t = n1; n1 = n2; n2 = t;
              }
              if (fl & 4) {
                int t = y1; y1 = y2; y2 = t;
                t = m1; m1 = m2; m2 = t;
              }
              const double nf =
                  cart_norm(LA, A0 + a) * cart_norm(LB, B0 + b) * cart_norm(LC, c) * cart_norm(LD, d);
              ints[off + ((x1 * n2 + x2) * m1 + y1) * m2 + y2] =
                  (double)out[((a * NBS + b) * C::NC + c) * C::ND + d] * nf;
            }
      return;
    }
    const bool same = bra == ket;
#pragma unroll
    for (int a = 0; a < NAS; ++a)
#pragma unroll
      for (int b = 0; b < NBS; ++b)
#pragma unroll
        for (int c = 0; c < C::NC; ++c)
#pragma unroll
          for (int d = 0; d < C::ND; ++d) {
            const int i = bp.bfA + A0 + a, j = bp.bfB + B0 + b, k = kp.bfA + c, l = kp.bfB + d;
            if (!store_canonical(bp, kp, same, i, j, k, l)) continue;
            const double nf =
                cart_norm(LA, A0 + a) * cart_norm(LB, B0 + b) * cart_norm(LC, c) * cart_norm(LD, d);
            ints[ijkl2intindex(i, j, k, l)] = (double)out[((a * NBS + b) * C::NC + c) * C::ND + d] * nf;
          }
  } else {  // MODE_DIAG: Q = max over components sqrt|(ab|ab)|
    if (!active) return;
    double qmax = 0.0;
#pragma unroll
    for (int a = 0; a < NAS; ++a)
#pragma unroll
      for (int b = 0; b < NBS; ++b) {
        const double nf = cart_norm(LA, A0 + a) * cart_norm(LB, B0 + b);
        const int c = A0 + a, d = B0 + b;  // diagonal ket component (bra == ket class)
        const double v = fabs((double)out[((a * NBS + b) * C::NC + c) * C::ND + d]) * nf * nf;
        qmax = fmax(qmax, v);
      }
    atomic_max_nonneg(&qout[ket], sqrt(qmax));
  }
}

template <int I, int N, typename F>
__device__ __forceinline__ void static_switch(int i, F&& f) {
  if constexpr (I < N) {
    if (i == I) f(std::integral_constant<int, I>{});
    else static_switch<I + 1, N>(i, f);
  }
}

// Block layouts:
//   lds_mode 0: 4 wavefronts, 4 independent tiles, global atomics only;
//   lds_mode 1: one tile per block; unsplit classes run 1 wavefront, split
//     classes kWaveSplit wavefronts that take consecutive component splits of
//     the same tile (blockIdx.y picks the chunk), sharing the LDS K rows and
//     the ket-side J accumulator.
constexpr int kWaveSplit = 4;

// Root table of order NR staged in LDS by k_eri_group (padded interval
// stride), in doubles of LDS: the FP64 table for orders 1-3 (7.4, 17.2 and
// 29.8 KB), the FP32 table for every order (1.5-14.1 KB); 0 = read the
// global table.
template <int NR>
constexpr int kRysIntervals = NR == 1 ? 37 : NR == 2 ? 44 : NR == 3 ? 51 : NR == 4 ? 58 : 64;  // rys_nint
template <int NR>
constexpr int kRysIntervals32 = NR == 1 ? 19 : NR == 2 ? 25 : NR == 3 ? 31 : NR == 4 ? 37 : 42;
template <int NR, typename T>
constexpr int kLdsTableDoubles = std::is_same<T, double>::value
                                     ? (NR <= 3 ? kRysIntervals<NR> * kRysLdsStride<NR> : 0)
                                     : kRysIntervals32<NR> * kRysLdsStride32<NR> / 2;
// per class and mode: the Boys grid table for the HGP J/K classes
template <int LA, int LB, int LC, int LD, typename T, int MODE>
constexpr int kGroupTableDoubles = kHgpJK<LA, LB, LC, LD, T, MODE>
                                       ? BOYS_NG * BOYS_STRIDE
                                       : kLdsTableDoubles<Cls<LA, LB, LC, LD>::NR, T>;

// Copies the order-NR table into LDS with the padded interval stride.
template <int NR, typename T>
__device__ __forceinline__ void stage_rys_table(double* lds) {
  if constexpr (std::is_same<T, double>::value) {
    constexpr int S = kRysStride<NR>, N = kRysIntervals<NR> * S;
    const double* src = rys_coef + rys_off[NR - 1];
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      const int k = i / S;
      lds[i + k * (kRysLdsStride<NR> - S)] = src[i];
    }
  } else {
    constexpr int S = kRysStride32<NR>, N = kRysIntervals32<NR> * S;
    const float* src = rys_coef32 + rys_off32[NR - 1];
    float* t32 = reinterpret_cast<float*>(lds);
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      const int k = i / S;
      t32[i + k * (kRysLdsStride32<NR> - S)] = src[i];
    }
  }
}

// Unsplit classes with bra groups: one block of 4 wavefronts walks a run of
// tiles that share the bra pair (wavefront w takes tiles w, w+4, ...), so the
// LDS K rows are zeroed and flushed once per run.  stage_d: the bra rows of D
// are copied to LDS once per run as well.
// The body of one block (bra run gidx) of k_eri_group, also run by the
// fused multi-class launch (pq_eri_multi.hip); pq_lds is the block's
// dynamic LDS.
template <int LA, int LB, int LC, int LD, typename T, int MODE>
__device__ __forceinline__ void group_body(const EngineView& ev, const Tile* __restrict__ tiles,
                                           const int2* __restrict__ groups, int gidx, const int* __restrict__ kets,
                                           const double* __restrict__ D, double* Jacc, double* Kacc, int stage_d,
                                           double* pq_lds) {
  using C = Cls<LA, LB, LC, LD>;
  static_assert(C::SPLIT == 0, "bra groups are for unsplit classes");
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int n = ev.nbf;
  const int2 g = groups[gidx];
  // light FP64 classes read the root table from LDS (one copy per block)
  // instead of gathering it from L1/L2 per lane
  constexpr int NT = kGroupTableDoubles<LA, LB, LC, LD, T, MODE>;
  void* rtab = NT ? pq_lds : nullptr;
  if constexpr (kHgpJK<LA, LB, LC, LD, T, MODE>) stage_boys_table(pq_lds);
  else if constexpr (NT > 0) stage_rys_table<C::NR, T>(pq_lds);
  constexpr int ND = mode_ndens(MODE);
  constexpr int NL = mode_limbs(MODE);  // accumulator words per element
  double* krow = pq_lds + NT;
  for (int i = threadIdx.x; i < ND * NL * (C::NA + C::NB) * n; i += blockDim.x) krow[i] = 0.0;
  const PairInfo bp = ld_uniform(ev.pairs + tiles[g.x].bra);
  const int ms = bp.mstride;
  const int region = (C::NA + C::NB) * n;
  // J bra block accumulator after the K rows (and the staged D rows)
  double* jrow = krow + ND * region * (NL + (stage_d ? 1 : 0));
  for (int i = threadIdx.x; i < ND * NL * C::NA * C::NB; i += blockDim.x) jrow[i] = 0.0;
  double* drow = nullptr;
  if (stage_d) {
    drow = krow + ND * NL * region;
    for (int d = 0; d < ND; ++d)
      for (int r = 0; r < C::NA + C::NB; ++r) {
        const int grow = r < C::NA ? bp.bfA + r : bp.bfB + (r - C::NA);
        for (int col = threadIdx.x; col < ms; col += blockDim.x)
          drow[d * region + r * ms + col] = D[(size_t)d * ev.msize + bp.moff + grow * ms + col];
      }
  }
  __syncthreads();
  // the lanes' J_ab partial sums over the run (kRunJab classes)
  constexpr bool RJ = kRunJab<LA, LB, MODE>;
  double jab_run[RJ ? ND * C::NA * C::NB : 1];
#pragma unroll
  for (int i = 0; i < (RJ ? ND * C::NA * C::NB : 1); ++i) jab_run[i] = 0.0;
  for (int tt = wv; tt < g.y; tt += 4) {
    Tile tl = ld_uniform(tiles + g.x + tt);
    tl.bra = __builtin_amdgcn_readfirstlane(tl.bra);
    tl.ket0 = __builtin_amdgcn_readfirstlane(tl.ket0);
    tl.nket = __builtin_amdgcn_readfirstlane(tl.nket);
    tl.prim = __builtin_amdgcn_readfirstlane(tl.prim);
    const TileLane tn = tile_lane(tl.nket, lane);
    const int ket = kets[tl.ket0 + tn.slot];
    lane_body<LA, LB, LC, LD, T, MODE, 0, C::NA, 0, C::NB, (NT > 0), true>(
        ev, tl, tn.active, tl.bra, ket, D, Jacc, Kacc, nullptr, nullptr, krow, nullptr, rtab, tn.kc, tn.ks, tn.nk,
        drow, jrow, tl.ket0 + tn.slot, RJ ? jab_run : nullptr);
  }
  constexpr int NAB = C::NA * C::NB;
  if constexpr (RJ) {
#pragma unroll
    for (int i = 0; i < ND * NAB; ++i) {
      const double s = wave_sum(jab_run[i]);
      if (lane == 0) acc_add<MODE, false>(&jrow[(i / NAB) * NL * NAB + i % NAB], NAB, 2.0 * s);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ND * NAB; i += blockDim.x) {
    const int d = i / NAB, ab = i - d * NAB;
    acc_flush<MODE>(jrow, d * NL * NAB + ab, NAB,
                    &Jacc[(size_t)d * NL * ev.msize + bp.moff + (bp.bfA + ab / C::NB) * ms + bp.bfB + ab % C::NB],
                    ev.msize);
  }
  for (int d = 0; d < ND; ++d)
    for (int r = 0; r < C::NA + C::NB; ++r) {
      const int grow = r < C::NA ? bp.bfA + r : bp.bfB + (r - C::NA);
      for (int col = threadIdx.x; col < ms; col += blockDim.x)
        acc_flush<MODE>(krow, d * NL * region + r * ms + col, region,
                        &Kacc[(size_t)d * NL * ev.msize + bp.moff + grow * ms + col], ev.msize);
    }
}

// Minimum waves per SIMD that the J/K class kernels' registers must allow
// (PQ_WPE, set per class translation unit; 0 leaves it to the compiler).
#ifndef PQ_WPE
#define PQ_WPE 0
#endif
#if PQ_WPE > 0
#define PQ_WPE_ATTR __attribute__((amdgpu_waves_per_eu(PQ_WPE)))
#else
#define PQ_WPE_ATTR
#endif

template <int LA, int LB, int LC, int LD, typename T, int MODE>
__global__ __launch_bounds__(256) PQ_WPE_ATTR void k_eri_group(EngineView ev, const Tile* __restrict__ tiles,
                                                   const int2* __restrict__ groups, int ngroups,
                                                   const int* __restrict__ kets, const double* __restrict__ D,
                                                   double* Jacc, double* Kacc, int stage_d) {
  extern __shared__ double pq_lds[];
  group_body<LA, LB, LC, LD, T, MODE>(ev, tiles, groups, (int)blockIdx.x, kets, D, Jacc, Kacc, stage_d, pq_lds);
}

// Dynamic LDS of one bra-run block: the root (or Boys) table, the K rows, the
// bra J block and, with stage_d, the staged D rows.
template <int LA, int LB, int LC, int LD, typename T, int MODE>
__host__ __device__ constexpr size_t group_lds_bytes(int nbf, bool stage_d) {
  using C = Cls<LA, LB, LC, LD>;
  constexpr size_t NL = mode_limbs(MODE);
  const size_t rows = (size_t)mode_ndens(MODE) * (C::NA + C::NB) * nbf;
  return (kGroupTableDoubles<LA, LB, LC, LD, T, MODE> + rows * (NL + (stage_d ? 1 : 0)) +
          (size_t)mode_ndens(MODE) * NL * C::NA * C::NB) * sizeof(double);
}

template <int LA, int LB, int LC, int LD, typename T, int MODE>
__global__ __launch_bounds__(256) PQ_WPE_ATTR void k_eri_lane(EngineView ev, const Tile* __restrict__ tiles, int ntiles,
                                                  const int* __restrict__ kets, const double* __restrict__ D,
                                                  double* Jacc, double* Kacc, double* ints,
                                                  double* qout, int lds_mode) {
  using C = Cls<LA, LB, LC, LD>;
  extern __shared__ double pq_lds[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int n = ev.nbf;
  const int t = __builtin_amdgcn_readfirstlane(lds_mode ? (int)blockIdx.x : (int)(blockIdx.x * 4 + wv));
  double* krow = nullptr;
  double* jsum = nullptr;
  constexpr int ND = mode_ndens(MODE);
  constexpr int NL = mode_limbs(MODE);  // accumulator words per element
  if (lds_mode) {
    krow = pq_lds;
    if (C::SPLIT) jsum = pq_lds + ND * NL * (C::NA + C::NB) * n;
    const int tot = ND * NL * ((C::NA + C::NB) * n + (C::SPLIT ? 64 * C::NC * C::ND : 0));
    for (int i = threadIdx.x; i < tot; i += blockDim.x) pq_lds[i] = 0.0;
    __syncthreads();
  } else if (t >= ntiles) {
    return;  // uniform per wavefront, no block barrier in this layout
  }
  Tile tl = ld_uniform(tiles + t);
  tl.bra = __builtin_amdgcn_readfirstlane(tl.bra);
  tl.ket0 = __builtin_amdgcn_readfirstlane(tl.ket0);
  tl.nket = __builtin_amdgcn_readfirstlane(tl.nket);
  tl.prim = __builtin_amdgcn_readfirstlane(tl.prim);
  const TileLane tn = tile_lane(tl.nket, lane);
  const bool active = tn.active;
  const int ket = kets[tl.ket0 + tn.slot];
  const int bra = (MODE == MODE_DIAG) ? ket : tl.bra;
  if constexpr (C::SPLIT == 0) {
    lane_body<LA, LB, LC, LD, T, MODE, 0, C::NA, 0, C::NB>(ev, tl, active, bra, ket, D, Jacc, Kacc, ints, qout,
                                                          krow, jsum, nullptr, tn.kc, tn.ks, tn.nk, nullptr,
                                                          nullptr, tl.ket0 + tn.slot);
  } else {
    const int split = lds_mode ? (int)blockIdx.y * kWaveSplit + wv : (int)blockIdx.y;
    if constexpr (C::SPLIT == 1) {
      static_switch<0, C::NA>(split, [&](auto ia) {
        lane_body<LA, LB, LC, LD, T, MODE, decltype(ia)::value, 1, 0, C::NB>(
            ev, tl, active, bra, ket, D, Jacc, Kacc, ints, qout, krow, jsum, nullptr, tn.kc, tn.ks, tn.nk, nullptr,
            nullptr, tl.ket0 + tn.slot);
      });
    } else {
      static_switch<0, C::NA * C::NB>(split, [&](auto iab) {
        constexpr int v = decltype(iab)::value;
        lane_body<LA, LB, LC, LD, T, MODE, v / C::NB, 1, v % C::NB, 1>(ev, tl, active, bra, ket, D, Jacc, Kacc,
                                                                        ints, qout, krow, jsum, nullptr, tn.kc,
                                                                        tn.ks, tn.nk, nullptr, nullptr,
                                                                        tl.ket0 + tn.slot);
      });
    }
  }
  if (!lds_mode) return;
  __syncthreads();
  // flush the K rows (coalesced) and the ket-side J accumulator
  const PairInfo bp = ev.pairs[tl.bra];
  const int ms = bp.mstride;
  const int region = (C::NA + C::NB) * n;
  for (int d = 0; d < ND; ++d)
    for (int r = 0; r < C::NA + C::NB; ++r) {
      const int grow = r < C::NA ? bp.bfA + r : bp.bfB + (r - C::NA);
      for (int col = threadIdx.x; col < ms; col += blockDim.x)
        acc_flush<MODE>(krow, d * NL * region + r * ms + col, region,
                        &Kacc[(size_t)d * NL * ev.msize + bp.moff + grow * ms + col], ev.msize);
    }
  if constexpr (C::SPLIT != 0) {
    constexpr int NCD = C::NC * C::ND;
    constexpr int Y = pair_class_of(LC, LD);
    // component-major over the lanes: consecutive threads add into
    // consecutive ket positions of the position-indexed ket J
    for (int d = 0; d < ND; ++d)
      for (int i = threadIdx.x; i < 64 * NCD; i += blockDim.x) {
        const int cd = i >> 6, l = i & 63;
        const TileLane tl_l = tile_lane(tl.nket, l);
        if (!tl_l.active) continue;
        acc_flush<MODE>(jsum, d * NL * 64 * NCD + l * NCD + cd, 64 * NCD,
                        &ev.jpair[d * NL * ev.jp_size + ev.jp_off[Y] + cd * ev.jp_n[Y] + tl.ket0 + tl_l.slot],
                        ev.jp_size);
      }
  }
}

// ---------------------------------------------------------------------
// Launch
// ---------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename T, int MODE>
static hipError_t launch_one(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets,
                             const double* D, double* Jacc, double* Kacc, double* ints, double* qout,
                             hipStream_t st, const int2* groups, int ngroups, int gflags) {
  if (ntiles <= 0) return hipSuccess;
  using C = Cls<LA, LB, LC, LD>;
  // the K rows (and split classes' ket J sums), one word per element and
  // limb; the staged D rows take one word per element
  constexpr size_t NL = mode_limbs(MODE);
  const size_t drows = mode_ndens(MODE) * (size_t)(C::NA + C::NB) * ev.nbf * sizeof(double);
  const size_t lds = NL * mode_ndens(MODE) *
                     ((size_t)(C::NA + C::NB) * ev.nbf + (C::SPLIT ? 64 * C::NC * C::ND : 0)) * sizeof(double);
  if constexpr (C::SPLIT == 0 && mode_is_jk(MODE)) {
    const size_t glds = lds + kGroupTableDoubles<LA, LB, LC, LD, T, MODE> * sizeof(double);
    // plus the bra J block
    const size_t jlds = NL * mode_ndens(MODE) * C::NA * C::NB * sizeof(double);
    const size_t lmax = (size_t)ev.lds_max;
    const int stage_d = (gflags & kGroupStageD) && glds + jlds + drows <= lmax;
    if (groups && ngroups > 0 && glds + jlds <= lmax) {
      hipLaunchKernelGGL((k_eri_group<LA, LB, LC, LD, T, MODE>), dim3(ngroups), dim3(256),
                         glds + jlds + (stage_d ? drows : 0), st, ev, tiles, groups, ngroups, kets, D, Jacc, Kacc,
                         stage_d);
      return hipGetLastError();
    }
  }
  if (mode_is_jk(MODE) && lds <= (size_t)ev.lds_max) {
    const int waves = C::SPLIT ? kWaveSplit : 1;
    const int chunks = (C::NSPLIT + waves - 1) / waves;
    hipLaunchKernelGGL((k_eri_lane<LA, LB, LC, LD, T, MODE>), dim3(ntiles, chunks), dim3(64 * waves), lds, st, ev,
                       tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, 1);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((k_eri_lane<LA, LB, LC, LD, T, MODE>), dim3((ntiles + 3) / 4, C::NSPLIT), dim3(256), 0, st,
                     ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, 0);
  return hipGetLastError();
}

}  // namespace pq
