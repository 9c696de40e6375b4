// pq_grad_impl.h -- nuclear-gradient (backward) kernels of the two-electron
// energy E2 = sum_ijkl (ij|kl) [2 D_ij D_kl - D_ik D_jl] (RHF, D = C_occ C_occ^T).
//
// Restates the two-electron part of PyQuante's rhf_force (force.py:48-70,
// der2JmK force.py:318-357) without materialising derivative integrals
// (AnalyticDerivatives.py:457-675 builds them per primitive): for each
// primitive quartet the 1-D Rys integrals are built one angular momentum
// higher on A, B and C, the centre derivative of a primitive Cartesian
// Gaussian d/dA_x g(l) = 2 alpha g(l+1) - l g(l-1) is applied on the fly, and
// the result is contracted at once with the density factor
//   Gamma = f (16 D_ab D_cd - 4 (D_ac D_bd + D_ad D_bc))
// (f the permutational degeneracy of the unique quartet).  For an
// unrestricted wavefunction (uhf_force, force.py:72-89, whose two-electron
// energy is 1/2 sum (ij|kl) [P_ij P_kl - Pa_ik Pa_jl - Pb_ik Pb_jl] with
// P = Pa + Pb) the factor is
//   Gamma = f (4 P_ab P_cd - 2 sum_s (Ps_ac Ps_bd + Ps_ad Ps_bc)),
// which is the closed-shell one for Pa = Pb = D.  The D-centre derivative
// follows from translational invariance.
#pragma once
#include "pq_eri_impl.h"

namespace pq {

// Component split of the gradient kernels: up to 54 Cartesian integrals per
// lane (each split recomputes the roots and 1-D integrals, so fewer splits
// beat the occupancy their registers cost; measured against 27 and 81).
template <int LA, int LB, int LC, int LD>
struct GCls {
  using C = Cls<LA, LB, LC, LD>;
  static constexpr int NCOMP = C::NCOMP;
  static constexpr int SPLIT = NCOMP <= 54 ? 0 : (NCOMP / C::NA <= 54 ? 1 : 2);
  static constexpr int NSPLIT = SPLIT == 0 ? 1 : (SPLIT == 1 ? C::NA : C::NA * C::NB);
  static constexpr int NR = (C::L + 1) / 2 + 1;
};

template <int LA, int LB, int LC, int LD, int A0, int NAS, int B0, int NBS>
__device__ __forceinline__ void grad_lane(const EngineView& ev, const Tile& tl, bool active, int bra, int ket,
                                          const double* __restrict__ D, double* gcopy, double* gl, int kc,
                                          int ks) {
  using C = Cls<LA, LB, LC, LD>;
  constexpr int NR = GCls<LA, LB, LC, LD>::NR;
  constexpr int NO = NAS * NBS * C::NC * C::ND;
  const PairInfo bp = ld_uniform(ev.pairs + bra);
  const PairInfo kp = ev.pairs[ket];
  const int n = bp.mstride;
  D += bp.moff;
  const int a0 = bp.bfA + A0, b0 = bp.bfB + B0, c0 = kp.bfA, d0 = kp.bfB;
  // density factor per component (normalisation folded in); the
  // unrestricted form only when a beta density is given (one branch per
  // quartet, outside the component loops)
  double G[NO];
  {
    const double f = active ? quartet_degeneracy(bp, kp, bra, ket) : 0.0;
    auto fill = [&](auto unrestricted) __attribute__((always_inline)) {
      constexpr bool U = decltype(unrestricted)::value;
      const double* Db = ev.grad_db + (U ? bp.moff : 0);
#pragma unroll
      for (int a = 0; a < NAS; ++a)
#pragma unroll
        for (int b = 0; b < NBS; ++b)
#pragma unroll
          for (int c = 0; c < C::NC; ++c)
#pragma unroll
            for (int d = 0; d < C::ND; ++d) {
              const double nf =
                  cart_norm(LA, A0 + a) * cart_norm(LB, B0 + b) * cart_norm(LC, c) * cart_norm(LD, d);
              const int ab = (a0 + a) * n + b0 + b, cd = (c0 + c) * n + d0 + d;
              const int ac = (a0 + a) * n + c0 + c, bd = (b0 + b) * n + d0 + d;
              const int ad = (a0 + a) * n + d0 + d, bc = (b0 + b) * n + c0 + c;
              double g;
              if constexpr (U)
                g = 4.0 * (D[ab] + Db[ab]) * (D[cd] + Db[cd]) -
                    2.0 * (D[ac] * D[bd] + D[ad] * D[bc] + Db[ac] * Db[bd] + Db[ad] * Db[bc]);
              else
                g = 16.0 * D[ab] * D[cd] - 4.0 * (D[ac] * D[bd] + D[ad] * D[bc]);
              G[((a * NBS + b) * C::NC + c) * C::ND + d] = f * nf * g;
            }
    };
    if (ev.grad_db) fill(std::true_type{});
    else fill(std::false_type{});
  }
  double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
  if (active) {
    const int i0 = tl.prim & 0xffff, ni = tl.prim >> 16;
    const int i1 = ni ? i0 + ni : bp.pp_n;
    const double AB[3] = {bp.ABx, bp.ABy, bp.ABz};
    const double CD[3] = {kp.ABx, kp.ABy, kp.ABz};
    // the next ket primitive pair is loaded while the current one is
    // processed (its latency hides behind the bra loop)
    const int k0 = ket_chunk(kc, kp.pp_n, ks), k1 = ket_chunk(kc + 1, kp.pp_n, ks);
    PrimPair qn{};
    double gn = 0.0;
    if (k0 < k1) {
      qn = ev.prims[kp.pp_off + k0];
      gn = ev.prim_a[kp.pp_off + k0];
    }
#pragma unroll 1
    for (int k = k0; k < k1; ++k) {
      const PrimPair q = qn;
      const double gam = gn;
      const int kn = k + 1 < k1 ? k + 1 : k;
      qn = ev.prims[kp.pp_off + kn];
      gn = ev.prim_a[kp.pp_off + kn];
#pragma unroll 1
      for (int i = i0; i < i1; ++i) {
        const PrimPair b = ld_uniform(ev.prims + bp.pp_off + i);
        const double alp = ld_uniform(ev.prim_a + bp.pp_off + i), bet = b.p - alp;
        const PrimQ s = prim_setup(bp, b, kp, q);
        const RysPoint<NR> rp = rys_prepare<NR>(s.X);
        constexpr int UNR = NR <= 2 ? NR : 1;
#pragma unroll UNR
        for (int r = 0; r < NR; ++r) {
          double ur, wr;
          rys_root<NR, double>(rp, r, ur, wr);
          const double B00 = 0.5 * s.rpq * ur;
          const double B10 = s.i2p * (1.0 - s.qfac * ur);
          const double B01 = s.i2q * (1.0 - s.pfac * ur);
          double Ix[LA + 2][LB + 2][LC + 2][LD + 1], Iy[LA + 2][LB + 2][LC + 2][LD + 1],
              Iz[LA + 2][LB + 2][LC + 2][LD + 1];
          rys_1d<LA + 1, LB + 1, LC + 1, LD, double>(1.0, s.PA[0] - s.qfac * s.PQ[0] * ur,
                                                     s.QC[0] + s.pfac * s.PQ[0] * ur, B10, B01, B00, AB[0], CD[0],
                                                     Ix);
          rys_1d<LA + 1, LB + 1, LC + 1, LD, double>(1.0, s.PA[1] - s.qfac * s.PQ[1] * ur,
                                                     s.QC[1] + s.pfac * s.PQ[1] * ur, B10, B01, B00, AB[1], CD[1],
                                                     Iy);
          rys_1d<LA + 1, LB + 1, LC + 1, LD, double>(wr * s.pref, s.PA[2] - s.qfac * s.PQ[2] * ur,
                                                     s.QC[2] + s.pfac * s.PQ[2] * ur, B10, B01, B00, AB[2], CD[2],
                                                     Iz);
          const double ta = 2.0 * alp, tb = 2.0 * bet, tc = 2.0 * gam;
#pragma unroll
          for (int a = 0; a < NAS; ++a)
#pragma unroll
            for (int bb = 0; bb < NBS; ++bb)
#pragma unroll
              for (int c = 0; c < C::NC; ++c)
#pragma unroll
                for (int d = 0; d < C::ND; ++d) {
                  const double g = G[((a * NBS + bb) * C::NC + c) * C::ND + d];
                  const int ax = cart_pow(LA, A0 + a, 0), ay = cart_pow(LA, A0 + a, 1), az = cart_pow(LA, A0 + a, 2);
                  const int bx = cart_pow(LB, B0 + bb, 0), by = cart_pow(LB, B0 + bb, 1), bz = cart_pow(LB, B0 + bb, 2);
                  const int cx = cart_pow(LC, c, 0), cy = cart_pow(LC, c, 1), cz = cart_pow(LC, c, 2);
                  const int dx = cart_pow(LD, d, 0), dy = cart_pow(LD, d, 1), dz = cart_pow(LD, d, 2);
                  const double x0 = Ix[ax][bx][cx][dx], y0 = Iy[ay][by][cy][dy], z0 = Iz[az][bz][cz][dz];
                  const double gyz = g * y0 * z0, gxz = g * x0 * z0, gxy = g * x0 * y0;
                  // centre A
                  gA[0] += gyz * (ta * Ix[ax + 1][bx][cx][dx] - (ax ? ax * Ix[ax ? ax - 1 : 0][bx][cx][dx] : 0.0));
                  gA[1] += gxz * (ta * Iy[ay + 1][by][cy][dy] - (ay ? ay * Iy[ay ? ay - 1 : 0][by][cy][dy] : 0.0));
                  gA[2] += gxy * (ta * Iz[az + 1][bz][cz][dz] - (az ? az * Iz[az ? az - 1 : 0][bz][cz][dz] : 0.0));
                  // centre B
                  gB[0] += gyz * (tb * Ix[ax][bx + 1][cx][dx] - (bx ? bx * Ix[ax][bx ? bx - 1 : 0][cx][dx] : 0.0));
                  gB[1] += gxz * (tb * Iy[ay][by + 1][cy][dy] - (by ? by * Iy[ay][by ? by - 1 : 0][cy][dy] : 0.0));
                  gB[2] += gxy * (tb * Iz[az][bz + 1][cz][dz] - (bz ? bz * Iz[az][bz ? bz - 1 : 0][cz][dz] : 0.0));
                  // centre C
                  gC[0] += gyz * (tc * Ix[ax][bx][cx + 1][dx] - (cx ? cx * Ix[ax][bx][cx ? cx - 1 : 0][dx] : 0.0));
                  gC[1] += gxz * (tc * Iy[ay][by][cy + 1][dy] - (cy ? cy * Iy[ay][by][cy ? cy - 1 : 0][dy] : 0.0));
                  gC[2] += gxy * (tc * Iz[az][bz][cz + 1][dz] - (cz ? cz * Iz[az][bz][cz ? cz - 1 : 0][dz] : 0.0));
                }
        }
      }
    }
  }
  // A and B are shared by the wavefront: reduce, one atomic each
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const double sa = wave_sum(gA[x]), sb = wave_sum(gB[x]);
    if (lane == 0) {
      unsafeAtomicAdd(&gcopy[3 * bp.atA + x], sa);
      unsafeAtomicAdd(&gcopy[3 * bp.atB + x], sb);
    }
  }
  if (active) {
    // gl: the molecule's atoms, from its first atom
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      atomicAdd(&gl[3 * (kp.atA - bp.mat0) + x], gC[x]);
      atomicAdd(&gl[3 * (kp.atB - bp.mat0) + x], -(gA[x] + gB[x] + gC[x]));
    }
  }
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_eri_grad(EngineView ev, const Tile* __restrict__ tiles, int ntiles,
                                                 const int* __restrict__ kets, const double* __restrict__ D,
                                                 double* gacc, int natom) {
  using GC = GCls<LA, LB, LC, LD>;
  using C = Cls<LA, LB, LC, LD>;
  extern __shared__ double pq_lds[];
  const int lane = threadIdx.x;
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  double* gl = pq_lds;  // 3 x (atoms of the tile's molecule)
  Tile tl = ld_uniform(tiles + t);
  const PairInfo bpm = ld_uniform(ev.pairs + __builtin_amdgcn_readfirstlane(tl.bra));
  for (int i = lane; i < 3 * bpm.mnat; i += 64) gl[i] = 0.0;
  __syncthreads();
  tl.bra = __builtin_amdgcn_readfirstlane(tl.bra);
  tl.ket0 = __builtin_amdgcn_readfirstlane(tl.ket0);
  tl.nket = __builtin_amdgcn_readfirstlane(tl.nket);
  tl.prim = __builtin_amdgcn_readfirstlane(tl.prim);
  const TileLane tn = tile_lane(tl.nket, lane);
  const bool active = tn.active;
  const int ket = kets[tl.ket0 + tn.slot];
  double* gcopy = gacc + (size_t)((t * GC::NSPLIT + blockIdx.y) % kGradCopies) * 3 * natom;
  if constexpr (GC::SPLIT == 0) {
    grad_lane<LA, LB, LC, LD, 0, C::NA, 0, C::NB>(ev, tl, active, tl.bra, ket, D, gcopy, gl, tn.kc, tn.ks);
  } else if constexpr (GC::SPLIT == 1) {
    static_switch<0, C::NA>(blockIdx.y, [&](auto ia) {
      grad_lane<LA, LB, LC, LD, decltype(ia)::value, 1, 0, C::NB>(ev, tl, active, tl.bra, ket, D, gcopy, gl, tn.kc,
                                                                  tn.ks);
    });
  } else {
    static_switch<0, C::NA * C::NB>(blockIdx.y, [&](auto iab) {
      constexpr int v = decltype(iab)::value;
      grad_lane<LA, LB, LC, LD, v / C::NB, 1, v % C::NB, 1>(ev, tl, active, tl.bra, ket, D, gcopy, gl, tn.kc,
                                                            tn.ks);
    });
  }
  __syncthreads();
  for (int i = lane; i < 3 * bpm.mnat; i += 64) {
    const double v = gl[i];
    if (v != 0.0) unsafeAtomicAdd(&gcopy[3 * bpm.mat0 + i], v);
  }
}

template <int LA, int LB, int LC, int LD>
static hipError_t launch_grad_one(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets,
                                  const double* D, double* gacc, int natom, hipStream_t st) {
  if (ntiles <= 0) return hipSuccess;
  using GC = GCls<LA, LB, LC, LD>;
  hipLaunchKernelGGL((k_eri_grad<LA, LB, LC, LD>), dim3(ntiles, GC::NSPLIT), dim3(64),
                     (size_t)3 * ev.maxat * sizeof(double), st, ev, tiles, ntiles, kets, D, gacc, natom);
  return hipGetLastError();
}

}  // namespace pq
