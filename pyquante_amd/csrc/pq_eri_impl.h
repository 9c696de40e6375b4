// pq_eri_impl.h -- Rys-quadrature two-electron-integral kernels for gfx950.
//
// One template per angular-momentum class (la lb|lc ld), l <= 2, la >= lb,
// lc >= ld, pair class (la,lb) >= (lc,ld).  Work arrives as tiles: one bra
// shell pair and up to 64 ket shell pairs taken from the ket class's pair list
// (sorted by Schwarz bound), so a wavefront shares its bra pair.
//
//  * light classes (<= 81 Cartesian integrals): one lane per shell quartet,
//    the contracted block held in registers;
//  * heavy classes: one wavefront per shell quartet, the 1-D Rys integrals
//    of each primitive quartet staged in LDS and the Cartesian products spread
//    over the 64 lanes.
//
// Modes: MODE_JK digests the block into J and K (the reference's
// getJ/getK/get2JmK, Ints.py:195-238, done directly from the integrals),
// MODE_STORE writes the reference's flat unique-ERI array (Ints.py:146-167,
// index cints.c:413-425), MODE_DIAG produces Schwarz bounds from (ab|ab).
#pragma once
#include "pq_device.h"
#include "pq_kernels.h"

namespace pq {

constexpr double kTwoPi25 = 34.98683665524972497;  // 2 pi^(5/2)

template <int LA, int LB, int LC, int LD>
struct Cls {
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int NCOMP = NA * NB * NC * ND;
  static constexpr int L = LA + LB + LC + LD;
  static constexpr int NR = L / 2 + 1;
};

// Per-primitive-quartet scalars shared by all roots.
struct PrimQ {
  double X, pref, rpq, qfac, pfac, i2p, i2q;
  double PA[3], PQ[3], QC[3];
};

__device__ __forceinline__ PrimQ prim_setup(const PairInfo& bp, const PrimPair& b, const PairInfo& kp,
                                            const PrimPair& q) {
  PrimQ s;
  const double pq = b.p + q.p;
  s.PA[0] = b.Px - bp.Ax;
  s.PA[1] = b.Py - bp.Ay;
  s.PA[2] = b.Pz - bp.Az;
  s.QC[0] = q.Px - kp.Ax;
  s.QC[1] = q.Py - kp.Ay;
  s.QC[2] = q.Pz - kp.Az;
  s.PQ[0] = b.Px - q.Px;
  s.PQ[1] = b.Py - q.Py;
  s.PQ[2] = b.Pz - q.Pz;
  const double rpq = 1.0 / pq;
  const double rho = b.p * q.p * rpq;
  s.X = rho * (s.PQ[0] * s.PQ[0] + s.PQ[1] * s.PQ[1] + s.PQ[2] * s.PQ[2]);
  s.pref = kTwoPi25 * b.K * q.K / (b.p * q.p) * sqrt(rpq);
  s.rpq = rpq;
  s.qfac = q.p * rpq;
  s.pfac = b.p * rpq;
  s.i2p = 0.5 / b.p;
  s.i2q = 0.5 / q.p;
  return s;
}

// Contracted block of one shell quartet, components in PyQuante order, without
// the per-component normalisation factor.
template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ void eri_block(const EngineView& ev, const PairInfo& bp, const PairInfo& kp,
                                          T (&out)[Cls<LA, LB, LC, LD>::NCOMP]) {
  using C = Cls<LA, LB, LC, LD>;
  constexpr int NR = C::NR;
#pragma unroll
  for (int c = 0; c < C::NCOMP; ++c) out[c] = (T)0;
  const T AB[3] = {(T)bp.ABx, (T)bp.ABy, (T)bp.ABz};
  const T CD[3] = {(T)kp.ABx, (T)kp.ABy, (T)kp.ABz};
  for (int i = 0; i < bp.pp_n; ++i) {
    const PrimPair b = ev.prims[bp.pp_off + i];
    for (int k = 0; k < kp.pp_n; ++k) {
      const PrimPair q = ev.prims[kp.pp_off + k];
      const PrimQ s = prim_setup(bp, b, kp, q);
      T u[NR], w[NR];
      rys_roots<NR, T>(s.X, u, w);
      const T pref = (T)s.pref, qfac = (T)s.qfac, pfac = (T)s.pfac;
      const T half_rpq = (T)(0.5 * s.rpq), i2p = (T)s.i2p, i2q = (T)s.i2q;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const T ur = u[r];
        const T B00 = half_rpq * ur;
        const T B10 = i2p * ((T)1 - qfac * ur);
        const T B01 = i2q * ((T)1 - pfac * ur);
        T Ix[LA + 1][LB + 1][LC + 1][LD + 1], Iy[LA + 1][LB + 1][LC + 1][LD + 1],
            Iz[LA + 1][LB + 1][LC + 1][LD + 1];
        rys_1d<LA, LB, LC, LD, T>((T)1, (T)s.PA[0] - qfac * (T)s.PQ[0] * ur,
                                  (T)s.QC[0] + pfac * (T)s.PQ[0] * ur, B10, B01, B00, AB[0], CD[0], Ix);
        rys_1d<LA, LB, LC, LD, T>((T)1, (T)s.PA[1] - qfac * (T)s.PQ[1] * ur,
                                  (T)s.QC[1] + pfac * (T)s.PQ[1] * ur, B10, B01, B00, AB[1], CD[1], Iy);
        rys_1d<LA, LB, LC, LD, T>(w[r] * pref, (T)s.PA[2] - qfac * (T)s.PQ[2] * ur,
                                  (T)s.QC[2] + pfac * (T)s.PQ[2] * ur, B10, B01, B00, AB[2], CD[2], Iz);
#pragma unroll
        for (int a = 0; a < C::NA; ++a)
#pragma unroll
          for (int bb = 0; bb < C::NB; ++bb)
#pragma unroll
            for (int c = 0; c < C::NC; ++c)
#pragma unroll
              for (int d = 0; d < C::ND; ++d) {
                const int o = ((a * C::NB + bb) * C::NC + c) * C::ND + d;
                out[o] += Ix[cart_pow(LA, a, 0)][cart_pow(LB, bb, 0)][cart_pow(LC, c, 0)][cart_pow(LD, d, 0)] *
                          Iy[cart_pow(LA, a, 1)][cart_pow(LB, bb, 1)][cart_pow(LC, c, 1)][cart_pow(LD, d, 1)] *
                          Iz[cart_pow(LA, a, 2)][cart_pow(LB, bb, 2)][cart_pow(LC, c, 2)][cart_pow(LD, d, 2)];
              }
      }
    }
  }
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
template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ void digest_lane(const EngineView& ev, const PairInfo& bp, const PairInfo& kp,
                                            const T (&out)[Cls<LA, LB, LC, LD>::NCOMP], double f,
                                            bool active, const double* __restrict__ D, double* Jacc,
                                            double* Kacc) {
  using C = Cls<LA, LB, LC, LD>;
  const int n = ev.nbf;
  const int a0 = bp.bfA, b0 = bp.bfB, c0 = kp.bfA, d0 = kp.bfB;
  double jab[C::NA][C::NB], jcd[C::NC][C::ND], kac[C::NA][C::NC], kad[C::NA][C::ND],
      kbc[C::NB][C::NC], kbd[C::NB][C::ND];
  double Dab[C::NA][C::NB], Dcd[C::NC][C::ND], Dac[C::NA][C::NC], Dad[C::NA][C::ND],
      Dbc[C::NB][C::NC], Dbd[C::NB][C::ND];
#pragma unroll
  for (int a = 0; a < C::NA; ++a) {
#pragma unroll
    for (int b = 0; b < C::NB; ++b) { Dab[a][b] = D[(a0 + a) * n + b0 + b]; jab[a][b] = 0; }
#pragma unroll
    for (int c = 0; c < C::NC; ++c) { Dac[a][c] = D[(a0 + a) * n + c0 + c]; kac[a][c] = 0; }
#pragma unroll
    for (int d = 0; d < C::ND; ++d) { Dad[a][d] = D[(a0 + a) * n + d0 + d]; kad[a][d] = 0; }
  }
#pragma unroll
  for (int b = 0; b < C::NB; ++b) {
#pragma unroll
    for (int c = 0; c < C::NC; ++c) { Dbc[b][c] = D[(b0 + b) * n + c0 + c]; kbc[b][c] = 0; }
#pragma unroll
    for (int d = 0; d < C::ND; ++d) { Dbd[b][d] = D[(b0 + b) * n + d0 + d]; kbd[b][d] = 0; }
  }
#pragma unroll
  for (int c = 0; c < C::NC; ++c)
#pragma unroll
    for (int d = 0; d < C::ND; ++d) { Dcd[c][d] = D[(c0 + c) * n + d0 + d]; jcd[c][d] = 0; }

#pragma unroll
  for (int a = 0; a < C::NA; ++a)
#pragma unroll
    for (int b = 0; b < C::NB; ++b)
#pragma unroll
      for (int c = 0; c < C::NC; ++c)
#pragma unroll
        for (int d = 0; d < C::ND; ++d) {
          const double nf = cart_norm(LA, a) * cart_norm(LB, b) * cart_norm(LC, c) * cart_norm(LD, d);
          const double v = (double)out[((a * C::NB + b) * C::NC + c) * C::ND + d] * (nf * f);
          jab[a][b] += v * Dcd[c][d];
          jcd[c][d] += v * Dab[a][b];
          kac[a][c] += v * Dbd[b][d];
          kad[a][d] += v * Dbc[b][c];
          kbc[b][c] += v * Dad[a][d];
          kbd[b][d] += v * Dac[a][c];
        }
  // J_ab: every lane of the wave shares the bra pair -> reduce, one atomic
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int a = 0; a < C::NA; ++a)
#pragma unroll
    for (int b = 0; b < C::NB; ++b) {
      const double s = wave_sum(jab[a][b]);
      if (lane == 0) unsafeAtomicAdd(&Jacc[(a0 + a) * n + b0 + b], 2.0 * s);
    }
  if (!active) return;
#pragma unroll
  for (int c = 0; c < C::NC; ++c)
#pragma unroll
    for (int d = 0; d < C::ND; ++d) unsafeAtomicAdd(&Jacc[(c0 + c) * n + d0 + d], 2.0 * jcd[c][d]);
#pragma unroll
  for (int a = 0; a < C::NA; ++a) {
#pragma unroll
    for (int c = 0; c < C::NC; ++c) unsafeAtomicAdd(&Kacc[(a0 + a) * n + c0 + c], kac[a][c]);
#pragma unroll
    for (int d = 0; d < C::ND; ++d) unsafeAtomicAdd(&Kacc[(a0 + a) * n + d0 + d], kad[a][d]);
  }
#pragma unroll
  for (int b = 0; b < C::NB; ++b) {
#pragma unroll
    for (int c = 0; c < C::NC; ++c) unsafeAtomicAdd(&Kacc[(b0 + b) * n + c0 + c], kbc[b][c]);
#pragma unroll
    for (int d = 0; d < C::ND; ++d) unsafeAtomicAdd(&Kacc[(b0 + b) * n + d0 + d], kbd[b][d]);
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

// ---------------------------------------------------------------------
// Light classes: one lane per shell quartet.
// ---------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename T, int MODE>
__global__ __launch_bounds__(256) void k_eri_lane(EngineView ev, const Tile* __restrict__ tiles, int ntiles,
                                                  const int* __restrict__ kets, const double* __restrict__ D,
                                                  double* Jacc, double* Kacc, double* ints,
                                                  double* qout) {
  using C = Cls<LA, LB, LC, LD>;
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (t >= ntiles) return;  // uniform per wavefront
  const Tile tl = tiles[t];
  const bool active = lane < tl.nket;
  const int ket = kets[tl.ket0 + (active ? lane : 0)];
  const int bra = (MODE == MODE_DIAG) ? ket : tl.bra;
  const PairInfo bp = ev.pairs[bra];
  const PairInfo kp = ev.pairs[ket];
  T out[C::NCOMP];
  if (active) {
    eri_block<LA, LB, LC, LD, T>(ev, bp, kp, out);
  } else {
#pragma unroll
    for (int c = 0; c < C::NCOMP; ++c) out[c] = (T)0;
  }
  if (MODE == MODE_JK) {
    const double f = active ? quartet_degeneracy(bp, kp, bra, ket) : 0.0;
    digest_lane<LA, LB, LC, LD, T>(ev, bp, kp, out, f, active, D, Jacc, Kacc);
  } else if (MODE == MODE_STORE) {
    if (!active) return;
    const bool same = bra == ket;
#pragma unroll
    for (int a = 0; a < C::NA; ++a)
#pragma unroll
      for (int b = 0; b < C::NB; ++b)
#pragma unroll
        for (int c = 0; c < C::NC; ++c)
#pragma unroll
          for (int d = 0; d < C::ND; ++d) {
            const int i = bp.bfA + a, j = bp.bfB + b, k = kp.bfA + c, l = kp.bfB + d;
            if (!store_canonical(bp, kp, same, i, j, k, l)) continue;
            const double nf = cart_norm(LA, a) * cart_norm(LB, b) * cart_norm(LC, c) * cart_norm(LD, d);
            ints[ijkl2intindex(i, j, k, l)] = (double)out[((a * C::NB + b) * C::NC + c) * C::ND + d] * nf;
          }
  } else {  // MODE_DIAG: Q = max over components sqrt|(ab|ab)|
    if (!active) return;
    double qmax = 0.0;
#pragma unroll
    for (int a = 0; a < C::NA; ++a)
#pragma unroll
      for (int b = 0; b < C::NB; ++b) {
        const double nf = cart_norm(LA, a) * cart_norm(LB, b);
        const double v = fabs((double)out[((a * C::NB + b) * C::NC + a) * C::ND + b]) * nf * nf;
        qmax = fmax(qmax, v);
      }
    qout[ket] = sqrt(qmax);
  }
}

// ---------------------------------------------------------------------
// Heavy classes: one wavefront per shell quartet.  Lanes (r, dir) build the
// 1-D integrals of root r and direction dir for each primitive quartet in
// LDS; every lane then accumulates its share of Cartesian components.
// ---------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename T, int MODE>
__global__ __launch_bounds__(64) void k_eri_wave(EngineView ev, const Tile* __restrict__ tiles, int ntiles,
                                                 const int* __restrict__ kets, const double* __restrict__ D,
                                                 double* Jacc, double* Kacc, double* ints, double* qout) {
  using C = Cls<LA, LB, LC, LD>;
  constexpr int NR = C::NR;
  constexpr int N1 = (LA + 1) * (LB + 1) * (LC + 1) * (LD + 1);
  constexpr int CPL = (C::NCOMP + 63) / 64;
  __shared__ T sI[NR][3][N1];
  const int lane = threadIdx.x;
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  const Tile tl = tiles[t];

  // this lane's components and their offsets into the 1-D tables
  int offs[CPL][3];
  int comp[CPL];
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int c = lane + 64 * j;
    comp[j] = c;
    const int cc = c < C::NCOMP ? c : 0;
    const int id = cc % C::ND, ic = (cc / C::ND) % C::NC, ib = (cc / (C::ND * C::NC)) % C::NB,
              ia = cc / (C::ND * C::NC * C::NB);
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      int pa = 0, pb = 0, pc = 0, pd = 0;
#pragma unroll
      for (int q = 0; q < C::NA; ++q) if (q == ia) pa = cart_pow(LA, q, x);
#pragma unroll
      for (int q = 0; q < C::NB; ++q) if (q == ib) pb = cart_pow(LB, q, x);
#pragma unroll
      for (int q = 0; q < C::NC; ++q) if (q == ic) pc = cart_pow(LC, q, x);
#pragma unroll
      for (int q = 0; q < C::ND; ++q) if (q == id) pd = cart_pow(LD, q, x);
      offs[j][x] = ((pa * (LB + 1) + pb) * (LC + 1) + pc) * (LD + 1) + pd;
    }
  }

  for (int kk = 0; kk < tl.nket; ++kk) {
    const int ket = kets[tl.ket0 + kk];
    const int bra = (MODE == MODE_DIAG) ? ket : tl.bra;
    const PairInfo bp = ev.pairs[bra];
    const PairInfo kp = ev.pairs[ket];
    T acc[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) acc[j] = (T)0;
    const int r_me = lane / 3, x_me = lane % 3;
    const T ABm = (T)(x_me == 0 ? bp.ABx : (x_me == 1 ? bp.ABy : bp.ABz));
    const T CDm = (T)(x_me == 0 ? kp.ABx : (x_me == 1 ? kp.ABy : kp.ABz));
    for (int i = 0; i < bp.pp_n; ++i) {
      const PrimPair b = ev.prims[bp.pp_off + i];
      for (int k = 0; k < kp.pp_n; ++k) {
        const PrimPair q = ev.prims[kp.pp_off + k];
        if (lane < 3 * NR) {
          const PrimQ s = prim_setup(bp, b, kp, q);
          T u[NR], w[NR];
          rys_roots<NR, T>(s.X, u, w);
          T ur = u[0], wr = w[0];
#pragma unroll
          for (int r = 1; r < NR; ++r)
            if (r == r_me) { ur = u[r]; wr = w[r]; }
          const T qfac = (T)s.qfac, pfac = (T)s.pfac;
          const T B00 = (T)(0.5 * s.rpq) * ur;
          const T B10 = (T)s.i2p * ((T)1 - qfac * ur);
          const T B01 = (T)s.i2q * ((T)1 - pfac * ur);
          const T C00 = (T)s.PA[x_me] - qfac * (T)s.PQ[x_me] * ur;
          const T D00 = (T)s.QC[x_me] + pfac * (T)s.PQ[x_me] * ur;
          const T seed = x_me == 2 ? wr * (T)s.pref : (T)1;
          T I[LA + 1][LB + 1][LC + 1][LD + 1];
          rys_1d<LA, LB, LC, LD, T>(seed, C00, D00, B10, B01, B00, ABm, CDm, I);
          T* dst = &sI[r_me][x_me][0];
#pragma unroll
          for (int a = 0; a <= LA; ++a)
#pragma unroll
            for (int bb = 0; bb <= LB; ++bb)
#pragma unroll
              for (int c = 0; c <= LC; ++c)
#pragma unroll
                for (int d = 0; d <= LD; ++d)
                  dst[((a * (LB + 1) + bb) * (LC + 1) + c) * (LD + 1) + d] = I[a][bb][c][d];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          if (comp[j] < C::NCOMP) {
            T v = (T)0;
#pragma unroll
            for (int r = 0; r < NR; ++r) v += sI[r][0][offs[j][0]] * sI[r][1][offs[j][1]] * sI[r][2][offs[j][2]];
            acc[j] += v;
          }
        }
        __syncthreads();
      }
    }
    // epilogue per component
    const double f = quartet_degeneracy(bp, kp, bra, ket);
    const int n = ev.nbf;
    double qmax = 0.0;
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int c = comp[j];
      if (c >= C::NCOMP) continue;
      const int id = c % C::ND, ic = (c / C::ND) % C::NC, ib = (c / (C::ND * C::NC)) % C::NB,
                ia = c / (C::ND * C::NC * C::NB);
      double nf = 1.0;
#pragma unroll
      for (int q = 0; q < C::NA; ++q) if (q == ia) nf *= cart_norm(LA, q);
#pragma unroll
      for (int q = 0; q < C::NB; ++q) if (q == ib) nf *= cart_norm(LB, q);
#pragma unroll
      for (int q = 0; q < C::NC; ++q) if (q == ic) nf *= cart_norm(LC, q);
#pragma unroll
      for (int q = 0; q < C::ND; ++q) if (q == id) nf *= cart_norm(LD, q);
      const double v = (double)acc[j] * nf;
      const int ai = bp.bfA + ia, bi = bp.bfB + ib, ci = kp.bfA + ic, di = kp.bfB + id;
      if (MODE == MODE_JK) {
        const double fv = f * v;
        unsafeAtomicAdd(&Jacc[ai * n + bi], 2.0 * fv * D[ci * n + di]);
        unsafeAtomicAdd(&Jacc[ci * n + di], 2.0 * fv * D[ai * n + bi]);
        unsafeAtomicAdd(&Kacc[ai * n + ci], fv * D[bi * n + di]);
        unsafeAtomicAdd(&Kacc[ai * n + di], fv * D[bi * n + ci]);
        unsafeAtomicAdd(&Kacc[bi * n + ci], fv * D[ai * n + di]);
        unsafeAtomicAdd(&Kacc[bi * n + di], fv * D[ai * n + ci]);
      } else if (MODE == MODE_STORE) {
        if (store_canonical(bp, kp, bra == ket, ai, bi, ci, di)) ints[ijkl2intindex(ai, bi, ci, di)] = v;
      } else {
        if (ia == ic && ib == id) qmax = fmax(qmax, fabs(v));
      }
    }
    if (MODE == MODE_DIAG) {
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) qmax = fmax(qmax, __shfl_xor(qmax, off, 64));
      if (lane == 0) qout[ket] = sqrt(qmax);
    }
  }
}

// ---------------------------------------------------------------------
// Launch table
// ---------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename T, int MODE>
static hipError_t launch_one(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets,
                             const double* D, double* Jacc, double* Kacc, double* ints, double* qout,
                             hipStream_t st) {
  if (ntiles <= 0) return hipSuccess;
  if constexpr (Cls<LA, LB, LC, LD>::NCOMP <= kLaneMaxComp) {
    const int wpb = 4;
    hipLaunchKernelGGL((k_eri_lane<LA, LB, LC, LD, T, MODE>), dim3((ntiles + wpb - 1) / wpb), dim3(64 * wpb),
                       0, st, ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout);
  } else {
    hipLaunchKernelGGL((k_eri_wave<LA, LB, LC, LD, T, MODE>), dim3(ntiles), dim3(64), 0, st, ev, tiles,
                       ntiles, kets, D, Jacc, Kacc, ints, qout);
  }
  return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
static hipError_t launch_cls(int mode, int fp32, const EngineView& ev, const Tile* tiles, int ntiles,
                             const int* kets, const double* D, double* Jacc, double* Kacc, double* ints,
                             double* qout, hipStream_t st) {
  if (mode == MODE_JK) {
    if (fp32)
      return launch_one<LA, LB, LC, LD, float, MODE_JK>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st);
    return launch_one<LA, LB, LC, LD, double, MODE_JK>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st);
  }
  if (mode == MODE_STORE)
    return launch_one<LA, LB, LC, LD, double, MODE_STORE>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st);
  if constexpr (LA == LC && LB == LD)
    return launch_one<LA, LB, LC, LD, double, MODE_DIAG>(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st);
  return hipErrorInvalidValue;
}

}  // namespace pq
