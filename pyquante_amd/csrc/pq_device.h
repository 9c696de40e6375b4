// pq_device.h -- device-side building blocks of the MI355X ERI engine:
// Cartesian component tables, shell-pair records, Rys root evaluation and
// the 1-D Rys recurrence + transfer, shared by every integral kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "boys_table.h"
#include "pq_types.h"
#include "rys_table.h"
#include "rys_table32.h"

namespace pq {

// ---------------------------------------------------------------------
// Cartesian components in PyQuante order (Ints.py:23-29): for a shell of
// angular momentum L, x-power runs from L down, then y-power from L-lx down.
// D = xx, xy, xz, yy, yz, zz.
// ---------------------------------------------------------------------

__host__ __device__ constexpr int cart_pow(int L, int i, int d) {
  int idx = 0;
  for (int lx = L; lx >= 0; --lx)
    for (int ly = L - lx; ly >= 0; --ly) {
      if (idx == i) return d == 0 ? lx : (d == 1 ? ly : L - lx - ly);
      ++idx;
    }
  return 0;
}

__host__ __device__ constexpr int dfact(int n) { return n <= 1 ? 1 : n * dfact(n - 2); }

// Per-component part of the primitive normalisation (primitive-gto.c:76-83):
// N = Nrad(alpha, L) / sqrt((2l-1)!!(2m-1)!!(2n-1)!!).  Nrad is folded into
// the primitive-pair prefactor, this factor is applied per component.
// 1/sqrt(n) for the double-factorial products that occur up to L = 4; kept
// as literals so that the factor folds to a constant after loop unrolling.
__host__ __device__ constexpr double inv_sqrt_dfprod(int n) {
  return n == 1 ? 1.0
       : n == 3 ? 0.57735026918962576451
       : n == 9 ? 0.33333333333333333333
       : n == 15 ? 0.25819888974716112568
       : n == 105 ? 0.09759000729485331981
       : 0.0;
}
__host__ __device__ constexpr double cart_norm(int L, int i) {
  return inv_sqrt_dfprod(dfact(2 * cart_pow(L, i, 0) - 1) * dfact(2 * cart_pow(L, i, 1) - 1) *
                         dfact(2 * cart_pow(L, i, 2) - 1));
}

// Lanes of a tile: with nk kets and a ket-primitive split ks, lane l takes
// ket l % nk and the (l / nk)-th of ks near-equal chunks of that ket's
// primitive pairs; each lane's partial block digests on its own (J and K are
// linear in the integrals).  ks = 1 for stored integrals and Schwarz bounds.
struct TileLane {
  bool active;
  int slot;  // ket index within the tile
  int kc;    // ket-primitive chunk
  int ks;
  int nk;
};
__device__ __forceinline__ TileLane tile_lane(int nket_field, int lane) {
  const int nk = nket_field & 0xff, ks = 1 << ((nket_field >> 8) & 7);
  TileLane t;
  t.active = lane < nk * ks;
  // lane / nk without an integer division: (lane + 1/2) / nk lies at least
  // 1/(2 nk) >= 1/128 from an integer, far beyond the float reciprocal's
  // error, so the truncation is exact
  const int q = (int)(((float)lane + 0.5f) * __builtin_amdgcn_rcpf((float)nk));
  t.slot = t.active ? lane - q * nk : 0;
  t.kc = t.active ? q : 0;
  t.ks = ks;
  t.nk = nk;
  return t;
}

// First primitive pair of ket-primitive chunk kc of n pairs split ks ways
// (ks a power of two): kc * n / ks.
__device__ __forceinline__ int ket_chunk(int kc, int n, int ks) { return (kc * n) >> __builtin_ctz(ks); }

// Everything a kernel needs about the basis and the pair data.
struct EngineView {
  const PairInfo* pairs;
  const PrimPair* prims;
  const double* prim_a;  // alpha_a of each primitive pair (one-electron and derivative kernels)
  int nbf;               // largest molecule's function count (LDS row length)
  int64_t msize;         // elements of one packed matrix (sum of n_m^2)
  int maxat;             // largest molecule's atom count
  // ket-side J of the J/K kernels, indexed by ket position instead of matrix
  // element so that the lanes of a tile add into consecutive addresses: ket
  // pair class Y's block starts at jp_off[Y] and holds component cd of the
  // pair at position j of that class's ket order at cd * jp_n[Y] + j; one
  // density's array is jp_size long
  double* jpair;
  int64_t jp_size;
  int jp_off[6];
  int jp_n[6];
  // shell-quartet blocks (store kernels only, else null): for the lane at
  // list position j, blk[j] = block offset | (orientation flags << 61) (the
  // store branch of lane_body)
  const uint64_t* blk;
  // dynamic LDS a J/K launch may request, bytes (the device's per-workgroup
  // limit unless lowered with pq_engine_set_lds_limit)
  int lds_max;
  // gradient kernels: the beta density of an unrestricted wavefunction (the
  // kernels' D is then the alpha density), null for a closed shell
  const double* grad_db;
};

// pair class of shells with l1 >= l2 (ss, ps, pp, ds, dp, dd)
constexpr int pair_class_of(int l1, int l2) { return l1 * (l1 + 1) / 2 + l2; }

// ---------------------------------------------------------------------
// Rys roots u_i = t_i^2 and weights w_i for order NR at argument X.
// Piecewise polynomial in X below rys_xasym[NR], large-X limit above;
// the large-X branch is always evaluated in FP64.
// ---------------------------------------------------------------------
template <int NR, typename T>
__device__ __forceinline__ void rys_roots(double X, T* u, T* w) {
  if (X >= rys_xasym[NR]) {
    const double rx = 1.0 / X, rs = rsqrt(X);
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      u[i] = (T)(rys_asym[NR - 1][i] * rx);
      w[i] = (T)(rys_asym[NR - 1][NR + i] * rs);
    }
    return;
  }
  const int k = (int)X;  // interval width 1
  const T s = (T)(2.0 * (X - (double)k) - 1.0);
  const double* c = rys_coef + rys_off[NR - 1] + k * (2 * NR * (RYS_DEG + 1));
#pragma unroll
  for (int f = 0; f < 2 * NR; ++f) {
    T v = (T)c[f * (RYS_DEG + 1) + RYS_DEG];
#pragma unroll
    for (int j = RYS_DEG - 1; j >= 0; --j) v = v * s + (T)c[f * (RYS_DEG + 1) + j];
    if (f < NR) u[f] = v; else w[f - NR] = v;
  }
}

// One root r of the order-NR rule (large heavy-class bodies keep the root loop
// rolled to bound code size).  `asym` and `c`/`sv` are set by rys_prepare.
template <int NR>
struct RysPoint {
  bool asym;
  double rx, rs;        // 1/X, 1/sqrt(X) on the large-X branch
  const double* c;      // coefficient block of the interval
  double sv;            // local variable on the interval
};

// Interval stride of the order-NR tables in global memory and of their LDS
// copies.  The LDS copy pads every interval by one double (FP64) or one float4
// (FP32) so that the stride is 2 mod 64 dwords (FP64) or 4 mod 64 dwords
// (FP32): lanes whose X falls in different intervals then read different
// banks instead of the same two or four.
template <int NR>
constexpr int kRysStride = 2 * NR * (RYS_DEG + 1);
template <int NR>
constexpr int kRysLdsStride = kRysStride<NR> + 1;
template <int NR>
constexpr int kRysStride32 = 2 * NR * RYS_STRIDE32;
template <int NR>
constexpr int kRysLdsStride32 = kRysStride32<NR> + 4;

// LT: tab is the order-NR block of the table (rys_coef + rys_off[NR - 1])
// staged in LDS with the padded stride; otherwise the global table is read
template <int NR, bool LT = false>
__device__ __forceinline__ RysPoint<NR> rys_prepare(double X, const double* tab = nullptr) {
  RysPoint<NR> p;
  const double* base = LT ? tab : rys_coef + rys_off[NR - 1];
  p.asym = X >= rys_xasym[NR];
  if (p.asym) {
    // 1/X as the square of 1/sqrt(X): a few ulp, without a full division
    p.rs = rsqrt(X);
    p.rx = p.rs * p.rs;
    p.c = base;
    p.sv = 0.0;
  } else {
    const int k = (int)X;
    p.rx = p.rs = 0.0;
    p.sv = 2.0 * (X - (double)k) - 1.0;
    p.c = base + k * (LT ? kRysLdsStride<NR> : kRysStride<NR>);
  }
  return p;
}

template <int NR, typename T>
__device__ __forceinline__ void rys_root(const RysPoint<NR>& p, int r, T& u, T& w) {
  if (p.asym) {
    u = (T)(rys_asym[NR - 1][r] * p.rx);
    w = (T)(rys_asym[NR - 1][NR + r] * p.rs);
    return;
  }
  const T s = (T)p.sv;
  const double* cu = p.c + r * (RYS_DEG + 1);
  const double* cw = p.c + (NR + r) * (RYS_DEG + 1);
  T vu = (T)cu[RYS_DEG], vw = (T)cw[RYS_DEG];
#pragma unroll
  for (int j = RYS_DEG - 1; j >= 0; --j) {
    vu = vu * s + (T)cu[j];
    vw = vw * s + (T)cw[j];
  }
  u = vu;
  w = vw;
}

// FP32 rule (FP32 quartets of mixed-precision builds): degree-6 float table
// below rys_xasym32[NR], each function's coefficients two aligned float4
// loads; above it the large-X limit, evaluated in FP64 and rounded.
template <int NR>
struct RysPoint32 {
  bool asym;
  double rx, rs;
  const float* c;
  float sv;
};

template <int NR, bool LT = false>
__device__ __forceinline__ RysPoint32<NR> rys_prepare32(double X, const float* tab = nullptr) {
  RysPoint32<NR> p;
  const float* base = LT ? tab : rys_coef32 + rys_off32[NR - 1];
  p.asym = X >= rys_xasym32[NR];
  if (p.asym) {
    p.rs = rsqrt(X);
    p.rx = p.rs * p.rs;
    p.c = base;
    p.sv = 0.f;
  } else {
    const int k = (int)X;
    p.rx = p.rs = 0.0;
    p.sv = (float)(2.0 * (X - (double)k) - 1.0);
    p.c = base + k * (LT ? kRysLdsStride32<NR> : kRysStride32<NR>);
  }
  return p;
}

__device__ __forceinline__ float horner6(const float4 lo, const float4 hi, float s) {
  static_assert(RYS_DEG32 == 6 && RYS_STRIDE32 == 8, "table layout");
  float v = hi.z;
  v = v * s + hi.y;
  v = v * s + hi.x;
  v = v * s + lo.w;
  v = v * s + lo.z;
  v = v * s + lo.y;
  return v * s + lo.x;
}

template <int NR>
__device__ __forceinline__ void rys_root32(const RysPoint32<NR>& p, int r, float& u, float& w) {
  if (p.asym) {
    u = (float)(rys_asym[NR - 1][r] * p.rx);
    w = (float)(rys_asym[NR - 1][NR + r] * p.rs);
    return;
  }
  const float4* cu = reinterpret_cast<const float4*>(p.c + r * RYS_STRIDE32);
  const float4* cw = reinterpret_cast<const float4*>(p.c + (NR + r) * RYS_STRIDE32);
  u = horner6(cu[0], cu[1], p.sv);
  w = horner6(cw[0], cw[1], p.sv);
}

// Rule selection by arithmetic type: FP64 table for double, FP32 table for float.
template <int NR, typename T>
struct RysRule {
  using Point = RysPoint<NR>;
  template <bool LT>
  static __device__ __forceinline__ Point prepare(double X, const void* tab) {
    return rys_prepare<NR, LT>(X, static_cast<const double*>(tab));
  }
  static __device__ __forceinline__ void root(const Point& p, int r, T& u, T& w) { rys_root<NR, T>(p, r, u, w); }
};
template <int NR>
struct RysRule<NR, float> {
  using Point = RysPoint32<NR>;
  template <bool LT>
  static __device__ __forceinline__ Point prepare(double X, const void* tab) {
    return rys_prepare32<NR, LT>(X, static_cast<const float*>(tab));
  }
  static __device__ __forceinline__ void root(const Point& p, int r, float& u, float& w) {
    rys_root32<NR>(p, r, u, w);
  }
};

// Runtime-order variant (scalar entry points and one-electron kernels).
template <typename T>
__device__ inline void rys_roots_rt(int nr, double X, T* u, T* w) {
  switch (nr) {
    case 1: rys_roots<1, T>(X, u, w); break;
    case 2: rys_roots<2, T>(X, u, w); break;
    case 3: rys_roots<3, T>(X, u, w); break;
    case 4: rys_roots<4, T>(X, u, w); break;
    default: rys_roots<5, T>(X, u, w); break;
  }
}

// ---------------------------------------------------------------------
// 1-D Rys integrals for one root and one Cartesian direction.
// G(n,m) = I(n,0,m,0) by the ABD recurrence (crys.c:1481-1516 semantics,
// in terms of u = t^2), then the transfer to I(ia,ib,ic,id)
// (crys.c:1518-1533), done here as two horizontal recurrences
// I(a,b+1) = I(a+1,b) + (A-B) I(a,b) and likewise on the ket.
// G00 is the caller's seed (1, or the weight x prefactor for one direction).
// ---------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename T>
__device__ __forceinline__ void rys_1d(T G00, T C00, T D00, T B10, T B01, T B00, T AB, T CD,
                                       T (&I)[LA + 1][LB + 1][LC + 1][LD + 1]) {
  constexpr int N = LA + LB, M = LC + LD;
  T G[N + 1][M + 1];
  G[0][0] = G00;
  if (N > 0) G[1][0] = C00 * G00;
#pragma unroll
  for (int a = 2; a <= N; ++a) G[a][0] = (T)(a - 1) * B10 * G[a - 2][0] + C00 * G[a - 1][0];
  if (M > 0) {
    G[0][1] = D00 * G00;
#pragma unroll
    for (int b = 2; b <= M; ++b) G[0][b] = (T)(b - 1) * B01 * G[0][b - 2] + D00 * G[0][b - 1];
#pragma unroll
    for (int a = 1; a <= N; ++a) {
      G[a][1] = (T)a * B00 * G[a - 1][0] + D00 * G[a][0];
#pragma unroll
      for (int b = 2; b <= M; ++b)
        G[a][b] = (T)(b - 1) * B01 * G[a][b - 2] + (T)a * B00 * G[a - 1][b - 1] + D00 * G[a][b - 1];
    }
  }
  // bra transfer: H[b][a][m], valid for a <= N - b
  T H[LB + 1][N + 1][M + 1];
#pragma unroll
  for (int a = 0; a <= N; ++a)
#pragma unroll
    for (int m = 0; m <= M; ++m) H[0][a][m] = G[a][m];
#pragma unroll
  for (int b = 1; b <= LB; ++b)
#pragma unroll
    for (int a = 0; a <= N - b; ++a)
#pragma unroll
      for (int m = 0; m <= M; ++m) H[b][a][m] = H[b - 1][a + 1][m] + AB * H[b - 1][a][m];
  // ket transfer
#pragma unroll
  for (int ia = 0; ia <= LA; ++ia)
#pragma unroll
    for (int ib = 0; ib <= LB; ++ib) {
      T Kd[LD + 1][M + 1];
#pragma unroll
      for (int m = 0; m <= M; ++m) Kd[0][m] = H[ib][ia][m];
#pragma unroll
      for (int d = 1; d <= LD; ++d)
#pragma unroll
        for (int c = 0; c <= M - d; ++c) Kd[d][c] = Kd[d - 1][c + 1] + CD * Kd[d - 1][c];
#pragma unroll
      for (int ic = 0; ic <= LC; ++ic)
#pragma unroll
        for (int id = 0; id <= LD; ++id) I[ia][ib][ic][id] = Kd[id][ic];
    }
}

constexpr double kTwoPi25 = 34.98683665524972497;  // 2 pi^(5/2)

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
  const double r = rsqrt(pq);
  const double rpq = r * r;
  s.X = b.p * q.p * rpq * (s.PQ[0] * s.PQ[0] + s.PQ[1] * s.PQ[1] + s.PQ[2] * s.PQ[2]);
  s.pref = kTwoPi25 * b.Kp * q.Kp * r;
  s.rpq = rpq;
  s.qfac = q.p * rpq;
  s.pfac = b.p * rpq;
  s.i2p = 0.5 * b.rp;
  s.i2q = 0.5 * q.rp;
  return s;
}

// Boys function values F_0(T) .. F_L(T), FP64.  Below BOYS_TASYM: the Taylor
// series of order BOYS_K about the nearest grid point T_g (dF_m/dT =
// -F_{m+1}), from one row of the grid table (F_0 .. F_{L+K} at T_g):
//   F_m(T) = sum_k F_{m+k}(T_g) x^k / k!,  x = T_g - T,  |x| <= 1/8,
// truncation below 2.9e-18 relative.  Above: the asymptotic form
// F_0 = sqrt(pi / T) / 2, F_m = F_{m-1} (2m - 1) / (2T).  LT: tab is the
// table's LDS copy (same layout); otherwise the global table is read.
template <int L, bool LT = false>
__device__ __forceinline__ void boys_grid_eval(double T, double (&F)[L + 1], const double* tab = nullptr) {
  static_assert(L + BOYS_K < BOYS_MTAB, "Boys grid table too short for this order");
  if (T < (double)BOYS_TASYM) {
    const int g = (int)(T * BOYS_INV_H + 0.5);
    const double x = (double)g * (1.0 / BOYS_INV_H) - T;
    constexpr int NV = L + BOYS_K + 1;
    double c[NV];
    if constexpr (LT) {
      const double* row = tab + g * BOYS_STRIDE;
#pragma unroll
      for (int i = 0; i < NV; ++i) c[i] = row[i];
    } else {
      constexpr int NP = (NV + 1) / 2;
      const double2* row = reinterpret_cast<const double2*>(boys_grid + g * BOYS_STRIDE);
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        const double2 v = row[i];
        c[2 * i] = v.x;
        if (2 * i + 1 < NV) c[2 * i + 1] = v.y;
      }
    }
    double xk[BOYS_K];  // x / (k + 1)
#pragma unroll
    for (int k = 0; k < BOYS_K; ++k) xk[k] = x * (1.0 / (k + 1));
#pragma unroll
    for (int m = 0; m <= L; ++m) {
      double v = c[m + BOYS_K];
#pragma unroll
      for (int k = BOYS_K - 1; k >= 0; --k) v = fma(v, xk[k], c[m + k]);
      F[m] = v;
    }
  } else {
    const double rs = rsqrt(T), rx = rs * rs;
    F[0] = 0.88622692545275801365 * rs;  // sqrt(pi) / 2
#pragma unroll
    for (int m = 1; m <= L; ++m) F[m] = F[m - 1] * ((m - 0.5) * rx);
  }
}

// Copies the Boys grid table into LDS (same layout).
__device__ __forceinline__ void stage_boys_table(double* lds) {
  for (int i = threadIdx.x; i < BOYS_NG * BOYS_STRIDE; i += blockDim.x) lds[i] = boys_grid[i];
}

// Canonical unique-ERI index (cints.c:413-425) in 64-bit arithmetic.
__host__ __device__ inline uint64_t ijkl2intindex(int64_t i, int64_t j, int64_t k, int64_t l) {
  if (i < j) { int64_t t = i; i = j; j = t; }
  if (k < l) { int64_t t = k; k = l; l = t; }
  uint64_t ij = (uint64_t)(i * (i + 1) / 2 + j), kl = (uint64_t)(k * (k + 1) / 2 + l);
  if (ij < kl) { uint64_t t = ij; ij = kl; kl = t; }
  return ij * (ij + 1) / 2 + kl;
}

// Wavefront-wide sum (64 lanes) by butterfly exchange.
// r with r(r+1)/2 <= t < (r+1)(r+2)/2: row of triangular index t.
__device__ __forceinline__ int64_t tri_root(uint64_t t) {
  int64_t r = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((uint64_t)(r + 1) * (uint64_t)(r + 2) / 2 <= t) ++r;
  while (r > 0 && (uint64_t)r * (uint64_t)(r + 1) / 2 > t) --r;
  return r;
}

// Load of read-only data at a wavefront-uniform address through the constant
// address space, so that the compiler can use a scalar load (no vector memory
// instruction per lane).  Only for memory no kernel writes while it runs.
template <typename T>
__device__ __forceinline__ T ld_uniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "word-sized records only");
  T r;
  const __attribute__((address_space(4))) uint32_t* src = (const __attribute__((address_space(4))) uint32_t*)p;
  uint32_t* dst = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = src[i];
  return r;
}

// v moved across lanes by the DPP control CTRL; lanes of rows outside
// ROW_MASK (and lanes without a source) receive 0
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wavefront (every lane active), returned to all
// lanes: quad permutations, half-row and row mirrors, then the row
// broadcasts 15 and 31 gather the total in lane 63 without leaving the VALU.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0xb1>(v);        // quad_perm [1,0,3,2]
  v += dpp_move<0x4e>(v);        // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);       // row_half_mirror: 8-lane sums
  v += dpp_move<0x140>(v);       // row_mirror: 16-lane row sums
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

}  // namespace pq
