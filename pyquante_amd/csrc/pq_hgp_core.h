// pq_hgp_core.h -- the Head-Gordon-Pople recurrences shared by the stored-
// integral HGP kernels (pq_hgp_impl.h) and the J/K kernels of the light
// classes (pq_eri_impl.h): the [e0|f0]^(m) table layout of a class, the
// Obara-Saika vertical recurrence per primitive quartet (chgp.c:244-419) and
// the closed-form horizontal recurrence (chgp.c:78-170).
#pragma once
#include <type_traits>

#include "pq_device.h"

namespace pq {

// Cartesian triples of total order 0, 1, 2, ... in one flat index: order l
// starts at ncum(l - 1), components within an order in PyQuante order.
__host__ __device__ constexpr int ncum(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }
__host__ __device__ constexpr int tri_order(int t) {
  int l = 0;
  while (ncum(l) <= t) ++l;
  return l;
}
__host__ __device__ constexpr int tri_pow(int t, int d) {
  return cart_pow(tri_order(t), t - ncum(tri_order(t) - 1), d);
}
__host__ __device__ constexpr int tri_index(int lx, int ly, int lz) {
  const int l = lx + ly + lz;
  return ncum(l - 1) + (l - lx) * (l - lx + 1) / 2 + (l - lx - ly);
}
// t with its power along d lowered by one
__host__ __device__ constexpr int tri_lower(int t, int d) {
  return tri_index(tri_pow(t, 0) - (d == 0), tri_pow(t, 1) - (d == 1), tri_pow(t, 2) - (d == 2));
}
// the direction the recurrences step along: the first with a nonzero power
__host__ __device__ constexpr int tri_dir(int t) { return tri_pow(t, 0) > 0 ? 0 : (tri_pow(t, 1) > 0 ? 1 : 2); }
__host__ __device__ constexpr int binom_c(int n, int k) {
  return (k < 0 || k > n) ? 0 : (k == 0 || k == n) ? 1 : binom_c(n - 1, k - 1) + binom_c(n - 1, k);
}

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// Table layout of one class.  [e0|f0]^(m) is kept for m = 0 .. L-|e| at
// f = 0 (the bra stage feeds every ket level) and for m = 0 .. LCD-|f| at
// f > 0 (each ket step uses order m + 1 of the level below); at ket level
// |f| only the e with |e| >= LA - (LCD - |f|) feed the targets.
template <int LA, int LB, int LC, int LD>
struct HgpCls {
  static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD, NR = L / 2 + 1;
  static constexpr int NE = ncum(LAB), NF = ncum(LCD);
  static constexpr int E0 = ncum(LA - 1), F0 = ncum(LC - 1);  // first target triples
  static constexpr int NEA = NE - E0, NFC = NF - F0;
  // targets, and their count rounded up to the 8 the rolled variant adds per batch
  static constexpr int NT = NEA * NFC, NTP = (NT + 7) / 8 * 8;
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr bool need(int e, int f) { return f == 0 || tri_order(e) >= LA - (LCD - tri_order(f)); }
  static constexpr int nm(int e, int f) {
    return !need(e, f) ? 0 : (f == 0 ? L - tri_order(e) + 1 : LCD - tri_order(f) + 1);
  }
  static constexpr int off(int e, int f) {
    int s = 0;
    for (int i = 0; i < NE; ++i)
      for (int j = 0; j < NF; ++j) {
        if (i == e && j == f) return s;
        s += nm(i, j);
      }
    return s;
  }
  static constexpr int SIZE = off(NE, 0);
  static_assert(off(0, 0) == 0, "the [00|00] orders start the table");
};

// Classes up to this many table entries run the unrolled recurrence (tables
// in registers): the s/p classes up to (pp|ps), and (ds|ps).  With 70
// ((dd|ss) as well) the shared store launch needs 256 registers; at 52 it
// takes (H2O)_8 get2ints from 19.7 to 16.9 ms against 40.
constexpr int kHgpUnrollMax = 52;

// Classes above kHgpUnrollMax and up to kHgpAloneMax entries ((dd|ss),
// (dp|ps), (ds|pp), (ds|ds), (pp|pp), (dd|ps), (dp|ds), (dp|pp)) run the
// unrolled recurrence too, each in a launch of its own: unrolled they need
// 224-256 registers (1-2 waves per SIMD; (dp|pp) spills 310 of them), which
// inside the shared launch would set the occupancy of every class, but their
// recurrence tables leave private memory for registers.  Measured on (H2O)_8
// 6-31G** get2ints: 16.9 -> 8.7 ms (with (dd|ps), (dp|ds) and (dp|pp) kept
// on step tables 11.8 ms).
constexpr int kHgpAloneMax = 300;
template <int LA, int LB, int LC, int LD>
constexpr bool kHgpUnrolled = HgpCls<LA, LB, LC, LD>::SIZE <= kHgpAloneMax;
template <int LA, int LB, int LC, int LD>
constexpr bool kHgpAlone = HgpCls<LA, LB, LC, LD>::SIZE > kHgpUnrollMax && kHgpUnrolled<LA, LB, LC, LD>;

// Step tables of the rolled variant.  A VRR step writes nm orders at o from
// the entries at o1 (and o2, oe); dir is the recurrence direction, c1 the
// power of the lower entry along it (bra e_i - 1, ket f_j - 1), c2 the bra
// power e_j of a ket step.
struct HgpStep {
  short o, o1, o2, oe;
  unsigned char nm, dir, c1, c2;
};
// One term of the closed-form horizontal recurrence: target row/column idx,
// binomial coefficient cb and the powers of (A-B) or (C-D) along x, y, z.
struct HgpTerm {
  short idx;
  unsigned char cb, p0, p1, p2;
};

template <int LA, int LB, int LC, int LD>
struct HgpSteps {
  using H = HgpCls<LA, LB, LC, LD>;
  static constexpr int count_ket() {
    int n = 0;
    for (int f = 1; f < H::NF; ++f)
      for (int e = 0; e < H::NE; ++e) n += H::need(e, f);
    return n;
  }
  static constexpr int NBRA = H::NE - 1, NKET = count_ket() > 0 ? count_ket() : 1;
  static constexpr int NAB = H::NA * H::NB, NCD = H::NC * H::ND, MAXT = 4;
  HgpStep bra[NBRA > 0 ? NBRA : 1];
  HgpStep ket[NKET];
  short tgt[H::NTP];  // entries past NT stay 0 (they read V[0] into acc's padding)
  unsigned char nbt[NAB], nkt[NCD];
  HgpTerm bt[NAB][MAXT], kt[NCD][MAXT];

  // terms of one side: the shell pair's components (l1 = first, l2 = second)
  static constexpr void terms(int l1, int l2, int x0, HgpTerm (&t)[MAXT], unsigned char& n, int i1, int i2) {
    n = 0;
    const int p1[3] = {cart_pow(l1, i1, 0), cart_pow(l1, i1, 1), cart_pow(l1, i1, 2)};
    const int p2[3] = {cart_pow(l2, i2, 0), cart_pow(l2, i2, 1), cart_pow(l2, i2, 2)};
    for (int kx = 0; kx <= p2[0]; ++kx)
      for (int ky = 0; ky <= p2[1]; ++ky)
        for (int kz = 0; kz <= p2[2]; ++kz) {
          HgpTerm& r = t[n++];
          r.idx = (short)(tri_index(p1[0] + kx, p1[1] + ky, p1[2] + kz) - x0);
          r.cb = (unsigned char)(binom_c(p2[0], kx) * binom_c(p2[1], ky) * binom_c(p2[2], kz));
          r.p0 = (unsigned char)(p2[0] - kx);
          r.p1 = (unsigned char)(p2[1] - ky);
          r.p2 = (unsigned char)(p2[2] - kz);
        }
  }
  static constexpr HgpSteps make() {
    HgpSteps t{};
    // entry offsets, one pass (H::off walks the table per call)
    int ot[H::NE + 1][H::NF] = {};
    int acc = 0;
    for (int i = 0; i < H::NE; ++i)
      for (int j = 0; j < H::NF; ++j) {
        ot[i][j] = acc;
        acc += H::nm(i, j);
      }
    for (int e = 1; e < H::NE; ++e) {
      const int i = tri_dir(e), e1 = tri_lower(e, i), ei = tri_pow(e, i) - 1;
      HgpStep& s = t.bra[e - 1];
      s.o = (short)ot[e][0];
      s.o1 = (short)ot[e1][0];
      s.o2 = (short)(ei >= 1 ? ot[tri_lower(e1, i)][0] : 0);
      s.oe = 0;
      s.nm = (unsigned char)H::nm(e, 0);
      s.dir = (unsigned char)i;
      s.c1 = (unsigned char)ei;
      s.c2 = 0;
    }
    int n = 0;
    for (int f = 1; f < H::NF; ++f) {
      const int j = tri_dir(f), f1 = tri_lower(f, j), fj = tri_pow(f, j) - 1;
      for (int e = 0; e < H::NE; ++e) {
        if (!H::need(e, f)) continue;
        const int ej = tri_pow(e, j);
        HgpStep& s = t.ket[n++];
        s.o = (short)ot[e][f];
        s.o1 = (short)ot[e][f1];
        s.o2 = (short)(fj >= 1 ? ot[e][tri_lower(f1, j)] : 0);
        s.oe = (short)(ej >= 1 ? ot[tri_lower(e, j)][f1] : 0);
        s.nm = (unsigned char)H::nm(e, f);
        s.dir = (unsigned char)j;
        s.c1 = (unsigned char)fj;
        s.c2 = (unsigned char)ej;
      }
    }
    for (int e = H::E0; e < H::NE; ++e)
      for (int f = H::F0; f < H::NF; ++f) t.tgt[(e - H::E0) * H::NFC + (f - H::F0)] = (short)ot[e][f];
    for (int a = 0; a < H::NA; ++a)
      for (int b = 0; b < H::NB; ++b) terms(LA, LB, H::E0, t.bt[a * H::NB + b], t.nbt[a * H::NB + b], a, b);
    for (int c = 0; c < H::NC; ++c)
      for (int d = 0; d < H::ND; ++d) terms(LC, LD, H::F0, t.kt[c * H::ND + d], t.nkt[c * H::ND + d], c, d);
    return t;
  }
};

template <int LA, int LB, int LC, int LD>
__constant__ HgpSteps<LA, LB, LC, LD> kHgpSteps = HgpSteps<LA, LB, LC, LD>::make();

__device__ __forceinline__ double pick3(const double (&v)[3], int d) { return d == 0 ? v[0] : (d == 1 ? v[1] : v[2]); }

// The vertical recurrence of hgp_vrr driven by the step tables.  The table
// lives in per-lane private memory, so each step first loads the orders of
// its operands that it uses (behind uniform branches, so that the loads are
// still issued back to back; V is padded so that the reads stay inside it)
// and then computes and stores its nm orders: the loads of one step are in
// flight together instead of one order at a time.  acc (NTP entries) takes the targets in batches of 8 for
// the same reason.  (Loading the orders two at a time, from entries aligned
// to even offsets, measured no faster.)
// V: the caller's table (HgpCls::SIZE + L + 1 entries); with accumulate
// false the targets stay in V (a quartet of one primitive quartet, whose
// horizontal recurrence then reads them there).
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void hgp_vrr_rolled(const PairInfo& bp, const PrimPair& b, const PairInfo& kp,
                                               const PrimPair& q, double* acc, double* V, bool accumulate) {
  using H = HgpCls<LA, LB, LC, LD>;
  using S = HgpSteps<LA, LB, LC, LD>;
  const S& T = kHgpSteps<LA, LB, LC, LD>;
  constexpr int L = H::L, LCD = H::LCD;
  const PrimQ s = prim_setup(bp, b, kp, q);
  double u[H::NR], w[H::NR];
  rys_roots<H::NR, double>(s.X, u, w);
#pragma unroll
  for (int r = 0; r < H::NR; ++r) w[r] *= s.pref;
#pragma unroll
  for (int m = 0; m <= L; ++m) {
    double f = 0.0;
#pragma unroll
    for (int r = 0; r < H::NR; ++r) {
      f += w[r];
      w[r] *= u[r];
    }
    V[m] = f;  // [00|00] entries start the table
  }
  const double PA[3] = {s.PA[0], s.PA[1], s.PA[2]}, QC[3] = {s.QC[0], s.QC[1], s.QC[2]};
  const double WP[3] = {-s.qfac * s.PQ[0], -s.qfac * s.PQ[1], -s.qfac * s.PQ[2]};
  const double WQ[3] = {s.pfac * s.PQ[0], s.pfac * s.PQ[1], s.pfac * s.PQ[2]};
  const double i2pq = 0.5 * s.rpq;
  // bra steps: nm = L - |e| + 1 <= L orders, operands up to order nm
#pragma unroll 1
  for (int n = 0; n < S::NBRA; ++n) {
    const HgpStep st = T.bra[n];
    const double pa = pick3(PA, st.dir), wp = pick3(WP, st.dir), c2 = st.c1 * s.i2p;
    double x1[L + 1], x2[L + 1];
#pragma unroll
    for (int m = 0; m <= L; ++m)
      if (m <= st.nm) x1[m] = V[st.o1 + m];
    if (st.c1) {
#pragma unroll
      for (int m = 0; m <= L; ++m)
        if (m <= st.nm) x2[m] = V[st.o2 + m];
    }
#pragma unroll
    for (int m = 0; m < L; ++m) {
      if (m >= st.nm) break;
      double v = pa * x1[m] + wp * x1[m + 1];
      if (st.c1) v += c2 * (x2[m] - s.qfac * x2[m + 1]);
      V[st.o + m] = v;
    }
  }
  // ket steps: nm = LCD - |f| + 1 <= LCD orders
#pragma unroll 1
  for (int n = 0; n < S::NKET; ++n) {
    const HgpStep st = T.ket[n];
    const double qc = pick3(QC, st.dir), wq = pick3(WQ, st.dir), c2 = st.c1 * s.i2q, c3 = st.c2 * i2pq;
    double x1[LCD + 1], x2[LCD + 1], xe[LCD > 0 ? LCD : 1];
#pragma unroll
    for (int m = 0; m <= LCD; ++m)
      if (m <= st.nm) x1[m] = V[st.o1 + m];
    if (st.c1) {
#pragma unroll
      for (int m = 0; m <= LCD; ++m)
        if (m <= st.nm) x2[m] = V[st.o2 + m];
    }
    if (st.c2) {
#pragma unroll
      for (int m = 0; m < LCD; ++m)
        if (m < st.nm) xe[m] = V[st.oe + m + 1];
    }
#pragma unroll
    for (int m = 0; m < LCD; ++m) {
      if (m >= st.nm) break;
      double v = qc * x1[m] + wq * x1[m + 1];
      if (st.c1) v += c2 * (x2[m] - s.pfac * x2[m + 1]);
      if (st.c2) v += c3 * xe[m];
      V[st.o + m] = v;
    }
  }
  if (accumulate) {
#pragma unroll 1
    for (int t0 = 0; t0 < H::NTP; t0 += 8) {
      double x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = V[T.tgt[t0 + j]];
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[t0 + j] += x[j];
    }
  }
}

// One primitive quartet: the vertical recurrence, targets added into acc
// ([e - E0][f - F0], NEA x NFC).  GRID: the Boys values from the grid table
// (boys_grid_eval; LT: its LDS copy btab); otherwise from the Rys rule of
// the class.
template <int LA, int LB, int LC, int LD, bool GRID = false, bool LT = false>
__device__ __forceinline__ void hgp_vrr(const PairInfo& bp, const PrimPair& b, const PairInfo& kp,
                                        const PrimPair& q, double* acc, const double* btab = nullptr) {
  using H = HgpCls<LA, LB, LC, LD>;
  const PrimQ s = prim_setup(bp, b, kp, q);
  double V[H::SIZE];
  // [00|00]^(m) = 2 pi^(5/2) K_ab K_cd / (p q sqrt(p+q)) F_m(X)
  if constexpr (GRID) {
    double F[H::L + 1];
    boys_grid_eval<H::L, LT>(s.X, F, btab);
#pragma unroll
    for (int m = 0; m <= H::L; ++m) V[m] = s.pref * F[m];  // [00|00] entries start the table
  } else {
    double u[H::NR], w[H::NR];
    rys_roots<H::NR, double>(s.X, u, w);
#pragma unroll
    for (int r = 0; r < H::NR; ++r) w[r] *= s.pref;
    static_for<0, H::L + 1>([&](auto M) {
      double f = 0.0;
#pragma unroll
      for (int r = 0; r < H::NR; ++r) {
        f += w[r];
        w[r] *= u[r];
      }
      V[decltype(M)::value] = f;
    });
  }
  // W - P = -q/(p+q) (P - Q),  W - Q = p/(p+q) (P - Q)
  const double WP[3] = {-s.qfac * s.PQ[0], -s.qfac * s.PQ[1], -s.qfac * s.PQ[2]};
  const double WQ[3] = {s.pfac * s.PQ[0], s.pfac * s.PQ[1], s.pfac * s.PQ[2]};
  const double i2pq = 0.5 * s.rpq;
  // bra: [e+1_i,0|00]^(m) = PA_i [e]^(m) + WP_i [e]^(m+1)
  //                         + e_i/(2p) ([e-1_i]^(m) - q/(p+q) [e-1_i]^(m+1))
  static_for<1, H::NE>([&](auto E) {
    constexpr int e = decltype(E)::value, i = tri_dir(e), e1 = tri_lower(e, i);
    constexpr int ei = tri_pow(e, i) - 1;  // power of e1 along i
    constexpr int o = H::off(e, 0), o1 = H::off(e1, 0);
    constexpr int o2 = ei >= 1 ? H::off(tri_lower(e1, i), 0) : 0;
    constexpr int nm = H::nm(e, 0);
    const double c2 = ei * s.i2p;
#pragma unroll
    for (int m = 0; m < nm; ++m) {
      double v = s.PA[i] * V[o1 + m] + WP[i] * V[o1 + m + 1];
      if constexpr (ei >= 1) v += c2 * (V[o2 + m] - s.qfac * V[o2 + m + 1]);
      V[o + m] = v;
    }
  });
  // ket: [e0|f+1_j,0]^(m) = QC_j [e|f]^(m) + WQ_j [e|f]^(m+1)
  //        + f_j/(2q) ([e|f-1_j]^(m) - p/(p+q) [e|f-1_j]^(m+1))
  //        + e_j/(2(p+q)) [e-1_j|f]^(m+1)
  static_for<1, H::NF>([&](auto F) {
    constexpr int f = decltype(F)::value, j = tri_dir(f), f1 = tri_lower(f, j);
    constexpr int fj = tri_pow(f, j) - 1;  // power of f1 along j
    constexpr int f2 = fj >= 1 ? tri_lower(f1, j) : 0;
    const double c2 = fj * s.i2q;
    static_for<0, H::NE>([&](auto E) {
      constexpr int e = decltype(E)::value;
      if constexpr (H::need(e, f)) {
        constexpr int ej = tri_pow(e, j);
        constexpr int em = ej >= 1 ? tri_lower(e, j) : 0;
        static_assert(H::need(e, f1) && (fj < 1 || H::need(e, f2)) && (ej < 1 || H::need(em, f1)),
                      "HGP ket recurrence reads a dropped entry");
        constexpr int o = H::off(e, f), o1 = H::off(e, f1);
        constexpr int o2 = fj >= 1 ? H::off(e, f2) : 0;
        constexpr int oe = ej >= 1 ? H::off(em, f1) : 0;
        constexpr int nm = H::nm(e, f);
#pragma unroll
        for (int m = 0; m < nm; ++m) {
          double v = s.QC[j] * V[o1 + m] + WQ[j] * V[o1 + m + 1];
          if constexpr (fj >= 1) v += c2 * (V[o2 + m] - s.pfac * V[o2 + m + 1]);
          if constexpr (ej >= 1) v += (ej * i2pq) * V[oe + m + 1];
          V[o + m] = v;
        }
      }
    });
  });
  static_for<H::E0, H::NE>([&](auto E) {
    static_for<H::F0, H::NF>([&](auto F) {
      // offsets as constants: off() left to the optimiser is a loop over the
      // table that, rolled, puts V in private memory
      constexpr int e = decltype(E)::value, f = decltype(F)::value, o = H::off(e, f);
      acc[(e - H::E0) * H::NFC + (f - H::F0)] += V[o];
    });
  });
}

// One (c, d) ket component pair of the closed-form horizontal recurrence for
// bra triple row er (a + k): sum over l <= d of C(d,l) CD^(d-l) [er|c+l].
template <int LC, int LD, int NFC, int F0, int C, int Dc>
__device__ __forceinline__ double hrr_ket(const double* row, const double (&CDp)[3][3]) {
  constexpr int cx = cart_pow(LC, C, 0), cy = cart_pow(LC, C, 1), cz = cart_pow(LC, C, 2);
  constexpr int dx = cart_pow(LD, Dc, 0), dy = cart_pow(LD, Dc, 1), dz = cart_pow(LD, Dc, 2);
  double v = 0.0;
  static_for<0, dx + 1>([&](auto LX) {
    static_for<0, dy + 1>([&](auto LY) {
      static_for<0, dz + 1>([&](auto LZ) {
        constexpr int lx = decltype(LX)::value, ly = decltype(LY)::value, lz = decltype(LZ)::value;
        constexpr double cb = (double)(binom_c(dx, lx) * binom_c(dy, ly) * binom_c(dz, lz));
        constexpr int f = tri_index(cx + lx, cy + ly, cz + lz);
        double t = cb * row[f - F0];
        if constexpr (dx - lx > 0) t *= CDp[0][dx - lx];
        if constexpr (dy - ly > 0) t *= CDp[1][dy - ly];
        if constexpr (dz - lz > 0) t *= CDp[2][dz - lz];
        v += t;
      });
    });
  });
  return v;
}

// The horizontal recurrence of one contracted block in closed form: out
// (NA x NB x NC x ND, PyQuante component order, no normalisation factors)
// from the contracted targets acc ([e - E0][f - F0]):
//   (ab|cd) = sum_{k<=b} C(b,k) AB^(b-k) sum_{l<=d} C(d,l) CD^(d-l) [a+k,0|c+l,0]
// (chgp.c:78-170, contr_hrr, as one sum instead of a recursion).
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void hgp_hrr_block(const double* acc, const double (&AB)[3], const double (&CD)[3],
                                              double (&out)[ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD)]) {
  using H = HgpCls<LA, LB, LC, LD>;
  double ABp[3][3], CDp[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ABp[d][0] = CDp[d][0] = 1.0;
    ABp[d][1] = AB[d];
    ABp[d][2] = AB[d] * AB[d];
    CDp[d][1] = CD[d];
    CDp[d][2] = CD[d] * CD[d];
  }
  static_for<0, H::NA>([&](auto A_) {
    static_for<0, H::NB>([&](auto B_) {
      constexpr int a = decltype(A_)::value, bb = decltype(B_)::value;
      constexpr int ax = cart_pow(LA, a, 0), ay = cart_pow(LA, a, 1), az = cart_pow(LA, a, 2);
      constexpr int bx = cart_pow(LB, bb, 0), by = cart_pow(LB, bb, 1), bz = cart_pow(LB, bb, 2);
      static_for<0, H::NC>([&](auto C_) {
        static_for<0, H::ND>([&](auto D_) {
          constexpr int c = decltype(C_)::value, d = decltype(D_)::value;
          double v = 0.0;
          static_for<0, bx + 1>([&](auto KX) {
            static_for<0, by + 1>([&](auto KY) {
              static_for<0, bz + 1>([&](auto KZ) {
                constexpr int kx = decltype(KX)::value, ky = decltype(KY)::value, kz = decltype(KZ)::value;
                constexpr double cb = (double)(binom_c(bx, kx) * binom_c(by, ky) * binom_c(bz, kz));
                constexpr int e = tri_index(ax + kx, ay + ky, az + kz);
                double t = cb * hrr_ket<LC, LD, H::NFC, H::F0, c, d>(acc + (e - H::E0) * H::NFC, CDp);
                if constexpr (bx - kx > 0) t *= ABp[0][bx - kx];
                if constexpr (by - ky > 0) t *= ABp[1][by - ky];
                if constexpr (bz - kz > 0) t *= ABp[2][bz - kz];
                v += t;
              });
            });
          });
          out[((a * H::NB + bb) * H::NC + c) * H::ND + d] = v;
        });
      });
    });
  });
}

}  // namespace pq
