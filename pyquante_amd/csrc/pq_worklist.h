// pq_worklist.h -- host-side construction of the class work lists: the tiles
// of each angular-momentum class (one bra pair, up to 64 kets of equal
// contraction length, a bra primitive range), their split between the FP64
// and FP32 buckets, and the bra runs of the run-walking kernels.  Plain C++
// (no HIP), so that tests can build it with a host compiler under the
// address and undefined-behaviour sanitizers; the engine uploads the results.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/pq_abi.h"
#include "pq_types.h"

namespace pq {

// (l1, l2) of the six pair classes ss, ps, pp, ds, dp, dd (l1 >= l2)
constexpr int kPairL[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};

// Bra primitive pairs per tile are capped so that one lane evaluates at most
// about kPrimBudget primitive quartets: heavily contracted quartets are
// spread over several tiles (their partial sums digest independently).
constexpr int kPrimBudget = 36;
constexpr int kTileKets = 64;

// A bra run: tiles [x, x + y) of one bra pair (the int2 the kernels read).
struct Run {
  int x, y;
};

// Work list of one class (X|Y), X >= Y pair classes: FP64 and FP32 tiles,
// the bra runs, and the counts of what the tiles cover.
struct ClassTiles {
  int X = 0, Y = 0;
  std::vector<Tile> t64, t32;
  std::vector<Run> g64, g32;    // runs of tiles with one bra pair
  int cap[2] = {0, 0};          // tiles per run (0 = no runs) for FP64 / FP32 tiles, | PQ_CAP_STAGE_D
  uint64_t nq64 = 0, nq32 = 0;  // shell quartets
  double np64 = 0, np32 = 0;    // primitive quartets
  double nf64 = 0, nf32 = 0;    // unique function quartets
};

// What the builder reads: P the pair copy whose primitive counts the kernels
// run (screened), P0 the unscreened pairs (ket groups keep the unscreened
// contraction lengths), S per pair class the pair ids in list order (by
// molecule, unscreened contraction length, then Q descending), Q the Schwarz
// bound of each pair, sh_mol the molecule of each shell.
struct WorkInput {
  const std::vector<PairInfo>& P;
  const std::vector<PairInfo>& P0;
  const std::vector<std::vector<int>>& S;
  const std::vector<double>& Q;
  const std::vector<int>& sh_mol;
  int nmol;
};

// Rotates the tiles [first, end) of bra number i by a bra-dependent offset.
inline void rotate_tiles(std::vector<Tile>& t, size_t first, size_t i) {
  const size_t cnt = t.size() - first;
  if (cnt < 2) return;
  const size_t r = (i * 2654435761u) % cnt;
  std::rotate(t.begin() + first, t.begin() + first + r, t.end());
}

// The tiles of class (X|Y): every (bra, ket) with Q_bra Q_ket >= cut (all
// when cut = 0), bra at or after the ket within one class (X == Y), kets of
// the bra's own molecule; quartets with Q_bra Q_ket < fp32cut go to the FP32
// tiles.  split_prims: the J/K and gradient layout (bra primitive sums split
// to the budget, ket primitives spread over idle lanes); otherwise one tile
// per (bra, 64 kets) with every primitive, as the stored-integral kernels
// take them.
inline void build_class_tiles(const WorkInput& in, int X, int Y, double cut, double fp32cut, bool split_prims,
                              ClassTiles& cw) {
  cw = ClassTiles();
  cw.X = X;
  cw.Y = Y;
  const auto& P = in.P;
  const auto& bras = in.S[X];
  const auto& kets = in.S[Y];
  if (bras.empty() || kets.empty()) return;
  const int la = kPairL[X][0], lb = kPairL[X][1], lc = kPairL[Y][0], ld = kPairL[Y][1];
  const int tile = kTileKets;
  std::vector<double> kq(kets.size());
  std::vector<double> kpre(kets.size() + 1, 0.0), fpre(kets.size() + 1, 0.0);
  auto ufun = [&](int pid, int l1, int l2) -> double {  // unique function pairs of a shell pair
    const PairInfo& p = P[pid];
    return p.shA == p.shB ? ncart(l1) * (ncart(l1) + 1) / 2.0 : (double)ncart(l1) * ncart(l2);
  };
  for (size_t j = 0; j < kets.size(); ++j) {
    kq[j] = in.Q[kets[j]];
    kpre[j + 1] = kpre[j] + P[kets[j]].pp_n;
    fpre[j + 1] = fpre[j] + ufun(kets[j], lc, ld);
  }
  // ket groups: runs of one molecule and equal (unscreened) contraction
  // length, each sorted by Q descending; gmol[m] = first group of molecule m
  auto kmol = [&](int j) { return in.sh_mol[P[kets[j]].shA]; };
  std::vector<int> gs;
  for (size_t j = 0; j < kets.size(); ++j)
    if (j == 0 || in.P0[kets[j]].pp_n != in.P0[kets[j - 1]].pp_n || kmol(j) != kmol(j - 1))
      gs.push_back((int)j);
  gs.push_back((int)kets.size());
  std::vector<int> gmol(in.nmol + 1, (int)gs.size() - 1);
  for (int g = (int)gs.size() - 2; g >= 0; --g) gmol[kmol(gs[g])] = g;
  for (int m = in.nmol - 1; m >= 0; --m)
    if (gmol[m] > gmol[m + 1]) gmol[m] = gmol[m + 1];
  auto count_ge = [&](int g0, int g1, double thr) -> int {  // entries of [g0,g1) with Q >= thr
    return (int)(std::upper_bound(kq.begin() + g0, kq.begin() + g1, thr,
                                  [](double t, double q) { return t > q; }) -
                 (kq.begin() + g0));
  };
  for (size_t i = 0; i < bras.size(); ++i) {
    const double qi = in.Q[bras[i]];
    const int bpn = P[bras[i]].pp_n;
    const double ua = ufun(bras[i], la, lb);
    const int bm = in.sh_mol[P[bras[i]].shA];
    const size_t t64_0 = cw.t64.size(), t32_0 = cw.t32.size();
    for (int g = gmol[bm]; g < gmol[bm + 1]; ++g) {
      const int g0 = gs[g], g1 = gs[g + 1];
      if (X == Y && g0 > (int)i) break;
      int c;
      if (cut <= 0.0) c = g1 - g0;
      else if (qi <= 0.0) c = 0;
      else c = count_ge(g0, g1, cut / qi);
      int end = g0 + c;
      if (X == Y) end = std::min(end, (int)i + 1);
      int e64 = end;
      if (fp32cut > 0.0) e64 = qi <= 0.0 ? g0 : std::min(end, g0 + count_ge(g0, g1, fp32cut / qi));
      // split the bra primitive sum when the quartets are heavily contracted
      const int kpn = P[kets[g0]].pp_n;
      int chunk = bpn;
      if (split_prims) chunk = std::max(1, std::min(bpn, kPrimBudget / std::max(1, kpn)));
      // a tile with fewer than 64 kets spreads each ket's primitive pairs
      // over ks lanes (ks a power of two, ks * kets <= 64), whose blocks are
      // summed by shuffles before digestion; (ss|ss) has too little work per
      // lane for that to pay (measured)
      const bool ksplit = split_prims && la + lb + lc + ld > 0;
      auto nket_field = [&](int cnt) {
        int lg = 0;
        while (ksplit && lg < 6 && (cnt << (lg + 1)) <= tile && (1 << (lg + 1)) <= kpn) ++lg;
        return cnt | (lg << 8);
      };
      for (int p0 = 0; p0 < bpn; p0 += chunk) {
        const int np = std::min(chunk, bpn - p0);
        const int prim = np == bpn ? 0 : (p0 | (np << 16));
        for (int k = g0; k < e64; k += tile) {
          cw.t64.push_back(Tile{bras[i], k, nket_field(std::min(tile, e64 - k)), prim});
        }
        for (int k = e64; k < end; k += tile) {
          cw.t32.push_back(Tile{bras[i], k, nket_field(std::min(tile, end - k)), prim});
        }
      }
      cw.nq64 += e64 - g0;
      cw.nq32 += end - e64;
      cw.np64 += bpn * (kpre[e64] - kpre[g0]);
      cw.np32 += bpn * (kpre[end] - kpre[e64]);
      // unique function quartets; the diagonal quartet (bra == ket) has
      // ua(ua+1)/2 instead of ua^2
      double f64 = ua * (fpre[e64] - fpre[g0]), f32 = ua * (fpre[end] - fpre[e64]);
      if (X == Y && end == (int)i + 1 && (int)i >= g0) {
        const double corr = ua * ua - ua * (ua + 1) / 2.0;
        if ((int)i < e64) f64 -= corr;
        else f32 -= corr;
      }
      cw.nf64 += f64;
      cw.nf32 += f32;
    }
    // start each bra's tiles at a different ket tile: blocks that run at the
    // same time then add into different J[c,d] entries instead of all
    // contending for the atomics of the first kets
    rotate_tiles(cw.t64, t64_0, i);
    rotate_tiles(cw.t32, t32_0, i);
  }
}

// The bra runs of one precision of a class for a run cap (0 = none; the
// PQ_CAP_STAGE_D flag is ignored here), in tile order.  (Ordering the runs
// heaviest first, by an estimate of their work, gained nothing on one
// molecule and cost 15% on the 512-molecule batch, whose concurrent blocks
// then touch many molecules' matrices at once.)
inline void make_runs(ClassTiles& cw, int prec, int cap) {
  const std::vector<Tile>& t = prec ? cw.t32 : cw.t64;
  std::vector<Run>& g = prec ? cw.g32 : cw.g64;
  g.clear();
  cw.cap[prec] = cap;
  const int run = cap > 0 ? cap & (PQ_CAP_STAGE_D - 1) : 0;
  if (run <= 0) return;
  for (int i = 0; i < (int)t.size(); ++i) {
    if (g.empty() || t[i].bra != t[g.back().x].bra || g.back().y >= run) g.push_back(Run{i, 0});
    g.back().y += 1;
  }
}

}  // namespace pq
