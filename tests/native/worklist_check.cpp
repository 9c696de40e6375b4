// Host harness over pyquante_amd/csrc/pq_worklist.h for the CPU tests, built
// with the address and undefined-behaviour sanitizers: reads a pair set
// (binary file written by tests/test_worklist.py), builds the work lists of
// every class as the engine does, and checks them against what they must
// cover.  Prints one line per class and the failures; exit status 0 when
// every check passes.
//
// Input (little-endian): int32 npair, nshell, nmol; per pair int32 class,
// shA, shB, pp_n (screened), pp_n (unscreened); int32 sh_mol[nshell];
// float64 Q[npair]; per pair class c = 0..5: int32 n, then n pair ids (the
// list order); float64 cut, fp32cut; int32 split_prims, cap64, cap32.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../pyquante_amd/csrc/pq_worklist.h"

using namespace pq;

namespace {

int g_fail = 0;

bool check(bool ok, const char* what, int X, int Y, long a = -1, long b = -1) {
  if (ok) return true;
  if (g_fail < 50) std::printf("FAIL class (%d,%d): %s [%ld %ld]\n", X, Y, what, a, b);
  ++g_fail;
  return false;
}

template <typename T>
bool rd(FILE* f, T* p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: worklist_check input.bin\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[3];
  if (!rd(f, hdr, 3)) return 2;
  const int npair = hdr[0], nshell = hdr[1], nmol = hdr[2];
  std::vector<PairInfo> P(npair), P0(npair);
  std::vector<int> pcls(npair);
  for (int i = 0; i < npair; ++i) {
    int32_t r[5];
    if (!rd(f, r, 5)) return 2;
    std::memset(&P[i], 0, sizeof(PairInfo));
    pcls[i] = r[0];
    P[i].shA = r[1];
    P[i].shB = r[2];
    P[i].pp_n = r[3];
    P0[i] = P[i];
    P0[i].pp_n = r[4];
  }
  std::vector<int> sh_mol(nshell);
  std::vector<double> Q(npair);
  if (!rd(f, sh_mol.data(), nshell) || !rd(f, Q.data(), npair)) return 2;
  std::vector<std::vector<int>> S(6);
  for (int c = 0; c < 6; ++c) {
    int32_t n;
    if (!rd(f, &n, 1)) return 2;
    S[c].resize(n);
    if (n && !rd(f, S[c].data(), n)) return 2;
  }
  double cuts[2];
  int32_t opt[3];
  if (!rd(f, cuts, 2) || !rd(f, opt, 3)) return 2;
  std::fclose(f);
  const double cut = cuts[0], fp32cut = cuts[1];
  const bool split = opt[0] != 0;
  const int caps[2] = {opt[1], opt[2]};

  const WorkInput in{P, P0, S, Q, sh_mol, nmol};
  for (int X = 0; X < 6; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      ClassTiles cw;
      build_class_tiles(in, X, Y, cut, fp32cut, split, cw);
      for (int prec = 0; prec < 2; ++prec) make_runs(cw, prec, caps[prec]);
      const auto& bras = S[X];
      const auto& kets = S[Y];
      const size_t nb = bras.size(), nk = kets.size();
      std::vector<int> bpos(npair, -1);
      for (size_t i = 0; i < nb; ++i) bpos[bras[i]] = (int)i;
      // bits of the bra primitive range covered, per (bra, ket) and bucket
      std::vector<uint64_t> mask(nb * nk, 0);
      std::vector<signed char> bucket(nb * nk, -1);
      uint64_t nq[2] = {0, 0};
      double np[2] = {0, 0};
      for (int prec = 0; prec < 2; ++prec) {
        const auto& t = prec ? cw.t32 : cw.t64;
        for (size_t ti = 0; ti < t.size(); ++ti) {
          const Tile& tl = t[ti];
          const int nkt = tl.nket & 0xff, lg = (tl.nket >> 8) & 7;
          bool ok = check(nkt >= 1 && nkt <= kTileKets, "kets per tile in 1..64", X, Y, (long)ti, nkt);
          ok = check((nkt << lg) <= kTileKets, "kets x ket-primitive split <= 64", X, Y, (long)ti, lg) && ok;
          ok = check(tl.bra >= 0 && tl.bra < npair && bpos[tl.bra] >= 0, "bra in the bra list", X, Y, (long)ti) && ok;
          ok = check(tl.ket0 >= 0 && tl.ket0 + nkt <= (int)nk, "kets inside the ket list", X, Y, (long)ti) && ok;
          if (!ok) continue;
          const int bi = bpos[tl.bra];
          const int bpn = P[tl.bra].pp_n;
          const int i0 = tl.prim & 0xffff, ni = tl.prim >> 16;
          const int p0 = tl.prim ? i0 : 0, pn = tl.prim ? ni : bpn;
          check(pn >= 1 && p0 + pn <= bpn, "bra primitive range inside the pair", X, Y, (long)ti, tl.prim);
          check(!split ? tl.prim == 0 && lg == 0 : true, "stored-integral tiles take every primitive", X, Y);
          const uint64_t bits = (pn >= 64 ? ~0ull : ((1ull << pn) - 1)) << p0;
          for (int k = tl.ket0; k < tl.ket0 + nkt; ++k) {
            check(P0[kets[k]].pp_n == P0[kets[tl.ket0]].pp_n, "one contraction length per tile", X, Y, (long)ti, k);
            check(sh_mol[P[kets[k]].shA] == sh_mol[P[tl.bra].shA], "bra and ket of one molecule", X, Y, (long)ti, k);
            const size_t idx = (size_t)bi * nk + k;
            check((mask[idx] & bits) == 0, "primitive ranges do not overlap", X, Y, bi, k);
            if (mask[idx] == 0) {
              bucket[idx] = (signed char)prec;
              ++nq[prec];
              np[prec] += (double)bpn * P[kets[k]].pp_n;
            }
            check(bucket[idx] == prec, "one bucket per quartet", X, Y, bi, k);
            mask[idx] |= bits;
          }
        }
        // bra runs: a partition of the tiles, in order, into runs of one
        // bra at most the cap long
        const int run = caps[prec] > 0 ? caps[prec] & (PQ_CAP_STAGE_D - 1) : 0;
        const auto& g = prec ? cw.g32 : cw.g64;
        if (run <= 0) {
          check(g.empty(), "no runs without a cap", X, Y);
          continue;
        }
        std::vector<int> seen(t.size(), 0);
        int next = 0;
        for (size_t r = 0; r < g.size(); ++r) {
          if (!check(g[r].y >= 1 && g[r].y <= run && g[r].x >= 0 && g[r].x + g[r].y <= (int)t.size(),
                     "run inside the tile list and at most the cap long", X, Y, (long)r))
            break;
          check(g[r].x == next, "runs in tile order", X, Y, (long)r, g[r].x);
          next = g[r].x + g[r].y;
          for (int i = g[r].x; i < g[r].x + g[r].y; ++i) {
            ++seen[i];
            check(t[i].bra == t[g[r].x].bra, "one bra per run", X, Y, (long)r, i);
          }
        }
        for (size_t i = 0; i < t.size(); ++i) check(seen[i] == 1, "every tile in one run", X, Y, (long)i);
      }
      // completeness: every (bra, ket) the screening keeps is covered once,
      // whole, in the bucket its bound selects
      for (size_t bi = 0; bi < nb; ++bi) {
        const int b = bras[bi];
        const double qb = Q[b];
        const int bpn = P[b].pp_n;
        const uint64_t full = bpn >= 64 ? ~0ull : (1ull << bpn) - 1;
        for (size_t k = 0; k < nk; ++k) {
          const int kp = kets[k];
          bool keep = sh_mol[P[kp].shA] == sh_mol[P[b].shA] && (X != Y || k <= bi);
          if (keep && cut > 0.0) keep = qb > 0.0 && Q[kp] >= cut / qb;
          const size_t idx = bi * nk + k;
          check(keep ? mask[idx] == full : mask[idx] == 0, keep ? "kept quartet covered whole" : "screened quartet absent",
                X, Y, (long)bi, (long)k);
          if (keep && fp32cut > 0.0) {
            const int want = (qb > 0.0 && Q[kp] >= fp32cut / qb) ? 0 : 1;
            check(bucket[idx] == want, "bucket from the bound", X, Y, (long)bi, (long)k);
          }
        }
      }
      check(nq[0] == cw.nq64 && nq[1] == cw.nq32, "shell-quartet counts", X, Y, (long)nq[0], (long)cw.nq64);
      check(np[0] == cw.np64 && np[1] == cw.np32, "primitive-quartet counts", X, Y);
      std::printf("class %d %d quartets %llu %llu tiles %zu %zu runs %zu %zu\n", X, Y, (unsigned long long)cw.nq64,
                  (unsigned long long)cw.nq32, cw.t64.size(), cw.t32.size(), cw.g64.size(), cw.g32.size());
    }
  std::printf("failures %d\n", g_fail);
  return g_fail ? 1 : 0;
}
