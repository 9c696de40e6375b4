// pq_engine.hip -- host side of libpqhip.so: the engine handle (basis,
// shell-pair data in HBM, Schwarz-sorted work lists) and the C ABI declared in
// include/pq_abi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/pq_abi.h"
#include "pq_device.h"
#include "pq_kernels.h"
#include "pq_worklist.h"

using namespace pq;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define PQ_HIP(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(PQ_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
  } while (0)

template <typename T>
hipError_t dev_alloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) return hipSuccess;
  return hipMalloc((void**)p, n * sizeof(T));
}

template <typename T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int pair_class(int l1, int l2) { return l1 * (l1 + 1) / 2 + l2; }  // l1 >= l2

double fact2(int n) { return n <= 1 ? 1.0 : n * fact2(n - 2); }

// radial part of the primitive norm, primitive-gto.c:76-83 without the
// per-component double factorials
double nrad(double a, int L) {
  return std::sqrt(std::pow(2.0, 2 * L + 1.5) * std::pow(a, L + 1.5) / std::pow(M_PI, 1.5));
}

// Rys operation count of one primitive quartet of class (la lb|lc ld) and of
// digesting one contracted quartet (FMA = 2 flops).
double rys_flops_prim(int la, int lb, int lc, int ld) {
  const int n = la + lb, m = lc + ld, nr = (n + m) / 2 + 1;
  const int Nc = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
  double sb = 0, sd = 0;
  for (int j = 1; j <= lb; ++j) sb += n - j + 1;
  for (int l = 1; l <= ld; ++l) sd += m - l + 1;
  return 30.0 + nr * (60.0 + 3.0 * (4.0 * (n + 1) * (m + 1) + 2.0 * (m + 1) * sb +
                                    2.0 * (la + 1) * (lb + 1) * sd) + 4.0 * Nc) + 2.0 * Nc;
}
double rys_flops_dig(int la, int lb, int lc, int ld) {
  return 12.0 * ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
}

// A class work list with its device copies (tiles and bra runs).
struct ClassWork : ClassTiles {
  Tile* d_t64 = nullptr;
  Tile* d_t32 = nullptr;
  int2* d_g64 = nullptr;
  int2* d_g32 = nullptr;
  int gflags(int prec) const { return (cap[prec] & PQ_CAP_STAGE_D) ? kGroupStageD : 0; }
};
static_assert(sizeof(Run) == sizeof(int2), "bra runs are uploaded as int2");

}  // namespace

struct pq_engine {
  int nshell = 0, natom = 0, nbf = 0;
  std::vector<int> sh_l, sh_np, sh_off, sh_atom, sh_bf;
  std::vector<double> exps, coefs, atom_xyz, atom_z;

  // pair data
  bool pairs_ok = false;
  std::vector<PairInfo> pairs;
  std::vector<PrimPair> prims;
  std::vector<double> prim_a;
  std::vector<int> pcls;                 // pair class of each pair
  std::vector<std::vector<int>> members;  // per pair class: pair ids (unsorted)
  std::vector<double> Q;                 // Schwarz bound per pair
  std::vector<std::vector<int>> sorted;   // per pair class: sorted by Q desc
  PairInfo* d_pairs = nullptr;
  PrimPair* d_prims = nullptr;
  double* d_prim_a = nullptr;
  int* d_members[6] = {};
  int* d_sorted[6] = {};
  // primitive screening: Schwarz factor of every primitive pair (each pair's
  // primitives are stored in decreasing order of it), and the J/K and
  // gradient copy of the pairs whose pp_n counts only the primitives that
  // survive the current cut, with its own per-class ket order
  std::vector<double> qprim;
  double qprim_max = 0;
  std::vector<PairInfo> pairs_s;
  std::vector<std::vector<int>> sorted_s;
  // each dropped primitive quartet is bounded by prim_rel * cut, below the
  // quartets the shell-level cut drops
  double prim_rel = 1e-4;
  PairInfo* d_pairs_s = nullptr;
  int* d_sorted_s[6] = {};
  // position-indexed ket J (EngineView::jpair), room for two densities
  double* d_jpair = nullptr;
  int64_t jp_size = 0;
  int jp_off[6] = {}, jp_n[6] = {};
  double* d_atom_xyz = nullptr;
  double* d_atom_z = nullptr;

  // work lists
  bool work_ok = false;
  double cut = 0, fp32cut = 0;
  std::vector<ClassWork> work;
  std::vector<ClassWork> store_work;  // unscreened, for get2ints
  bool store_ok = false;
  std::vector<ClassWork> hgp_work;  // unscreened, primitive splits, for the HGP get2ints
  bool hgp_ok = false;
  double* d_Jacc = nullptr;  // 2 n^2: room for two densities
  double* d_Kacc = nullptr;
  double* d_D2 = nullptr;    // two stacked densities (allocated on first use)
  double* d_gacc = nullptr;  // private gradient copies
  size_t gacc_n = 0;

  // side streams (default priority) for running class kernels concurrently
  static constexpr int kSide = 4;
  hipStream_t side[kSide] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kSide] = {};
  bool streams_ok = false;
  int nside = kSide;  // streams in use (0 = caller's stream only)
  bool fused = false;  // one launch for the light classes' bra runs (pq_engine_set_fused)
  // deterministic J/K (pq_engine_set_deterministic): fixed-point limbs of J,
  // K and the ket J, 3 words per element each (allocated on first use)
  bool det = false;
  double* d_fx = nullptr;
  size_t fx_cap = 0;

  // tiles per bra run for each (bra class, ket class, precision): -1 until
  // set by pq_engine_autotune or pq_engine_set_group_cap
  int group_cap[6][6][2];
  // bumped whenever device arrays that a captured build reads (pair data,
  // tiles, bra runs, the ket J) are replaced or freed
  uint64_t gen = 1;
  pq_engine() {
    for (auto& a : group_cap)
      for (auto& b : a) b[0] = b[1] = -1;
  }

  // molecules: atoms and shells of molecule m are contiguous; matrices are
  // stored as per-molecule blocks (molecule m: mol_nbf[m]^2 elements from
  // mol_moff[m]); one molecule (the usual case) is the plain n x n layout
  int nmol = 1;
  std::vector<int> atom_mol, sh_mol, mol_nbf, mol_fbase, mol_at0, mol_nat, mol_sh0;
  std::vector<int64_t> mol_moff;
  int64_t msize = 0;  // elements of one packed matrix
  int maxn = 0;       // largest molecule's function count
  int* d_mol_moff = nullptr;
  int* d_mol_n = nullptr;
  int* d_atom_mol = nullptr;
  int* d_mol_at0 = nullptr;  // first atom and atom count of each molecule
  int* d_mol_nat = nullptr;

  int maxat = 0;      // largest molecule's atom count
  int lds_dev = 65536;  // the device's LDS per workgroup, bytes
  int lds_max = 65536;  // the J/K launches' limit (lds_dev unless lowered)
  EngineView view() const {
    EngineView v{d_pairs, d_prims, d_prim_a, maxn, msize, maxat};
    v.lds_max = lds_max;
    return v;
  }
  // pairs with the screened primitive counts and the ket J (J/K and
  // gradient kernels)
  EngineView view_s() const {
    EngineView v{d_pairs_s, d_prims, d_prim_a, maxn, msize, maxat};
    v.lds_max = lds_max;
    v.jpair = d_jpair;
    v.jp_size = jp_size;
    for (int y = 0; y < 6; ++y) {
      v.jp_off[y] = jp_off[y];
      v.jp_n[y] = jp_n[y];
    }
    return v;
  }

  // per-function primitive data for the HGP path (built on first use, dropped
  // with the geometry)
  int* d_fnp = nullptr;
  int* d_foff = nullptr;
  int* d_fpw = nullptr;
  double* d_forg = nullptr;
  double* d_fexps = nullptr;
  double* d_fcoefs = nullptr;
  double* d_fnorms = nullptr;
  void free_functions() {
    dev_free(d_fnp);
    dev_free(d_foff);
    dev_free(d_fpw);
    dev_free(d_forg);
    dev_free(d_fexps);
    dev_free(d_fcoefs);
    dev_free(d_fnorms);
  }

  void free_work(std::vector<ClassWork>& w) {
    for (auto& c : w) {
      dev_free(c.d_t64);
      dev_free(c.d_t32);
      dev_free(c.d_g64);
      dev_free(c.d_g32);
    }
    w.clear();
  }
  void free_pairs() {
    ++gen;
    dev_free(d_pairs);
    dev_free(d_prims);
    dev_free(d_prim_a);
    dev_free(d_pairs_s);
    dev_free(d_jpair);
    jp_size = 0;
    for (int i = 0; i < 6; ++i) {
      dev_free(d_members[i]);
      dev_free(d_sorted[i]);
      dev_free(d_sorted_s[i]);
    }
    free_work(work);
    free_work(store_work);
    free_work(hgp_work);
    pairs_ok = work_ok = store_ok = hgp_ok = false;
  }
  ~pq_engine() {
    if (streams_ok) {
      for (int i = 0; i < kSide; ++i) {
        (void)hipStreamDestroy(side[i]);
        (void)hipEventDestroy(join_ev[i]);
      }
      (void)hipEventDestroy(fork_ev);
    }
    free_pairs();
    free_functions();
    dev_free(d_atom_xyz);
    dev_free(d_atom_z);
    dev_free(d_Jacc);
    dev_free(d_Kacc);
    dev_free(d_fx);
    dev_free(d_D2);
    dev_free(d_gacc);
    dev_free(d_mol_moff);
    dev_free(d_mol_n);
    dev_free(d_atom_mol);
    dev_free(d_mol_at0);
    dev_free(d_mol_nat);
  }
};

namespace {

// Schwarz factors max over components sqrt|(ab|ab)| of the pairs ids[c] (per
// pair class c, indices into the device pair array d_list of n entries)
// into q (n entries, 0 for pairs not listed).
int diag_factors(pq_engine* e, const PairInfo* d_list, size_t n, const std::vector<std::vector<int>>& ids,
                 std::vector<double>& q) {
  EngineView v = e->view();
  v.pairs = d_list;
  double* d_q = nullptr;
  PQ_HIP(dev_alloc(&d_q, n));
  PQ_HIP(launch_fill(d_q, (int64_t)n, 0.0, 0));
  int rc = PQ_OK;
  for (int c = 0; c < 6 && rc == PQ_OK; ++c) {
    const int np = (int)ids[c].size();
    if (!np) continue;
    std::vector<Tile> tiles;
    for (int k = 0; k < np; k += 64) tiles.push_back(Tile{-1, k, std::min(64, np - k), 0});
    Tile* d_tiles = nullptr;
    int* d_ids = nullptr;
    hipError_t err = dev_alloc(&d_tiles, tiles.size());
    if (err == hipSuccess) err = dev_alloc(&d_ids, (size_t)np);
    if (err == hipSuccess)
      err = hipMemcpy(d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(d_ids, ids[c].data(), np * sizeof(int), hipMemcpyHostToDevice);
    if (err == hipSuccess)
      err = launch_eri_class(kPairL[c][0], kPairL[c][1], kPairL[c][0], kPairL[c][1], MODE_DIAG, 0, v, d_tiles,
                             (int)tiles.size(), d_ids, nullptr, nullptr, nullptr, nullptr, d_q, 0);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    dev_free(d_tiles);
    dev_free(d_ids);
    if (err != hipSuccess) rc = fail(PQ_E_HIP, hipGetErrorString(err));
  }
  q.assign(n, 0.0);
  if (rc == PQ_OK) {
    hipError_t err = hipMemcpy(q.data(), d_q, n * sizeof(double), hipMemcpyDeviceToHost);
    if (err != hipSuccess) rc = fail(PQ_E_HIP, hipGetErrorString(err));
  }
  dev_free(d_q);
  return rc;
}

// Per pair class, the pair ids sorted by molecule (a bra's kets are its own
// molecule's), then contraction length in P (uniform primitive loops within a
// tile), then Q descending; uploaded to d_sorted.
int sort_pairs(pq_engine* e, const std::vector<PairInfo>& P, std::vector<std::vector<int>>& sorted,
               int* (&d_sorted)[6]) {
  sorted.assign(6, {});
  for (int c = 0; c < 6; ++c) {
    auto v = e->members[c];
    std::stable_sort(v.begin(), v.end(), [&](int a, int b) {
      const int ma = e->sh_mol[P[a].shA], mb = e->sh_mol[P[b].shA];
      if (ma != mb) return ma < mb;
      if (P[a].pp_n != P[b].pp_n) return P[a].pp_n > P[b].pp_n;
      return e->Q[a] > e->Q[b];
    });
    sorted[c] = v;
    dev_free(d_sorted[c]);
    PQ_HIP(dev_alloc(&d_sorted[c], v.size()));
    if (!v.empty()) PQ_HIP(hipMemcpy(d_sorted[c], v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  return PQ_OK;
}

// The screened pair copy for a shell-quartet cut: primitive pairs are
// dropped from the J/K and gradient work when their Schwarz factor times the
// largest one is below e->prim_rel * cut (cut = 0 or prim_rel = 0 keeps every
// primitive); pp_n = the primitives that pass, at least one.
int build_screened(pq_engine* e, double cut) {
  ++e->gen;
  const double tau = cut > 0.0 ? cut * e->prim_rel : 0.0;
  e->pairs_s = e->pairs;
  if (tau > 0.0)
    for (auto& p : e->pairs_s) {
      int n = 1;
      while (n < p.pp_n && e->qprim[p.pp_off + n] * e->qprim_max >= tau) ++n;
      p.pp_n = n;
    }
  dev_free(e->d_pairs_s);
  PQ_HIP(dev_alloc(&e->d_pairs_s, e->pairs_s.size()));
  PQ_HIP(hipMemcpy(e->d_pairs_s, e->pairs_s.data(), e->pairs_s.size() * sizeof(PairInfo), hipMemcpyHostToDevice));
  // ket groups keep the unscreened contraction lengths: grouping by the
  // screened ones splits them into many short groups, and the extra partly
  // filled tiles cost more than the skipped primitives save (measured)
  int rc = sort_pairs(e, e->pairs, e->sorted_s, e->d_sorted_s);
  if (rc) return rc;
  // the ket J by ket position: class Y holds ncart(l1) * ncart(l2) components
  // for each of its pairs
  int64_t off = 0;
  for (int y = 0; y < 6; ++y) {
    e->jp_off[y] = (int)off;
    e->jp_n[y] = (int)e->sorted_s[y].size();
    off += (int64_t)e->jp_n[y] * ncart(kPairL[y][0]) * ncart(kPairL[y][1]);
  }
  if (off > INT32_MAX / 2) return fail(PQ_E_ARG, "too many shell pairs for the ket J array");
  dev_free(e->d_jpair);
  e->jp_size = off;
  PQ_HIP(dev_alloc(&e->d_jpair, (size_t)std::max<int64_t>(2 * off, 1)));
  return PQ_OK;
}

int build_pair_data(pq_engine* e) {
  e->pairs.clear();
  e->prims.clear();
  e->prim_a.clear();
  e->pcls.clear();
  e->members.assign(6, {});
  for (int A = 0; A < e->nshell; ++A)
    for (int B = e->mol_sh0[e->sh_mol[A]]; B <= A; ++B) {  // pairs within A's molecule
      const int m = e->sh_mol[A];
      int s1 = A, s2 = B;
      if (e->sh_l[B] > e->sh_l[A]) std::swap(s1, s2);
      const double* c1 = &e->atom_xyz[3 * e->sh_atom[s1]];
      const double* c2 = &e->atom_xyz[3 * e->sh_atom[s2]];
      PairInfo pi{};
      pi.Ax = c1[0];
      pi.Ay = c1[1];
      pi.Az = c1[2];
      pi.ABx = c1[0] - c2[0];
      pi.ABy = c1[1] - c2[1];
      pi.ABz = c1[2] - c2[2];
      pi.pp_off = (int)e->prims.size();
      pi.bfA = e->sh_bf[s1] - e->mol_fbase[m];
      pi.bfB = e->sh_bf[s2] - e->mol_fbase[m];
      pi.shA = s1;
      pi.shB = s2;
      pi.atA = e->sh_atom[s1];
      pi.atB = e->sh_atom[s2];
      pi.moff = (int)e->mol_moff[m];
      pi.mstride = e->mol_nbf[m];
      pi.mat0 = e->mol_at0[m];
      pi.mnat = e->mol_nat[m];
      const double ab2 = pi.ABx * pi.ABx + pi.ABy * pi.ABy + pi.ABz * pi.ABz;
      const int l1 = e->sh_l[s1], l2 = e->sh_l[s2];
      for (int i = 0; i < e->sh_np[s1]; ++i)
        for (int j = 0; j < e->sh_np[s2]; ++j) {
          const double a = e->exps[e->sh_off[s1] + i], b = e->exps[e->sh_off[s2] + j];
          const double ca = e->coefs[e->sh_off[s1] + i], cb = e->coefs[e->sh_off[s2] + j];
          PrimPair pp{};
          pp.p = a + b;
          pp.rp = 1.0 / pp.p;
          pp.Px = (a * c1[0] + b * c2[0]) / pp.p;
          pp.Py = (a * c1[1] + b * c2[1]) / pp.p;
          pp.Pz = (a * c1[2] + b * c2[2]) / pp.p;
          pp.Kp = ca * cb * nrad(a, l1) * nrad(b, l2) * std::exp(-a * b / pp.p * ab2) / pp.p;
          e->prims.push_back(pp);
          e->prim_a.push_back(a);
        }
      pi.pp_n = (int)e->prims.size() - pi.pp_off;
      const int pc = pair_class(l1, l2);
      e->members[pc].push_back((int)e->pairs.size());
      e->pcls.push_back(pc);
      e->pairs.push_back(pi);
    }
  dev_free(e->d_pairs);
  dev_free(e->d_prims);
  dev_free(e->d_prim_a);
  PQ_HIP(dev_alloc(&e->d_pairs, e->pairs.size()));
  PQ_HIP(dev_alloc(&e->d_prims, e->prims.size()));
  PQ_HIP(dev_alloc(&e->d_prim_a, e->prim_a.size()));
  PQ_HIP(hipMemcpy(e->d_prim_a, e->prim_a.data(), e->prim_a.size() * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(e->d_pairs, e->pairs.data(), e->pairs.size() * sizeof(PairInfo), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(e->d_prims, e->prims.data(), e->prims.size() * sizeof(PrimPair), hipMemcpyHostToDevice));
  for (int c = 0; c < 6; ++c) {
    dev_free(e->d_members[c]);
    PQ_HIP(dev_alloc(&e->d_members[c], e->members[c].size()));
    if (!e->members[c].empty())
      PQ_HIP(hipMemcpy(e->d_members[c], e->members[c].data(), e->members[c].size() * sizeof(int),
                       hipMemcpyHostToDevice));
  }
  // Schwarz bounds from the diagonal quartets (ab|ab), of the contracted
  // pairs and of every primitive pair
  {
    int rc = diag_factors(e, e->d_pairs, e->pairs.size(), e->members, e->Q);
    if (rc) return rc;
    std::vector<PairInfo> plist;
    std::vector<std::vector<int>> pids(6);
    plist.reserve(e->prims.size());
    for (size_t i = 0; i < e->pairs.size(); ++i)
      for (int k = 0; k < e->pairs[i].pp_n; ++k) {
        PairInfo pk = e->pairs[i];
        pk.pp_off += k;
        pk.pp_n = 1;
        pids[e->pcls[i]].push_back((int)plist.size());
        plist.push_back(pk);
      }
    PairInfo* d_plist = nullptr;
    PQ_HIP(dev_alloc(&d_plist, plist.size()));
    PQ_HIP(hipMemcpy(d_plist, plist.data(), plist.size() * sizeof(PairInfo), hipMemcpyHostToDevice));
    rc = diag_factors(e, d_plist, plist.size(), pids, e->qprim);
    dev_free(d_plist);
    if (rc) return rc;
  }
  // each pair's primitives in decreasing order of their Schwarz factor, so
  // that the ones a cut keeps are a prefix
  e->qprim_max = 0;
  for (const auto& p : e->pairs) {
    std::vector<int> perm(p.pp_n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int a, int b) { return e->qprim[p.pp_off + a] > e->qprim[p.pp_off + b]; });
    std::vector<PrimPair> pp(p.pp_n);
    std::vector<double> pa(p.pp_n), pq(p.pp_n);
    for (int k = 0; k < p.pp_n; ++k) {
      pp[k] = e->prims[p.pp_off + perm[k]];
      pa[k] = e->prim_a[p.pp_off + perm[k]];
      pq[k] = e->qprim[p.pp_off + perm[k]];
    }
    for (int k = 0; k < p.pp_n; ++k) {
      e->prims[p.pp_off + k] = pp[k];
      e->prim_a[p.pp_off + k] = pa[k];
      e->qprim[p.pp_off + k] = pq[k];
    }
    if (p.pp_n) e->qprim_max = std::max(e->qprim_max, pq[0]);
  }
  PQ_HIP(hipMemcpy(e->d_prim_a, e->prim_a.data(), e->prim_a.size() * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(e->d_prims, e->prims.data(), e->prims.size() * sizeof(PrimPair), hipMemcpyHostToDevice));
  {
    int rc = sort_pairs(e, e->pairs, e->sorted, e->d_sorted);
    if (rc) return rc;
  }
  e->pairs_ok = true;
  return PQ_OK;
}

// Default tiles per bra run when no tuned value is set: a run saves
// (tiles - 1) flushes of the bra K rows, which pays off when the rows are long.
int default_group_cap(int nbf) { return nbf >= 300 ? 16 : 0; }

// Rebuilds the bra runs of one precision of a class for a new cap (0 = none)
// and uploads them.
int set_groups(ClassWork& cw, int prec, int cap) {
  int2*& d = prec ? cw.d_g32 : cw.d_g64;
  dev_free(d);
  make_runs(cw, prec, cap);
  const std::vector<Run>& g = prec ? cw.g32 : cw.g64;
  PQ_HIP(dev_alloc(&d, g.size()));
  if (!g.empty()) PQ_HIP(hipMemcpy(d, g.data(), g.size() * sizeof(int2), hipMemcpyHostToDevice));
  return PQ_OK;
}

// Work lists over the pair copy P (its pp_n) and its per-class order S.
int build_work(pq_engine* e, std::vector<ClassWork>& out, double cut, double fp32cut, bool split_prims,
               const std::vector<PairInfo>& P, const std::vector<std::vector<int>>& S) {
  if (&out == &e->work) ++e->gen;  // the J/K lists (the stored-integral lists are not captured)
  for (auto& c : out) {
    dev_free(c.d_t64);
    dev_free(c.d_t32);
    dev_free(c.d_g64);
    dev_free(c.d_g32);
  }
  out.clear();
  const WorkInput in{P, e->pairs, S, e->Q, e->sh_mol, e->nmol};
  for (int X = 0; X < 6; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      ClassWork cw;
      build_class_tiles(in, X, Y, cut, fp32cut, split_prims, cw);
      for (int prec = 0; prec < 2 && split_prims; ++prec) {
        const int tuned = e->group_cap[X][Y][prec];
        const int rc = set_groups(cw, prec, tuned >= 0 ? tuned : default_group_cap(e->maxn));
        if (rc) return rc;
      }
      PQ_HIP(dev_alloc(&cw.d_t64, cw.t64.size()));
      PQ_HIP(dev_alloc(&cw.d_t32, cw.t32.size()));
      if (!cw.t64.empty())
        PQ_HIP(hipMemcpy(cw.d_t64, cw.t64.data(), cw.t64.size() * sizeof(Tile), hipMemcpyHostToDevice));
      if (!cw.t32.empty())
        PQ_HIP(hipMemcpy(cw.d_t32, cw.t32.data(), cw.t32.size() * sizeof(Tile), hipMemcpyHostToDevice));
      out.push_back(std::move(cw));
    }
  return PQ_OK;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

extern "C" {

const char* pq_last_error(void) { return g_err.c_str(); }
int pq_version(void) { return 100; }

int pq_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int pq_ijkl2intindex(int64_t i, int64_t j, int64_t k, int64_t l, uint64_t* out) {
  if (!out) return fail(PQ_E_ARG, "null output");
  if (i < 0 || j < 0 || k < 0 || l < 0) return fail(PQ_E_ARG, "negative index");
  *out = ijkl2intindex(i, j, k, l);
  return PQ_OK;
}

int pq_boys_batch(int n, const double* m_host, const double* x_host, int compat_nr, double* out_host) {
  if (n < 0 || (n > 0 && (!m_host || !x_host || !out_host))) return fail(PQ_E_ARG, "bad arguments");
  if (n == 0) return PQ_OK;
  for (int i = 0; i < n; ++i)
    if (!(m_host[i] >= 0.0) || !(x_host[i] >= 0.0)) return fail(PQ_E_ARG, "Boys function needs m, x >= 0");
  DevBuf bm, bx, bo;
  PQ_HIP(hipMalloc(&bm.p, n * sizeof(double)));
  PQ_HIP(hipMalloc(&bx.p, n * sizeof(double)));
  PQ_HIP(hipMalloc(&bo.p, n * sizeof(double)));
  PQ_HIP(hipMemcpy(bm.p, m_host, n * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(bx.p, x_host, n * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(launch_boys(n, (double*)bm.p, (double*)bx.p, compat_nr ? 3.0e-7 : 1e-17, (double*)bo.p, 0));
  PQ_HIP(hipMemcpy(out_host, bo.p, n * sizeof(double), hipMemcpyDeviceToHost));
  return PQ_OK;
}

int pq_rys_roots_batch(int n, int nr, const double* X_host, double* u_host, double* w_host) {
  if (n < 0 || (n > 0 && (!X_host || !u_host || !w_host))) return fail(PQ_E_ARG, "bad arguments");
  if (nr < 1 || nr > 5) return fail(PQ_E_LTOTAL, "Rys order must be 1..5");
  if (n == 0) return PQ_OK;
  for (int i = 0; i < n; ++i)
    if (!(X_host[i] >= 0.0)) return fail(PQ_E_ARG, "X must be >= 0");
  DevBuf bx, bu, bw;
  PQ_HIP(hipMalloc(&bx.p, n * sizeof(double)));
  PQ_HIP(hipMalloc(&bu.p, (size_t)n * nr * sizeof(double)));
  PQ_HIP(hipMalloc(&bw.p, (size_t)n * nr * sizeof(double)));
  PQ_HIP(hipMemcpy(bx.p, X_host, n * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(launch_rys_roots(n, nr, (double*)bx.p, (double*)bu.p, (double*)bw.p, 0));
  PQ_HIP(hipMemcpy(u_host, bu.p, (size_t)n * nr * sizeof(double), hipMemcpyDeviceToHost));
  PQ_HIP(hipMemcpy(w_host, bw.p, (size_t)n * nr * sizeof(double), hipMemcpyDeviceToHost));
  return PQ_OK;
}

}  // extern "C"

namespace {
// Checks and uploads one batch of contracted quartets (layout of
// pq_contr_coulomb_batch) and runs `launch` on it: the Rys and THO batches.
template <typename Launch>
int contr_batch(int n, const int* nprim_host, const int* poff_host, int nprim_total, const double* exps_host,
                const double* coefs_host, const double* norms_host, const double* origins_host,
                const int* powers_host, double* out_host, Launch launch) {
  if (n < 0 || nprim_total < 0) return fail(PQ_E_ARG, "negative size");
  if (n == 0) return PQ_OK;
  if (!nprim_host || !poff_host || !exps_host || !coefs_host || !norms_host || !origins_host || !powers_host ||
      !out_host)
    return fail(PQ_E_ARG, "null pointer");
  for (int q = 0; q < n; ++q) {
    int L = 0;
    for (int f = 0; f < 4; ++f) {
      const int np = nprim_host[4 * q + f], po = poff_host[4 * q + f];
      if (np < 1 || po < 0 || po + np > nprim_total) return fail(PQ_E_ARG, "primitive range out of bounds");
      for (int d = 0; d < 3; ++d) {
        const int pw = powers_host[12 * q + 3 * f + d];
        if (pw < 0) return fail(PQ_E_ARG, "negative power");
        L += pw;
      }
    }
    if (L > 9) return fail(PQ_E_LTOTAL, "L_total > 9 has no Rys rule here (reference: Root6 stub)");
  }
  DevBuf b_np, b_po, b_e, b_c, b_n, b_o, b_p, b_out;
  PQ_HIP(hipMalloc(&b_np.p, 4 * n * sizeof(int)));
  PQ_HIP(hipMalloc(&b_po.p, 4 * n * sizeof(int)));
  PQ_HIP(hipMalloc(&b_e.p, nprim_total * sizeof(double)));
  PQ_HIP(hipMalloc(&b_c.p, nprim_total * sizeof(double)));
  PQ_HIP(hipMalloc(&b_n.p, nprim_total * sizeof(double)));
  PQ_HIP(hipMalloc(&b_o.p, 12 * n * sizeof(double)));
  PQ_HIP(hipMalloc(&b_p.p, 12 * n * sizeof(int)));
  PQ_HIP(hipMalloc(&b_out.p, n * sizeof(double)));
  PQ_HIP(hipMemcpy(b_np.p, nprim_host, 4 * n * sizeof(int), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_po.p, poff_host, 4 * n * sizeof(int), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_e.p, exps_host, nprim_total * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_c.p, coefs_host, nprim_total * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_n.p, norms_host, nprim_total * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_o.p, origins_host, 12 * n * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_p.p, powers_host, 12 * n * sizeof(int), hipMemcpyHostToDevice));
  PQ_HIP(launch(n, (int*)b_np.p, (int*)b_po.p, (double*)b_e.p, (double*)b_c.p, (double*)b_n.p, (double*)b_o.p,
                (int*)b_p.p, (double*)b_out.p));
  PQ_HIP(hipMemcpy(out_host, b_out.p, n * sizeof(double), hipMemcpyDeviceToHost));
  return PQ_OK;
}
}  // namespace

extern "C" {

int pq_contr_coulomb_batch(int n, const int* nprim_host, const int* poff_host, int nprim_total,
                           const double* exps_host, const double* coefs_host, const double* norms_host,
                           const double* origins_host, const int* powers_host, double* out_host) {
  return contr_batch(n, nprim_host, poff_host, nprim_total, exps_host, coefs_host, norms_host, origins_host,
                     powers_host, out_host,
                     [](int n_, const int* np, const int* po, const double* e, const double* c, const double* nr,
                        const double* o, const int* pw, double* out) {
                       return launch_contr_eri(n_, np, po, e, c, nr, o, pw, out, 0);
                     });
}

int pq_tho_contr_coulomb_batch(int n, const int* nprim_host, const int* poff_host, int nprim_total,
                               const double* exps_host, const double* coefs_host, const double* norms_host,
                               const double* origins_host, const int* powers_host, int compat_nr, double* out_host) {
  return contr_batch(n, nprim_host, poff_host, nprim_total, exps_host, coefs_host, norms_host, origins_host,
                     powers_host, out_host,
                     [compat_nr](int n_, const int* np, const int* po, const double* e, const double* c,
                                 const double* nr, const double* o, const int* pw, double* out) {
                       return launch_tho_contr(n_, np, po, e, c, nr, o, pw, compat_nr, out, 0);
                     });
}

int pq_hgp_contr_coulomb_batch(int n, const int* nprim_host, const int* poff_host, int nprim_total,
                               const double* exps_host, const double* coefs_host, const double* norms_host,
                               const double* origins_host, const int* powers_host, const int* mext_host,
                               int compat_nr, double* out_host) {
  if (n < 0 || nprim_total < 0) return fail(PQ_E_ARG, "negative size");
  if (n == 0) return PQ_OK;
  if (!nprim_host || !poff_host || !exps_host || !coefs_host || !norms_host || !origins_host || !powers_host ||
      !out_host)
    return fail(PQ_E_ARG, "null pointer");
  for (int q = 0; q < n; ++q) {
    int pw[4][3];
    for (int f = 0; f < 4; ++f) {
      const int np = nprim_host[4 * q + f], po = poff_host[4 * q + f];
      if (np < 1 || po < 0 || po + np > nprim_total) return fail(PQ_E_ARG, "primitive range out of bounds");
      for (int d = 0; d < 3; ++d) {
        pw[f][d] = powers_host[12 * q + 3 * f + d];
        if (pw[f][d] < 0) return fail(PQ_E_ARG, "negative power");
      }
    }
    const int mext = mext_host ? mext_host[q] : 0;
    if (mext < 0) return fail(PQ_E_ARG, "negative auxiliary order");
    // the largest leaves are a+b and c+d: box sizes and total order
    int boxe = 1, boxf = 1, ltot = mext;
    for (int d = 0; d < 3; ++d) {
      boxe *= pw[0][d] + pw[1][d] + 1;
      boxf *= pw[2][d] + pw[3][d] + 1;
      ltot += pw[0][d] + pw[1][d] + pw[2][d] + pw[3][d];
    }
    if (boxe > 12 || boxf > 12 || ltot > 8)
      return fail(PQ_E_AM, "HGP leaf beyond the supported size (box <= 12 per side, total order <= 8)");
  }
  DevBuf b_np, b_po, b_e, b_c, b_n, b_o, b_p, b_m, b_out;
  PQ_HIP(hipMalloc(&b_np.p, 4 * n * sizeof(int)));
  PQ_HIP(hipMalloc(&b_po.p, 4 * n * sizeof(int)));
  PQ_HIP(hipMalloc(&b_e.p, nprim_total * sizeof(double)));
  PQ_HIP(hipMalloc(&b_c.p, nprim_total * sizeof(double)));
  PQ_HIP(hipMalloc(&b_n.p, nprim_total * sizeof(double)));
  PQ_HIP(hipMalloc(&b_o.p, 12 * n * sizeof(double)));
  PQ_HIP(hipMalloc(&b_p.p, 12 * n * sizeof(int)));
  PQ_HIP(hipMalloc(&b_out.p, n * sizeof(double)));
  if (mext_host) PQ_HIP(hipMalloc(&b_m.p, n * sizeof(int)));
  PQ_HIP(hipMemcpy(b_np.p, nprim_host, 4 * n * sizeof(int), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_po.p, poff_host, 4 * n * sizeof(int), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_e.p, exps_host, nprim_total * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_c.p, coefs_host, nprim_total * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_n.p, norms_host, nprim_total * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_o.p, origins_host, 12 * n * sizeof(double), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_p.p, powers_host, 12 * n * sizeof(int), hipMemcpyHostToDevice));
  if (mext_host) PQ_HIP(hipMemcpy(b_m.p, mext_host, n * sizeof(int), hipMemcpyHostToDevice));
  PQ_HIP(launch_hgp_contr(n, (int*)b_np.p, (int*)b_po.p, (double*)b_e.p, (double*)b_c.p, (double*)b_n.p,
                          (double*)b_o.p, (int*)b_p.p, (int*)b_m.p, compat_nr, (double*)b_out.p, 0));
  PQ_HIP(hipMemcpy(out_host, b_out.p, n * sizeof(double), hipMemcpyDeviceToHost));
  return PQ_OK;
}

int pq_engine_create(int nshell, const int* shell_l, const int* shell_nprim, const int* shell_poff,
                     const int* shell_atom, const double* exps, const double* coefs, int natom,
                     const double* atom_xyz, const double* atom_z, pq_engine** out) {
  return pq_engine_create_batch(nshell, shell_l, shell_nprim, shell_poff, shell_atom, exps, coefs, natom, atom_xyz,
                                atom_z, nullptr, out);
}

int pq_engine_create_batch(int nshell, const int* shell_l, const int* shell_nprim, const int* shell_poff,
                           const int* shell_atom, const double* exps, const double* coefs, int natom,
                           const double* atom_xyz, const double* atom_z, const int* atom_mol, pq_engine** out) {
  if (!out || nshell <= 0 || natom <= 0 || !shell_l || !shell_nprim || !shell_poff || !shell_atom || !exps ||
      !coefs || !atom_xyz || !atom_z)
    return fail(PQ_E_ARG, "bad engine arguments");
  if (atom_mol) {
    if (atom_mol[0] != 0) return fail(PQ_E_ARG, "molecule ids must start at 0");
    for (int a = 1; a < natom; ++a)
      if (atom_mol[a] != atom_mol[a - 1] && atom_mol[a] != atom_mol[a - 1] + 1)
        return fail(PQ_E_ARG, "atoms of a molecule must be contiguous, molecules numbered 0, 1, ...");
    for (int s = 1; s < nshell; ++s)
      if (shell_atom[s] >= 0 && shell_atom[s] < natom && shell_atom[s - 1] >= 0 && shell_atom[s - 1] < natom &&
          atom_mol[shell_atom[s]] < atom_mol[shell_atom[s - 1]])
        return fail(PQ_E_ARG, "shells must be ordered by molecule");
  }
  *out = nullptr;
  int nprim_total = 0;
  for (int s = 0; s < nshell; ++s) {
    if (shell_l[s] < 0 || shell_l[s] > 2) return fail(PQ_E_AM, "shell angular momentum must be 0..2");
    if (shell_nprim[s] < 1 || shell_poff[s] < 0) return fail(PQ_E_ARG, "bad primitive range");
    if (shell_atom[s] < 0 || shell_atom[s] >= natom) return fail(PQ_E_ARG, "bad shell atom");
    nprim_total = std::max(nprim_total, shell_poff[s] + shell_nprim[s]);
  }
  pq_engine* e = new pq_engine();
  e->nshell = nshell;
  e->natom = natom;
  e->sh_l.assign(shell_l, shell_l + nshell);
  e->sh_np.assign(shell_nprim, shell_nprim + nshell);
  e->sh_off.assign(shell_poff, shell_poff + nshell);
  e->sh_atom.assign(shell_atom, shell_atom + nshell);
  e->exps.assign(exps, exps + nprim_total);
  e->coefs.assign(coefs, coefs + nprim_total);
  e->atom_xyz.assign(atom_xyz, atom_xyz + 3 * natom);
  e->atom_z.assign(atom_z, atom_z + natom);
  e->sh_bf.resize(nshell);
  int nbf = 0;
  for (int s = 0; s < nshell; ++s) {
    e->sh_bf[s] = nbf;
    nbf += ncart(shell_l[s]);
  }
  e->nbf = nbf;
  // molecules
  e->atom_mol.assign(natom, 0);
  if (atom_mol) e->atom_mol.assign(atom_mol, atom_mol + natom);
  e->nmol = e->atom_mol[natom - 1] + 1;
  e->sh_mol.resize(nshell);
  e->mol_nbf.assign(e->nmol, 0);
  e->mol_fbase.assign(e->nmol, -1);
  e->mol_sh0.assign(e->nmol, -1);
  e->mol_at0.assign(e->nmol, -1);
  e->mol_nat.assign(e->nmol, 0);
  for (int a = 0; a < natom; ++a) {
    const int m = e->atom_mol[a];
    if (e->mol_at0[m] < 0) e->mol_at0[m] = a;
    e->mol_nat[m] += 1;
  }
  for (int sidx = 0; sidx < nshell; ++sidx) {
    const int m = e->atom_mol[shell_atom[sidx]];
    e->sh_mol[sidx] = m;
    if (e->mol_sh0[m] < 0) {
      e->mol_sh0[m] = sidx;
      e->mol_fbase[m] = e->sh_bf[sidx];
    }
    e->mol_nbf[m] += ncart(shell_l[sidx]);
  }
  e->mol_moff.assign(e->nmol, 0);
  e->msize = 0;
  e->maxn = 0;
  std::vector<int> moff32(e->nmol);
  for (int m = 0; m < e->nmol; ++m) {
    if (e->mol_sh0[m] < 0) {
      delete e;
      return fail(PQ_E_ARG, "a molecule has no shells");
    }
    e->mol_moff[m] = e->msize;
    moff32[m] = (int)e->msize;
    e->msize += (int64_t)e->mol_nbf[m] * e->mol_nbf[m];
    e->maxn = std::max(e->maxn, e->mol_nbf[m]);
    e->maxat = std::max(e->maxat, e->mol_nat[m]);
  }
  if (e->msize >= (int64_t)1 << 31) {
    delete e;
    return fail(PQ_E_ARG, "packed matrices beyond 2^31 elements");
  }
  if (dev_alloc(&e->d_atom_xyz, 3 * natom) != hipSuccess || dev_alloc(&e->d_atom_z, natom) != hipSuccess ||
      dev_alloc(&e->d_Jacc, 2 * (size_t)e->msize) != hipSuccess ||
      dev_alloc(&e->d_Kacc, 2 * (size_t)e->msize) != hipSuccess ||
      dev_alloc(&e->d_mol_moff, e->nmol) != hipSuccess || dev_alloc(&e->d_mol_n, e->nmol) != hipSuccess ||
      dev_alloc(&e->d_atom_mol, natom) != hipSuccess || dev_alloc(&e->d_mol_at0, e->nmol) != hipSuccess ||
      dev_alloc(&e->d_mol_nat, e->nmol) != hipSuccess) {
    delete e;
    return fail(PQ_E_ALLOC, "device allocation failed");
  }
  if (hipMemcpy(e->d_mol_moff, moff32.data(), e->nmol * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_mol_n, e->mol_nbf.data(), e->nmol * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_atom_mol, e->atom_mol.data(), natom * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_mol_at0, e->mol_at0.data(), e->nmol * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_mol_nat, e->mol_nat.data(), e->nmol * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    delete e;
    return fail(PQ_E_HIP, "copy of the molecule layout failed");
  }
  if (hipMemcpy(e->d_atom_xyz, atom_xyz, 3 * natom * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_atom_z, atom_z, natom * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    delete e;
    return fail(PQ_E_HIP, "copy of nuclei failed");
  }
  // LDS per workgroup of the current device (160 KB on gfx950): the J/K
  // launches keep the bra rows of K in LDS up to that size
  {
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0) {
      e->lds_dev = e->lds_max = lds;
    }
  }
  *out = e;
  return PQ_OK;
}

int pq_engine_destroy(pq_engine* eng) {
  delete eng;
  return PQ_OK;
}

int pq_engine_layout(const pq_engine* eng, int* nmol, int64_t* msize, int* mol_nbf, int64_t* mol_moff,
                     int* mol_first_atom, int* mol_natom) {
  if (!eng || !nmol || !msize) return fail(PQ_E_ARG, "null argument");
  *nmol = eng->nmol;
  *msize = eng->msize;
  for (int m = 0; m < eng->nmol; ++m) {
    if (mol_nbf) mol_nbf[m] = eng->mol_nbf[m];
    if (mol_moff) mol_moff[m] = eng->mol_moff[m];
    if (mol_first_atom) mol_first_atom[m] = eng->mol_at0[m];
    if (mol_natom) mol_natom[m] = eng->mol_nat[m];
  }
  return PQ_OK;
}

int pq_engine_generation(const pq_engine* eng, uint64_t* gen) {
  if (!eng || !gen) return fail(PQ_E_ARG, "null argument");
  *gen = eng->gen;
  return PQ_OK;
}

int pq_engine_nbf(const pq_engine* eng, int* nbf) {
  if (!eng || !nbf) return fail(PQ_E_ARG, "null argument");
  *nbf = eng->nbf;
  return PQ_OK;
}

int pq_engine_set_geometry(pq_engine* eng, const double* atom_xyz_host) {
  if (!eng || !atom_xyz_host) return fail(PQ_E_ARG, "null argument");
  eng->atom_xyz.assign(atom_xyz_host, atom_xyz_host + 3 * eng->natom);
  PQ_HIP(hipMemcpy(eng->d_atom_xyz, atom_xyz_host, 3 * eng->natom * sizeof(double), hipMemcpyHostToDevice));
  eng->free_pairs();
  eng->free_functions();
  return PQ_OK;
}

int pq_build_pairs(pq_engine* eng, double schwarz_cut, double fp32_cut) {
  if (!eng) return fail(PQ_E_ARG, "null engine");
  if (!(schwarz_cut >= 0.0) || !(fp32_cut >= 0.0)) return fail(PQ_E_ARG, "cutoffs must be >= 0");
  if (!eng->pairs_ok) {
    int rc = build_pair_data(eng);
    if (rc) return rc;
  }
  int rc = build_screened(eng, schwarz_cut);
  if (rc) return rc;
  rc = build_work(eng, eng->work, schwarz_cut, fp32_cut, true, eng->pairs_s, eng->sorted_s);
  if (rc) return rc;
  eng->cut = schwarz_cut;
  eng->fp32cut = fp32_cut;
  eng->work_ok = true;
  return PQ_OK;
}

int pq_engine_set_prim_screen(pq_engine* eng, double rel) {
  if (!eng) return fail(PQ_E_ARG, "null engine");
  if (!(rel >= 0.0) || rel > 1.0) return fail(PQ_E_ARG, "rel must be in [0, 1]");
  eng->prim_rel = rel;
  return PQ_OK;
}

int pq_engine_set_lds_limit(pq_engine* eng, int bytes, int* device_bytes) {
  if (!eng) return fail(PQ_E_ARG, "null engine");
  if (bytes < 0) return fail(PQ_E_ARG, "bytes must be >= 0");
  if (device_bytes) *device_bytes = eng->lds_dev;
  const int lim = bytes == 0 ? eng->lds_dev : std::min(bytes, eng->lds_dev);
  if (lim != eng->lds_max) {
    eng->lds_max = lim;
    ++eng->gen;  // captured builds hold launches sized for the old limit
  }
  return PQ_OK;
}

int pq_engine_set_fused(pq_engine* eng, int on) {
  if (!eng) return fail(PQ_E_ARG, "null engine");
  if ((on != 0) != eng->fused) {
    eng->fused = on != 0;
    ++eng->gen;  // a captured build holds the other launch pattern
  }
  return PQ_OK;
}

int pq_engine_set_deterministic(pq_engine* eng, int on) {
  if (!eng) return fail(PQ_E_ARG, "null engine");
  if ((on != 0) != eng->det) {
    eng->det = on != 0;
    ++eng->gen;  // a captured build holds the other kernels and accumulators
  }
  return PQ_OK;
}

int pq_engine_set_streams(pq_engine* eng, int nstreams) {
  if (!eng) return fail(PQ_E_ARG, "null engine");
  if (nstreams < 0 || nstreams > pq_engine::kSide) return fail(PQ_E_ARG, "nstreams must be 0..4");
  eng->nside = nstreams;
  return PQ_OK;
}

// -1, 0, or a run length 1..1024 optionally or-ed with PQ_CAP_STAGE_D
static bool valid_cap(int cap) {
  if (cap == -1 || cap == 0) return true;
  const int run = cap & ~PQ_CAP_STAGE_D;
  return cap > 0 && run >= 1 && run <= 1024;
}

int pq_engine_set_group_cap(pq_engine* eng, int cap) {
  if (!eng) return fail(PQ_E_ARG, "null engine");
  if (!valid_cap(cap)) return fail(PQ_E_ARG, "cap must be -1, 0 or 1..1024 (optionally | PQ_CAP_STAGE_D)");
  ++eng->gen;
  for (auto& a : eng->group_cap)
    for (auto& b : a) b[0] = b[1] = cap;
  if (eng->work_ok)
    for (auto& c : eng->work)
      for (int prec = 0; prec < 2; ++prec) {
        int rc = set_groups(c, prec, cap >= 0 ? cap : default_group_cap(eng->maxn));
        if (rc) return rc;
      }
  return PQ_OK;
}

int pq_engine_set_group_caps(pq_engine* eng, const int* caps) {
  if (!eng || !caps) return fail(PQ_E_ARG, "null argument");
  for (int i = 0; i < 42; ++i)
    if (!valid_cap(caps[i])) return fail(PQ_E_ARG, "caps must be -1, 0 or 1..1024 (optionally | PQ_CAP_STAGE_D)");
  ++eng->gen;
  int ci = 0;
  for (int X = 0; X < 6; ++X)
    for (int Y = 0; Y <= X; ++Y, ++ci)
      for (int prec = 0; prec < 2; ++prec) eng->group_cap[X][Y][prec] = caps[2 * ci + prec];
  if (eng->work_ok)
    for (auto& c : eng->work)
      for (int prec = 0; prec < 2; ++prec) {
        const int cap = eng->group_cap[c.X][c.Y][prec];
        int rc = set_groups(c, prec, cap >= 0 ? cap : default_group_cap(eng->maxn));
        if (rc) return rc;
      }
  return PQ_OK;
}

int pq_engine_autotune(pq_engine* eng, const double* D_dev, int reps, void* stream, int* caps_out) {
  if (!eng || !D_dev) return fail(PQ_E_ARG, "null argument");
  if (!eng->work_ok) return fail(PQ_E_STATE, "pq_build_pairs not called");
  if (reps < 1 || reps > 100) return fail(PQ_E_ARG, "reps must be 1..100");
  static const int kCand[] = {0,  4,  8,  16,  32, 4 | PQ_CAP_STAGE_D, 8 | PQ_CAP_STAGE_D, 16 | PQ_CAP_STAGE_D,
                              32 | PQ_CAP_STAGE_D};
  hipStream_t st = (hipStream_t)stream;
  ++eng->gen;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PQ_HIP(hipEventCreate(&e0));
  PQ_HIP(hipEventCreate(&e1));
  int rc = PQ_OK;
  for (size_t ci = 0; ci < eng->work.size() && rc == PQ_OK; ++ci) {
    ClassWork& c = eng->work[ci];
    const int la = kPairL[c.X][0], lb = kPairL[c.X][1], lc = kPairL[c.Y][0], ld = kPairL[c.Y][1];
    const bool unsplit = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) <= kLaneMaxComp;
    for (int prec = 0; prec < 2 && rc == PQ_OK; ++prec) {
      const auto& tl = prec ? c.t32 : c.t64;
      if (caps_out) caps_out[2 * ci + prec] = c.cap[prec];
      if (tl.empty() || !unsplit) continue;
      float best = 1e30f;
      int best_cap = c.cap[prec];
      for (int cap : kCand) {
        if ((rc = set_groups(c, prec, cap)) != PQ_OK) break;
        hipError_t err = hipSuccess;
        for (int r = 0; r <= reps && err == hipSuccess; ++r) {
          if (r == 1) err = hipEventRecord(e0, st);
          if (err == hipSuccess)
            err = launch_eri_class(la, lb, lc, ld, MODE_JK, prec, eng->view_s(), prec ? c.d_t32 : c.d_t64,
                                   (int)tl.size(), eng->d_sorted_s[c.Y], D_dev, eng->d_Jacc, eng->d_Kacc, nullptr,
                                   nullptr, st, prec ? c.d_g32 : c.d_g64, (int)(prec ? c.g32 : c.g64).size(),
                                   c.gflags(prec));
        }
        if (err == hipSuccess) err = hipEventRecord(e1, st);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float t = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
        if (err != hipSuccess) { rc = fail(PQ_E_HIP, hipGetErrorString(err)); break; }
        if (t < best) { best = t; best_cap = cap; }
      }
      if (rc == PQ_OK) rc = set_groups(c, prec, best_cap);
      eng->group_cap[c.X][c.Y][prec] = best_cap;
      if (caps_out) caps_out[2 * ci + prec] = best_cap;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int pq_engine_stats(const pq_engine* eng, uint64_t counts[6], double flops[2]) {
  if (!eng || !counts || !flops) return fail(PQ_E_ARG, "null argument");
  if (!eng->work_ok) return fail(PQ_E_STATE, "pq_build_pairs not called");
  for (int i = 0; i < 6; ++i) counts[i] = 0;
  flops[0] = flops[1] = 0;
  for (const auto& c : eng->work) {
    const int la = kPairL[c.X][0], lb = kPairL[c.X][1], lc = kPairL[c.Y][0], ld = kPairL[c.Y][1];
    counts[0] += c.nq64;
    counts[1] += c.nq32;
    counts[2] += (uint64_t)c.np64;
    counts[3] += (uint64_t)c.np32;
    counts[4] += (uint64_t)c.nf64;
    counts[5] += (uint64_t)c.nf32;
    flops[0] += c.np64 * rys_flops_prim(la, lb, lc, ld) + c.nq64 * rys_flops_dig(la, lb, lc, ld);
    flops[1] += c.np32 * rys_flops_prim(la, lb, lc, ld) + c.nq32 * rys_flops_dig(la, lb, lc, ld);
  }
  return PQ_OK;
}

int pq_get2ints(pq_engine* eng, double* out_dev, uint64_t len, void* stream) {
  if (!eng || !out_dev) return fail(PQ_E_ARG, "null argument");
  if (eng->nmol > 1) return fail(PQ_E_STATE, "stored integrals need a single-molecule engine");
  const uint64_t n = eng->nbf, npair = n * (n + 1) / 2;
  if (len != npair * (npair + 1) / 2) return fail(PQ_E_ARG, "output length must be npair*(npair+1)/2");
  if (!eng->pairs_ok) {
    int rc = build_pair_data(eng);
    if (rc) return rc;
  }
  if (!eng->store_ok) {
    int rc = build_work(eng, eng->store_work, 0.0, 0.0, false, eng->pairs, eng->sorted);
    if (rc) return rc;
    eng->store_ok = true;
  }
  hipStream_t st = (hipStream_t)stream;
  PQ_HIP(launch_fill(out_dev, (int64_t)len, 0.0, st));
  for (const auto& c : eng->store_work) {
    if (c.t64.empty()) continue;
    PQ_HIP(launch_eri_class(kPairL[c.X][0], kPairL[c.X][1], kPairL[c.Y][0], kPairL[c.Y][1], MODE_STORE, 0,
                            eng->view(), c.d_t64, (int)c.t64.size(), eng->d_sorted[c.Y], nullptr, nullptr,
                            nullptr, out_dev, nullptr, st));
  }
  return PQ_OK;
}

namespace {
// Per-function primitive arrays for the HGP kernels: exponents, raw
// coefficients and the full primitive norm for the function's own powers
// (primitive-gto.c:76-83), with the function's centre and powers.
int build_functions(pq_engine* e) {
  if (e->d_fnp) return PQ_OK;
  std::vector<int> fnp, foff, fpw;
  std::vector<double> forg, fexps, fcoefs, fnorms;
  for (int s = 0; s < e->nshell; ++s) {
    const int L = e->sh_l[s];
    for (int c = 0; c < ncart(L); ++c) {
      fnp.push_back(e->sh_np[s]);
      foff.push_back((int)fexps.size());
      int df = 1;
      for (int d = 0; d < 3; ++d) {
        fpw.push_back(cart_pow(L, c, d));
        forg.push_back(e->atom_xyz[3 * e->sh_atom[s] + d]);
        df *= dfact(2 * cart_pow(L, c, d) - 1);
      }
      for (int i = 0; i < e->sh_np[s]; ++i) {
        const double a = e->exps[e->sh_off[s] + i];
        fexps.push_back(a);
        fcoefs.push_back(e->coefs[e->sh_off[s] + i]);
        fnorms.push_back(nrad(a, L) / std::sqrt((double)df));
      }
    }
  }
  auto up = [](auto** d, const auto& v) -> hipError_t {
    hipError_t r = hipMalloc((void**)d, std::max<size_t>(1, v.size()) * sizeof(v[0]));
    if (r == hipSuccess && !v.empty()) r = hipMemcpy(*d, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice);
    return r;
  };
  hipError_t r;
  if ((r = up(&e->d_fnp, fnp)) != hipSuccess || (r = up(&e->d_foff, foff)) != hipSuccess ||
      (r = up(&e->d_fpw, fpw)) != hipSuccess || (r = up(&e->d_forg, forg)) != hipSuccess ||
      (r = up(&e->d_fexps, fexps)) != hipSuccess || (r = up(&e->d_fcoefs, fcoefs)) != hipSuccess ||
      (r = up(&e->d_fnorms, fnorms)) != hipSuccess) {
    e->free_functions();
    return fail(PQ_E_HIP, hipGetErrorString(r));
  }
  return PQ_OK;
}
}  // namespace

static hipError_t fork_side(pq_engine* eng, hipStream_t st);
static hipError_t join_side(pq_engine* eng, hipStream_t st);

int pq_get2ints_hgp(pq_engine* eng, double* out_dev, uint64_t len, int compat_nr, void* stream) {
  if (!eng || !out_dev) return fail(PQ_E_ARG, "null argument");
  if (eng->nmol > 1) return fail(PQ_E_STATE, "stored integrals need a single-molecule engine");
  const uint64_t n = eng->nbf, npair = n * (n + 1) / 2;
  if (len != npair * (npair + 1) / 2) return fail(PQ_E_ARG, "output length must be npair*(npair+1)/2");
  hipStream_t st = (hipStream_t)stream;
  if (!compat_nr) {
    // one shell quartet per lane over the store work lists
    if (!eng->pairs_ok) {
      int rc = build_pair_data(eng);
      if (rc) return rc;
    }
    if (!eng->hgp_ok) {
      // tiles split over bra primitives and ket-primitive lanes, as for
      // J/K: the split tiles add their partial blocks into the zeroed array
      int rc = build_work(eng, eng->hgp_work, 0.0, 0.0, true, eng->pairs, eng->sorted);
      if (rc) return rc;
      eng->hgp_ok = true;
    }
    PQ_HIP(launch_fill(out_dev, (int64_t)len, 0.0, st));
    // all classes on the caller's stream: the d-shell classes keep their
    // recurrence tables in per-lane private memory, whose backing store the
    // runtime allocates per hardware queue; spread over the side streams'
    // queues as well, it outgrew the runtime's scratch limit and the next
    // large-scratch dispatch failed (HSA_STATUS_ERROR_OUT_OF_RESOURCES)
    // the step-table classes and the small unrolled ones share one launch
    // (heaviest class first, so that its long-running waves start first);
    // the large unrolled classes (hgp_multi_class false: 1-2 waves per SIMD
    // of registers) run in launches of their own, as does a list beyond the
    // launch's segment table.  Those of them with no private memory go to
    // the side streams, beside the shared launch; the others stay on st.
    PQ_HIP(fork_side(eng, st));
    int next = 0;
    HgpMulti mm{};
    int64_t nb = 0;
    for (auto it = eng->hgp_work.rbegin(); it != eng->hgp_work.rend(); ++it) {
      const auto& c = *it;
      if (c.t64.empty()) continue;
      const int la = kPairL[c.X][0], lb = kPairL[c.X][1], lc = kPairL[c.Y][0], ld = kPairL[c.Y][1];
      int split = 1;
      const int64_t blocks = (int64_t)c.t64.size();
      if (hgp_multi_class(la, lb, lc, ld, &split) && mm.nseg < kHgpMaxSeg && nb + blocks * split < (1ll << 31)) {
        mm.seg[mm.nseg++] = HgpSeg{c.d_t64, eng->d_sorted[c.Y], hgp_code(la, lb, lc, ld), split, (int)nb, 0};
        nb += blocks * split;
      } else {
        hipStream_t ks = st;
        if (eng->nside > 0 && hgp_class_private_bytes(la, lb, lc, ld) == 0) {
          ks = eng->side[next];
          next = (next + 1) % eng->nside;
        }
        PQ_HIP(launch_hgp_class(la, lb, lc, ld, eng->view(), c.d_t64, (int)c.t64.size(), eng->d_sorted[c.Y],
                                out_dev, ks));
      }
    }
    PQ_HIP(launch_hgp_multi(eng->view(), mm, (int)nb, out_dev, st));
    PQ_HIP(join_side(eng, st));
    return PQ_OK;
  }
  int rc = build_functions(eng);
  if (rc) return rc;
  PQ_HIP(launch_hgp_store(len, eng->d_fnp, eng->d_foff, eng->d_forg, eng->d_fpw, eng->d_fexps, eng->d_fcoefs,
                          eng->d_fnorms, compat_nr, out_dev, st));
  return PQ_OK;
}

int pq_engine_class_stats(const pq_engine* eng, int c, int l[4], uint64_t counts[8], double flops[2]) {
  if (!eng || !l || !counts || !flops) return fail(PQ_E_ARG, "null argument");
  if (!eng->work_ok) return fail(PQ_E_STATE, "pq_build_pairs not called");
  if (c < 0 || c >= (int)eng->work.size()) return fail(PQ_E_ARG, "class index out of range");
  const ClassWork& w = eng->work[c];
  l[0] = kPairL[w.X][0];
  l[1] = kPairL[w.X][1];
  l[2] = kPairL[w.Y][0];
  l[3] = kPairL[w.Y][1];
  counts[0] = w.nq64;
  counts[1] = w.nq32;
  counts[2] = (uint64_t)w.np64;
  counts[3] = (uint64_t)w.np32;
  counts[4] = (uint64_t)w.nf64;
  counts[5] = (uint64_t)w.nf32;
  counts[6] = w.t64.size();
  counts[7] = w.t32.size();
  flops[0] = w.np64 * rys_flops_prim(l[0], l[1], l[2], l[3]) + w.nq64 * rys_flops_dig(l[0], l[1], l[2], l[3]);
  flops[1] = w.np32 * rys_flops_prim(l[0], l[1], l[2], l[3]) + w.nq32 * rys_flops_dig(l[0], l[1], l[2], l[3]);
  return PQ_OK;
}

static int fock_jk_impl(pq_engine* eng, int nd, const double* const* Ds, double* const* Js, double* const* Ks,
                        int precision, void* stream, float* ms);

// Makes the engine's side streams (default priority) wait for st; creates
// them and their events on first use.  No-op with nside = 0.
constexpr int kEagerSide = 3;

static hipError_t fork_side(pq_engine* eng, hipStream_t st) {
  if (eng->nside <= 0) return hipSuccess;
  hipError_t err = hipSuccess;
  if (!eng->streams_ok) {
    for (int i = 0; i < pq_engine::kSide && err == hipSuccess; ++i) {
      err = hipStreamCreateWithFlags(&eng->side[i], hipStreamNonBlocking);
      if (err == hipSuccess) err = hipEventCreateWithFlags(&eng->join_ev[i], hipEventDisableTiming);
    }
    if (err == hipSuccess) err = hipEventCreateWithFlags(&eng->fork_ev, hipEventDisableTiming);
    if (err != hipSuccess) return err;
    eng->streams_ok = true;
  }
  err = hipEventRecord(eng->fork_ev, st);
  for (int i = 0; i < eng->nside && err == hipSuccess; ++i) err = hipStreamWaitEvent(eng->side[i], eng->fork_ev, 0);
  return err;
}

// Makes st wait for everything enqueued on the side streams.
static hipError_t join_side(pq_engine* eng, hipStream_t st) {
  hipError_t err = hipSuccess;
  for (int i = 0; i < eng->nside && err == hipSuccess; ++i) {
    err = hipEventRecord(eng->join_ev[i], eng->side[i]);
    if (err == hipSuccess) err = hipStreamWaitEvent(st, eng->join_ev[i], 0);
  }
  return err;
}

int pq_fock_jk(pq_engine* eng, const double* D_dev, double* J_dev, double* K_dev, int precision, void* stream) {
  return fock_jk_impl(eng, 1, &D_dev, &J_dev, &K_dev, precision, stream, nullptr);
}

int pq_fock_jk_timed(pq_engine* eng, const double* D_dev, double* J_dev, double* K_dev, int precision,
                     void* stream, float* ms) {
  if (!ms) return fail(PQ_E_ARG, "null timing array");
  return fock_jk_impl(eng, 1, &D_dev, &J_dev, &K_dev, precision, stream, ms);
}

int pq_fock_jk_multi(pq_engine* eng, int ndens, const double* const* D_dev, double* const* J_dev,
                     double* const* K_dev, int precision, void* stream) {
  if (!D_dev || !J_dev || !K_dev) return fail(PQ_E_ARG, "null pointer array");
  if (ndens < 1 || ndens > 2) return fail(PQ_E_ARG, "ndens must be 1 or 2");
  if (ndens == 2 && precision != PQ_PREC_FP64) return fail(PQ_E_ARG, "two-density builds are FP64");
  return fock_jk_impl(eng, ndens, D_dev, J_dev, K_dev, precision, stream, nullptr);
}

static int fock_jk_impl(pq_engine* eng, int nd, const double* const* Ds, double* const* Js, double* const* Ks,
                        int precision, void* stream, float* ms) {
  if (!eng || !Ds[0] || (nd == 2 && !Ds[1])) return fail(PQ_E_ARG, "null argument");
  if (!eng->work_ok) return fail(PQ_E_STATE, "pq_build_pairs must be called before pq_fock_jk");
  if (precision != PQ_PREC_FP64 && precision != PQ_PREC_MIXED) return fail(PQ_E_ARG, "bad precision");
  if (precision == PQ_PREC_MIXED && !(eng->fp32cut > 0.0))
    return fail(PQ_E_STATE, "mixed precision needs pq_build_pairs with fp32_cut > 0");
  if (precision == PQ_PREC_FP64 && eng->fp32cut > 0.0)
    return fail(PQ_E_STATE, "work lists were built for mixed precision; rebuild with fp32_cut = 0");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nn = eng->msize;  // one packed matrix (n^2 for one molecule)
  // timing: one event pair per class launch, recorded on the stream the
  // kernel runs on (so concurrent launches are timed as they run)
  std::vector<hipEvent_t> ev;
  std::vector<int> slot;  // ms[] entry of each event pair after the first (whole build)
  auto mark = [&](hipStream_t s) -> hipError_t {
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    ev.push_back(e);
    return hipEventRecord(e, s);
  };
  // Calls outside a graph capture use at most kEagerSide side streams: with
  // all four beside the caller's stream, five streams share the process's
  // four hardware queues, and eager calls measured 1.21 ms against 1.02 ms
  // with three ((H2O)_8); graph replays are fastest with four (0.96 ms
  // against 1.02).  Timed calls (the per-class autotuner) keep the set count.
  const int nside_set = eng->nside;
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  if (!ms && hipStreamIsCapturing(st, &capture) == hipSuccess && capture == hipStreamCaptureStatusNone &&
      eng->nside > kEagerSide)
    eng->nside = kEagerSide;
  int rc = PQ_OK;
  do {
    hipError_t err = hipSuccess;
#define PQ_TRY(x)                                  \
  if ((err = (x)) != hipSuccess) {                 \
    rc = fail(PQ_E_HIP, hipGetErrorString(err));   \
    break;                                         \
  }
    if (ms) PQ_TRY(mark(st));
    // two densities: stacked copies, digested by the MODE_JK2 kernels
    const double* D_dev = Ds[0];
    if (nd == 2) {
      if (!eng->d_D2) {
        PQ_TRY(dev_alloc(&eng->d_D2, 2 * (size_t)nn));
        ++eng->gen;
      }
      PQ_TRY(hipMemcpyAsync(eng->d_D2, Ds[0], nn * sizeof(double), hipMemcpyDeviceToDevice, st));
      PQ_TRY(hipMemcpyAsync(eng->d_D2 + nn, Ds[1], nn * sizeof(double), hipMemcpyDeviceToDevice, st));
      D_dev = eng->d_D2;
    }
    // the deterministic build accumulates into fixed-point limbs instead
    const bool det = eng->det;
    EngineView kview = eng->view_s();
    double *Jacc = eng->d_Jacc, *Kacc = eng->d_Kacc;
    if (det) {
      // per density three limb arrays each of J, K and the ket J:
      // [J d0 | J d1 | K d0 | K d1 | ket J d0 | ket J d1]
      const size_t need = (size_t)nd * (6 * (size_t)nn + 3 * (size_t)eng->jp_size);
      if (eng->fx_cap < need) {
        dev_free(eng->d_fx);
        eng->fx_cap = 0;
        PQ_TRY(dev_alloc(&eng->d_fx, need));
        eng->fx_cap = need;
        ++eng->gen;
      }
      Jacc = eng->d_fx;
      Kacc = eng->d_fx + 3 * nd * nn;
      kview.jpair = eng->d_fx + 6 * nd * nn;
      PQ_TRY(launch_zero3(Jacc, 3 * nd * nn, Kacc, 3 * nd * nn, kview.jpair, 3 * nd * eng->jp_size, st));
    } else {
      PQ_TRY(launch_zero3(Jacc, nd * nn, Kacc, nd * nn, eng->d_jpair, nd * eng->jp_size, st));
    }
    // fork: class kernels spread over the side streams, joined before the
    // symmetrisation (all of them only add into Jacc/Kacc)
    const int nside = eng->nside;
    PQ_TRY(fork_side(eng, st));
    int next = 0;
    // fused launch: the FP64 bra runs of the light classes in one grid,
    // segments ordered by their Rys operation count (heaviest first)
    std::vector<char> in_multi(eng->work.size(), 0);
    if (eng->fused && nd == 1 && !det) {
      std::vector<size_t> mc;
      for (size_t ci = 0; ci < eng->work.size(); ++ci) {
        const auto& c = eng->work[ci];
        const int la = kPairL[c.X][0], lb = kPairL[c.X][1], lc = kPairL[c.Y][0], ld = kPairL[c.Y][1];
        if (jk_multi_class(la, lb, lc, ld) && !c.t64.empty() && !c.g64.empty()) mc.push_back(ci);
      }
      std::sort(mc.begin(), mc.end(), [&](size_t a, size_t b) {
        const auto &x = eng->work[a], &y = eng->work[b];
        return x.np64 * rys_flops_prim(kPairL[x.X][0], kPairL[x.X][1], kPairL[x.Y][0], kPairL[x.Y][1]) >
               y.np64 * rys_flops_prim(kPairL[y.X][0], kPairL[y.X][1], kPairL[y.Y][0], kPairL[y.Y][1]);
      });
      JkMulti mm{};
      int nb = 0;
      size_t lds = 0;
      for (size_t ci : mc) {
        if (mm.nseg >= kJkMaxSeg) break;
        const auto& c = eng->work[ci];
        const int la = kPairL[c.X][0], lb = kPairL[c.X][1], lc = kPairL[c.Y][0], ld = kPairL[c.Y][1];
        const int sd = c.gflags(0) & kGroupStageD;
        size_t need = jk_multi_lds(la, lb, lc, ld, eng->maxn, sd);
        int stage = sd;
        if (need > (size_t)eng->lds_max) {  // as the per-class launch: drop the staged D rows first
          stage = 0;
          need = jk_multi_lds(la, lb, lc, ld, eng->maxn, 0);
        }
        if (need > (size_t)eng->lds_max) continue;
        mm.seg[mm.nseg++] = JkSeg{c.d_t64, c.d_g64, eng->d_sorted_s[c.Y], hgp_code(la, lb, lc, ld), nb, stage, 0};
        nb += (int)c.g64.size();
        lds = std::max(lds, need);
        in_multi[ci] = 1;
      }
      if (mm.nseg > 0) {
        hipStream_t ks = nside > 0 ? eng->side[next] : st;
        if (nside > 0) next = (next + 1) % nside;
        PQ_TRY(launch_jk_multi(eng->view_s(), mm, nb, lds, D_dev, eng->d_Jacc, eng->d_Kacc, ks));
      }
    }
    for (size_t ci = 0; ci < eng->work.size() && rc == PQ_OK; ++ci) {
      const auto& c = eng->work[ci];
      const int la = kPairL[c.X][0], lb = kPairL[c.X][1], lc = kPairL[c.Y][0], ld = kPairL[c.Y][1];
      for (int prec = 0; prec < 2; ++prec) {
        const auto& tl = prec ? c.t32 : c.t64;
        if (tl.empty() || (prec == 0 && in_multi[ci])) continue;
        hipStream_t ks = nside > 0 ? eng->side[next] : st;
        if (nside > 0) next = (next + 1) % nside;
        if (ms) PQ_TRY(mark(ks));
        const int mode = det ? (nd == 2 ? MODE_JK2D : MODE_JKD) : (nd == 2 ? MODE_JK2 : MODE_JK);
        PQ_TRY(launch_eri_class(la, lb, lc, ld, mode, prec, kview,
                                prec ? c.d_t32 : c.d_t64,
                                (int)tl.size(), eng->d_sorted_s[c.Y], D_dev, Jacc, Kacc, nullptr,
                                nullptr, ks, prec ? c.d_g32 : c.d_g64, (int)(prec ? c.g32 : c.g64).size(),
                                c.gflags(prec)));
        if (ms) {
          PQ_TRY(mark(ks));
          slot.push_back((int)(2 * ci + prec));
        }
      }
    }
    if (rc != PQ_OK) break;
    PQ_TRY(join_side(eng, st));
    // the ket-side J, accumulated by ket position, into Jacc
    JpairMap jmap;
    for (int y = 0; y < 6; ++y) {
      jmap.ids[y] = eng->d_sorted_s[y];
      jmap.ndf[y] = ncart(kPairL[y][1]);
    }
    if (det) {
      for (int d = 0; d < nd && rc == PQ_OK; ++d) {
        EngineView dv = kview;
        dv.jpair = kview.jpair + (size_t)d * 3 * eng->jp_size;
        long long* Jfx = reinterpret_cast<long long*>(Jacc + (size_t)d * 3 * nn);
        const long long* Kfx = reinterpret_cast<const long long*>(Kacc + (size_t)d * 3 * nn);
        PQ_TRY(launch_jpair_scatter_fx(dv, jmap, Jfx, st));
        PQ_TRY(launch_symmetrize_fx(eng->nmol, eng->maxn, eng->nmol == 1 ? nullptr : eng->d_mol_moff, eng->d_mol_n,
                                    nn, Jfx, Kfx, Js[d], Ks[d], st));
      }
      if (rc != PQ_OK) break;
    } else {
      PQ_TRY(launch_jpair_scatter(eng->view_s(), jmap, nd, eng->d_Jacc, st));
    }
    for (int d = 0; d < nd && rc == PQ_OK && !det; ++d)
      PQ_TRY(eng->nmol == 1
                 ? launch_symmetrize(eng->nbf, eng->d_Jacc + d * nn, eng->d_Kacc + d * nn, Js[d], Ks[d], 1.0, 1.0,
                                     st)
                 : launch_symmetrize_packed(eng->nmol, eng->maxn, eng->d_mol_moff, eng->d_mol_n,
                                            eng->d_Jacc + d * nn, eng->d_Kacc + d * nn, Js[d], Ks[d], st));
    if (rc != PQ_OK) break;
    if (ms) {
      PQ_TRY(mark(st));
      PQ_TRY(hipStreamSynchronize(st));
      for (int i = 0; i < 43; ++i) ms[i] = 0.f;
      for (size_t k = 0; k < slot.size(); ++k) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, ev[1 + 2 * k], ev[2 + 2 * k]);
        ms[slot[k]] += t;
      }
      float tot = 0.f;
      (void)hipEventElapsedTime(&tot, ev.front(), ev.back());
      ms[42] = tot;
    }
#undef PQ_TRY
  } while (0);
  eng->nside = nside_set;
  for (auto e : ev) (void)hipEventDestroy(e);
  return rc;
}

int pq_digest_stored(int nbf, const double* ints_dev, uint64_t len, const double* D_dev, double* J_dev,
                     double* K_dev, void* stream) {
  if (nbf <= 0 || !ints_dev || !D_dev) return fail(PQ_E_ARG, "bad arguments");
  const uint64_t n = nbf, npair = n * (n + 1) / 2;
  if (len != npair * (npair + 1) / 2) return fail(PQ_E_ARG, "integral array length mismatch");
  hipStream_t st = (hipStream_t)stream;
  double *Jacc = nullptr, *Kacc = nullptr;
  PQ_HIP(dev_alloc(&Jacc, (size_t)nbf * nbf));
  PQ_HIP(dev_alloc(&Kacc, (size_t)nbf * nbf));
  hipError_t err = launch_fill(Jacc, (int64_t)nbf * nbf, 0.0, st);
  if (err == hipSuccess) err = launch_fill(Kacc, (int64_t)nbf * nbf, 0.0, st);
  if (err == hipSuccess) err = launch_digest_stored(nbf, ints_dev, len, D_dev, Jacc, Kacc, st);
  if (err == hipSuccess) err = launch_symmetrize(nbf, Jacc, Kacc, J_dev, K_dev, 1.0, 1.0, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  dev_free(Jacc);
  dev_free(Kacc);
  PQ_HIP(err);
  return PQ_OK;
}

}  // extern "C"

// Scratch of pq_digest_stored_noalloc: the J and K accumulators of one call,
// 2 * kNoAllocMaxNbf^2 doubles (4 MiB).  Each GPU holds its own copy for the
// life of the process.  In host code `scratch` names a host-side stand-in,
// not device memory: kernels get the address from scratch_on_device().
constexpr int kNoAllocMaxNbf = 512;
static __device__ __align__(256) double scratch[2 * kNoAllocMaxNbf * kNoAllocMaxNbf];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Device address of the current device's scratch copy, looked up once per
// device (hipGetSymbolAddress with that device current) and kept.
int scratch_on_device(double** out) {
  int dev = 0;
  PQ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(PQ_E_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PQ_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return PQ_OK;
}
}  // namespace

extern "C" {

int pq_digest_stored_noalloc(int nbf, const double* ints_dev, uint64_t len, const double* D_dev, double* J_dev,
                             double* K_dev, void* stream) {
  if (nbf <= 0 || !ints_dev || !D_dev) return fail(PQ_E_ARG, "bad arguments");
  if (nbf > kNoAllocMaxNbf)
    return fail(PQ_E_ARG, "nbf " + std::to_string(nbf) + " exceeds the no-allocation limit of " +
                              std::to_string(kNoAllocMaxNbf) + " functions");
  const uint64_t n = nbf, npair = n * (n + 1) / 2;
  if (len != npair * (npair + 1) / 2) return fail(PQ_E_ARG, "integral array length mismatch");
  double* s = nullptr;
  const int rc = scratch_on_device(&s);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  double* Jacc = s;
  double* Kacc = s + (size_t)nbf * nbf;
  // a fill kernel, not a memset, so that the call can be graph-captured
  PQ_HIP(launch_fill(s, 2 * (int64_t)nbf * nbf, 0.0, st));
  PQ_HIP(launch_digest_stored(nbf, ints_dev, len, D_dev, Jacc, Kacc, st));
  PQ_HIP(launch_symmetrize(nbf, Jacc, Kacc, J_dev, K_dev, 1.0, 1.0, st));
  return PQ_OK;
}

int pq_unpack_eri(int nbf, const double* ints_dev, uint64_t len, int mu0, int nmu, double* out_dev, void* stream) {
  if (nbf <= 0 || !ints_dev || !out_dev) return fail(PQ_E_ARG, "bad arguments");
  const uint64_t n = nbf, npair = n * (n + 1) / 2;
  if (len != npair * (npair + 1) / 2) return fail(PQ_E_ARG, "integral array length mismatch");
  if (mu0 < 0 || nmu < 0 || mu0 + nmu > nbf) return fail(PQ_E_ARG, "row range out of bounds");
  PQ_HIP(launch_unpack_eri(nbf, ints_dev, mu0, nmu, out_dev, (hipStream_t)stream));
  return PQ_OK;
}

int pq_shell_quartet_blocks(pq_engine* eng, int nq, const int* shells_host, double* out_dev, uint64_t out_len,
                            void* stream) {
  if (!eng || (nq > 0 && (!shells_host || !out_dev))) return fail(PQ_E_ARG, "null argument");
  if (nq < 0) return fail(PQ_E_ARG, "negative quartet count");
  if (eng->nmol > 1) return fail(PQ_E_STATE, "shell-quartet blocks need a single-molecule engine");
  if (nq == 0) return PQ_OK;
  std::vector<uint64_t> off(nq);
  uint64_t tot = 0;
  for (int q = 0; q < nq; ++q) {
    uint64_t sz = 1;
    for (int s = 0; s < 4; ++s) {
      const int sh = shells_host[4 * q + s];
      if (sh < 0 || sh >= eng->nshell) return fail(PQ_E_ARG, "shell index out of range");
      sz *= ncart(eng->sh_l[sh]);
    }
    off[q] = tot;
    tot += sz;
  }
  if (out_len < tot) return fail(PQ_E_ARG, "output buffer too small for the requested blocks");
  if (tot >= (1ull << 61)) return fail(PQ_E_ARG, "requested blocks too large");
  if (!eng->pairs_ok) {
    int rc = build_pair_data(eng);
    if (rc) return rc;
  }
  // pair record of two shells (stored once per unordered pair)
  const int ns = eng->nshell;
  std::vector<int> pid((size_t)ns * ns, -1);
  for (size_t p = 0; p < eng->pairs.size(); ++p) {
    const PairInfo& pr = eng->pairs[p];
    pid[(size_t)pr.shA * ns + pr.shB] = pid[(size_t)pr.shB * ns + pr.shA] = (int)p;
  }
  // one entry per request: kernel bra pair (class >= ket class), ket pair,
  // output code; grouped by class, then bra pair
  struct Entry {
    int cls, bra, ket;
    uint64_t code;
  };
  std::vector<int> cidx(36, -1);
  for (int X = 0, ci = 0; X < 6; ++X)
    for (int Y = 0; Y <= X; ++Y, ++ci) cidx[X * 6 + Y] = ci;
  std::vector<Entry> ent(nq);
  for (int q = 0; q < nq; ++q) {
    const int* s = shells_host + 4 * q;
    const int p1 = pid[(size_t)s[0] * ns + s[1]], p2 = pid[(size_t)s[2] * ns + s[3]];
    const int c1 = eng->pcls[p1], c2 = eng->pcls[p2];
    const bool swap = c1 < c2;  // the request's first pair becomes the ket
    const int bra = swap ? p2 : p1, ket = swap ? p1 : p2;
    // does the request list a pair's shells in reverse of the stored order?
    const bool r1 = s[0] != s[1] && s[0] != eng->pairs[p1].shA;
    const bool r2 = s[2] != s[3] && s[2] != eng->pairs[p2].shA;
    uint64_t fl = (swap ? 1u : 0u) | (r1 ? 2u : 0u) | (r2 ? 4u : 0u);
    ent[q] = Entry{cidx[eng->pcls[bra] * 6 + eng->pcls[ket]], bra, ket, off[q] | (fl << 61)};
  }
  std::stable_sort(ent.begin(), ent.end(), [](const Entry& x, const Entry& y) {
    return x.cls != y.cls ? x.cls < y.cls : x.bra < y.bra;
  });
  std::vector<int> kets(nq);
  std::vector<uint64_t> codes(nq);
  std::vector<Tile> tiles;
  std::vector<int> cls_t0(22, 0);
  for (int q = 0; q < nq; ++q) {
    kets[q] = ent[q].ket;
    codes[q] = ent[q].code;
    const bool start = q == 0 || ent[q].cls != ent[q - 1].cls || ent[q].bra != ent[q - 1].bra ||
                       (tiles.back().nket & 0xff) >= 64;
    if (start) tiles.push_back(Tile{ent[q].bra, q, 0, 0});
    tiles.back().nket += 1;
  }
  for (size_t t = 0, c = 0; c <= 21; ++c) {
    while (t < tiles.size() && ent[tiles[t].ket0].cls < (int)c) ++t;
    cls_t0[c] = (int)t;
  }
  DevBuf b_t, b_k, b_c;
  PQ_HIP(hipMalloc(&b_t.p, tiles.size() * sizeof(Tile)));
  PQ_HIP(hipMalloc(&b_k.p, kets.size() * sizeof(int)));
  PQ_HIP(hipMalloc(&b_c.p, codes.size() * sizeof(uint64_t)));
  PQ_HIP(hipMemcpy(b_t.p, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_k.p, kets.data(), kets.size() * sizeof(int), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(b_c.p, codes.data(), codes.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  EngineView v = eng->view();
  v.blk = (const uint64_t*)b_c.p;
  hipStream_t st = (hipStream_t)stream;
  for (int X = 0, ci = 0; X < 6; ++X)
    for (int Y = 0; Y <= X; ++Y, ++ci) {
      const int t0 = cls_t0[ci], nt = cls_t0[ci + 1] - t0;
      if (nt <= 0) continue;
      PQ_HIP(launch_eri_class(kPairL[X][0], kPairL[X][1], kPairL[Y][0], kPairL[Y][1], MODE_STORE, 0, v,
                              (const Tile*)b_t.p + t0, nt, (const int*)b_k.p, nullptr, nullptr, nullptr, out_dev,
                              nullptr, st));
    }
  // the request arrays are freed on return
  PQ_HIP(hipStreamSynchronize(st));
  return PQ_OK;
}

// D_dev, W_dev: the one-electron terms' density and energy-weighted density;
// Da_dev, Db_dev: the spin densities of the two-electron term (both null for
// a closed shell, whose two-electron term uses D_dev).
static int forces_impl(pq_engine* eng, const double* D_dev, const double* W_dev, const double* Da_dev,
                       const double* Db_dev, double* grad_dev, void* stream, float* ms) {
  if (!eng || !D_dev || !W_dev || !grad_dev) return fail(PQ_E_ARG, "null argument");
  if ((Da_dev == nullptr) != (Db_dev == nullptr)) return fail(PQ_E_ARG, "both spin densities are needed");
  if (!eng->work_ok) return fail(PQ_E_STATE, "pq_build_pairs must be called before pq_forces");
  hipStream_t st = (hipStream_t)stream;
  const int ncopy = kGradCopies;
  const size_t gsize = (size_t)ncopy * 3 * eng->natom;
  if (eng->gacc_n < gsize) {
    dev_free(eng->d_gacc);
    PQ_HIP(dev_alloc(&eng->d_gacc, gsize));
    eng->gacc_n = gsize;
  }
  // timing: one event pair per launch on the stream the kernel runs on
  std::vector<hipEvent_t> ev;
  std::vector<int> slot;
  auto mark = [&](hipStream_t s) -> hipError_t {
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    ev.push_back(e);
    return hipEventRecord(e, s);
  };
  int rc = PQ_OK;
  do {
    hipError_t err = hipSuccess;
#define PQ_TRY(x)                                \
  if ((err = (x)) != hipSuccess) {               \
    rc = fail(PQ_E_HIP, hipGetErrorString(err)); \
    break;                                       \
  }
    if (ms) PQ_TRY(mark(st));
    PQ_TRY(launch_fill(eng->d_gacc, (int64_t)gsize, 0.0, st));
    // class kernels round-robin over the side streams, as in the Fock build
    // (they only add into the gradient copies)
    PQ_TRY(fork_side(eng, st));
    int next = 0;
    auto ks = [&]() {
      hipStream_t s = eng->nside > 0 ? eng->side[next] : st;
      if (eng->nside > 0) next = (next + 1) % eng->nside;
      return s;
    };
    for (size_t ci = 0; ci < eng->work.size() && rc == PQ_OK; ++ci) {
      const auto& c = eng->work[ci];
      const int la = kPairL[c.X][0], lb = kPairL[c.X][1], lc = kPairL[c.Y][0], ld = kPairL[c.Y][1];
      for (int prec = 0; prec < 2; ++prec) {
        const auto& tl = prec ? c.t32 : c.t64;
        if (tl.empty()) continue;
        hipStream_t s = ks();
        if (ms) PQ_TRY(mark(s));
        EngineView gview = eng->view_s();
        gview.grad_db = Db_dev;
        PQ_TRY(launch_grad_class(la, lb, lc, ld, gview, prec ? c.d_t32 : c.d_t64, (int)tl.size(),
                                 eng->d_sorted_s[c.Y], Da_dev ? Da_dev : D_dev, eng->d_gacc, eng->natom, s));
        if (ms) {
          PQ_TRY(mark(s));
          slot.push_back((int)(2 * ci + prec));
        }
      }
    }
    if (rc != PQ_OK) break;
    for (int c = 0; c < 6; ++c) {
      if (eng->members[c].empty()) continue;
      hipStream_t s = ks();
      if (ms) PQ_TRY(mark(s));
      PQ_TRY(launch_onee_grad(kPairL[c][0], kPairL[c][1], eng->view(), eng->d_members[c],
                              (int)eng->members[c].size(), eng->d_atom_xyz, eng->d_atom_z, eng->natom, D_dev, W_dev,
                              eng->d_gacc, s));
      if (ms) {
        PQ_TRY(mark(s));
        slot.push_back(42 + c);
      }
    }
    if (rc != PQ_OK) break;
    PQ_TRY(join_side(eng, st));
    PQ_TRY(launch_grad_finish(eng->d_gacc, ncopy, eng->natom, eng->d_atom_xyz, eng->d_atom_z,
                              eng->nmol > 1 ? eng->d_atom_mol : nullptr, eng->d_mol_at0, eng->d_mol_nat,
                              grad_dev, st));
    if (ms) {
      PQ_TRY(mark(st));
      PQ_TRY(hipStreamSynchronize(st));
      for (int i = 0; i < 49; ++i) ms[i] = 0.f;
      for (size_t k = 0; k < slot.size(); ++k) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, ev[1 + 2 * k], ev[2 + 2 * k]);
        ms[slot[k]] += t;
      }
      float tot = 0.f;
      (void)hipEventElapsedTime(&tot, ev.front(), ev.back());
      ms[48] = tot;
    }
#undef PQ_TRY
  } while (0);
  for (auto e : ev) (void)hipEventDestroy(e);
  return rc;
}

int pq_forces(pq_engine* eng, const double* D_dev, const double* W_dev, double* grad_dev, void* stream) {
  return forces_impl(eng, D_dev, W_dev, nullptr, nullptr, grad_dev, stream, nullptr);
}

int pq_forces_uhf(pq_engine* eng, const double* Da_dev, const double* Db_dev, const double* D_dev,
                  const double* W_dev, double* grad_dev, void* stream) {
  if (!Da_dev || !Db_dev) return fail(PQ_E_ARG, "null argument");
  return forces_impl(eng, D_dev, W_dev, Da_dev, Db_dev, grad_dev, stream, nullptr);
}

int pq_forces_timed(pq_engine* eng, const double* D_dev, const double* W_dev, double* grad_dev, void* stream,
                    float* ms) {
  if (!ms) return fail(PQ_E_ARG, "null timing array");
  return forces_impl(eng, D_dev, W_dev, nullptr, nullptr, grad_dev, stream, ms);
}

int pq_get1ints(pq_engine* eng, double* S_dev, double* T_dev, double* V_dev, int compat_nr, void* stream) {
  if (!eng || !S_dev || !T_dev || !V_dev) return fail(PQ_E_ARG, "null argument");
  if (!eng->pairs_ok) {
    int rc = build_pair_data(eng);
    if (rc) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  for (int c = 0; c < 6; ++c) {
    if (eng->members[c].empty()) continue;
    PQ_HIP(launch_onee(kPairL[c][0], kPairL[c][1], eng->view(), eng->d_members[c], (int)eng->members[c].size(),
                       eng->d_atom_xyz, eng->d_atom_z, eng->natom, S_dev, T_dev, V_dev, compat_nr ? 1 : 0, st));
  }
  return PQ_OK;
}

}  // extern "C"
