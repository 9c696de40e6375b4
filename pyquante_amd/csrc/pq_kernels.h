// pq_kernels.h -- launch interface between the host engine and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "pq_device.h"

namespace pq {

enum { MODE_JK = 0, MODE_STORE = 1, MODE_DIAG = 2, MODE_JK2 = 3, MODE_JKD = 4, MODE_JK2D = 5 };
// J/K modes and the number of densities they digest (MODE_JK2: two stacked
// n x n densities into two stacked J/K accumulators).  MODE_JKD and
// MODE_JK2D are the deterministic builds of one and two densities: every
// accumulator element (global J/K, position-indexed ket J, LDS K rows and J
// sums) is three int64 fixed-point limbs (pq_fixed.h) at a stride of one
// array apart, so the sums do not depend on the order of the atomics; the
// limbs of density d start NL arrays after those of density d - 1.
constexpr bool mode_is_jk(int m) { return m == MODE_JK || m == MODE_JK2 || m == MODE_JKD || m == MODE_JK2D; }
constexpr int mode_ndens(int m) { return m == MODE_JK2 || m == MODE_JK2D ? 2 : 1; }
constexpr int mode_limbs(int m) { return m == MODE_JKD || m == MODE_JK2D ? 3 : 1; }

// classes with at most this many Cartesian integrals run one lane per quartet
constexpr int kLaneMaxComp = 108;

// private copies of the natom x 3 gradient that the gradient kernels spread
// their atomics over (summed by launch_grad_finish)
constexpr int kGradCopies = 256;

// groups (J/K mode, unsplit classes): ngroups entries of (first tile, tile
// count), each a run of tiles with the same bra pair processed by one block;
// gflags & kGroupStageD: those blocks also stage the bra rows of D in LDS
// (when they fit)
constexpr int kGroupStageD = 1;
hipError_t launch_eri_class(int la, int lb, int lc, int ld, int mode, int fp32, const EngineView& ev,
                            const Tile* tiles, int ntiles, const int* kets, const double* D, double* Jacc,
                            double* Kacc, double* ints, double* qout, hipStream_t st,
                            const int2* groups = nullptr, int ngroups = 0, int gflags = 0);
// Shell-quartet Head-Gordon-Pople kernel of one class over store tiles
// (the tiles and kets of MODE_STORE), writing the flat unique-ERI array.
hipError_t launch_hgp_class(int la, int lb, int lc, int ld, const EngineView& ev, const Tile* tiles, int ntiles,
                            const int* kets, double* ints, hipStream_t st);
// Per-lane private memory of that class's kernel (SIZE_MAX if unknown).
size_t hgp_class_private_bytes(int la, int lb, int lc, int ld);
// The HGP classes run in one launch (hgp_multi_class: the class is in it;
// split = its output parts per tile): segment i covers blocks [b0, next b0)
// of class code = hgp_code(la, lb, lc, ld), block b0 + tile * split + part.
constexpr int hgp_code(int a, int b, int c, int d) { return ((a * 3 + b) * 3 + c) * 3 + d; }
struct HgpSeg {
  const Tile* tiles;
  const int* kets;
  int code, split, b0, pad;
};
constexpr int kHgpMaxSeg = 24;
struct HgpMulti {
  HgpSeg seg[kHgpMaxSeg];
  int nseg;
};
bool hgp_multi_class(int la, int lb, int lc, int ld, int* split);

// One launch over the bra runs (FP64 tiles) of the light J/K classes
// (jk_multi_class): segment i covers blocks [b0, next b0) of class code =
// hgp_code(la, lb, lc, ld), its bra runs `groups` over `tiles` and ket list.
struct JkSeg {
  const Tile* tiles;
  const int2* groups;
  const int* kets;
  int code, b0, stage_d, pad;
};
constexpr int kJkMaxSeg = 8;
struct JkMulti {
  JkSeg seg[kJkMaxSeg];
  int nseg;
};
bool jk_multi_class(int la, int lb, int lc, int ld);
// dynamic LDS one block of that class's bra runs needs
size_t jk_multi_lds(int la, int lb, int lc, int ld, int nbf, int stage_d);
hipError_t launch_jk_multi(const EngineView& ev, const JkMulti& m, int nblocks, size_t lds, const double* D,
                           double* Jacc, double* Kacc, hipStream_t st);
hipError_t launch_hgp_multi(const EngineView& ev, const HgpMulti& m, int nblocks, double* ints, hipStream_t st);

// ket order of each pair class (ids[Y][j] = pair at position j) and the
// function count of its second shell
struct JpairMap {
  const int* ids[6];
  int ndf[6];
};
// adds the position-indexed ket J of every pair class (ev.jpair) into nd
// stacked Jacc matrices, one launch
hipError_t launch_jpair_scatter(const EngineView& ev, const JpairMap& map, int nd, double* Jacc, hipStream_t st);

// zeroes three arrays in one launch (fill kernel, safe under graph capture)
hipError_t launch_zero3(double* a, int64_t na, double* b, int64_t nb, double* c, int64_t nc, hipStream_t st);

// two-electron nuclear gradient into kGradCopies private copies of the
// natom x 3 gradient (gacc), summed by launch_grad_finish
hipError_t launch_grad_class(int la, int lb, int lc, int ld, const EngineView& ev, const Tile* tiles, int ntiles,
                             const int* kets, const double* D, double* gacc, int natom, hipStream_t st);

// one-electron gradient (2 tr(D dh) - 2 tr(W dS)) of the pairs of one class
hipError_t launch_onee_grad(int la, int lb, const EngineView& ev, const int* pair_ids, int npairs,
                            const double* atom_xyz, const double* atom_z, int natom, const double* D,
                            const double* W, double* gacc, hipStream_t st);

// grad = sum of the copies + nuclear-repulsion gradient
hipError_t launch_grad_finish(const double* gacc, int ncopies, int natom, const double* atom_xyz,
                              const double* atom_z, const int* atom_mol, const int* mol_at0, const int* mol_nat,
                              double* grad, hipStream_t st);

// per-molecule blocks of packed matrices (block m at moff[m], n = mn[m] <= maxn)
hipError_t launch_symmetrize_packed(int nmol, int maxn, const int* moff, const int* mn, const double* Jacc,
                                    const double* Kacc, double* J, double* K, hipStream_t st);
hipError_t launch_symmetrize(int n, const double* Jacc, const double* Kacc, double* J, double* K,
                             double jscale, double kscale, hipStream_t st);
// MODE_JKD: the ket J limbs into the J limbs, and J/K from the limb sums
// (moff null: one molecule of maxn functions)
hipError_t launch_jpair_scatter_fx(const EngineView& ev, const JpairMap& map, long long* Jfx, hipStream_t st);
hipError_t launch_symmetrize_fx(int nmol, int maxn, const int* moff, const int* mn, int64_t msize,
                                const long long* Jfx, const long long* Kfx, double* J, double* K, hipStream_t st);

// J/K from a stored flat unique-integral array (Ints.py:195-223 over Ints)
hipError_t launch_digest_stored(int n, const double* ints, uint64_t len, const double* D, double* Jacc,
                                double* Kacc, hipStream_t st);

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t st);

// one-electron S, T, V (full matrices, row-major n x n); compat: V with the
// reference's EPS = 3e-7 Boys function (its values)
hipError_t launch_onee(int la, int lb, const EngineView& ev, const int* pair_ids, int npairs,
                       const double* atom_xyz, const double* atom_z, int natom, double* S, double* T,
                       double* V, int compat, hipStream_t st);

// scalar-compatibility batches
// HGP contracted ERIs per function quartet (batch layout of launch_contr_eri);
// mext (nullable) selects the auxiliary order of a primitive VRR leaf
hipError_t launch_hgp_contr(int n, const int* nprim, const int* poff, const double* exps, const double* coefs,
                            const double* norms, const double* origins, const int* powers, const int* mext,
                            int compat_nr, double* out, hipStream_t st);
// HGP for every unique function quartet of a basis, reference flat layout
hipError_t launch_hgp_store(uint64_t len, const int* fnp, const int* foff, const double* forg, const int* fpw,
                            const double* fexps, const double* fcoefs, const double* fnorms, int compat_nr,
                            double* out, hipStream_t st);
// rows [mu0, mu0+nmu) of the full (mu nu|lam sig) tensor from the flat array
hipError_t launch_unpack_eri(int n, const double* ints, int mu0, int nmu, double* out, hipStream_t st);
hipError_t launch_boys(int n, const double* m, const double* x, double eps, double* out, hipStream_t st);
hipError_t launch_rys_roots(int n, int nr, const double* X, double* u, double* w, hipStream_t st);
hipError_t launch_contr_eri(int n, const int* nprim, const int* poff, const double* exps,
                            const double* coefs, const double* norms, const double* origins,
                            const int* powers, double* out, hipStream_t st);
// the same batch by the coefficients of the integrand in u contracted with
// Boys values (compat_nr = 1: the reference's EPS = 3e-7 Boys function, so
// the values of its THO cints path); L_total <= 9
hipError_t launch_tho_contr(int n, const int* nprim, const int* poff, const double* exps, const double* coefs,
                            const double* norms, const double* origins, const int* powers, int compat_nr,
                            double* out, hipStream_t st);

}  // namespace pq
