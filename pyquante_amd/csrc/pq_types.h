// pq_types.h -- the plain records shared by the device kernels and the host
// work-list builder: shell-pair and primitive-pair records as they lie in
// HBM, and the tiles of the class work lists.  No HIP dependency, so that the
// work-list builder (pq_worklist.h) also compiles with a host compiler.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define PQ_HD __host__ __device__
#else
#define PQ_HD
#endif

namespace pq {

// Cartesian components of a shell of angular momentum L.
PQ_HD constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }

// ---------------------------------------------------------------------
// Shell-pair records in HBM.  A pair (A,B) is stored with l_A >= l_B.
// ---------------------------------------------------------------------
struct alignas(16) PrimPair {
  double p;    // alpha_a + alpha_b
  double rp;   // 1 / p
  double Kp;   // K / p, K = c_a c_b Nrad_a Nrad_b exp(-alpha_a alpha_b |AB|^2 / p)
  double Px, Py, Pz;
};

// Matrices (D, J, K, S, ...) are stored per molecule: molecule m's n_m x n_m
// block starts at element moff and has row stride mstride = n_m; bfA/bfB are
// function indices within the molecule.  A single-molecule engine has
// moff = 0 and mstride = nbf.  Pairs never mix molecules.
struct alignas(16) PairInfo {
  double Ax, Ay, Az;     // centre of shell A
  double ABx, ABy, ABz;  // A - B
  int pp_off, pp_n;      // primitive pairs [pp_off, pp_off + pp_n)
  int bfA, bfB;          // first basis function of each shell (within its molecule)
  int shA, shB;          // shell indices
  int atA, atB;          // atom indices
  int moff, mstride;     // the molecule's matrix block: first element, row stride
  int mat0, mnat;        // the molecule's atoms [mat0, mat0 + mnat)
};

struct Tile {
  int bra;   // pair id
  int ket0;  // first entry of the ket pair list covered by this tile
  int nket;  // number of kets (bits 0-7) and log2 of the ket-primitive split ks (bits 8-10)
  int prim;  // bra primitive-pair range: first (low 16 bits), count (high 16 bits; 0 = all)
};

}  // namespace pq
