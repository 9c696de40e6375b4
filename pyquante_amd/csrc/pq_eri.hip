// pq_eri.hip -- dispatch of the per-class Rys ERI kernels (each class is
// compiled in its own translation unit from pq_eri_cls.hip), the stored-array
// digestion kernel and the J/K symmetrisation.
#include "pq_device.h"
#include "pq_fixed.h"
#include "pq_kernels.h"

namespace pq {

#define PQ_DECL1(A, B, C_, D_, M)                                                                    \
  hipError_t launch_cls_##A##B##C_##D_##_m##M(const EngineView& ev, const Tile* tiles, int ntiles,         \
                                              const int* kets, const double* D, double* Jacc, double* Kacc, \
                                              double* ints, double* qout, hipStream_t st, const int2* groups, \
                                              int ngroups, int gflags);
#define PQ_DECL(A, B, C_, D_) \
  PQ_DECL1(A, B, C_, D_, 0) PQ_DECL1(A, B, C_, D_, 1) PQ_DECL1(A, B, C_, D_, 2) PQ_DECL1(A, B, C_, D_, 5) \
  PQ_DECL1(A, B, C_, D_, 6) PQ_DECL1(A, B, C_, D_, 7) PQ_DECL1(A, B, C_, D_, 8)
PQ_DECL(0, 0, 0, 0)
PQ_DECL(1, 0, 0, 0)
PQ_DECL(1, 0, 1, 0)
PQ_DECL(1, 1, 0, 0)
PQ_DECL(1, 1, 1, 0)
PQ_DECL(1, 1, 1, 1)
PQ_DECL(2, 0, 0, 0)
PQ_DECL(2, 0, 1, 0)
PQ_DECL(2, 0, 1, 1)
PQ_DECL(2, 0, 2, 0)
PQ_DECL(2, 1, 0, 0)
PQ_DECL(2, 1, 1, 0)
PQ_DECL(2, 1, 1, 1)
PQ_DECL(2, 1, 2, 0)
PQ_DECL(2, 1, 2, 1)
PQ_DECL(2, 2, 0, 0)
PQ_DECL(2, 2, 1, 0)
PQ_DECL(2, 2, 1, 1)
PQ_DECL(2, 2, 2, 0)
PQ_DECL(2, 2, 2, 1)
PQ_DECL(2, 2, 2, 2)
PQ_DECL1(0, 0, 0, 0, 3)
PQ_DECL1(1, 0, 1, 0, 3)
PQ_DECL1(1, 1, 1, 1, 3)
PQ_DECL1(2, 0, 2, 0, 3)
PQ_DECL1(2, 1, 2, 1, 3)
PQ_DECL1(2, 2, 2, 2, 3)
#undef PQ_DECL

#define PQ_GDECL(A, B, C_, D_)                                                                     \
  hipError_t launch_cls_##A##B##C_##D_##_m4(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets, \
                                            const double* D, double* gacc, int natom, hipStream_t st);
PQ_GDECL(0, 0, 0, 0)
PQ_GDECL(1, 0, 0, 0)
PQ_GDECL(1, 0, 1, 0)
PQ_GDECL(1, 1, 0, 0)
PQ_GDECL(1, 1, 1, 0)
PQ_GDECL(1, 1, 1, 1)
PQ_GDECL(2, 0, 0, 0)
PQ_GDECL(2, 0, 1, 0)
PQ_GDECL(2, 0, 1, 1)
PQ_GDECL(2, 0, 2, 0)
PQ_GDECL(2, 1, 0, 0)
PQ_GDECL(2, 1, 1, 0)
PQ_GDECL(2, 1, 1, 1)
PQ_GDECL(2, 1, 2, 0)
PQ_GDECL(2, 1, 2, 1)
PQ_GDECL(2, 2, 0, 0)
PQ_GDECL(2, 2, 1, 0)
PQ_GDECL(2, 2, 1, 1)
PQ_GDECL(2, 2, 2, 0)
PQ_GDECL(2, 2, 2, 1)
PQ_GDECL(2, 2, 2, 2)
#undef PQ_GDECL
#undef PQ_DECL1

// ---------------------------------------------------------------------
// J = Jacc + Jacc^T, K = Kacc + Kacc^T (outputs may be null).
// ---------------------------------------------------------------------
__global__ void k_symmetrize(int n, const double* __restrict__ Jacc, const double* __restrict__ Kacc,
                             double* J, double* K, double jscale, double kscale) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int i = (int)(idx / n), j = (int)(idx % n);
  if (J) J[idx] = jscale * (Jacc[idx] + Jacc[(int64_t)j * n + i]);
  if (K) K[idx] = kscale * (Kacc[idx] + Kacc[(int64_t)j * n + i]);
}

// The same per molecule block of a packed set of matrices: block m starts
// at moff[m] with n = mn[m] (blockIdx.y = m).
__global__ void k_symmetrize_packed(const int* __restrict__ moff, const int* __restrict__ mn,
                                    const double* __restrict__ Jacc, const double* __restrict__ Kacc, double* J,
                                    double* K) {
  const int m = blockIdx.y, n = mn[m];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int i = idx / n, j = idx % n;
  const int64_t o = moff[m];
  if (J) J[o + idx] = Jacc[o + idx] + Jacc[o + j * n + i];
  if (K) K[o + idx] = Kacc[o + idx] + Kacc[o + j * n + i];
}

hipError_t launch_symmetrize_packed(int nmol, int maxn, const int* moff, const int* mn, const double* Jacc,
                                    const double* Kacc, double* J, double* K, hipStream_t st) {
  if (nmol <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_symmetrize_packed, dim3((maxn * maxn + 255) / 256, nmol), dim3(256), 0, st, moff, mn, Jacc,
                     Kacc, J, K);
  return hipGetLastError();
}

// ---------------------------------------------------------------------
// Digestion of a stored flat array Ints[ijkl2intindex(i,j,k,l)]: entry t
// is decoded back to (ij, kl) with ij >= kl and then to i >= j, k >= l; the
// same accumulation rule as the direct path applies with the function-level
// degeneracy factor.
// ---------------------------------------------------------------------

__global__ void k_digest_stored(int n, const double* __restrict__ ints, uint64_t len, const double* __restrict__ D,
                                double* Jacc, double* Kacc) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; t < len; t += stride) {
    const double v = ints[t];
    if (v == 0.0) continue;
    const int64_t ij = tri_root(t), kl = (int64_t)(t - (uint64_t)ij * (ij + 1) / 2);
    const int64_t i = tri_root(ij), j = ij - i * (i + 1) / 2;
    const int64_t k = tri_root(kl), l = kl - k * (k + 1) / 2;
    double f = 1.0;
    if (i == j) f *= 0.5;
    if (k == l) f *= 0.5;
    if (ij == kl) f *= 0.5;
    const double fv = f * v;
    unsafeAtomicAdd(&Jacc[i * n + j], 2.0 * fv * D[k * n + l]);
    unsafeAtomicAdd(&Jacc[k * n + l], 2.0 * fv * D[i * n + j]);
    unsafeAtomicAdd(&Kacc[i * n + k], fv * D[j * n + l]);
    unsafeAtomicAdd(&Kacc[i * n + l], fv * D[j * n + k]);
    unsafeAtomicAdd(&Kacc[j * n + k], fv * D[i * n + l]);
    unsafeAtomicAdd(&Kacc[j * n + l], fv * D[i * n + k]);
  }
}

hipError_t launch_digest_stored(int n, const double* ints, uint64_t len, const double* D, double* Jacc,
                                double* Kacc, hipStream_t st) {
  if (len == 0) return hipSuccess;
  uint64_t blocks = (len + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_digest_stored, dim3((unsigned)blocks), dim3(256), 0, st, n, ints, len, D, Jacc, Kacc);
  return hipGetLastError();
}

// Rows mu0 .. mu0+nmu-1 of the full 4-index tensor (mu nu|lam sig), row-major,
// from the flat unique array: out[((mu-mu0) n + nu) n^2 + lam n + sig] =
// ints[ijkl2intindex(mu, nu, lam, sig)] (the reference's unpacking,
// MP.py:38-46).  Writes are coalesced; reads gather.
__global__ void k_unpack_eri(int n, const double* __restrict__ ints, int mu0, uint64_t total, double* out) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t n2 = (uint64_t)n * n, n3 = n2 * n;
  for (; t < total; t += stride) {
    const int mu = mu0 + (int)(t / n3);
    const uint64_t r = t % n3;
    const int nu = (int)(r / n2), lam = (int)((r / n) % n), sig = (int)(r % n);
    out[t] = ints[ijkl2intindex(mu, nu, lam, sig)];
  }
}

hipError_t launch_unpack_eri(int n, const double* ints, int mu0, int nmu, double* out, hipStream_t st) {
  const uint64_t total = (uint64_t)nmu * n * n * n;
  if (total == 0) return hipSuccess;
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_unpack_eri, dim3((unsigned)blocks), dim3(256), 0, st, n, ints, mu0, total, out);
  return hipGetLastError();
}

static hipError_t launch_diag_0000(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets, const double* D,
                                   double* Jacc, double* Kacc, double* ints, double* qout, hipStream_t st) {
  return launch_cls_0000_m3(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, nullptr, 0, 0);
}
static hipError_t launch_diag_1000(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_1010(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets, const double* D,
                                   double* Jacc, double* Kacc, double* ints, double* qout, hipStream_t st) {
  return launch_cls_1010_m3(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, nullptr, 0, 0);
}
static hipError_t launch_diag_1100(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_1110(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_1111(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets, const double* D,
                                   double* Jacc, double* Kacc, double* ints, double* qout, hipStream_t st) {
  return launch_cls_1111_m3(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, nullptr, 0, 0);
}
static hipError_t launch_diag_2000(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2010(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2011(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2020(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets, const double* D,
                                   double* Jacc, double* Kacc, double* ints, double* qout, hipStream_t st) {
  return launch_cls_2020_m3(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, nullptr, 0, 0);
}
static hipError_t launch_diag_2100(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2110(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2111(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2120(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2121(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets, const double* D,
                                   double* Jacc, double* Kacc, double* ints, double* qout, hipStream_t st) {
  return launch_cls_2121_m3(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, nullptr, 0, 0);
}
static hipError_t launch_diag_2200(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2210(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2211(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2220(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2221(const EngineView&, const Tile*, int, const int*, const double*, double*, double*,
                                   double*, double*, hipStream_t) {
  return hipErrorInvalidValue;
}
static hipError_t launch_diag_2222(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets, const double* D,
                                   double* Jacc, double* Kacc, double* ints, double* qout, hipStream_t st) {
  return launch_cls_2222_m3(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, nullptr, 0, 0);
}

// Adds the position-indexed ket J of every pair class into Jacc for nd
// stacked densities, one thread per (class, component, ket position).  Every
// (pair, component) owns one matrix element, so plain read-modify-writes
// suffice once the J/K kernels have finished.
__global__ void k_jpair_scatter(EngineView ev, JpairMap map, int nd, double* Jacc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ev.jp_size) return;
  // class blocks are consecutive; an empty class shares its offset with the
  // next one, so the last class starting at or before i owns it
  int Y = 0;
#pragma unroll
  for (int y = 1; y < 6; ++y)
    if (i >= ev.jp_off[y]) Y = y;
  const int np = ev.jp_n[Y], ndf = map.ndf[Y];
  const int local = i - ev.jp_off[Y];
  const int cd = local / np, j = local - cd * np;
  const PairInfo p = ev.pairs[map.ids[Y][j]];
  const int64_t e = p.moff + (int64_t)(p.bfA + cd / ndf) * p.mstride + p.bfB + cd % ndf;
  for (int d = 0; d < nd; ++d) {
    const double v = ev.jpair[d * ev.jp_size + i];
    if (v != 0.0) Jacc[d * ev.msize + e] += v;
  }
}

hipError_t launch_jpair_scatter(const EngineView& ev, const JpairMap& map, int nd, double* Jacc, hipStream_t st) {
  if (ev.jp_size <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jpair_scatter, dim3((unsigned)((ev.jp_size + 255) / 256)), dim3(256), 0, st, ev, map, nd,
                     Jacc);
  return hipGetLastError();
}

hipError_t launch_eri_class(int la, int lb, int lc, int ld, int mode, int fp32, const EngineView& ev,
                            const Tile* tiles, int ntiles, const int* kets, const double* D, double* Jacc,
                            double* Kacc, double* ints, double* qout, hipStream_t st, const int2* groups,
                            int ngroups, int gflags) {
#define PQ_CASE(A, B, C_, D_)                                                                      \
  if (la == A && lb == B && lc == C_ && ld == D_) {                                                \
    if (mode == MODE_JK && !fp32)                                                                  \
      return launch_cls_##A##B##C_##D_##_m0(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, groups, ngroups, gflags); \
    if (mode == MODE_JK)                                                                           \
      return launch_cls_##A##B##C_##D_##_m1(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, groups, ngroups, gflags); \
    if (mode == MODE_STORE)                                                                        \
      return launch_cls_##A##B##C_##D_##_m2(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, nullptr, 0, 0); \
    if (mode == MODE_JK2)                                                                          \
      return launch_cls_##A##B##C_##D_##_m5(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, groups, ngroups, gflags); \
    if (mode == MODE_JKD && !fp32)                                                                 \
      return launch_cls_##A##B##C_##D_##_m6(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, groups, ngroups, gflags); \
    if (mode == MODE_JKD)                                                                          \
      return launch_cls_##A##B##C_##D_##_m7(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, groups, ngroups, gflags); \
    if (mode == MODE_JK2D)                                                                         \
      return launch_cls_##A##B##C_##D_##_m8(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st, groups, ngroups, gflags); \
    return launch_diag_##A##B##C_##D_(ev, tiles, ntiles, kets, D, Jacc, Kacc, ints, qout, st);      \
  }
  PQ_CASE(0, 0, 0, 0)
  PQ_CASE(1, 0, 0, 0)
  PQ_CASE(1, 0, 1, 0)
  PQ_CASE(1, 1, 0, 0)
  PQ_CASE(1, 1, 1, 0)
  PQ_CASE(1, 1, 1, 1)
  PQ_CASE(2, 0, 0, 0)
  PQ_CASE(2, 0, 1, 0)
  PQ_CASE(2, 0, 1, 1)
  PQ_CASE(2, 0, 2, 0)
  PQ_CASE(2, 1, 0, 0)
  PQ_CASE(2, 1, 1, 0)
  PQ_CASE(2, 1, 1, 1)
  PQ_CASE(2, 1, 2, 0)
  PQ_CASE(2, 1, 2, 1)
  PQ_CASE(2, 2, 0, 0)
  PQ_CASE(2, 2, 1, 0)
  PQ_CASE(2, 2, 1, 1)
  PQ_CASE(2, 2, 2, 0)
  PQ_CASE(2, 2, 2, 1)
  PQ_CASE(2, 2, 2, 2)
#undef PQ_CASE
  return hipErrorInvalidValue;
}


#define PQ_HDECL(A, B, C_, D_)                                                                          \
  hipError_t launch_hgp_##A##B##C_##D_(const EngineView& ev, const Tile* tiles, int ntiles, const int* kets, \
                                       double* ints, hipStream_t st);                                     \
  size_t hgp_private_##A##B##C_##D_();
PQ_HDECL(0, 0, 0, 0)
PQ_HDECL(1, 0, 0, 0)
PQ_HDECL(1, 0, 1, 0)
PQ_HDECL(1, 1, 0, 0)
PQ_HDECL(1, 1, 1, 0)
PQ_HDECL(1, 1, 1, 1)
PQ_HDECL(2, 0, 0, 0)
PQ_HDECL(2, 0, 1, 0)
PQ_HDECL(2, 0, 1, 1)
PQ_HDECL(2, 0, 2, 0)
PQ_HDECL(2, 1, 0, 0)
PQ_HDECL(2, 1, 1, 0)
PQ_HDECL(2, 1, 1, 1)
PQ_HDECL(2, 1, 2, 0)
PQ_HDECL(2, 1, 2, 1)
PQ_HDECL(2, 2, 0, 0)
PQ_HDECL(2, 2, 1, 0)
PQ_HDECL(2, 2, 1, 1)
PQ_HDECL(2, 2, 2, 0)
PQ_HDECL(2, 2, 2, 1)
PQ_HDECL(2, 2, 2, 2)
#undef PQ_HDECL

hipError_t launch_hgp_class(int la, int lb, int lc, int ld, const EngineView& ev, const Tile* tiles, int ntiles,
                            const int* kets, double* ints, hipStream_t st) {
#define PQ_HCASE(A, B, C_, D_) \
  if (la == A && lb == B && lc == C_ && ld == D_) return launch_hgp_##A##B##C_##D_(ev, tiles, ntiles, kets, ints, st);
  PQ_HCASE(0, 0, 0, 0)
  PQ_HCASE(1, 0, 0, 0)
  PQ_HCASE(1, 0, 1, 0)
  PQ_HCASE(1, 1, 0, 0)
  PQ_HCASE(1, 1, 1, 0)
  PQ_HCASE(1, 1, 1, 1)
  PQ_HCASE(2, 0, 0, 0)
  PQ_HCASE(2, 0, 1, 0)
  PQ_HCASE(2, 0, 1, 1)
  PQ_HCASE(2, 0, 2, 0)
  PQ_HCASE(2, 1, 0, 0)
  PQ_HCASE(2, 1, 1, 0)
  PQ_HCASE(2, 1, 1, 1)
  PQ_HCASE(2, 1, 2, 0)
  PQ_HCASE(2, 1, 2, 1)
  PQ_HCASE(2, 2, 0, 0)
  PQ_HCASE(2, 2, 1, 0)
  PQ_HCASE(2, 2, 1, 1)
  PQ_HCASE(2, 2, 2, 0)
  PQ_HCASE(2, 2, 2, 1)
  PQ_HCASE(2, 2, 2, 2)
#undef PQ_HCASE
  return hipErrorInvalidValue;
}

size_t hgp_class_private_bytes(int la, int lb, int lc, int ld) {
#define PQ_PCASE(A, B, C_, D_) \
  if (la == A && lb == B && lc == C_ && ld == D_) return hgp_private_##A##B##C_##D_();
  PQ_PCASE(0, 0, 0, 0)
  PQ_PCASE(1, 0, 0, 0)
  PQ_PCASE(1, 0, 1, 0)
  PQ_PCASE(1, 1, 0, 0)
  PQ_PCASE(1, 1, 1, 0)
  PQ_PCASE(1, 1, 1, 1)
  PQ_PCASE(2, 0, 0, 0)
  PQ_PCASE(2, 0, 1, 0)
  PQ_PCASE(2, 0, 1, 1)
  PQ_PCASE(2, 0, 2, 0)
  PQ_PCASE(2, 1, 0, 0)
  PQ_PCASE(2, 1, 1, 0)
  PQ_PCASE(2, 1, 1, 1)
  PQ_PCASE(2, 1, 2, 0)
  PQ_PCASE(2, 1, 2, 1)
  PQ_PCASE(2, 2, 0, 0)
  PQ_PCASE(2, 2, 1, 0)
  PQ_PCASE(2, 2, 1, 1)
  PQ_PCASE(2, 2, 2, 0)
  PQ_PCASE(2, 2, 2, 1)
  PQ_PCASE(2, 2, 2, 2)
#undef PQ_PCASE
  return SIZE_MAX;
}

hipError_t launch_grad_class(int la, int lb, int lc, int ld, const EngineView& ev, const Tile* tiles, int ntiles,
                             const int* kets, const double* D, double* gacc, int natom, hipStream_t st) {
#define PQ_GCASE(A, B, C_, D_)                            \
  if (la == A && lb == B && lc == C_ && ld == D_)         \
    return launch_cls_##A##B##C_##D_##_m4(ev, tiles, ntiles, kets, D, gacc, natom, st);
  PQ_GCASE(0, 0, 0, 0)
  PQ_GCASE(1, 0, 0, 0)
  PQ_GCASE(1, 0, 1, 0)
  PQ_GCASE(1, 1, 0, 0)
  PQ_GCASE(1, 1, 1, 0)
  PQ_GCASE(1, 1, 1, 1)
  PQ_GCASE(2, 0, 0, 0)
  PQ_GCASE(2, 0, 1, 0)
  PQ_GCASE(2, 0, 1, 1)
  PQ_GCASE(2, 0, 2, 0)
  PQ_GCASE(2, 1, 0, 0)
  PQ_GCASE(2, 1, 1, 0)
  PQ_GCASE(2, 1, 1, 1)
  PQ_GCASE(2, 1, 2, 0)
  PQ_GCASE(2, 1, 2, 1)
  PQ_GCASE(2, 2, 0, 0)
  PQ_GCASE(2, 2, 1, 0)
  PQ_GCASE(2, 2, 1, 1)
  PQ_GCASE(2, 2, 2, 0)
  PQ_GCASE(2, 2, 2, 1)
  PQ_GCASE(2, 2, 2, 2)
#undef PQ_GCASE
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------
// The deterministic build's fixed-point accumulators (pq_fixed.h): limb k of
// J/K element e at acc[k * msize + e], of ket position i at
// jpair[k * jp_size + i].  Integer adds throughout, so the result does not
// depend on the order of any accumulation.
// ---------------------------------------------------------------------
__global__ void k_jpair_scatter_fx(EngineView ev, JpairMap map, long long* Jfx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ev.jp_size) return;
  int Y = 0;
#pragma unroll
  for (int y = 1; y < 6; ++y)
    if (i >= ev.jp_off[y]) Y = y;
  const int np = ev.jp_n[Y], ndf = map.ndf[Y];
  const int local = i - ev.jp_off[Y];
  const int cd = local / np, j = local - cd * np;
  const PairInfo p = ev.pairs[map.ids[Y][j]];
  const int64_t e = p.moff + (int64_t)(p.bfA + cd / ndf) * p.mstride + p.bfB + cd % ndf;
  const long long* jp = reinterpret_cast<const long long*>(ev.jpair);
  for (int k = 0; k < 3; ++k) {
    const long long v = jp[k * ev.jp_size + i];
    if (v) Jfx[k * ev.msize + e] += v;
  }
}

hipError_t launch_jpair_scatter_fx(const EngineView& ev, const JpairMap& map, long long* Jfx, hipStream_t st) {
  if (ev.jp_size <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jpair_scatter_fx, dim3((unsigned)((ev.jp_size + 255) / 256)), dim3(256), 0, st, ev, map,
                     Jfx);
  return hipGetLastError();
}

// J = Jacc + Jacc^T, K = Kacc + Kacc^T from the limb sums (added as integers,
// then converted once), per molecule block m (blockIdx.y): block m starts at
// moff[m] with n = mn[m], or at 0 with n = n1 when moff is null.
__global__ void k_symmetrize_fx(const int* __restrict__ moff, const int* __restrict__ mn, int n1, int64_t msize,
                                const long long* __restrict__ Jfx, const long long* __restrict__ Kfx, double* J,
                                double* K) {
  const int m = blockIdx.y;
  const int n = moff ? mn[m] : n1;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int i = (int)(idx / n), j = (int)(idx % n);
  const int64_t o = moff ? moff[m] : 0;
  const int64_t a = o + idx, b = o + (int64_t)j * n + i;
  if (J)
    J[a] = fx_value(Jfx[a] + Jfx[b], Jfx[msize + a] + Jfx[msize + b], Jfx[2 * msize + a] + Jfx[2 * msize + b]);
  if (K)
    K[a] = fx_value(Kfx[a] + Kfx[b], Kfx[msize + a] + Kfx[msize + b], Kfx[2 * msize + a] + Kfx[2 * msize + b]);
}

hipError_t launch_symmetrize_fx(int nmol, int maxn, const int* moff, const int* mn, int64_t msize,
                                const long long* Jfx, const long long* Kfx, double* J, double* K, hipStream_t st) {
  if (nmol <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_symmetrize_fx, dim3((unsigned)(((int64_t)maxn * maxn + 255) / 256), nmol), dim3(256), 0, st,
                     moff, mn, maxn, msize, Jfx, Kfx, J, K);
  return hipGetLastError();
}

hipError_t launch_symmetrize(int n, const double* Jacc, const double* Kacc, double* J, double* K, double jscale,
                             double kscale, hipStream_t st) {
  const int64_t tot = (int64_t)n * n;
  hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, Jacc, Kacc, J, K,
                     jscale, kscale);
  return hipGetLastError();
}

}  // namespace pq
