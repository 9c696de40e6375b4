// pq_onee.hip -- one-electron integrals (S, T, V) and the scalar-compatibility
// kernels behind the drop-in cints/crys/chgp functions.
#include <math.h>

#include <algorithm>

#include "pq_device.h"
#include "pq_kernels.h"

namespace pq {

// ---------------------------------------------------------------------
// Runtime-angular-momentum 1-D Rys integrals.  The Rys integrand of a
// primitive ERI (or nuclear-attraction integral) is, after the Gaussian
// transform, prod_d I_d(u) e^(-X u) integrated over t in [0, 1], u = t^2,
// and every I_d is a polynomial in u of degree <= its total power.  The
// same recurrence therefore gives either I_d at one Rys root (S = double)
// or the coefficients of I_d in u (S = UPoly); contracting the product's
// coefficients c_k with Boys values F_k(X) = int_0^1 t^(2k) e^(-X t^2) dt
// gives sum_k c_k F_k, the form in which the reference's THO expressions
// (cints.c:110-187 ERIs, 241-307 nuclear attraction) consume their Boys
// function.  With the reference's EPS = 3e-7 Boys values that reproduces
// its numbers to rounding.
// ---------------------------------------------------------------------
constexpr int kRtMaxL = 9;   // largest total angular momentum of these paths
constexpr int kRtMaxG = 30;  // (N+1)(M+1) for N + M <= kRtMaxL

struct UPoly {
  double c[kRtMaxL + 1];
};
__device__ __forceinline__ UPoly up_lin(double a, double b) {
  UPoly r;
#pragma unroll
  for (int k = 0; k <= kRtMaxL; ++k) r.c[k] = 0.0;
  r.c[0] = a;
  r.c[1] = b;
  return r;
}
__device__ __forceinline__ UPoly operator+(const UPoly& a, const UPoly& b) {
  UPoly r;
#pragma unroll
  for (int k = 0; k <= kRtMaxL; ++k) r.c[k] = a.c[k] + b.c[k];
  return r;
}
__device__ __forceinline__ UPoly operator*(double s, const UPoly& a) {
  UPoly r;
#pragma unroll
  for (int k = 0; k <= kRtMaxL; ++k) r.c[k] = s * a.c[k];
  return r;
}
// product truncated at degree kRtMaxL (the callers never exceed it)
__device__ __forceinline__ UPoly operator*(const UPoly& a, const UPoly& b) {
  UPoly r;
#pragma unroll
  for (int k = 0; k <= kRtMaxL; ++k) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i <= k; ++i) s += a.c[i] * b.c[k - i];
    r.c[k] = s;
  }
  return r;
}
template <typename S>
__device__ __forceinline__ S rt_one();
template <>
__device__ __forceinline__ double rt_one<double>() { return 1.0; }
template <>
__device__ __forceinline__ UPoly rt_one<UPoly>() { return up_lin(1.0, 0.0); }

// I(i, j, k, l) of one Cartesian direction, the runtime form of rys_1d:
// G(n, m) = I(n, 0, m, 0) for n <= i + j, m <= k + l by the same vertical
// recurrence, then the horizontal transfers I(a, b+1) = I(a+1, b) + AB I(a, b)
// on the bra (in place, rows of G) and likewise on the ket (row i).  Needs
// i + j + k + l <= kRtMaxL.
template <typename S>
__device__ S rys_1d_rt(int i, int j, int k, int l, const S& C00, const S& D00, const S& B10, const S& B01,
                       const S& B00, double AB, double CD) {
  const int N = i + j, M = k + l, W = M + 1;
  S G[kRtMaxG];
  G[0] = rt_one<S>();
  if (N > 0) G[W] = C00;
  for (int a = 2; a <= N; ++a) G[a * W] = (double)(a - 1) * (B10 * G[(a - 2) * W]) + C00 * G[(a - 1) * W];
  if (M > 0) {
    G[1] = D00;
    for (int b = 2; b <= M; ++b) G[b] = (double)(b - 1) * (B01 * G[b - 2]) + D00 * G[b - 1];
    for (int a = 1; a <= N; ++a) {
      G[a * W + 1] = (double)a * (B00 * G[(a - 1) * W]) + D00 * G[a * W];
      for (int b = 2; b <= M; ++b)
        G[a * W + b] = (double)(b - 1) * (B01 * G[a * W + b - 2]) + (double)a * (B00 * G[(a - 1) * W + b - 1]) +
                       D00 * G[a * W + b - 1];
    }
  }
  for (int b = 1; b <= j; ++b)
    for (int a = 0; a <= N - b; ++a)
      for (int m = 0; m <= M; ++m) G[a * W + m] = G[(a + 1) * W + m] + AB * G[a * W + m];
  S* R = G + i * W;
  for (int d = 1; d <= l; ++d)
    for (int c = 0; c <= M - d; ++c) R[c] = R[c + 1] + CD * R[c];
  return R[k];
}

__device__ double fgamma_dev(double m, double x, double eps, int itmax);

// ---------------------------------------------------------------------
// Fill (graph-capturable replacement for hipMemsetAsync)
// ---------------------------------------------------------------------
__global__ void k_fill(double* p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_zero3(double* a, int64_t na, double* b, int64_t nb, double* c, int64_t nc) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na + nb + nc; i += stride) {
    if (i < na) a[i] = 0.0;
    else if (i < na + nb) b[i - na] = 0.0;
    else c[i - na - nb] = 0.0;
  }
}

hipError_t launch_zero3(double* a, int64_t na, double* b, int64_t nb, double* c, int64_t nc, hipStream_t st) {
  const int64_t n = na + nb + nc;
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_zero3, dim3((unsigned)blocks), dim3(256), 0, st, a, na, b, nb, c, nc);
  return hipGetLastError();
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

// ---------------------------------------------------------------------
// One-electron integrals over one shell pair (l_A >= l_B), all components:
//   S_ab (THO eq. 2.12, cints.c:214-239) by the Obara-Saika overlap
//   recurrence; T_ab by the reference's expression in shifted overlaps
//   (cints.c:188-212); V_ab = sum_C Z_C <a|-1/r_C|b> (cints.c:241-275) by Rys
//   quadrature with X = p |P - C|^2.  Full matrices are written (both
//   triangles).  compat: V from the coefficients of its integrand in u
//   contracted with the reference's EPS = 3e-7 Boys values F_0..F_{LA+LB}
//   (the Fgamma calls of cints.c:241-275), which reproduces its values.
// ---------------------------------------------------------------------
template <int LA, int LB>
__global__ __launch_bounds__(256) void k_onee(EngineView ev, const int* __restrict__ pair_ids, int npairs,
                                              const double* __restrict__ atom_xyz,
                                              const double* __restrict__ atom_z, int natom, double* S,
                                              double* T, double* V, int compat) {
  constexpr int NA = ncart(LA), NB = ncart(LB);
  constexpr int NRV = (LA + LB) / 2 + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs) return;
  const PairInfo bp = ev.pairs[pair_ids[t]];
  const double A[3] = {bp.Ax, bp.Ay, bp.Az};
  const double B[3] = {bp.Ax - bp.ABx, bp.Ay - bp.ABy, bp.Az - bp.ABz};
  double s[NA][NB], tk[NA][NB], v[NA][NB];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) s[a][b] = tk[a][b] = v[a][b] = 0.0;

  for (int ip = 0; ip < bp.pp_n; ++ip) {
    const PrimPair pp = ev.prims[bp.pp_off + ip];
    const double p = pp.p, beta = pp.p - ev.prim_a[bp.pp_off + ip];
    const double Kab = pp.Kp * pp.p;
    const double P[3] = {pp.Px, pp.Py, pp.Pz};
    const double i2p = 0.5 / p;
    // 1-D overlaps O[d][i][j], i <= LA, j <= LB + 2
    double O[3][LA + 1][LB + 3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double PA = P[d] - A[d], PB = P[d] - B[d];
      O[d][0][0] = 1.0;
#pragma unroll
      for (int i = 0; i <= LA; ++i) {
        if (i > 0) {
          O[d][i][0] = PA * O[d][i - 1][0] + (i > 1 ? (i - 1) * i2p * O[d][i - 2][0] : 0.0);
        }
#pragma unroll
        for (int j = 1; j <= LB + 2; ++j)
          O[d][i][j] = PB * O[d][i][j - 1] + i2p * ((i > 0 ? i * O[d][i - 1][j - 1] : 0.0) +
                                                    (j > 1 ? (j - 1) * O[d][i][j - 2] : 0.0));
      }
    }
    const double pre = Kab * pow(M_PI / p, 1.5);
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ax = cart_pow(LA, a, 0), ay = cart_pow(LA, a, 1), az = cart_pow(LA, a, 2);
        const int bx = cart_pow(LB, b, 0), by = cart_pow(LB, b, 1), bz = cart_pow(LB, b, 2);
        const double sxyz = O[0][ax][bx] * O[1][ay][by] * O[2][az][bz];
        s[a][b] += pre * sxyz;
        double t0 = beta * (2 * (bx + by + bz) + 3) * sxyz;
        double t1 = O[0][ax][bx + 2] * O[1][ay][by] * O[2][az][bz] +
                    O[0][ax][bx] * O[1][ay][by + 2] * O[2][az][bz] +
                    O[0][ax][bx] * O[1][ay][by] * O[2][az][bz + 2];
        double t2 = 0.0;
        if (bx >= 2) t2 += bx * (bx - 1) * O[0][ax][bx - 2] * O[1][ay][by] * O[2][az][bz];
        if (by >= 2) t2 += by * (by - 1) * O[0][ax][bx] * O[1][ay][by - 2] * O[2][az][bz];
        if (bz >= 2) t2 += bz * (bz - 1) * O[0][ax][bx] * O[1][ay][by] * O[2][az][bz - 2];
        tk[a][b] += pre * (t0 - 2.0 * beta * beta * t1 - 0.5 * t2);
      }
    // nuclear attraction (the nuclei of the pair's own molecule)
    const double vpre = -2.0 * M_PI * pp.Kp;
    for (int c = bp.mat0; c < bp.mat0 + bp.mnat; ++c) {
      const double Cc[3] = {atom_xyz[3 * c], atom_xyz[3 * c + 1], atom_xyz[3 * c + 2]};
      const double PC[3] = {P[0] - Cc[0], P[1] - Cc[1], P[2] - Cc[2]};
      const double X = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
      const double zc = atom_z[c] * vpre;
      if (compat) {
        constexpr int L = LA + LB;
        double F[L + 1];
#pragma unroll
        for (int k = 0; k <= L; ++k) F[k] = fgamma_dev((double)k, X, 3.0e-7, 100);
        // I_d(u) for every (ia, ib): C00 = PA - u PC, B10 = (1 - u) / (2p)
        UPoly I[3][LA + 1][LB + 1];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const UPoly C00 = up_lin(P[d] - A[d], -PC[d]), B10 = up_lin(i2p, -i2p);
#pragma unroll
          for (int ia = 0; ia <= LA; ++ia)
#pragma unroll
            for (int ib = 0; ib <= LB; ++ib)
              I[d][ia][ib] = rys_1d_rt<UPoly>(ia, ib, 0, 0, C00, C00, B10, B10, B10, A[d] - B[d], 0.0);
        }
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const UPoly pr = I[0][cart_pow(LA, a, 0)][cart_pow(LB, b, 0)] *
                             I[1][cart_pow(LA, a, 1)][cart_pow(LB, b, 1)] *
                             I[2][cart_pow(LA, a, 2)][cart_pow(LB, b, 2)];
            double s = 0.0;
#pragma unroll
            for (int k = 0; k <= L; ++k) s += pr.c[k] * F[k];
            v[a][b] += zc * s;
          }
        continue;
      }
      double u[NRV], w[NRV];
      rys_roots<NRV, double>(X, u, w);
#pragma unroll
      for (int r = 0; r < NRV; ++r) {
        const double B10 = i2p * (1.0 - u[r]);
        double I[3][LA + 1][LB + 1];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double C00 = (P[d] - A[d]) - u[r] * PC[d];
          const double AB = A[d] - B[d];
          double G[LA + LB + 1];
          G[0] = 1.0;
          if (LA + LB > 0) G[1] = C00;
#pragma unroll
          for (int n = 2; n <= LA + LB; ++n) G[n] = C00 * G[n - 1] + (n - 1) * B10 * G[n - 2];
          double H[LB + 1][LA + LB + 1];
#pragma unroll
          for (int n = 0; n <= LA + LB; ++n) H[0][n] = G[n];
#pragma unroll
          for (int j = 1; j <= LB; ++j)
#pragma unroll
            for (int n = 0; n <= LA + LB - j; ++n) H[j][n] = H[j - 1][n + 1] + AB * H[j - 1][n];
#pragma unroll
          for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int j = 0; j <= LB; ++j) I[d][i][j] = H[j][i];
        }
        const double wz = w[r] * zc;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b)
            v[a][b] += wz * I[0][cart_pow(LA, a, 0)][cart_pow(LB, b, 0)] *
                       I[1][cart_pow(LA, a, 1)][cart_pow(LB, b, 1)] *
                       I[2][cart_pow(LA, a, 2)][cart_pow(LB, b, 2)];
      }
    }
  }
  const int n = bp.mstride;
  S += bp.moff;
  T += bp.moff;
  V += bp.moff;
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const double nf = cart_norm(LA, a) * cart_norm(LB, b);
      const int i = bp.bfA + a, j = bp.bfB + b;
      S[i * n + j] = S[j * n + i] = s[a][b] * nf;
      T[i * n + j] = T[j * n + i] = tk[a][b] * nf;
      V[i * n + j] = V[j * n + i] = v[a][b] * nf;
    }
}

hipError_t launch_onee(int la, int lb, const EngineView& ev, const int* pair_ids, int npairs,
                       const double* atom_xyz, const double* atom_z, int natom, double* S, double* T,
                       double* V, int compat, hipStream_t st) {
  if (npairs <= 0) return hipSuccess;
  dim3 g((npairs + 127) / 128), b(128);
#define PQ_OCASE(A, B_)                                                                                 \
  if (la == A && lb == B_) {                                                                            \
    hipLaunchKernelGGL((k_onee<A, B_>), g, b, 0, st, ev, pair_ids, npairs, atom_xyz, atom_z, natom, S, T, \
                       V, compat);                                                                      \
    return hipGetLastError();                                                                           \
  }
  PQ_OCASE(0, 0)
  PQ_OCASE(1, 0)
  PQ_OCASE(1, 1)
  PQ_OCASE(2, 0)
  PQ_OCASE(2, 1)
  PQ_OCASE(2, 2)
#undef PQ_OCASE
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------
// One-electron part of the RHF nuclear gradient (force.py:48-70: 2 tr(D dH)
// - 2 tr(Q dS), der_Hcore_element / der_overlap_element of
// AnalyticDerivatives.py:27-455): centre derivatives of S, T and V by the
// shift rule d/dA_x g(l) = 2 alpha g(l+1) - l g(l-1) on the primitives; the
// nucleus-C derivative of V from translational invariance of the
// three-centre integral; d/dB of S and T = -d/dA.  One lane per (shell pair,
// nucleus of its molecule): the lane of the molecule's first nucleus also does
// the S and T part; contracted at once with D and W; results go to
// kGradCopies private copies.
// ---------------------------------------------------------------------
template <int LA, int LB>
__global__ __launch_bounds__(128) void k_onee_grad(EngineView ev, const int* __restrict__ pair_ids, int npairs,
                                                   const double* __restrict__ atom_xyz,
                                                   const double* __restrict__ atom_z, int natom,
                                                   const double* __restrict__ D, const double* __restrict__ W,
                                                   double* gacc) {
  constexpr int NA = ncart(LA), NB = ncart(LB);
  constexpr int NRV = (LA + LB + 1) / 2 + 1;
  const int tt = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = tt / ev.maxat, slot = tt - t * ev.maxat;
  if (t >= npairs) return;
  const PairInfo bp = ev.pairs[pair_ids[t]];
  if (slot >= bp.mnat) return;
  const int n = bp.mstride;
  D += bp.moff;
  W += bp.moff;
  double* gcopy = gacc + (size_t)(tt % kGradCopies) * 3 * natom;
  const double A[3] = {bp.Ax, bp.Ay, bp.Az};
  const double B[3] = {bp.Ax - bp.ABx, bp.Ay - bp.ABy, bp.Az - bp.ABz};
  const double sym = bp.shA == bp.shB ? 1.0 : 2.0;
  double dw[NA][NB], ww[NA][NB];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const double nf = cart_norm(LA, a) * cart_norm(LB, b);
      dw[a][b] = 2.0 * sym * nf * D[(bp.bfA + a) * n + bp.bfB + b];
      ww[a][b] = 2.0 * sym * nf * W[(bp.bfA + a) * n + bp.bfB + b];
    }
  double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
  for (int ip = 0; ip < (slot == 0 ? bp.pp_n : 0); ++ip) {
    const PrimPair pp = ev.prims[bp.pp_off + ip];
    const double p = pp.p, alp = ev.prim_a[bp.pp_off + ip], beta = p - alp;
    const double P[3] = {pp.Px, pp.Py, pp.Pz};
    const double i2p = 0.5 / p;
    // 1-D overlaps, i <= LA + 1, j <= LB + 2
    double O[3][LA + 2][LB + 3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double PA = P[d] - A[d], PB = P[d] - B[d];
      O[d][0][0] = 1.0;
#pragma unroll
      for (int i = 0; i <= LA + 1; ++i) {
        if (i > 0) O[d][i][0] = PA * O[d][i - 1][0] + (i > 1 ? (i - 1) * i2p * O[d][i - 2][0] : 0.0);
#pragma unroll
        for (int j = 1; j <= LB + 2; ++j)
          O[d][i][j] = PB * O[d][i][j - 1] + i2p * ((i > 0 ? i * O[d][i - 1][j - 1] : 0.0) +
                                                    (j > 1 ? (j - 1) * O[d][i][j - 2] : 0.0));
      }
    }
    const double pre = pp.Kp * p * pow(M_PI / p, 1.5);
    // kinetic integral with bra powers (ax, ay, az) and ket powers (bx, by, bz)
    auto tk = [&](int ax, int ay, int az, int bx, int by, int bz) -> double {
      const double s0 = O[0][ax][bx] * O[1][ay][by] * O[2][az][bz];
      double t1 = O[0][ax][bx + 2] * O[1][ay][by] * O[2][az][bz] + O[0][ax][bx] * O[1][ay][by + 2] * O[2][az][bz] +
                  O[0][ax][bx] * O[1][ay][by] * O[2][az][bz + 2];
      double t2 = 0.0;
      if (bx >= 2) t2 += bx * (bx - 1) * O[0][ax][bx - 2] * O[1][ay][by] * O[2][az][bz];
      if (by >= 2) t2 += by * (by - 1) * O[0][ax][bx] * O[1][ay][by - 2] * O[2][az][bz];
      if (bz >= 2) t2 += bz * (bz - 1) * O[0][ax][bx] * O[1][ay][by] * O[2][az][bz - 2];
      return beta * (2 * (bx + by + bz) + 3) * s0 - 2.0 * beta * beta * t1 - 0.5 * t2;
    };
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ap[3] = {cart_pow(LA, a, 0), cart_pow(LA, a, 1), cart_pow(LA, a, 2)};
        const int bq[3] = {cart_pow(LB, b, 0), cart_pow(LB, b, 1), cart_pow(LB, b, 2)};
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          int up[3] = {ap[0], ap[1], ap[2]}, dn[3] = {ap[0], ap[1], ap[2]};
          up[x] += 1;
          dn[x] -= 1;
          const double sup = O[0][up[0]][bq[0]] * O[1][up[1]][bq[1]] * O[2][up[2]][bq[2]];
          const double sdn = ap[x] ? O[0][dn[0] < 0 ? 0 : dn[0]][bq[0]] * O[1][dn[1] < 0 ? 0 : dn[1]][bq[1]] *
                                         O[2][dn[2] < 0 ? 0 : dn[2]][bq[2]]
                                   : 0.0;
          const double dS = pre * (2.0 * alp * sup - ap[x] * sdn);
          const double tup = tk(up[0], up[1], up[2], bq[0], bq[1], bq[2]);
          const double tdn = ap[x] ? tk(dn[0] < 0 ? 0 : dn[0], dn[1] < 0 ? 0 : dn[1], dn[2] < 0 ? 0 : dn[2], bq[0],
                                        bq[1], bq[2])
                                   : 0.0;
          const double dT = pre * (2.0 * alp * tup - ap[x] * tdn);
          const double g = dw[a][b] * dT - ww[a][b] * dS;
          gA[x] += g;
          gB[x] -= g;
        }
      }
  }
  // nuclear attraction of this lane's nucleus: derivatives on A, B and C
  const int c = bp.mat0 + slot;
  double gC[3] = {0, 0, 0};
  for (int ip = 0; ip < bp.pp_n; ++ip) {
    const PrimPair pp = ev.prims[bp.pp_off + ip];
    const double p = pp.p, alp = ev.prim_a[bp.pp_off + ip], beta = p - alp;
    const double P[3] = {pp.Px, pp.Py, pp.Pz};
    const double i2p = 0.5 / p;
    const double vpre = -2.0 * M_PI * pp.Kp;
    {
      const double Cc[3] = {atom_xyz[3 * c], atom_xyz[3 * c + 1], atom_xyz[3 * c + 2]};
      const double PC[3] = {P[0] - Cc[0], P[1] - Cc[1], P[2] - Cc[2]};
      const double X = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
      double u[NRV], w[NRV];
      rys_roots<NRV, double>(X, u, w);
      const double zc = atom_z[c] * vpre;
      double vA[3] = {0, 0, 0}, vB[3] = {0, 0, 0};
#pragma unroll
      for (int r = 0; r < NRV; ++r) {
        const double B10 = i2p * (1.0 - u[r]);
        double I[3][LA + 2][LB + 2];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double C00 = (P[d] - A[d]) - u[r] * PC[d];
          const double AB = A[d] - B[d];
          constexpr int NN = LA + LB + 2;
          double G[NN + 1];
          G[0] = 1.0;
          G[1] = C00;
#pragma unroll
          for (int m = 2; m <= NN; ++m) G[m] = C00 * G[m - 1] + (m - 1) * B10 * G[m - 2];
          double H[LB + 2][NN + 1];
#pragma unroll
          for (int m = 0; m <= NN; ++m) H[0][m] = G[m];
#pragma unroll
          for (int j = 1; j <= LB + 1; ++j)
#pragma unroll
            for (int m = 0; m <= NN - j; ++m) H[j][m] = H[j - 1][m + 1] + AB * H[j - 1][m];
#pragma unroll
          for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
            for (int j = 0; j <= LB + 1; ++j) I[d][i][j] = H[j][i];
        }
        const double wz = w[r] * zc;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const int ap[3] = {cart_pow(LA, a, 0), cart_pow(LA, a, 1), cart_pow(LA, a, 2)};
            const int bq[3] = {cart_pow(LB, b, 0), cart_pow(LB, b, 1), cart_pow(LB, b, 2)};
            const double i0[3] = {I[0][ap[0]][bq[0]], I[1][ap[1]][bq[1]], I[2][ap[2]][bq[2]]};
            const double g = dw[a][b] * wz;
#pragma unroll
            for (int x = 0; x < 3; ++x) {
              const int y = (x + 1) % 3, z = (x + 2) % 3;
              const double oth = i0[y] * i0[z];
              const double da = 2.0 * alp * I[x][ap[x] + 1][bq[x]] - (ap[x] ? ap[x] * I[x][ap[x] ? ap[x] - 1 : 0][bq[x]] : 0.0);
              const double db = 2.0 * beta * I[x][ap[x]][bq[x] + 1] - (bq[x] ? bq[x] * I[x][ap[x]][bq[x] ? bq[x] - 1 : 0] : 0.0);
              vA[x] += g * da * oth;
              vB[x] += g * db * oth;
            }
          }
      }
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        gA[x] += vA[x];
        gB[x] += vB[x];
        gC[x] -= vA[x] + vB[x];
      }
    }
  }
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    unsafeAtomicAdd(&gcopy[3 * c + x], gC[x]);
    unsafeAtomicAdd(&gcopy[3 * bp.atA + x], gA[x]);
    unsafeAtomicAdd(&gcopy[3 * bp.atB + x], gB[x]);
  }
}

hipError_t launch_onee_grad(int la, int lb, const EngineView& ev, const int* pair_ids, int npairs,
                            const double* atom_xyz, const double* atom_z, int natom, const double* D,
                            const double* W, double* gacc, hipStream_t st) {
  if (npairs <= 0) return hipSuccess;
  if (ev.maxat <= 0) return hipErrorInvalidValue;
  const int64_t lanes = (int64_t)npairs * ev.maxat;  // one lane per (pair, nucleus slot)
  if (lanes > INT32_MAX - 128) return hipErrorInvalidValue;
  dim3 g((unsigned)((lanes + 127) / 128)), b(128);
#define PQ_OGCASE(A, B_)                                                                                    \
  if (la == A && lb == B_) {                                                                                \
    hipLaunchKernelGGL((k_onee_grad<A, B_>), g, b, 0, st, ev, pair_ids, npairs, atom_xyz, atom_z, natom, D, W, \
                       gacc);                                                                               \
    return hipGetLastError();                                                                               \
  }
  PQ_OGCASE(0, 0)
  PQ_OGCASE(1, 0)
  PQ_OGCASE(1, 1)
  PQ_OGCASE(2, 0)
  PQ_OGCASE(2, 1)
  PQ_OGCASE(2, 2)
#undef PQ_OGCASE
  return hipErrorInvalidValue;
}

// grad[i] = sum over the private copies + d/dR of sum_{A<B} Z_A Z_B / R_AB
// atom_mol (nullable): molecule of each atom; nuclear repulsion acts within a
// molecule only
// One wavefront per gradient component: its lanes sum the private copies
// and the nuclear-repulsion terms of the atom's molecule (atoms [at0, at0 +
// nat) when the engine holds several molecules, else every atom), then
// reduce.  dV_nn/dR_a = -Z_a sum_b Z_b (R_a - R_b) / |R_a - R_b|^3
// (force.py:172-193, der_enuke).
__global__ __launch_bounds__(256) void k_grad_finish(const double* __restrict__ gacc, int ncopies, int natom,
                                                     const double* __restrict__ xyz, const double* __restrict__ z,
                                                     const int* __restrict__ atom_mol, const int* __restrict__ mol_at0,
                                                     const int* __restrict__ mol_nat, double* grad) {
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= 3 * natom) return;  // uniform per wavefront
  double s = 0.0;
  for (int c = lane; c < ncopies; c += 64) s += gacc[(size_t)c * 3 * natom + i];
  const int a = i / 3, x = i % 3;
  int b0 = 0, b1 = natom;
  if (atom_mol) {
    const int m = atom_mol[a];
    b0 = mol_at0[m];
    b1 = b0 + mol_nat[m];
  }
  for (int b = b0 + lane; b < b1; b += 64) {
    if (b == a) continue;
    const double dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1], dz = xyz[3 * a + 2] - xyz[3 * b + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    const double dd = x == 0 ? dx : (x == 1 ? dy : dz);
    s -= z[a] * z[b] * dd / (r2 * r);
  }
  s = wave_sum(s);
  if (lane == 0) grad[i] = s;
}

hipError_t launch_grad_finish(const double* gacc, int ncopies, int natom, const double* atom_xyz,
                              const double* atom_z, const int* atom_mol, const int* mol_at0, const int* mol_nat,
                              double* grad, hipStream_t st) {
  const int waves = 3 * natom;
  hipLaunchKernelGGL(k_grad_finish, dim3((waves + 3) / 4), dim3(256), 0, st, gacc, ncopies, natom, atom_xyz, atom_z,
                     atom_mol, mol_at0, mol_nat, grad);
  return hipGetLastError();
}

// ---------------------------------------------------------------------
// Boys function F_m(x) = 1/2 x^(-m-1/2) gamma(m+1/2, x) by the incomplete
// gamma series (x < m + 3/2) or Lentz continued fraction, as the
// reference's Fgamma does (cints.c:340-411); eps is the relative stopping
// tolerance (3e-7 reproduces the reference, 1e-16 gives full precision).
// ---------------------------------------------------------------------
__device__ double fgamma_dev(double m, double x, double eps, int itmax) {
  if (x < 0.0) x = 0.0;
  if (eps < 1e-10 && (x < 50.0 || m == floor(m))) {
    // full-precision mode: F_m(x) = e^-x sum_k (2x)^k / prod_{j<=k}(2m+2j+1)
    // below x = 50; upward recursion from F_0 = sqrt(pi/x) erf(sqrt x)/2 above
    // (stable there: (2m+1)/(2x) < 1 for the orders in use)
    if (x < 50.0) {
      double term = 1.0 / (2.0 * m + 1.0), sum = term;
      for (int k = 1; k < 400; ++k) {
        term *= 2.0 * x / (2.0 * m + 2.0 * k + 1.0);
        sum += term;
        if (term < sum * 1e-17) break;
      }
      return exp(-x) * sum;
    } else {
      const double ex = exp(-x);
      double f = 0.5 * sqrt(M_PI / x) * erf(sqrt(x));
      for (int k = 0; k < (int)m; ++k) f = ((2.0 * k + 1.0) * f - ex) / (2.0 * x);
      return f;
    }
  }
  if (x < 1e-8) x = 1e-8;  // the reference's clamp (cints.c:340-411)
  const double a = m + 0.5;
  const double gln = lgamma(a);
  double gam;
  if (x < a + 1.0) {
    double ap = a, del = 1.0 / a, sum = del;
    for (int k = 1; k <= itmax; ++k) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (fabs(del) < fabs(sum) * eps) break;
    }
    gam = exp(gln) * sum * exp(-x + a * log(x) - gln);
  } else {
    const double fpmin = 1e-30;
    double b = x + 1.0 - a, c = 1.0 / fpmin, d = 1.0 / b, h = d;
    for (int i = 1; i <= itmax; ++i) {
      const double an = -i * (i - a);
      b += 2.0;
      d = an * d + b;
      if (fabs(d) < fpmin) d = fpmin;
      c = b + an / c;
      if (fabs(c) < fpmin) c = fpmin;
      d = 1.0 / d;
      const double del = d * c;
      h *= del;
      if (fabs(del - 1.0) < eps) break;
    }
    gam = exp(gln) * (1.0 - exp(-x + a * log(x) - gln) * h);
  }
  return 0.5 * pow(x, -m - 0.5) * gam;
}

__global__ void k_boys(int n, const double* m_, const double* x_, double eps, int itmax, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[t] = fgamma_dev(m_[t], x_[t], eps, itmax);
}

hipError_t launch_boys(int n, const double* m, const double* x, double eps, double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int itmax = eps < 1e-10 ? 2000 : 100;
  hipLaunchKernelGGL(k_boys, dim3((n + 127) / 128), dim3(128), 0, st, n, m, x, eps, itmax, out);
  return hipGetLastError();
}

__global__ void k_rys_roots(int n, int nr, const double* X, double* u, double* w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double uu[5], ww[5];
  rys_roots_rt<double>(nr, X[t], uu, ww);
  for (int i = 0; i < nr; ++i) {
    u[t * nr + i] = uu[i];
    w[t * nr + i] = ww[i];
  }
}

hipError_t launch_rys_roots(int n, int nr, const double* X, double* u, double* w, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rys_roots, dim3((n + 127) / 128), dim3(128), 0, st, n, nr, X, u, w);
  return hipGetLastError();
}

// ---------------------------------------------------------------------
// Contracted ERI of four arbitrary Cartesian functions: the batch behind the
// drop-in contr_coulomb / coulomb_repulsion.  Runtime angular momentum,
// L_total <= kRtMaxL.  Function f of quartet t has nprim[4t+f] primitives
// from poff[4t+f] in exps/coefs/norms, origin org[12t+3f..], powers pw[..].
//   k_contr_eri: Rys quadrature with the engine's roots (crys.c:34-97
//     semantics, one rys_1d_rt per root and direction);
//   k_tho_contr: the integrand's coefficients in u contracted with Boys
//     values F_0..F_L of eps (3e-7: the reference's Fgamma, so the values of
//     cints.c:80-153; 1e-17: full precision).
// ---------------------------------------------------------------------
struct ContrQuartet {
  int np[4], po[4], p[4][3], L;
  double o[4][3];
};

__device__ __forceinline__ ContrQuartet load_quartet(int t, const int* __restrict__ nprim,
                                                     const int* __restrict__ poff, const double* __restrict__ org,
                                                     const int* __restrict__ pw) {
  ContrQuartet q;
  q.L = 0;
  for (int f = 0; f < 4; ++f) {
    q.np[f] = nprim[4 * t + f];
    q.po[f] = poff[4 * t + f];
    for (int d = 0; d < 3; ++d) {
      q.p[f][d] = pw[12 * t + 3 * f + d];
      q.o[f][d] = org[12 * t + 3 * f + d];
      q.L += q.p[f][d];
    }
  }
  return q;
}

// Per-primitive-quartet quantities of the recurrence coefficients (in u):
// C00 = PA - qfac PQ u, D00 = QC + pfac PQ u, B10 = (1 - qfac u) / (2A),
// B01 = (1 - pfac u) / (2B), B00 = u / (2 (A + B)); T = A B / (A + B) |PQ|^2;
// pre = 2 pi^(5/2) / (A B sqrt(A + B)) exp(-ab/A |AB|^2 - cd/B |CD|^2).
struct PrimCoefs {
  double PA[3], QC[3], PQ[3], qfac, pfac, i2p, i2q, hr, T, pre;
};

__device__ __forceinline__ PrimCoefs prim_coefs(const ContrQuartet& q, double a0, double a1, double a2, double a3) {
  PrimCoefs c;
  const double A = a0 + a1, B = a2 + a3, r = 1.0 / (A + B);
  double rpq2 = 0.0, kab = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double P = (a0 * q.o[0][d] + a1 * q.o[1][d]) / A, Q = (a2 * q.o[2][d] + a3 * q.o[3][d]) / B;
    c.PA[d] = P - q.o[0][d];
    c.QC[d] = Q - q.o[2][d];
    c.PQ[d] = P - Q;
    rpq2 += c.PQ[d] * c.PQ[d];
    kab += a0 * a1 / A * (q.o[0][d] - q.o[1][d]) * (q.o[0][d] - q.o[1][d]) +
           a2 * a3 / B * (q.o[2][d] - q.o[3][d]) * (q.o[2][d] - q.o[3][d]);
  }
  c.qfac = B * r;
  c.pfac = A * r;
  c.i2p = 0.5 / A;
  c.i2q = 0.5 / B;
  c.hr = 0.5 * r;
  c.T = A * B * r * rpq2;
  c.pre = 2.0 * pow(M_PI, 2.5) / (A * B * sqrt(A + B)) * exp(-kab);
  return c;
}

__device__ __forceinline__ double quartet_prim_scale(const ContrQuartet& q, const double* __restrict__ coefs,
                                                     const double* __restrict__ norms, int i, int j, int k,
                                                     int l) {
  const int ix[4] = {q.po[0] + i, q.po[1] + j, q.po[2] + k, q.po[3] + l};
  double s = 1.0;
  for (int f = 0; f < 4; ++f) s *= coefs[ix[f]] * norms[ix[f]];
  return s;
}

__global__ void k_contr_eri(int n, const int* __restrict__ nprim, const int* __restrict__ poff,
                            const double* __restrict__ exps, const double* __restrict__ coefs,
                            const double* __restrict__ norms, const double* __restrict__ org,
                            const int* __restrict__ pw, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const ContrQuartet q = load_quartet(t, nprim, poff, org, pw);
  const int nr = q.L / 2 + 1;
  double val = 0.0;
  for (int i = 0; i < q.np[0]; ++i)
    for (int j = 0; j < q.np[1]; ++j)
      for (int k = 0; k < q.np[2]; ++k)
        for (int l = 0; l < q.np[3]; ++l) {
          const PrimCoefs c = prim_coefs(q, exps[q.po[0] + i], exps[q.po[1] + j], exps[q.po[2] + k],
                                         exps[q.po[3] + l]);
          double u[5], w[5];
          rys_roots_rt<double>(nr, c.T, u, w);
          double sum = 0.0;
          for (int r = 0; r < nr; ++r) {
            double prod = w[r];
            for (int d = 0; d < 3; ++d)
              prod *= rys_1d_rt<double>(q.p[0][d], q.p[1][d], q.p[2][d], q.p[3][d],
                                        c.PA[d] - c.qfac * c.PQ[d] * u[r], c.QC[d] + c.pfac * c.PQ[d] * u[r],
                                        c.i2p * (1.0 - c.qfac * u[r]), c.i2q * (1.0 - c.pfac * u[r]),
                                        c.hr * u[r], q.o[0][d] - q.o[1][d], q.o[2][d] - q.o[3][d]);
            sum += prod;
          }
          val += quartet_prim_scale(q, coefs, norms, i, j, k, l) * c.pre * sum;
        }
  out[t] = val;
}

__global__ void k_tho_contr(int n, const int* __restrict__ nprim, const int* __restrict__ poff,
                            const double* __restrict__ exps, const double* __restrict__ coefs,
                            const double* __restrict__ norms, const double* __restrict__ org,
                            const int* __restrict__ pw, double eps, int itmax, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const ContrQuartet q = load_quartet(t, nprim, poff, org, pw);
  double val = 0.0;
  for (int i = 0; i < q.np[0]; ++i)
    for (int j = 0; j < q.np[1]; ++j)
      for (int k = 0; k < q.np[2]; ++k)
        for (int l = 0; l < q.np[3]; ++l) {
          const PrimCoefs c = prim_coefs(q, exps[q.po[0] + i], exps[q.po[1] + j], exps[q.po[2] + k],
                                         exps[q.po[3] + l]);
          UPoly prod = up_lin(1.0, 0.0);
          for (int d = 0; d < 3; ++d)
            prod = prod * rys_1d_rt<UPoly>(q.p[0][d], q.p[1][d], q.p[2][d], q.p[3][d],
                                           up_lin(c.PA[d], -c.qfac * c.PQ[d]), up_lin(c.QC[d], c.pfac * c.PQ[d]),
                                           up_lin(c.i2p, -c.i2p * c.qfac), up_lin(c.i2q, -c.i2q * c.pfac),
                                           up_lin(0.0, c.hr), q.o[0][d] - q.o[1][d], q.o[2][d] - q.o[3][d]);
          double sum = 0.0;
          for (int m = 0; m <= q.L; ++m) sum += prod.c[m] * fgamma_dev((double)m, c.T, eps, itmax);
          val += quartet_prim_scale(q, coefs, norms, i, j, k, l) * c.pre * sum;
        }
  out[t] = val;
}

// ---------------------------------------------------------------------
// Head-Gordon-Pople ERIs (chgp.c:78-433), one contracted function quartet
// per lane.  (ab|cd) is expanded by the horizontal recurrence
// (a, b+1_i| = (a+1_i, b| + (A-B)_i (a, b| in closed form,
//   (ab|cd) = sum_{k<=b, l<=d} prod_i C(b_i,k_i) AB_i^(b_i-k_i) C(d_i,l_i) CD_i^(d_i-l_i) [a+k,0|c+l,0],
// and every leaf [e0|f0] is a contracted Obara-Saika vertical recurrence
// over its own box e' <= e, f' <= f with the Boys values of order |e|+|f|
// (chgp.c:244-419): the reference starts the downward Boys recursion at each
// leaf's own total order, and so does this kernel, which keeps its values
// equal to the reference's also with the reference's EPS = 3e-7 Boys
// function (eps).  Leaves are limited to |e|, |f| <= 4 (box <= 12 entries)
// and |e| + |f| + mext <= 8.
// mext (optional, per quartet) asks for the auxiliary order m of a single
// primitive leaf [a0|c0]^(m) instead (chgp.c:857-890, vrr).
// ---------------------------------------------------------------------
constexpr int kHgpBox = 12, kHgpM = 9;

__device__ __forceinline__ double binom(int n, int k) {
  double c = 1.0;
  for (int i = 1; i <= k; ++i) c = c * (n - k + i) / i;
  return c;
}

__device__ double hgp_leaf(const int e[3], const int f[3], int mext, const int np[4], const int po[4],
                           const double o[4][3], const double* __restrict__ exps,
                           const double* __restrict__ coefs, const double* __restrict__ norms, double eps,
                           int itmax, double (&V)[kHgpM][kHgpBox][kHgpBox]) {
  const int ne[3] = {e[0] + 1, e[1] + 1, e[2] + 1}, nf[3] = {f[0] + 1, f[1] + 1, f[2] + 1};
  const int le = e[0] + e[1] + e[2], lf = f[0] + f[1] + f[2];
  const int mtot = le + lf + mext;
  const int nE = ne[0] * ne[1] * ne[2], nF = nf[0] * nf[1] * nf[2];
  const int se[3] = {ne[1] * ne[2], ne[2], 1}, sf[3] = {nf[1] * nf[2], nf[2], 1};
  double acc = 0.0;
  for (int i = 0; i < np[0]; ++i)
    for (int j = 0; j < np[1]; ++j)
      for (int k = 0; k < np[2]; ++k)
        for (int l = 0; l < np[3]; ++l) {
          const double aa = exps[po[0] + i], ab = exps[po[1] + j], ac = exps[po[2] + k], ad = exps[po[3] + l];
          const double zeta = aa + ab, eta = ac + ad, ze = zeta + eta, rho = zeta * eta / ze;
          double P[3], Q[3], W[3], rab2 = 0.0, rcd2 = 0.0, rpq2 = 0.0;
          for (int d = 0; d < 3; ++d) {
            P[d] = (aa * o[0][d] + ab * o[1][d]) / zeta;
            Q[d] = (ac * o[2][d] + ad * o[3][d]) / eta;
            W[d] = (zeta * P[d] + eta * Q[d]) / ze;
            rab2 += (o[0][d] - o[1][d]) * (o[0][d] - o[1][d]);
            rcd2 += (o[2][d] - o[3][d]) * (o[2][d] - o[3][d]);
            rpq2 += (P[d] - Q[d]) * (P[d] - Q[d]);
          }
          const double T = rho * rpq2;
          // K = sqrt(2) pi^(5/4) / zeta exp(-ab/zeta |AB|^2) (chgp.c:273-277)
          const double kpi = 5.914967172795613;  // sqrt(2) pi^(5/4)
          const double Kab = kpi / zeta * exp(-aa * ab / zeta * rab2);
          const double Kcd = kpi / eta * exp(-ac * ad / eta * rcd2);
          const double pref = norms[po[0] + i] * norms[po[1] + j] * norms[po[2] + k] * norms[po[3] + l] * Kab *
                              Kcd / sqrt(ze);
          // Boys F_m(T), m = 0..mtot: F_mtot, then downward (chgp.c:287-291)
          double F[kHgpM];
          F[mtot] = fgamma_dev((double)mtot, T, eps, itmax);
          const double ex = exp(-T);
          for (int m = mtot - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + ex) / (2.0 * m + 1.0);
          for (int m = 0; m <= mtot; ++m) V[m][0][0] = pref * F[m];
          // bra: [e'0|00]^m for e' in the box, lexicographic order; the step
          // goes along the first nonzero direction of e'
          for (int ex_ = 0; ex_ < ne[0]; ++ex_)
            for (int ey = 0; ey < ne[1]; ++ey)
              for (int ez = 0; ez < ne[2]; ++ez) {
                const int ep[3] = {ex_, ey, ez};
                const int tot = ex_ + ey + ez;
                if (tot == 0) continue;
                const int dir = ex_ > 0 ? 0 : (ey > 0 ? 1 : 2);
                const int ie = ex_ * se[0] + ey * se[1] + ez;
                const int i1 = ie - se[dir], i2 = ie - 2 * se[dir];
                const double PA = P[dir] - o[0][dir], WP = W[dir] - P[dir];
                const double c2 = (ep[dir] - 1) / (2.0 * zeta);
                for (int m = 0; m <= mtot - tot; ++m) {
                  double v = PA * V[m][i1][0] + WP * V[m + 1][i1][0];
                  if (ep[dir] >= 2) v += c2 * (V[m][i2][0] - rho / zeta * V[m + 1][i2][0]);
                  V[m][ie][0] = v;
                }
              }
          // ket: [e'0|f'0]^m, f' lexicographic outside, every e' inside
          for (int fx = 0; fx < nf[0]; ++fx)
            for (int fy = 0; fy < nf[1]; ++fy)
              for (int fz = 0; fz < nf[2]; ++fz) {
                const int fp[3] = {fx, fy, fz};
                const int ftot = fx + fy + fz;
                if (ftot == 0) continue;
                const int dir = fx > 0 ? 0 : (fy > 0 ? 1 : 2);
                const int jf = fx * sf[0] + fy * sf[1] + fz;
                const int j1 = jf - sf[dir], j2 = jf - 2 * sf[dir];
                const double QC = Q[dir] - o[2][dir], WQ = W[dir] - Q[dir];
                const double c2 = (fp[dir] - 1) / (2.0 * eta), c3 = 1.0 / (2.0 * ze);
                for (int ex_ = 0; ex_ < ne[0]; ++ex_)
                  for (int ey = 0; ey < ne[1]; ++ey)
                    for (int ez = 0; ez < ne[2]; ++ez) {
                      const int ep[3] = {ex_, ey, ez};
                      const int ie = ex_ * se[0] + ey * se[1] + ez;
                      for (int m = 0; m <= mtot - ftot - (ex_ + ey + ez); ++m) {
                        double v = QC * V[m][ie][j1] + WQ * V[m + 1][ie][j1];
                        if (fp[dir] >= 2) v += c2 * (V[m][ie][j2] - rho / eta * V[m + 1][ie][j2]);
                        if (ep[dir] >= 1) v += ep[dir] * c3 * V[m + 1][ie - se[dir]][j1];
                        V[m][ie][jf] = v;
                      }
                    }
              }
          acc += coefs[po[0] + i] * coefs[po[1] + j] * coefs[po[2] + k] * coefs[po[3] + l] *
                 V[mext][nE - 1][nF - 1];
        }
  return acc;
}

// One contracted function quartet: leaves a+k, c+l of the horizontal expansion.
__device__ double hgp_quartet(const int np[4], const int po[4], const double o[4][3], const int p[4][3], int mext,
                              const double* __restrict__ exps, const double* __restrict__ coefs,
                              const double* __restrict__ norms, double eps, int itmax,
                              double (&V)[kHgpM][kHgpBox][kHgpBox]) {
  double val = 0.0;
  for (int kx = 0; kx <= p[1][0]; ++kx)
    for (int ky = 0; ky <= p[1][1]; ++ky)
      for (int kz = 0; kz <= p[1][2]; ++kz) {
        const int kk[3] = {kx, ky, kz};
        double hb = 1.0;
        int e[3];
        for (int d = 0; d < 3; ++d) {
          hb *= binom(p[1][d], kk[d]);
          for (int r = kk[d]; r < p[1][d]; ++r) hb *= o[0][d] - o[1][d];
          e[d] = p[0][d] + kk[d];
        }
        for (int lx = 0; lx <= p[3][0]; ++lx)
          for (int ly = 0; ly <= p[3][1]; ++ly)
            for (int lz = 0; lz <= p[3][2]; ++lz) {
              const int ll[3] = {lx, ly, lz};
              double hk = 1.0;
              int f[3];
              for (int d = 0; d < 3; ++d) {
                hk *= binom(p[3][d], ll[d]);
                for (int r = ll[d]; r < p[3][d]; ++r) hk *= o[2][d] - o[3][d];
                f[d] = p[2][d] + ll[d];
              }
              val += hb * hk * hgp_leaf(e, f, mext, np, po, o, exps, coefs, norms, eps, itmax, V);
            }
      }
  return val;
}

__global__ void k_hgp_contr(int n, const int* __restrict__ nprim, const int* __restrict__ poff,
                            const double* __restrict__ exps, const double* __restrict__ coefs,
                            const double* __restrict__ norms, const double* __restrict__ org,
                            const int* __restrict__ pw, const int* __restrict__ mext_, double eps, int itmax,
                            double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int np[4], po[4], p[4][3];
  double o[4][3];
  for (int f = 0; f < 4; ++f) {
    np[f] = nprim[4 * t + f];
    po[f] = poff[4 * t + f];
    for (int d = 0; d < 3; ++d) {
      p[f][d] = pw[12 * t + 3 * f + d];
      o[f][d] = org[12 * t + 3 * f + d];
    }
  }
  double V[kHgpM][kHgpBox][kHgpBox];
  out[t] = hgp_quartet(np, po, o, p, mext_ ? mext_[t] : 0, exps, coefs, norms, eps, itmax, V);
}

// Every unique function quartet of a basis in the reference's flat layout
// (index ij(ij+1)/2 + kl, Ints.py:146-167), by HGP.  Function f has fnp[f]
// primitives at foff[f] in fexps/fcoefs/fnorms (norms for its own powers).
__global__ void k_hgp_store(uint64_t len, const int* __restrict__ fnp, const int* __restrict__ foff,
                            const double* __restrict__ forg, const int* __restrict__ fpw,
                            const double* __restrict__ fexps, const double* __restrict__ fcoefs,
                            const double* __restrict__ fnorms, double eps, int itmax, double* out) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  double V[kHgpM][kHgpBox][kHgpBox];
  for (; t < len; t += stride) {
    const int64_t ij = tri_root(t), kl = (int64_t)(t - (uint64_t)ij * (ij + 1) / 2);
    const int64_t i = tri_root(ij), j = ij - i * (i + 1) / 2;
    const int64_t k = tri_root(kl), l = kl - k * (k + 1) / 2;
    const int64_t fs[4] = {i, j, k, l};
    int np[4], po[4], p[4][3];
    double o[4][3];
    for (int f = 0; f < 4; ++f) {
      np[f] = fnp[fs[f]];
      po[f] = foff[fs[f]];
      for (int d = 0; d < 3; ++d) {
        p[f][d] = fpw[3 * fs[f] + d];
        o[f][d] = forg[3 * fs[f] + d];
      }
    }
    out[t] = hgp_quartet(np, po, o, p, 0, fexps, fcoefs, fnorms, eps, itmax, V);
  }
}

hipError_t launch_hgp_store(uint64_t len, const int* fnp, const int* foff, const double* forg, const int* fpw,
                            const double* fexps, const double* fcoefs, const double* fnorms, int compat_nr,
                            double* out, hipStream_t st) {
  if (len == 0) return hipSuccess;
  const double eps = compat_nr ? 3.0e-7 : 1e-17;
  const int itmax = compat_nr ? 100 : 2000;
  const uint64_t blocks = std::min<uint64_t>((len + 63) / 64, 1u << 16);
  hipLaunchKernelGGL(k_hgp_store, dim3((unsigned)blocks), dim3(64), 0, st, len, fnp, foff, forg, fpw, fexps,
                     fcoefs, fnorms, eps, itmax, out);
  return hipGetLastError();
}

hipError_t launch_hgp_contr(int n, const int* nprim, const int* poff, const double* exps, const double* coefs,
                            const double* norms, const double* origins, const int* powers, const int* mext,
                            int compat_nr, double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const double eps = compat_nr ? 3.0e-7 : 1e-17;
  const int itmax = compat_nr ? 100 : 2000;
  hipLaunchKernelGGL(k_hgp_contr, dim3((n + 63) / 64), dim3(64), 0, st, n, nprim, poff, exps, coefs, norms,
                     origins, powers, mext, eps, itmax, out);
  return hipGetLastError();
}

hipError_t launch_contr_eri(int n, const int* nprim, const int* poff, const double* exps, const double* coefs,
                            const double* norms, const double* origins, const int* powers, double* out,
                            hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_contr_eri, dim3((n + 63) / 64), dim3(64), 0, st, n, nprim, poff, exps, coefs, norms,
                     origins, powers, out);
  return hipGetLastError();
}

hipError_t launch_tho_contr(int n, const int* nprim, const int* poff, const double* exps, const double* coefs,
                            const double* norms, const double* origins, const int* powers, int compat_nr,
                            double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const double eps = compat_nr ? 3.0e-7 : 1e-17;
  const int itmax = compat_nr ? 100 : 2000;
  hipLaunchKernelGGL(k_tho_contr, dim3((n + 63) / 64), dim3(64), 0, st, n, nprim, poff, exps, coefs, norms,
                     origins, powers, eps, itmax, out);
  return hipGetLastError();
}

}  // namespace pq
