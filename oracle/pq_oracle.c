/*
 * pq_oracle.c -- CPU restatement of PyQuante's two-electron-integral and
 * Fock-build path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg load it through oracle.py.
 * The product library (pyquante_amd/csrc) never links, loads or calls it.
 *
 * Every function names the reference file:line whose behaviour it restates.
 * Parity of each function is pinned by tests/test_oracle.py against:
 *   - golden vectors written by the compiled reference C (oracle/_ref),
 *   - the Boys known-answer tests of Src/lib/utils/tests/TestMath.c:19-20,
 *   - closed forms (F_0 via erf, (ss|ss), F_m(0)=1/(2m+1)), moment
 *     invariants of the Rys rule, and brute-force einsum for J and K.
 *
 * Rys roots/weights: the reference (crys.c:99-1443) evaluates GAMESS
 * polynomial fits.  This restatement computes the same Gauss rule by a
 * different published method -- the discretized Stieltjes procedure
 * (W. Gautschi, SIAM J. Sci. Stat. Comput. 3, 289 (1982)) on a
 * Gauss-Legendre discretisation of exp(-X t^2) dt on [0,1] -- and is
 * compared with the compiled crys.c roots and with the moment identities
 * sum_i w_i u_i^k = F_k(X), k < 2n.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* Boys function, Numerical-Recipes compatible (cints.c:340-411)       */
/* ------------------------------------------------------------------ */
#define NR_ITMAX 100
#define NR_EPS 3.0e-7
#define NR_FPMIN 1.0e-30
#define NR_SMALL 0.00000001

static void nr_gser(double *gamser, double a, double x, double *gln) {
  /* cints.c:369-389 */
  *gln = lgamma(a);
  if (x <= 0.0) { *gamser = 0.0; return; }
  double ap = a, sum = 1.0 / a, del = sum;
  for (int n = 1; n <= NR_ITMAX; n++) {
    ++ap;
    del *= x / ap;
    sum += del;
    if (fabs(del) < fabs(sum) * NR_EPS) {
      *gamser = sum * exp(-x + a * log(x) - (*gln));
      return;
    }
  }
  /* the reference prints and leaves gamser unset here; 0 is returned */
  *gamser = 0.0;
}

static void nr_gcf(double *gammcf, double a, double x, double *gln) {
  /* cints.c:391-411 (modified Lentz continued fraction) */
  *gln = lgamma(a);
  double b = x + 1.0 - a, c = 1.0 / NR_FPMIN, d = 1.0 / b, h = d;
  for (int i = 1; i <= NR_ITMAX; i++) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < NR_FPMIN) d = NR_FPMIN;
    c = b + an / c;
    if (fabs(c) < NR_FPMIN) c = NR_FPMIN;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < NR_EPS) break;
  }
  *gammcf = exp(-x + a * log(x) - (*gln)) * h;
}

double pqo_fgamma_nr(double m, double x) {
  /* cints.c:340-346 Fgamma / 348-362 gamm_inc */
  double gam, gln;
  if (fabs(x) < NR_SMALL) x = NR_SMALL;
  double a = m + 0.5;
  if (x < a + 1.0) {
    nr_gser(&gam, a, x, &gln);
    gam = exp(gln) * gam;
  } else {
    nr_gcf(&gam, a, x, &gln);
    gam = exp(gln) * (1.0 - gam);
  }
  return 0.5 * pow(x, -m - 0.5) * gam;
}

/* ------------------------------------------------------------------ */
/* Boys function, accurate: F_m(x) = int_0^1 t^(2m) exp(-x t^2) dt     */
/* Series e^{-x} sum_k (2x)^k / prod_{j<=k}(2m+2j+1) in long double    */
/* for x < 100, upward recursion from F_0 = sqrt(pi/x) erf(sqrt x)/2   */
/* above (stable there since (2m+1)/(2x) < 1 for the orders used).     */
/* ------------------------------------------------------------------ */
double pqo_boys(double m, double x) {
  if (x < 0) x = 0;
  if (x < 100.0) {
    long double xl = x, term = 1.0L / (2.0L * m + 1.0L), sum = term;
    for (int k = 1; k < 2000; k++) {
      term *= 2.0L * xl / (2.0L * m + 2.0L * k + 1.0L);
      sum += term;
      if (term < sum * 1e-21L) break;
    }
    return (double)(expl(-xl) * sum);
  }
  double mi = floor(m);
  if (m == mi && m < 60) {
    long double xl = x, f = 0.5L * sqrtl(M_PI / xl) * erfl(sqrtl(xl));
    long double ex = expl(-xl);
    for (int k = 0; k < (int)m; k++) f = ((2.0L * k + 1.0L) * f - ex) / (2.0L * xl);
    return (double)f;
  }
  /* non-integer order at large x: the incomplete-gamma tail is below
     exp(-100) relative for the small orders this path is used with */
  return (double)(0.5L * tgammal(m + 0.5L) / powl((long double)x, m + 0.5L));
}

/* ------------------------------------------------------------------ */
/* Rys roots and weights: Gauss rule for dmu(u) = exp(-X u)/(2 sqrt u) */
/* on u in [0,1], i.e. sum_i w_i f(u_i) ~ int_0^1 f(t^2) e^{-X t^2} dt  */
/* Returned roots are u = t^2 (crys.c returns u/(1-u), RecurFactors-   */
/* Gamess crys.c:1468-1479 converts back).                             */
/* ------------------------------------------------------------------ */
#define GL_M 64
static long double gl_x[GL_M], gl_w[GL_M];
static int gl_ready = 0;

static void gl_init(void) {
  /* Gauss-Legendre nodes on [-1,1] by Newton on P_M */
  for (int i = 0; i < GL_M; i++) {
    long double z = cosl(M_PI * (i + 0.75L) / (GL_M + 0.5L)), pp = 0;
    for (int it = 0; it < 100; it++) {
      long double p1 = 1, p2 = 0;
      for (int j = 0; j < GL_M; j++) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2.0L * j + 1.0L) * z * p2 - j * p3) / (j + 1.0L);
      }
      pp = GL_M * (z * p1 - p2) / (z * z - 1.0L);
      long double dz = p1 / pp;
      z -= dz;
      if (fabsl(dz) < 1e-20L) break;
    }
    gl_x[i] = z;
    gl_w[i] = 2.0L / ((1.0L - z * z) * pp * pp);
  }
  gl_ready = 1;
}

int pqo_rys_roots(int n, double X, double *u_out, double *w_out) {
  if (n < 1 || n > 10) return -1;
  if (!gl_ready) gl_init();
  /* discretise t in [0,T]: beyond T the weight is below exp(-60) */
  long double T = 1.0L;
  if (X > 60.0) T = sqrtl(60.0L / X);
  long double uj[2 * GL_M], mj[2 * GL_M];
  int M = 0;
  /* two panels [0,T/2],[T/2,T] keep the rule accurate near t=0 */
  for (int panel = 0; panel < 2; panel++) {
    long double a = panel * T / 2, b = a + T / 2;
    for (int i = 0; i < GL_M; i++) {
      long double t = 0.5L * (b - a) * gl_x[i] + 0.5L * (a + b);
      uj[M] = t * t;
      mj[M] = 0.5L * (b - a) * gl_w[i] * expl(-(long double)X * t * t);
      M++;
    }
  }
  long double pprev[2 * GL_M], pcur[2 * GL_M], alpha[12], beta[12], h[12];
  for (int j = 0; j < M; j++) { pprev[j] = 0; pcur[j] = 1; }
  for (int k = 0; k < n; k++) {
    long double nrm = 0, un = 0;
    for (int j = 0; j < M; j++) {
      long double v = mj[j] * pcur[j] * pcur[j];
      nrm += v;
      un += v * uj[j];
    }
    h[k] = nrm;
    alpha[k] = un / nrm;
    beta[k] = (k == 0) ? nrm : nrm / h[k - 1];
    for (int j = 0; j < M; j++) {
      long double nxt = (uj[j] - alpha[k]) * pcur[j] - ((k == 0) ? 0 : beta[k]) * pprev[j];
      pprev[j] = pcur[j];
      pcur[j] = nxt;
    }
  }
  /* roots of p_n by bisection with Sturm counts on the Jacobi matrix */
  for (int r = 0; r < n; r++) {
    long double lo = 0, hi = 1;
    for (int it = 0; it < 48; it++) {
      long double mid = 0.5L * (lo + hi);
      /* number of eigenvalues < mid */
      int cnt = 0;
      long double q = alpha[0] - mid;
      if (q < 0) cnt++;
      for (int k = 1; k < n; k++) {
        if (q == 0) q = 1e-300L;
        q = (alpha[k] - mid) - beta[k] / q;
        if (q < 0) cnt++;
      }
      if (cnt > r) hi = mid; else lo = mid;
    }
    long double u = 0.5L * (lo + hi);
    /* Newton polish on p_n(u), derivative by the differentiated recurrence */
    for (int it = 0; it < 3; it++) {
      long double p0 = 1, pm = 0, d0 = 0, dm = 0;
      for (int k = 0; k < n; k++) {
        long double p1 = (u - alpha[k]) * p0 - ((k == 0) ? 0 : beta[k]) * pm;
        long double d1 = p0 + (u - alpha[k]) * d0 - ((k == 0) ? 0 : beta[k]) * dm;
        pm = p0; p0 = p1; dm = d0; d0 = d1;
      }
      if (d0 != 0) {
        long double un = u - p0 / d0;
        if (un > lo && un < hi) u = un;
      }
    }
    /* Christoffel number w = 1 / sum_k p_k(u)^2 / h_k */
    long double p0 = 1, pm = 0, s = 1.0L / h[0];
    for (int k = 0; k + 1 < n; k++) {
      long double p1 = (u - alpha[k]) * p0 - ((k == 0) ? 0 : beta[k]) * pm;
      pm = p0;
      p0 = p1;
      s += p0 * p0 / h[k + 1];
    }
    u_out[r] = (double)u;
    w_out[r] = (double)(1.0L / s);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Rys ERI over primitives (crys.c:56-97, Int1d 1445-1453, Recur       */
/* 1481-1516, Shift 1518-1533, ABD eqs 5, 9, 11-16)                   */
/* ------------------------------------------------------------------ */
static int ifact(int n) { return n <= 1 ? 1 : n * ifact(n - 1); }
static int binom(int a, int b) { return ifact(a) / (ifact(b) * ifact(a - b)); }
static int fact2(int n) { return n <= 1 ? 1 : n * fact2(n - 2); }

static double rys_int1d(double u, int i, int j, int k, int l, double xi, double xj,
                        double xk, double xl, double ai, double aj, double ak,
                        double al) {
  /* Recur: G(n,m) = I(n,0,m,0) for one root u = t^2 */
  double G[12][12];
  int n = i + j, m = k + l;
  double A = ai + aj, B = ak + al;
  double Px = (ai * xi + aj * xj) / A, Qx = (ak * xk + al * xl) / B;
  double fff = u / (A + B);
  double B00 = 0.5 * fff, B1 = (1 - B * fff) / (2 * A), B1p = (1 - A * fff) / (2 * B);
  double C = (Px - xi) + B * (Qx - Px) * fff, Cp = (Qx - xk) + A * (Px - Qx) * fff;
  G[0][0] = M_PI * exp(-ai * aj * (xi - xj) * (xi - xj) / A - ak * al * (xk - xl) * (xk - xl) / B) /
            sqrt(A * B);
  if (n > 0) G[1][0] = C * G[0][0];
  if (m > 0) G[0][1] = Cp * G[0][0];
  for (int a = 2; a <= n; a++) G[a][0] = B1 * (a - 1) * G[a - 2][0] + C * G[a - 1][0];
  for (int b = 2; b <= m; b++) G[0][b] = B1p * (b - 1) * G[0][b - 2] + Cp * G[0][b - 1];
  if (m > 0 && n > 0) {
    for (int a = 1; a <= n; a++) {
      G[a][1] = a * B00 * G[a - 1][0] + Cp * G[a][0];
      for (int b = 2; b <= m; b++)
        G[a][b] = B1p * (b - 1) * G[a][b - 2] + a * B00 * G[a - 1][b - 1] + Cp * G[a][b - 1];
    }
  }
  /* Shift: I(i,j,k,l) from G */
  double xij = xi - xj, xkl = xk - xl, ijkl = 0;
  for (int mm = 0; mm <= l; mm++) {
    double ijm0 = 0;
    for (int nn = 0; nn <= j; nn++) ijm0 += binom(j, nn) * pow(xij, j - nn) * G[nn + i][mm + k];
    ijkl += binom(l, mm) * pow(xkl, l - mm) * ijm0;
  }
  return ijkl;
}

double pqo_rys_coulomb_repulsion(const double *A, double na, const int *pa, double aa,
                                 const double *B, double nb, const int *pb, double ab,
                                 const double *Cc, double nc, const int *pc, double ac,
                                 const double *D, double nd, const int *pd, double ad) {
  int L = pa[0] + pa[1] + pa[2] + pb[0] + pb[1] + pb[2] + pc[0] + pc[1] + pc[2] + pd[0] +
          pd[1] + pd[2];
  int norder = L / 2 + 1;
  double Ap = aa + ab, Bq = ac + ad, rho = Ap * Bq / (Ap + Bq);
  double rpq2 = 0;
  for (int x = 0; x < 3; x++) {
    double P = (aa * A[x] + ab * B[x]) / Ap, Q = (ac * Cc[x] + ad * D[x]) / Bq;
    rpq2 += (P - Q) * (P - Q);
  }
  double X = rpq2 * rho, u[12], w[12];
  pqo_rys_roots(norder, X, u, w);
  double sum = 0;
  for (int r = 0; r < norder; r++) {
    double I[3];
    for (int x = 0; x < 3; x++)
      I[x] = rys_int1d(u[r], pa[x], pb[x], pc[x], pd[x], A[x], B[x], Cc[x], D[x], aa, ab, ac, ad);
    sum += I[0] * I[1] * I[2] * w[r];
  }
  return 2 * sqrt(rho / M_PI) * na * nb * nc * nd * sum;
}

/* ------------------------------------------------------------------ */
/* THO / Huzinaga ERI (cints.c:110-187)                                */
/* ------------------------------------------------------------------ */
static double binomial_prefactor(int s, int ia, int ib, double xpa, double xpb) {
  /* cints.c:430-437 */
  double sum = 0;
  for (int t = 0; t < s + 1; t++)
    if ((s - ia <= t) && (t <= ib))
      sum += binom(ia, s - t) * binom(ib, t) * pow(xpa, ia - s + t) * pow(xpb, ib - t);
  return sum;
}
static double fact_ratio2(int a, int b) { return (double)ifact(a) / ifact(b) / ifact(a - 2 * b); }
static double tho_fB(int i, int l1, int l2, double px, double ax, double bx, int r, double g) {
  return binomial_prefactor(i, l1, l2, px - ax, px - bx) * fact_ratio2(i, r) * pow(4 * g, r - i);
}
static double tho_Bterm(int i1, int i2, int r1, int r2, int u, int l1, int l2, int l3, int l4,
                        double Px, double Ax, double Bx, double Qx, double Cx, double Dx,
                        double g1, double g2, double delta) {
  /* THO eq. 2.22, cints.c:176-187 */
  return tho_fB(i1, l1, l2, Px, Ax, Bx, r1, g1) * pow(-1, i2) *
         tho_fB(i2, l3, l4, Qx, Cx, Dx, r2, g2) * pow(-1, u) * fact_ratio2(i1 + i2 - 2 * (r1 + r2), u) *
         pow(Qx - Px, i1 + i2 - 2 * (r1 + r2) - 2 * u) / pow(delta, i1 + i2 - 2 * (r1 + r2) - u);
}
static void tho_Barray(double *Bv, int l1, int l2, int l3, int l4, double p, double a, double b,
                       double q, double c, double d, double g1, double g2, double delta) {
  /* cints.c:155-174 */
  int Imax = l1 + l2 + l3 + l4 + 1;
  for (int i = 0; i < Imax; i++) Bv[i] = 0;
  for (int i1 = 0; i1 < l1 + l2 + 1; i1++)
    for (int i2 = 0; i2 < l3 + l4 + 1; i2++)
      for (int r1 = 0; r1 < i1 / 2 + 1; r1++)
        for (int r2 = 0; r2 < i2 / 2 + 1; r2++)
          for (int u = 0; u < (i1 + i2) / 2 - r1 - r2 + 1; u++) {
            int I = i1 + i2 - 2 * (r1 + r2) - u;
            Bv[I] += tho_Bterm(i1, i2, r1, r2, u, l1, l2, l3, l4, p, a, b, q, c, d, g1, g2, delta);
          }
}

double pqo_tho_coulomb_repulsion(const double *A, double na, const int *pa, double aa,
                                 const double *B, double nb, const int *pb, double ab,
                                 const double *Cc, double nc, const int *pc, double ac,
                                 const double *D, double nd, const int *pd, double ad,
                                 int compat_nr) {
  double rab2 = 0, rcd2 = 0, rpq2 = 0, P[3], Q[3];
  double g1 = aa + ab, g2 = ac + ad, delta = (1. / g1 + 1. / g2) / 4.;
  for (int x = 0; x < 3; x++) {
    P[x] = (aa * A[x] + ab * B[x]) / g1;
    Q[x] = (ac * Cc[x] + ad * D[x]) / g2;
    rab2 += (A[x] - B[x]) * (A[x] - B[x]);
    rcd2 += (Cc[x] - D[x]) * (Cc[x] - D[x]);
    rpq2 += (P[x] - Q[x]) * (P[x] - Q[x]);
  }
  double Bx[24], By[24], Bz[24];
  tho_Barray(Bx, pa[0], pb[0], pc[0], pd[0], P[0], A[0], B[0], Q[0], Cc[0], D[0], g1, g2, delta);
  tho_Barray(By, pa[1], pb[1], pc[1], pd[1], P[1], A[1], B[1], Q[1], Cc[1], D[1], g1, g2, delta);
  tho_Barray(Bz, pa[2], pb[2], pc[2], pd[2], P[2], A[2], B[2], Q[2], Cc[2], D[2], g1, g2, delta);
  double sum = 0;
  for (int I = 0; I < pa[0] + pb[0] + pc[0] + pd[0] + 1; I++)
    for (int J = 0; J < pa[1] + pb[1] + pc[1] + pd[1] + 1; J++)
      for (int K = 0; K < pa[2] + pb[2] + pc[2] + pd[2] + 1; K++) {
        double F = compat_nr ? pqo_fgamma_nr(I + J + K, 0.25 * rpq2 / delta)
                             : pqo_boys(I + J + K, 0.25 * rpq2 / delta);
        sum += Bx[I] * By[J] * Bz[K] * F;
      }
  return 2. * pow(M_PI, 2.5) / (g1 * g2 * sqrt(g1 + g2)) * exp(-aa * ab * rab2 / g1) *
         exp(-ac * ad * rcd2 / g2) * sum * na * nb * nc * nd;
}

/* ------------------------------------------------------------------ */
/* Head-Gordon-Pople ERI (chgp.c:172-419): HRR on B then D, VRR leaf   */
/* ------------------------------------------------------------------ */
#define HGP_MAXAM 9
#define HGP_MAXM 24
static double hgp_vrr(const double *A, double na, const int *pa, double aa, const double *B,
                      double nb, double ab, const double *Cc, double nc, const int *pc, double ac,
                      const double *D, double nd, double ad, int compat_nr) {
  /* chgp.c:244-419: Obara-Saika VRR on (a0|c0)^(m) */
  double P[3], Q[3], W[3], rab2 = 0, rcd2 = 0, rpq2 = 0;
  double zeta = aa + ab, eta = ac + ad;
  for (int x = 0; x < 3; x++) {
    P[x] = (aa * A[x] + ab * B[x]) / zeta;
    Q[x] = (ac * Cc[x] + ad * D[x]) / eta;
    W[x] = (zeta * P[x] + eta * Q[x]) / (zeta + eta);
    rab2 += (A[x] - B[x]) * (A[x] - B[x]);
    rcd2 += (Cc[x] - D[x]) * (Cc[x] - D[x]);
    rpq2 += (P[x] - Q[x]) * (P[x] - Q[x]);
  }
  double Kab = sqrt(2.) * pow(M_PI, 1.25) / zeta * exp(-aa * ab / zeta * rab2);
  double Kcd = sqrt(2.) * pow(M_PI, 1.25) / eta * exp(-ac * ad / eta * rcd2);
  double T = zeta * eta / (zeta + eta) * rpq2;
  int la = pa[0], ma = pa[1], naa = pa[2], lc = pc[0], mc = pc[1], nc_ = pc[2];
  int mtot = la + ma + naa + lc + mc + nc_;
  double Fg[HGP_MAXM];
  Fg[mtot] = compat_nr ? pqo_fgamma_nr(mtot, T) : pqo_boys(mtot, T);
  for (int im = mtot - 1; im >= 0; im--) Fg[im] = (2. * T * Fg[im + 1] + exp(-T)) / (2. * im + 1);
  /* table v[i][j][k][q][r][s][m], dims (la+1)(ma+1)(na+1)(lc+1)(mc+1)(nc+1)(mtot+1) */
  int d0 = la + 1, d1 = ma + 1, d2 = naa + 1, d3 = lc + 1, d4 = mc + 1, d5 = nc_ + 1, d6 = mtot + 1;
  size_t tot = (size_t)d0 * d1 * d2 * d3 * d4 * d5 * d6;
  double *v = (double *)calloc(tot, sizeof(double));
#define VI(i, j, k, q, r, s, m) \
  (((((((size_t)(m) * d5 + (s)) * d4 + (r)) * d3 + (q)) * d2 + (k)) * d1 + (j)) * d0 + (i))
  for (int im = 0; im <= mtot; im++)
    v[VI(0, 0, 0, 0, 0, 0, im)] = na * nb * nc * nd * Kab * Kcd / sqrt(zeta + eta) * Fg[im];
  double rz = eta / (zeta + eta), re = zeta / (zeta + eta);
  for (int i = 0; i < la; i++)
    for (int im = 0; im < mtot - i; im++) {
      double t = (P[0] - A[0]) * v[VI(i, 0, 0, 0, 0, 0, im)] + (W[0] - P[0]) * v[VI(i, 0, 0, 0, 0, 0, im + 1)];
      if (i > 0) t += i / 2. / zeta * (v[VI(i - 1, 0, 0, 0, 0, 0, im)] - rz * v[VI(i - 1, 0, 0, 0, 0, 0, im + 1)]);
      v[VI(i + 1, 0, 0, 0, 0, 0, im)] = t;
    }
  for (int j = 0; j < ma; j++)
    for (int i = 0; i < la + 1; i++)
      for (int im = 0; im < mtot - i - j; im++) {
        double t = (P[1] - A[1]) * v[VI(i, j, 0, 0, 0, 0, im)] + (W[1] - P[1]) * v[VI(i, j, 0, 0, 0, 0, im + 1)];
        if (j > 0) t += j / 2. / zeta * (v[VI(i, j - 1, 0, 0, 0, 0, im)] - rz * v[VI(i, j - 1, 0, 0, 0, 0, im + 1)]);
        v[VI(i, j + 1, 0, 0, 0, 0, im)] = t;
      }
  for (int k = 0; k < naa; k++)
    for (int j = 0; j < ma + 1; j++)
      for (int i = 0; i < la + 1; i++)
        for (int im = 0; im < mtot - i - j - k; im++) {
          double t = (P[2] - A[2]) * v[VI(i, j, k, 0, 0, 0, im)] + (W[2] - P[2]) * v[VI(i, j, k, 0, 0, 0, im + 1)];
          if (k > 0) t += k / 2. / zeta * (v[VI(i, j, k - 1, 0, 0, 0, im)] - rz * v[VI(i, j, k - 1, 0, 0, 0, im + 1)]);
          v[VI(i, j, k + 1, 0, 0, 0, im)] = t;
        }
  for (int q = 0; q < lc; q++)
    for (int k = 0; k < naa + 1; k++)
      for (int j = 0; j < ma + 1; j++)
        for (int i = 0; i < la + 1; i++)
          for (int im = 0; im < mtot - i - j - k - q; im++) {
            double t = (Q[0] - Cc[0]) * v[VI(i, j, k, q, 0, 0, im)] + (W[0] - Q[0]) * v[VI(i, j, k, q, 0, 0, im + 1)];
            if (q > 0) t += q / 2. / eta * (v[VI(i, j, k, q - 1, 0, 0, im)] - re * v[VI(i, j, k, q - 1, 0, 0, im + 1)]);
            if (i > 0) t += i / 2. / (zeta + eta) * v[VI(i - 1, j, k, q, 0, 0, im + 1)];
            v[VI(i, j, k, q + 1, 0, 0, im)] = t;
          }
  for (int r = 0; r < mc; r++)
    for (int q = 0; q < lc + 1; q++)
      for (int k = 0; k < naa + 1; k++)
        for (int j = 0; j < ma + 1; j++)
          for (int i = 0; i < la + 1; i++)
            for (int im = 0; im < mtot - i - j - k - q - r; im++) {
              double t = (Q[1] - Cc[1]) * v[VI(i, j, k, q, r, 0, im)] + (W[1] - Q[1]) * v[VI(i, j, k, q, r, 0, im + 1)];
              if (r > 0) t += r / 2. / eta * (v[VI(i, j, k, q, r - 1, 0, im)] - re * v[VI(i, j, k, q, r - 1, 0, im + 1)]);
              if (j > 0) t += j / 2. / (zeta + eta) * v[VI(i, j - 1, k, q, r, 0, im + 1)];
              v[VI(i, j, k, q, r + 1, 0, im)] = t;
            }
  for (int s = 0; s < nc_; s++)
    for (int r = 0; r < mc + 1; r++)
      for (int q = 0; q < lc + 1; q++)
        for (int k = 0; k < naa + 1; k++)
          for (int j = 0; j < ma + 1; j++)
            for (int i = 0; i < la + 1; i++)
              for (int im = 0; im < mtot - i - j - k - q - r - s; im++) {
                double t = (Q[2] - Cc[2]) * v[VI(i, j, k, q, r, s, im)] + (W[2] - Q[2]) * v[VI(i, j, k, q, r, s, im + 1)];
                if (s > 0) t += s / 2. / eta * (v[VI(i, j, k, q, r, s - 1, im)] - re * v[VI(i, j, k, q, r, s - 1, im + 1)]);
                if (k > 0) t += k / 2. / (zeta + eta) * v[VI(i, j, k - 1, q, r, s, im + 1)];
                v[VI(i, j, k, q, r, s + 1, im)] = t;
              }
  double val = v[VI(la, ma, naa, lc, mc, nc_, 0)];
#undef VI
  free(v);
  return val;
}

double pqo_hgp_coulomb_repulsion(const double *A, double na, const int *pa, double aa,
                                 const double *B, double nb, const int *pb, double ab,
                                 const double *Cc, double nc, const int *pc, double ac,
                                 const double *D, double nd, const int *pd, double ad,
                                 int compat_nr) {
  /* chgp.c:172-242: move angular momentum off B (x,y,z) then off D */
  for (int x = 0; x < 3; x++) {
    if (pb[x] > 0) {
      int pa1[3] = {pa[0], pa[1], pa[2]}, pb1[3] = {pb[0], pb[1], pb[2]};
      pa1[x]++;
      pb1[x]--;
      double t1 = pqo_hgp_coulomb_repulsion(A, na, pa1, aa, B, nb, pb1, ab, Cc, nc, pc, ac, D, nd, pd, ad, compat_nr);
      double t2 = pqo_hgp_coulomb_repulsion(A, na, pa, aa, B, nb, pb1, ab, Cc, nc, pc, ac, D, nd, pd, ad, compat_nr);
      return t1 + (A[x] - B[x]) * t2;
    }
  }
  for (int x = 0; x < 3; x++) {
    if (pd[x] > 0) {
      int pc1[3] = {pc[0], pc[1], pc[2]}, pd1[3] = {pd[0], pd[1], pd[2]};
      pc1[x]++;
      pd1[x]--;
      double t1 = pqo_hgp_coulomb_repulsion(A, na, pa, aa, B, nb, pb, ab, Cc, nc, pc1, ac, D, nd, pd1, ad, compat_nr);
      double t2 = pqo_hgp_coulomb_repulsion(A, na, pa, aa, B, nb, pb, ab, Cc, nc, pc, ac, D, nd, pd1, ad, compat_nr);
      return t1 + (Cc[x] - D[x]) * t2;
    }
  }
  return hgp_vrr(A, na, pa, aa, B, nb, ab, Cc, nc, pc, ac, D, nd, ad, compat_nr);
}

/* ------------------------------------------------------------------ */
/* Contracted ERI (crys.c:34-54 / cints.c:80-108 / chgp.c:78-169)      */
/* method 0 = THO (cints), 1 = Rys (crys), 2 = HGP (chgp)               */
/* ------------------------------------------------------------------ */
typedef struct {
  int n;
  const double *exps, *coefs, *norms;
  double org[3];
  int pw[3];
} cgbf_t;

static double prim_eri(int method, int compat, const cgbf_t *a, int i, const cgbf_t *b, int j,
                       const cgbf_t *c, int k, const cgbf_t *d, int l) {
  switch (method) {
    case 0:
      return pqo_tho_coulomb_repulsion(a->org, a->norms[i], a->pw, a->exps[i], b->org, b->norms[j], b->pw,
                                       b->exps[j], c->org, c->norms[k], c->pw, c->exps[k], d->org,
                                       d->norms[l], d->pw, d->exps[l], compat);
    case 2:
      return pqo_hgp_coulomb_repulsion(a->org, a->norms[i], a->pw, a->exps[i], b->org, b->norms[j], b->pw,
                                       b->exps[j], c->org, c->norms[k], c->pw, c->exps[k], d->org,
                                       d->norms[l], d->pw, d->exps[l], compat);
    default:
      return pqo_rys_coulomb_repulsion(a->org, a->norms[i], a->pw, a->exps[i], b->org, b->norms[j], b->pw,
                                       b->exps[j], c->org, c->norms[k], c->pw, c->exps[k], d->org,
                                       d->norms[l], d->pw, d->exps[l]);
  }
}

static double contr_eri(int method, int compat, const cgbf_t *a, const cgbf_t *b, const cgbf_t *c,
                        const cgbf_t *d) {
  double val = 0;
  for (int i = 0; i < a->n; i++)
    for (int j = 0; j < b->n; j++)
      for (int k = 0; k < c->n; k++)
        for (int l = 0; l < d->n; l++)
          val += a->coefs[i] * b->coefs[j] * c->coefs[k] * d->coefs[l] *
                 prim_eri(method, compat, a, i, b, j, c, k, d, l);
  return val;
}

double pqo_contr_coulomb(int method, int compat_nr, const int *nprim, const double *exps,
                         const double *coefs, const double *norms, const double *origins,
                         const int *powers) {
  /* four functions packed back to back: nprim[4], exps/coefs/norms
     concatenated, origins[12], powers[12] */
  cgbf_t f[4];
  int off = 0;
  for (int q = 0; q < 4; q++) {
    f[q].n = nprim[q];
    f[q].exps = exps + off;
    f[q].coefs = coefs + off;
    f[q].norms = norms + off;
    for (int x = 0; x < 3; x++) {
      f[q].org[x] = origins[3 * q + x];
      f[q].pw[x] = powers[3 * q + x];
    }
    off += nprim[q];
  }
  return contr_eri(method, compat_nr, &f[0], &f[1], &f[2], &f[3]);
}

/* ------------------------------------------------------------------ */
/* ijkl2intindex (cints.c:413-425) with 64-bit arithmetic              */
/* ------------------------------------------------------------------ */
uint64_t pqo_ijkl2intindex(int64_t i, int64_t j, int64_t k, int64_t l) {
  if (i < j) { int64_t t = i; i = j; j = t; }
  if (k < l) { int64_t t = k; k = l; l = t; }
  uint64_t ij = (uint64_t)(i * (i + 1) / 2 + j), kl = (uint64_t)(k * (k + 1) / 2 + l);
  if (ij < kl) { uint64_t t = ij; ij = kl; kl = t; }
  return ij * (ij + 1) / 2 + kl;
}

/* ------------------------------------------------------------------ */
/* Basis description used by the drivers below: function f has        */
/* nprim[f] primitives starting at poff[f] in exps/coefs/norms,        */
/* origin org[3f..3f+2] and powers pw[3f..3f+2].                       */
/* ------------------------------------------------------------------ */
static void load_bf(cgbf_t *f, int idx, const int *nprim, const int *poff, const double *exps,
                    const double *coefs, const double *norms, const double *org, const int *pw) {
  f->n = nprim[idx];
  f->exps = exps + poff[idx];
  f->coefs = coefs + poff[idx];
  f->norms = norms + poff[idx];
  for (int x = 0; x < 3; x++) {
    f->org[x] = org[3 * idx + x];
    f->pw[x] = pw[3 * idx + x];
  }
}

/* get2ints (Ints.py:146-167): loops i, j<=i, k, l<=k, keeps ij <= kl */
int pqo_get2ints(int method, int compat_nr, int nbf, const int *nprim, const int *poff,
                 const double *exps, const double *coefs, const double *norms, const double *org,
                 const int *pw, double *ints) {
  cgbf_t fi, fj, fk, fl;
  for (int i = 0; i < nbf; i++) {
    load_bf(&fi, i, nprim, poff, exps, coefs, norms, org, pw);
    for (int j = 0; j <= i; j++) {
      load_bf(&fj, j, nprim, poff, exps, coefs, norms, org, pw);
      int64_t ij = (int64_t)i * (i + 1) / 2 + j;
      for (int k = 0; k < nbf; k++) {
        load_bf(&fk, k, nprim, poff, exps, coefs, norms, org, pw);
        for (int l = 0; l <= k; l++) {
          int64_t kl = (int64_t)k * (k + 1) / 2 + l;
          if (ij > kl) continue;
          load_bf(&fl, l, nprim, poff, exps, coefs, norms, org, pw);
          ints[pqo_ijkl2intindex(i, j, k, l)] = contr_eri(method, compat_nr, &fi, &fj, &fk, &fl);
        }
      }
    }
  }
  return 0;
}

/* getJ / getK / get2JmK (Ints.py:195-238): J_ij = sum_kl (ij|kl) D_kl,
   K_ij = sum_kl (ik|jl) D_kl, evaluated for i>=j and mirrored. */
void pqo_getJK(int nbf, const double *ints, const double *D, double *J, double *K, double *G) {
  for (int i = 0; i < nbf; i++)
    for (int j = 0; j <= i; j++) {
      double sj = 0, sk = 0;
      for (int k = 0; k < nbf; k++)
        for (int l = 0; l < nbf; l++) {
          double dkl = D[(size_t)k * nbf + l];
          sj += ints[pqo_ijkl2intindex(i, j, k, l)] * dkl;
          sk += ints[pqo_ijkl2intindex(i, k, j, l)] * dkl;
        }
      if (J) J[(size_t)i * nbf + j] = J[(size_t)j * nbf + i] = sj;
      if (K) K[(size_t)i * nbf + j] = K[(size_t)j * nbf + i] = sk;
      if (G) G[(size_t)i * nbf + j] = G[(size_t)j * nbf + i] = 2 * sj - sk;
    }
}

/* ------------------------------------------------------------------ */
/* One-electron integrals (cints.c:190-307, THO eqs 2.12, 2.14, 2.17)   */
/* ------------------------------------------------------------------ */
static double overlap_1D(int l1, int l2, double PAx, double PBx, double gamma) {
  double sum = 0;
  for (int i = 0; i < (1 + (int)floor(0.5 * (l1 + l2))); i++)
    sum += binomial_prefactor(2 * i, l1, l2, PAx, PBx) * fact2(2 * i - 1) / pow(2 * gamma, i);
  return sum;
}

double pqo_overlap(double a1, const int *p1, const double *A, double a2, const int *p2,
                   const double *B) {
  double rab2 = 0, P[3];
  double g = a1 + a2;
  for (int x = 0; x < 3; x++) {
    P[x] = (a1 * A[x] + a2 * B[x]) / g;
    rab2 += (A[x] - B[x]) * (A[x] - B[x]);
  }
  double pre = pow(M_PI / g, 1.5) * exp(-a1 * a2 * rab2 / g);
  return pre * overlap_1D(p1[0], p2[0], P[0] - A[0], P[0] - B[0], g) *
         overlap_1D(p1[1], p2[1], P[1] - A[1], P[1] - B[1], g) *
         overlap_1D(p1[2], p2[2], P[2] - A[2], P[2] - B[2], g);
}

double pqo_kinetic(double a1, const int *p1, const double *A, double a2, const int *p2,
                   const double *B) {
  /* cints.c:188-212 */
  int l2 = p2[0], m2 = p2[1], n2 = p2[2];
  double t0 = a2 * (2 * (l2 + m2 + n2) + 3) * pqo_overlap(a1, p1, A, a2, p2, B);
  double t1 = 0, t2 = 0;
  for (int x = 0; x < 3; x++) {
    int q[3] = {p2[0], p2[1], p2[2]};
    q[x] += 2;
    t1 += pqo_overlap(a1, p1, A, a2, q, B);
    if (p2[x] >= 2) {
      q[x] -= 4;
      t2 += p2[x] * (p2[x] - 1) * pqo_overlap(a1, p1, A, a2, q, B);
    }
  }
  return t0 - 2 * a2 * a2 * t1 - 0.5 * t2;
}

static double A_term(int i, int r, int u, int l1, int l2, double PAx, double PBx, double CPx,
                     double gamma) {
  /* THO eq. 2.18, cints.c:277-284 */
  return pow(-1, i) * binomial_prefactor(i, l1, l2, PAx, PBx) * pow(-1, u) * ifact(i) *
         pow(CPx, i - 2 * r - 2 * u) * pow(0.25 / gamma, r + u) / ifact(r) / ifact(u) /
         ifact(i - 2 * r - 2 * u);
}
static void A_array(double *Av, int l1, int l2, double PA, double PB, double CP, double g) {
  int Imax = l1 + l2 + 1;
  for (int i = 0; i < Imax; i++) Av[i] = 0;
  for (int i = 0; i < Imax; i++)
    for (int r = 0; r < (int)floor(i / 2) + 1; r++)
      for (int u = 0; u < (int)floor((i - 2 * r) / 2.) + 1; u++) {
        int I = i - 2 * r - u;
        Av[I] += A_term(i, r, u, l1, l2, PA, PB, CP, g);
      }
}

double pqo_nuclear_attraction(const double *A, double n1, const int *p1, double a1, const double *B,
                              double n2, const int *p2, double a2, const double *C, int compat_nr) {
  /* cints.c:241-275 */
  double g = a1 + a2, P[3], rab2 = 0, rcp2 = 0;
  for (int x = 0; x < 3; x++) {
    P[x] = (a1 * A[x] + a2 * B[x]) / g;
    rab2 += (A[x] - B[x]) * (A[x] - B[x]);
    rcp2 += (C[x] - P[x]) * (C[x] - P[x]);
  }
  double Ax[12], Ay[12], Az[12];
  A_array(Ax, p1[0], p2[0], P[0] - A[0], P[0] - B[0], P[0] - C[0], g);
  A_array(Ay, p1[1], p2[1], P[1] - A[1], P[1] - B[1], P[1] - C[1], g);
  A_array(Az, p1[2], p2[2], P[2] - A[2], P[2] - B[2], P[2] - C[2], g);
  double sum = 0;
  for (int I = 0; I < p1[0] + p2[0] + 1; I++)
    for (int J = 0; J < p1[1] + p2[1] + 1; J++)
      for (int K = 0; K < p1[2] + p2[2] + 1; K++) {
        double F = compat_nr ? pqo_fgamma_nr(I + J + K, rcp2 * g) : pqo_boys(I + J + K, rcp2 * g);
        sum += Ax[I] * Ay[J] * Az[K] * F;
      }
  return -n1 * n2 * 2 * M_PI / g * exp(-a1 * a2 * rab2 / g) * sum;
}

/* get1ints (Ints.py:73-87, CGBF.py:84-114): S and h = T + sum_A Z_A V_A */
void pqo_get1ints(int compat_nr, int nbf, const int *nprim, const int *poff, const double *exps,
                  const double *coefs, const double *norms, const double *org, const int *pw,
                  int natoms, const double *atom_xyz, const double *atom_z, double *S, double *h) {
  cgbf_t fi, fj;
  for (int i = 0; i < nbf; i++) {
    load_bf(&fi, i, nprim, poff, exps, coefs, norms, org, pw);
    for (int j = 0; j < nbf; j++) {
      load_bf(&fj, j, nprim, poff, exps, coefs, norms, org, pw);
      double s = 0, t = 0, v = 0;
      for (int p = 0; p < fi.n; p++)
        for (int q = 0; q < fj.n; q++) {
          double cc = fi.coefs[p] * fj.coefs[q];
          double nn = fi.norms[p] * fj.norms[q];
          s += cc * nn * pqo_overlap(fi.exps[p], fi.pw, fi.org, fj.exps[q], fj.pw, fj.org);
          t += cc * nn * pqo_kinetic(fi.exps[p], fi.pw, fi.org, fj.exps[q], fj.pw, fj.org);
          for (int a = 0; a < natoms; a++)
            v += atom_z[a] * cc *
                 pqo_nuclear_attraction(fi.org, fi.norms[p], fi.pw, fi.exps[p], fj.org, fj.norms[q],
                                        fj.pw, fj.exps[q], atom_xyz + 3 * a, compat_nr);
        }
      S[(size_t)i * nbf + j] = s;
      h[(size_t)i * nbf + j] = t + v;
    }
  }
}

/* primitive normalisation (primitive-gto.c:76-83) */
double pqo_prim_norm(double alpha, int l, int m, int n) {
  return sqrt(pow(2, 2 * (l + m + n) + 1.5) * pow(alpha, l + m + n + 1.5) / fact2(2 * l - 1) /
              fact2(2 * m - 1) / fact2(2 * n - 1) / pow(M_PI, 1.5));
}
