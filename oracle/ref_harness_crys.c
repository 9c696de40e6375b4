/*
 * ref_harness_crys.c -- compiles the reference's Src/crys.c numeric core in
 * place (the file is #included from the reference checkout, never copied)
 * and exposes it under plain C names for golden-vector generation and the
 * CPU baseline timing.  TEST INFRASTRUCTURE ONLY; output goes to _ref/.
 */
#include PQ_REF_SRC

/* crys.c:34-54 */
double pqref_crys_contr_coulomb(const int *np, const double *e, const double *c, const double *n,
                                const double *o, const int *p) {
  const double *e1 = e + np[0], *e2 = e1 + np[1], *e3 = e2 + np[2];
  const double *c1 = c + np[0], *c2 = c1 + np[1], *c3 = c2 + np[2];
  const double *n1 = n + np[0], *n2 = n1 + np[1], *n3 = n2 + np[2];
  return contr_coulomb(np[0], (double *)e, (double *)c, (double *)n, o[0], o[1], o[2], p[0], p[1], p[2],
                       np[1], (double *)e1, (double *)c1, (double *)n1, o[3], o[4], o[5], p[3], p[4], p[5],
                       np[2], (double *)e2, (double *)c2, (double *)n2, o[6], o[7], o[8], p[6], p[7], p[8],
                       np[3], (double *)e3, (double *)c3, (double *)n3, o[9], o[10], o[11], p[9], p[10], p[11]);
}

/* crys.c:56-97 */
double pqref_crys_coulomb_repulsion(const double *o, const double *nrm, const int *p, const double *a) {
  return coulomb_repulsion(o[0], o[1], o[2], nrm[0], p[0], p[1], p[2], a[0],
                           o[3], o[4], o[5], nrm[1], p[3], p[4], p[5], a[1],
                           o[6], o[7], o[8], nrm[2], p[6], p[7], p[8], a[2],
                           o[9], o[10], o[11], nrm[3], p[9], p[10], p[11], a[3]);
}

/* crys.c:99-109: roots (GAMESS convention u/(1-u)) and weights */
void pqref_crys_roots(int n, double X, double *r, double *w) {
  Roots(n, X);
  for (int i = 0; i < n; i++) {
    r[i] = roots[i];
    w[i] = weights[i];
  }
}

/* Batch over n contracted quartets (layout of pq_contr_coulomb_batch), for
 * timing the reference path without per-call Python overhead. */
void pqref_crys_contr_batch(int n, const int *nprim, const int *poff, const double *e, const double *c,
                            const double *nrm, const double *o, const int *p, double *out) {
  for (int q = 0; q < n; ++q) {
    const int *np = nprim + 4 * q, *po = poff + 4 * q;
    const double *oo = o + 12 * q;
    const int *pp = p + 12 * q;
    out[q] = contr_coulomb(np[0], (double *)e + po[0], (double *)c + po[0], (double *)nrm + po[0], oo[0], oo[1],
                           oo[2], pp[0], pp[1], pp[2], np[1], (double *)e + po[1], (double *)c + po[1],
                           (double *)nrm + po[1], oo[3], oo[4], oo[5], pp[3], pp[4], pp[5], np[2],
                           (double *)e + po[2], (double *)c + po[2], (double *)nrm + po[2], oo[6], oo[7], oo[8],
                           pp[6], pp[7], pp[8], np[3], (double *)e + po[3], (double *)c + po[3],
                           (double *)nrm + po[3], oo[9], oo[10], oo[11], pp[9], pp[10], pp[11]);
  }
}
