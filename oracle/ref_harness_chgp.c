/*
 * ref_harness_chgp.c -- compiles the reference's Src/chgp.c numeric core in
 * place (#included from the reference checkout) for golden vectors.
 * TEST INFRASTRUCTURE ONLY; output goes to _ref/.
 */
#include PQ_REF_SRC

/* chgp.c:78-145 contr_hrr */
double pqref_chgp_contr_coulomb(const int *np, const double *e, const double *c, const double *n,
                                const double *o, const int *p) {
  const double *e1 = e + np[0], *e2 = e1 + np[1], *e3 = e2 + np[2];
  const double *c1 = c + np[0], *c2 = c1 + np[1], *c3 = c2 + np[2];
  const double *n1 = n + np[0], *n2 = n1 + np[1], *n3 = n2 + np[2];
  return contr_hrr(np[0], o[0], o[1], o[2], (double *)n, p[0], p[1], p[2], (double *)e, (double *)c,
                   np[1], o[3], o[4], o[5], (double *)n1, p[3], p[4], p[5], (double *)e1, (double *)c1,
                   np[2], o[6], o[7], o[8], (double *)n2, p[6], p[7], p[8], (double *)e2, (double *)c2,
                   np[3], o[9], o[10], o[11], (double *)n3, p[9], p[10], p[11], (double *)e3, (double *)c3);
}

/* chgp.c:172-242 hrr */
double pqref_chgp_coulomb_repulsion(const double *o, const double *nrm, const int *p, const double *a) {
  return hrr(o[0], o[1], o[2], nrm[0], p[0], p[1], p[2], a[0],
             o[3], o[4], o[5], nrm[1], p[3], p[4], p[5], a[1],
             o[6], o[7], o[8], nrm[2], p[6], p[7], p[8], a[2],
             o[9], o[10], o[11], nrm[3], p[9], p[10], p[11], a[3]);
}

/* batch of contracted quartets in the pq_contr_coulomb_batch layout (CPU
 * baseline timing of the reference's HGP path) */
void pqref_chgp_contr_batch(int n, const int *nprim, const int *poff, const double *e, const double *c,
                            const double *nrm, const double *o, const int *p, double *out) {
  for (int q = 0; q < n; ++q) {
    const int *np = nprim + 4 * q, *po = poff + 4 * q;
    const double *oo = o + 12 * q;
    const int *pp = p + 12 * q;
    out[q] = contr_hrr(np[0], oo[0], oo[1], oo[2], (double *)nrm + po[0], pp[0], pp[1], pp[2], (double *)e + po[0],
                       (double *)c + po[0], np[1], oo[3], oo[4], oo[5], (double *)nrm + po[1], pp[3], pp[4], pp[5],
                       (double *)e + po[1], (double *)c + po[1], np[2], oo[6], oo[7], oo[8], (double *)nrm + po[2],
                       pp[6], pp[7], pp[8], (double *)e + po[2], (double *)c + po[2], np[3], oo[9], oo[10], oo[11],
                       (double *)nrm + po[3], pp[9], pp[10], pp[11], (double *)e + po[3], (double *)c + po[3]);
  }
}
