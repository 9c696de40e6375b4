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
