/*
 * ref_harness_cints.c -- compiles the reference's Src/cints.c numeric core
 * in place (#included from the reference checkout) for golden vectors.
 * TEST INFRASTRUCTURE ONLY; output goes to _ref/.
 */
#include PQ_REF_SRC

/* cints.c:80-108 */
double pqref_cints_contr_coulomb(const int *np, const double *e, const double *c, const double *n,
                                 const double *o, const int *p) {
  const double *e1 = e + np[0], *e2 = e1 + np[1], *e3 = e2 + np[2];
  const double *c1 = c + np[0], *c2 = c1 + np[1], *c3 = c2 + np[2];
  const double *n1 = n + np[0], *n2 = n1 + np[1], *n3 = n2 + np[2];
  return contr_coulomb(np[0], (double *)e, (double *)c, (double *)n, o[0], o[1], o[2], p[0], p[1], p[2],
                       np[1], (double *)e1, (double *)c1, (double *)n1, o[3], o[4], o[5], p[3], p[4], p[5],
                       np[2], (double *)e2, (double *)c2, (double *)n2, o[6], o[7], o[8], p[6], p[7], p[8],
                       np[3], (double *)e3, (double *)c3, (double *)n3, o[9], o[10], o[11], p[9], p[10], p[11]);
}

/* cints.c:110-153 */
double pqref_cints_coulomb_repulsion(const double *o, const double *nrm, const int *p, const double *a) {
  return coulomb_repulsion(o[0], o[1], o[2], nrm[0], p[0], p[1], p[2], a[0],
                           o[3], o[4], o[5], nrm[1], p[3], p[4], p[5], a[1],
                           o[6], o[7], o[8], nrm[2], p[6], p[7], p[8], a[2],
                           o[9], o[10], o[11], nrm[3], p[9], p[10], p[11], a[3]);
}

/* cints.c:340-346 */
double pqref_cints_fgamma(double m, double x) { return Fgamma(m, x); }

/* cints.c:413-425 (C int, as the reference has it) */
int pqref_cints_ijkl2intindex(int i, int j, int k, int l) { return ijkl2intindex(i, j, k, l); }

/* cints.c:190-275 */
double pqref_cints_overlap(double a1, const int *p1, const double *A, double a2, const int *p2,
                           const double *B) {
  return overlap(a1, p1[0], p1[1], p1[2], A[0], A[1], A[2], a2, p2[0], p2[1], p2[2], B[0], B[1], B[2]);
}
double pqref_cints_kinetic(double a1, const int *p1, const double *A, double a2, const int *p2,
                           const double *B) {
  return kinetic(a1, p1[0], p1[1], p1[2], A[0], A[1], A[2], a2, p2[0], p2[1], p2[2], B[0], B[1], B[2]);
}
double pqref_cints_nuclear_attraction(const double *A, double n1, const int *p1, double a1,
                                      const double *B, double n2, const int *p2, double a2,
                                      const double *C) {
  return nuclear_attraction(A[0], A[1], A[2], n1, p1[0], p1[1], p1[2], a1, B[0], B[1], B[2], n2,
                            p2[0], p2[1], p2[2], a2, C[0], C[1], C[2]);
}
