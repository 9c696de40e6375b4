/*
 * pq_abi.h -- C ABI of libpqhip.so, the MI355X two-electron-integral and
 * Fock-build engine for PyQuante.
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point
 * returns PQ_OK (0) or a negative PQ_E* code; pq_last_error() holds the
 * message of the last failure on the calling thread.  Device pointers
 * ("_dev") are HIP device addresses (e.g. a torch.cuda tensor's data_ptr());
 * host pointers ("_host") are ordinary memory.  A stream argument of NULL
 * means the null stream.  An engine handle is not safe to share between
 * threads without external locking; distinct handles are independent.
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef PQ_ABI_H
#define PQ_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PQ_OK = 0,
  PQ_E_ARG = -1,      /* bad argument (null pointer, size mismatch) */
  PQ_E_AM = -2,       /* angular momentum beyond what the entry point supports (shells above d, HGP leaves) */
  PQ_E_LTOTAL = -3,   /* L_total > 9: no Rys rule (the reference's Root6 stub, crys.c:1440-1443) */
  PQ_E_ALLOC = -4,    /* device allocation failed */
  PQ_E_HIP = -5,      /* HIP runtime error */
  PQ_E_STATE = -6,    /* call out of order (e.g. Fock build before pq_build_pairs) */
  PQ_E_NPRIM = -7     /* too many primitives for the scalar entry points */
};

enum { PQ_PREC_FP64 = 0, PQ_PREC_MIXED = 1 };

typedef struct pq_engine pq_engine;

/* Message of the last failure on this thread ("" if none). */
const char *pq_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int pq_version(void);
/* Number of HIP devices visible (0 when none). */
int pq_device_count(void);

/* ---------------- scalar / batched compatibility entry points --------------- */

/* cints.ijkl2intindex (Src/cints.c:413-425; PyQuante/Ints.py:16), widened to
 * 64 bits: the reference's C int overflows for nbf >= 304. Integer arithmetic
 * on the host; no device work. */
int pq_ijkl2intindex(int64_t i, int64_t j, int64_t k, int64_t l, uint64_t *out);

/* cints.Fgamma / chgp Fgamma (Src/cints.c:340-411): Boys function for n
 * (m, x) pairs, evaluated on the device.  compat_nr = 1 keeps the reference's
 * EPS = 3e-7 stopping rule; 0 converges to full double precision. */
int pq_boys_batch(int n, const double *m_host, const double *x_host, int compat_nr, double *out_host);

/* Rys roots u_i = t_i^2 and weights of order nr (1..5) at n arguments X,
 * evaluated by the device tables (Src/crys.c:99-1443 Roots; the reference
 * returns t^2/(1-t^2)).  u_host, w_host hold n*nr values. */
int pq_rys_roots_batch(int n, int nr, const double *X_host, double *u_host, double *w_host);

/* {crys,cints,chgp}.contr_coulomb (Src/crys.c:1556-1674, cints.c:572-690,
 * chgp.c:701-818) for n contracted quartets at once, on the device.
 * Quartet q uses functions 4q..4q+3: nprim[4q+f] primitives starting at
 * poff[4q+f] in exps/coefs/norms; origins[12q+3f+d], powers[12q+3f+d].
 * coulomb_repulsion (crys.c:1676-1711) is the case nprim = 1, coef = 1. */
int pq_contr_coulomb_batch(int n, const int *nprim_host, const int *poff_host, int nprim_total,
                           const double *exps_host, const double *coefs_host, const double *norms_host,
                           const double *origins_host, const int *powers_host, double *out_host);

/* cints.contr_coulomb / coulomb_repulsion (Src/cints.c:80-153 contr_coulomb,
 * coulomb_repulsion, B_array; wrappers cints.c:572-803): the
 * Taketa-Huzinaga-O-ohata path, same batch layout as pq_contr_coulomb_batch.
 * Each primitive quartet's Rys integrand is expanded in u = t^2 on the device
 * and its coefficients are contracted with Boys values F_0..F_L (the sum the
 * reference's THO expansion forms).  compat_nr = 1 takes those values from
 * the reference's EPS = 3e-7 Fgamma (cints.c:340-411) and so reproduces its
 * numbers; 0 uses the full-precision Boys function (then the value equals
 * pq_contr_coulomb_batch's).  L_total <= 9 (PQ_E_LTOTAL beyond). */
int pq_tho_contr_coulomb_batch(int n, const int *nprim_host, const int *poff_host, int nprim_total,
                               const double *exps_host, const double *coefs_host, const double *norms_host,
                               const double *origins_host, const int *powers_host, int compat_nr,
                               double *out_host);

/* Head-Gordon-Pople contracted ERIs, replacing chgp.contr_coulomb /
 * coulomb_repulsion (Src/chgp.c:78-169 contr_hrr/contr_vrr, 172-242 hrr,
 * 244-419 vrr) and, with mext, chgp.vrr (chgp.c:857-890).  Same batch layout
 * as pq_contr_coulomb_batch.  mext (NULL or n ints): the auxiliary order m of
 * a primitive leaf [a0|c0]^(m) (powers of b and d must be zero).
 * compat_nr = 1 uses the reference's Boys function (EPS = 3e-7, started at
 * each leaf's own total order) and reproduces chgp's values; 0 uses the
 * full-precision Boys function.  Limits (PQ_E_AM beyond): per side
 * (a+b, c+d) prod_i (p_i + 1) <= 12 and total order L + mext <= 8. */
int pq_hgp_contr_coulomb_batch(int n, const int *nprim, const int *poff, int nprim_total, const double *exps,
                               const double *coefs, const double *norms, const double *origins, const int *powers,
                               const int *mext, int compat_nr, double *out);

/* ---------------- engine: basis, geometry, pair data ---------------- */

/* Creates an engine for nshell shells (PyQuante/Basis/basis.py:11-50 order).
 * Shell s: angular momentum shell_l[s] (0..2), shell_nprim[s] primitives
 * whose exponents and raw contraction coefficients (CGBF.py:65-75; the
 * contraction is not renormalised) start at shell_poff[s] in exps/coefs,
 * centred on atom shell_atom[s].  Functions are numbered atom -> shell ->
 * Cartesian component in PyQuante order (Ints.py:23-29).  atom_xyz (bohr,
 * natom*3) and atom_z give the nuclei (for V and forces). */
int pq_engine_create(int nshell, const int *shell_l, const int *shell_nprim, const int *shell_poff,
                     const int *shell_atom, const double *exps, const double *coefs, int natom,
                     const double *atom_xyz, const double *atom_z, pq_engine **out);

/* A batch of independent molecules in one engine (one launch per class over
 * all molecules' quartets): as pq_engine_create, plus atom_mol[a] = molecule
 * of atom a (molecules numbered 0, 1, ... with contiguous atoms; shells
 * ordered by molecule).  Pairs and quartets never mix molecules.  Matrices
 * (D, J, K, S, T, V, W) are packed per molecule: molecule m's n_m x n_m
 * block, row-major, starts at element mol_moff[m] (pq_engine_layout).
 * pq_get2ints, pq_get2ints_hgp and pq_shell_quartet_blocks need a
 * single-molecule engine.  No reference counterpart (the reference handles
 * one molecule per call). */
int pq_engine_create_batch(int nshell, const int *shell_l, const int *shell_nprim, const int *shell_poff,
                           const int *shell_atom, const double *exps, const double *coefs, int natom,
                           const double *atom_xyz, const double *atom_z, const int *atom_mol, pq_engine **out);

/* Packed matrix layout: number of molecules, elements of one packed matrix,
 * and per molecule (arrays of nmol entries, each optional) its function
 * count, block offset, first atom and atom count. */
int pq_engine_layout(const pq_engine *eng, int *nmol, int64_t *msize, int *mol_nbf, int64_t *mol_moff,
                     int *mol_first_atom, int *mol_natom);
int pq_engine_destroy(pq_engine *eng);
/* Number of basis functions. */
int pq_engine_nbf(const pq_engine *eng, int *nbf);
/* Work-list generation: changes whenever the engine replaces or frees device
 * arrays that a Fock build reads (pair data, tiles, bra runs, scratch):
 * pq_engine_set_geometry, pq_build_pairs, pq_engine_set_group_cap(s),
 * pq_engine_autotune.  A HIP graph captured from pq_fock_jk holds those
 * addresses and must not be replayed once the generation has changed.  No
 * reference counterpart (graph-capture bookkeeping). */
int pq_engine_generation(const pq_engine *eng, uint64_t *gen);
/* Moves the nuclei (and the shells on them); invalidates pair data. */
int pq_engine_set_geometry(pq_engine *eng, const double *atom_xyz_host);

/* Builds shell-pair data and the class-sorted quartet work lists.  Quartets
 * whose Schwarz bound Q_ab*Q_cd is below schwarz_cut are skipped (the
 * reference has no screening: 0 keeps every quartet); in PQ_PREC_MIXED
 * builds, quartets with bound below fp32_cut run in FP32. */
int pq_build_pairs(pq_engine *eng, double schwarz_cut, double fp32_cut);

/* Primitive screening of the J/K and gradient work (takes effect at the
 * next pq_build_pairs): with a Schwarz cut > 0, primitive pairs whose factor
 * max sqrt|(ab|ab)| times the largest such factor is below rel * cut are
 * skipped, so every skipped primitive quartet is bounded by rel * cut.  The
 * default rel = 1e-4 keeps that bound below the shell quartets the cut
 * itself drops; rel = 0 keeps every primitive.  Stored integrals
 * (pq_get2ints) and one-electron integrals always use every primitive.  The
 * reference contracts every primitive (Ints.py:146-167 over
 * cints.contr_coulomb); this is the engine's own screening, like the cut. */
int pq_engine_set_prim_screen(pq_engine *eng, double rel);

/* Number of engine-owned streams (0..4, default 4) the class kernels of one
 * Fock build or gradient are spread over; 0 launches everything on the caller's stream
 * (one kernel at a time, e.g. for per-kernel profiling).  A Fock build
 * outside a graph capture uses at most 3 of them (with the caller's stream,
 * four streams for the process's four hardware queues); a captured one uses
 * all.  No reference counterpart (a launch knob). */
int pq_engine_set_streams(pq_engine *eng, int nstreams);

/* on = 1: the FP64 bra runs of the light classes (ss|ss), (ps|ss), (ps|ps),
 * (pp|ss) and (ds|ss) go into one launch (a segment table maps each block
 * to its class; heaviest class first) instead of one launch per class; 0
 * (default) restores the per-class launches.  The fused kernel carries the
 * registers and LDS of its largest member.  Single-density FP64 builds
 * only.  No reference counterpart (a launch knob). */
int pq_engine_set_fused(pq_engine *eng, int on);

/* on = 1: deterministic J/K.  pq_fock_jk then accumulates every contribution
 * as exact 64-bit integer limbs (resolution 2^-64, |contribution| < 2^62)
 * instead of FP64 atomics, and converts once at the end, so for one launch
 * configuration J and K are bit-identical from call to call whatever the
 * order of execution (streams, graph replay), as the reference's sequential
 * getJ/getK loops are (Ints.py:195-223).  Bra-run caps and the LDS limit
 * select the kernels, whose integrals may differ in the last bits.  Covers
 * one- and two-density builds (pq_fock_jk_multi) and the FP32 quartets of
 * mixed-precision builds; costs extra atomics and three times the LDS per
 * accumulator.  0 (default) restores the FP64-atomic build. */
int pq_engine_set_deterministic(pq_engine *eng, int on);

/* Dynamic LDS the J/K launches may use, bytes: 0 restores the device's
 * per-workgroup limit (queried at creation, 160 KB on gfx950), a smaller
 * value lowers it (clamped to the device's); device_bytes (optional)
 * receives the device limit.  Past the limit a class keeps the bra rows of
 * K in global memory (scattered atomics) instead of LDS.  No reference
 * counterpart (a launch knob). */
int pq_engine_set_lds_limit(pq_engine *eng, int bytes, int *device_bytes);

/* Bra runs: for the J/K kernels of classes with at most 108 Cartesian
 * components, one 256-thread block walks up to `cap` consecutive tiles that
 * share a bra shell pair and flushes its LDS copy of the bra K rows once.
 * cap = 0 disables runs, -1 restores the default (16 tiles when nbf >= 300,
 * else none).  A run length (1..1024) may be or-ed with PQ_CAP_STAGE_D: the
 * block then also copies the bra rows of the density to LDS (when they fit)
 * and reads them there.  Applies to every class and overrides tuned values.
 * No reference counterpart (a launch knob). */
enum { PQ_CAP_STAGE_D = 4096 };
int pq_engine_set_group_cap(pq_engine *eng, int cap);

/* Per-class bra-run caps: caps[2*c + prec] for class index c (the order of
 * pq_engine_class_stats) and prec 0 = FP64 tiles, 1 = FP32 tiles; -1 keeps
 * the default.  For replaying a table written by pq_engine_autotune.  No
 * reference counterpart (a launch knob). */
int pq_engine_set_group_caps(pq_engine *eng, const int *caps);

/* Autotuner: for each class and precision of the current work lists, times
 * the J/K kernel on D_dev (`reps` timed launches after one warm-up) for run
 * caps {0, 4, 8, 16, 32}, each run length with and without PQ_CAP_STAGE_D,
 * on `stream` and keeps the fastest.  The choice
 * persists across pq_build_pairs calls of this engine.  caps_out (optional,
 * 42 ints) receives the chosen cap per (class, precision); classes with more
 * than 108 components report their unused cap.  Overwrites the engine's J/K
 * scratch only.  No reference counterpart (a launch knob). */
int pq_engine_autotune(pq_engine *eng, const double *D_dev, int reps, void *stream, int *caps_out);

/* Statistics of the current work lists: counts[0..5] = FP64 shell quartets,
 * FP32 shell quartets, FP64 primitive quartets, FP32 primitive quartets,
 * FP64 and FP32 unique function quartets (i>=j, k>=l, ij>=kl: the unit the
 * reference's get2ints enumerates, Ints.py:146-167); flops[0..1] =
 * algorithmic Rys flops (FP64 part, FP32 part) of one build. */
int pq_engine_stats(const pq_engine *eng, uint64_t counts[6], double flops[2]);

/* Per-class statistics for class index c in 0..20 (pair classes X >= Y in
 * the order ss, ps, pp, ds, dp, dd; c enumerates (X,Y) row by row):
 * l[0..3] = (la, lb, lc, ld), counts[0..5] as in pq_engine_stats, counts[6]
 * and counts[7] the FP64 and FP32 tile counts; flops as in pq_engine_stats.
 * No reference counterpart (diagnostics for the roofline report). */
int pq_engine_class_stats(const pq_engine *eng, int c, int l[4], uint64_t counts[8], double flops[2]);

/* ---------------- integrals and Fock builds ---------------- */

/* Ints.get2ints (PyQuante/Ints.py:146-167): writes every unique ERI (ij|kl)
 * at index ijkl2intindex(i,j,k,l) of out_dev (len = npair(npair+1)/2,
 * npair = nbf(nbf+1)/2), FP64, unscreened. */
int pq_get2ints(pq_engine *eng, double *out_dev, uint64_t len, void *stream);

/* pq_get2ints by Head-Gordon-Pople recurrences: the reference's get2ints with
 * CGBF.coulomb bound to chgp (CGBF.py:23-25, Ints.py:146-167; recurrences
 * chgp.c:78-419).  compat_nr = 0: one shell quartet per lane, the vertical
 * recurrence per primitive quartet and the horizontal one once per
 * contracted block, full-precision Boys values (from the Rys rule of the
 * class), on the given stream.  compat_nr = 1: one function quartet per lane
 * in chgp's own order, with chgp's EPS = 3e-7 Boys function started at each
 * leaf's order, which reproduces chgp's values to 1e-12. */
int pq_get2ints_hgp(pq_engine *eng, double *out_dev, uint64_t len, int compat_nr, void *stream);

/* Shell-quartet blocks in the libint layout, the interface of the
 * reference's libint branch (Src/lib/clibint.h:37-52 shell_compute_eri,
 * Src/PyQuante/clibint.pyx:27-60, Ints.py:125-143): for each of nq quartets
 * of shell indices (shells_host[4q..4q+3], engine shell order) a row-major
 * block ((i*nb + j)*nc + k)*nd + l over the shells' functions in canonical
 * Cartesian order, blocks concatenated in request order into out_dev
 * (out_len doubles available).  Only the requested quartets are computed
 * (store-mode class kernels over tiles built from the request, each lane
 * writing its block in place), so no array of all integrals is formed and
 * the request arrays are freed before returning.  Synchronises the stream. */
int pq_shell_quartet_blocks(pq_engine *eng, int nq, const int *shells_host, double *out_dev, uint64_t out_len,
                            void *stream);

/* Ints.getJ / getK / get2JmK (PyQuante/Ints.py:195-238) in direct mode:
 * J_ij = sum_kl (ij|kl) D_kl and K_ij = sum_kl (ik|jl) D_kl from integrals
 * computed on the fly over the screened work lists.  D_dev is a symmetric
 * nbf x nbf row-major FP64 matrix; J_dev or K_dev may be NULL.
 * precision: PQ_PREC_FP64 or PQ_PREC_MIXED (needs pq_build_pairs with
 * fp32_cut > 0). */
int pq_fock_jk(pq_engine *eng, const double *D_dev, double *J_dev, double *K_dev, int precision,
               void *stream);

/* J and K of ndens (1 or 2) densities from one pass over the integrals:
 * D_dev, J_dev, K_dev are host arrays of ndens device pointers (entries of
 * J_dev / K_dev may be NULL).  Replaces the four getJ/getK calls of an
 * unrestricted SCF iteration (hartree_fock.py:233-236, PyQuante2.py:320-343)
 * with one integral evaluation.  ndens = 2 is FP64 only. */
int pq_fock_jk_multi(pq_engine *eng, int ndens, const double *const *D_dev, double *const *J_dev,
                     double *const *K_dev, int precision, void *stream);

/* pq_fock_jk with HIP events around every class launch, recorded on the
 * stream the kernel is launched on (the engine streams of
 * pq_engine_set_streams, so concurrent kernels are timed as they run; 0
 * streams times each kernel alone); ms[2c] / ms[2c+1] receive the FP64 / FP32
 * kernel time of class c (21 classes, 42 floats), ms[42] the whole build.
 * Synchronises the stream.  No reference counterpart (measurement). */
int pq_fock_jk_timed(pq_engine *eng, const double *D_dev, double *J_dev, double *K_dev, int precision,
                     void *stream, float *ms);

/* getJ / getK over a stored flat integral array (the reference's Ints object,
 * Ints.py:195-223), digesting every stored unique integral on the device. */
int pq_digest_stored(int nbf, const double *ints_dev, uint64_t len, const double *D_dev, double *J_dev,
                     double *K_dev, void *stream);

/* pq_digest_stored without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  The J and K accumulators
 * live in a library-owned static device array of 2 * 512^2 doubles (4 MiB,
 * held by each GPU the library runs on until the process ends), so nbf <= 512
 * (PQ_E_ARG beyond).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (different streams, or threads) race.  Asynchronous on
 * `stream`; J_dev or K_dev may be NULL. */
int pq_digest_stored_noalloc(int nbf, const double *ints_dev, uint64_t len, const double *D_dev, double *J_dev,
                             double *K_dev, void *stream);

/* Rows mu0 .. mu0+nmu-1 of the full 4-index tensor (mu nu|lam sig) from the
 * flat unique array (len = npair(npair+1)/2), row-major into out_dev
 * (nmu*nbf^3 doubles): the unpacking step of the reference's AO->MO transform
 * (PyQuante/MP.py:38-46), done once per row block instead of per element. */
int pq_unpack_eri(int nbf, const double *ints_dev, uint64_t len, int mu0, int nmu, double *out_dev, void *stream);

/* Ints.get1ints (PyQuante/Ints.py:73-87): S, T and V = sum_A Z_A V_A
 * (h = T + V), nbf x nbf row-major FP64, on the device.  compat_nr = 1
 * evaluates V with the reference's EPS = 3e-7 Boys function, as its
 * cints.nuclear_attraction does (Src/cints.c:241-275, 340-411; reached through
 * CGBF.py:108-114 and PGBF.py:90-96), and reproduces its values; 0 uses
 * full-precision Boys values (the V whose derivative pq_forces contracts). */
int pq_get1ints(pq_engine *eng, double *S_dev, double *T_dev, double *V_dev, int compat_nr, void *stream);

/* RHF nuclear gradient dE/dR (force.py:31-70 rhf_force; der2JmK
 * force.py:318-357; AnalyticDerivatives.py:27-675), natom x 3 row-major into
 * grad_dev (hartree/bohr): 2 tr(D dH) + sum (ij|kl)^x [2 D_ij D_kl - D_ik D_jl]
 * - 2 tr(W dS) + dV_nn, with D = C_occ C_occ^T and the energy-weighted
 * W = C_occ diag(e_occ) C_occ^T (Wavefunction.py:143-160).  Derivative
 * integrals are contracted on the fly over the current (FP64) work lists. */
int pq_forces(pq_engine *eng, const double *D_dev, const double *W_dev, double *grad_dev, void *stream);

/* UHF nuclear gradient (force.py:72-89 uhf_force: der_oneeE and der_dmat of
 * each spin plus der_twoeE_uhf and der_enuke), natom x 3 into grad_dev:
 * tr(P dH) + 1/2 sum (ij|kl)^x [P_ij P_kl - Pa_ik Pa_jl - Pb_ik Pb_jl]
 * - tr((Qa + Qb) dS) + dV_nn, with the spin densities Pa = Ca_occ Ca_occ^T,
 * Pb likewise, P = Pa + Pb and Qs = Cs_occ diag(es_occ) Cs_occ^T
 * (Wavefunction.py:143-160).  The caller passes Da_dev = Pa, Db_dev = Pb and,
 * for the one-electron terms, D_dev = P / 2 and W_dev = (Qa + Qb) / 2 (the
 * closed-shell convention of pq_forces, whose gradient this is for Pa = Pb). */
int pq_forces_uhf(pq_engine *eng, const double *Da_dev, const double *Db_dev, const double *D_dev,
                  const double *W_dev, double *grad_dev, void *stream);

/* pq_forces with HIP events around every kernel launch, recorded on the
 * stream it is launched on: ms[2c] / ms[2c+1] the two-electron derivative
 * kernel of class c over its FP64 / FP32-bucket tiles (21 classes, the order
 * of pq_engine_class_stats; all run in FP64), ms[42+p] the one-electron
 * derivative kernel of pair class p (ss, ps, pp, ds, dp, dd), ms[48] the
 * whole gradient (49 floats).  Synchronises the stream.  No reference
 * counterpart (measurement). */
int pq_forces_timed(pq_engine *eng, const double *D_dev, const double *W_dev, double *grad_dev, void *stream,
                    float *ms);

#ifdef __cplusplus
}
#endif
#endif /* PQ_ABI_H */
