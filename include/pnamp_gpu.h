/*
 * pnamp_gpu.h -- C-ABI of the MI355X (gfx950) PNAMP projected-kernel library.
 *
 * The library replaces the per-pair kernel step of the reference PNAMP program
 * (jmyao17/SCCM): for one or many (q, q') pairs of HFB states it evaluates, at
 * every Euler (alpha, beta, gamma) x gauge (phi_p, phi_n) mesh point, the
 * rotated norm overlap (Pfaffian), the transition densities, the particle
 * numbers and the energy, sums the gauge points (particle-number projection)
 * and projects onto angular momentum.  Everything runs in FP64 on one GPU.
 *
 * Conventions (all arrays owned by the caller; the library copies in/out):
 *  - m-scheme basis of n states, protons first (t = -1) then neutrons (t = +1),
 *    orbits contiguous, m = j .. -j inside an orbit (tnljm in the reference,
 *    Codes/GCM/HFBA_Hamiltonian.f90:42-68).  n even, n <= PNAMP_NMAX.
 *  - matrices are n x n, column-major (Fortran order), real U(i,k), V(i,k).
 *  - interaction list: 1-based (a,b,c,d) with a<b, c<d, a+b<=c+d and the
 *    value <ab|V|cd> (H%ka..kd, H%ME2BM of the reference).
 *  - complex outputs are interleaved (re, im) doubles.
 *  - kern layout: [PNAMP_NOBS][jmax+1][2 jmax+1][2 jmax+1] complex with
 *    index (io, J, K1 + jmax, K2 + jmax); io = 0 norm, 1 Z, 2 N, 3 E_kin,
 *    4 E + E0, 5 <J^2> (the ZRES slots 1, 2, 3, 4, 7, 8 of Kernel4Triaxial),
 *    each the projected *norm-weighted* sum, like ZsumJKK in the reference.
 *  - zres layout: [n_euler][PNAMP_NOBS] complex, Euler points enumerated
 *    beta outer, alpha, gamma inner (ieuler_ang of the reference).
 *
 * Return codes: 0 ok; < 0 API misuse or device error (see pnamp_last_error);
 * > 0 a singular matrix was met (LAPACK-style INFO: U of a state not
 * invertible in pnamp_set_states, or a zero pivot of T in a kernel call).
 * The reference STOPs in those cases (Codes/PNAMP/PNAMP_Proj.f90:33-46,174-184).
 *
 * Threading: a context belongs to one host thread and one device; calls are
 * ordered on the context's stream and return after the results are copied.
 */
#ifndef PNAMP_GPU_H
#define PNAMP_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PNAMP_NOBS 6
#define PNAMP_NMAX 64

enum {
  PNAMP_OK = 0,
  PNAMP_E_ARG = -1,    /* bad argument */
  PNAMP_E_SIZE = -2,   /* size beyond what the kernels support */
  PNAMP_E_DEVICE = -3, /* HIP runtime error */
  PNAMP_E_STATE = -4   /* call out of order (e.g. no states set) */
};

enum { PNAMP_FOLD_PLAIN = 0, PNAMP_FOLD_LITERAL = 1 };

/* Evaluation forms of pnamp_kernel_pair / pnamp_kernels_batched.
 * PLAIN     the general path: every gauge point of every Euler point in full
 *           (n x n densities, 2n x 2n Pfaffian), any U, V.
 * FACTORED  FP64, for states without proton-neutron mixing: per Euler point
 *           nphi proton and nphi neutron Pfaffians / density sets (h = n/2),
 *           the proton-neutron energy for all nphi^2 gauge pairs from the pn
 *           part of the interaction list.  Same results as PLAIN to rounding.
 * MIXED_FP32, MIXED_BF16  the FACTORED form with the two-body contraction
 *           (V times densities) as dense (h^2 x h^2) x (h^2 x columns) GEMMs
 *           on MFMA over all gauge points of a batch: FP32, or split BF16
 *           (hi + lo, three products, FP32 accumulation).  Pfaffian, the
 *           inversion of T, traces and <J^2> stay FP64; the GEMM output is
 *           dotted with FP64 densities.  Precision form of the field
 *           contraction HFB_FIELD_COMPLEX (Codes/GCM/VAPHFB_HFB_Field.f90:4-123). */
enum { PNAMP_FORM_PLAIN = 0, PNAMP_FORM_FACTORED = 1, PNAMP_FORM_MIXED_FP32 = 2, PNAMP_FORM_MIXED_BF16 = 3 };

typedef struct pnamp_ctx pnamp_ctx;

/* Single-particle basis (tnljm%t, n, l, twoj, twom of the reference). */
typedef struct {
  int n;
  const int32_t *t, *nq, *l, *twoj, *twom;
} pnamp_basis;

/* Hamiltonian (H%ka..kd, H%ME2BM, H%zME1BM, H%E0 of the reference). */
typedef struct {
  int64_t nme;
  const int32_t *ka, *kb, *kc, *kd;
  const double *me2b;
  const double *t1b; /* n x n column-major, real */
  double e0;
} pnamp_interaction;

/* Projection mesh (AMP%alpha/beta/gamma, AMP%wwalp/wwbet/wwgam, PNP%MPhi,
 * Input%Cscale, Input%iscale of the reference).  fold selects the
 * J-projection: PNAMP_FOLD_PLAIN  N^J = wJ sum w D^{J*}_{K1K2} ZRES, all J;
 *               PNAMP_FOLD_LITERAL the four-term mirror fold of
 *               Codes/PNAMP/PNAMP_Kernel3D.f90:415-526 (even J). */
typedef struct {
  int nalp, nbet, ngam;
  const double *alp, *walp, *bet, *wbet, *gam, *wgam;
  int nphi, jmax, fold;
  double cscale;
  int iscale;
} pnamp_mesh;

/* Context: uploads basis, interaction and mesh, builds the Wigner tables on
 * the device.  Replaces the setup of PNAMP_Main (Codes/PNAMP/PNAMP_Main.f90:24-50). */
int pnamp_create(const pnamp_basis *basis, const pnamp_interaction *inter, const pnamp_mesh *mesh,
                 int device, pnamp_ctx **out);

/* HFB states: U, V as [nstates][n*n] column-major.  Computes per state
 * Z = V U^{-1} and the vacuum overlap det(1 + Z^T Z)^{-1/4} on the device
 * (zbareover, Codes/PNAMP/PNAMP_Proj.f90:6-79); replaces the .wf reads of
 * Kernel4Triaxial (PNAMP_Kernel3D.f90:83-122). */
int pnamp_set_states(pnamp_ctx *ctx, int nstates, const double *U, const double *V);

/* Selects the evaluation form (default PNAMP_FORM_PLAIN).  No counterpart in
 * the reference, which has one form; callers choosing FACTORED or a MIXED
 * form must set states for which pnamp_states_separable returns 1, and an
 * interaction list whose entries conserve the charge, have a != b and c != d
 * (any order: each pair is brought to canonical order with the
 * antisymmetry sign) and, for proton-neutron entries, conserve M; otherwise
 * the kernel calls of those forms return PNAMP_E_ARG.  The PLAIN form takes
 * any list. */
int pnamp_set_form(pnamp_ctx *ctx, int form);

/* 1 if no state set mixes protons and neutrons (Z = V U^{-1} has no
 * proton-neutron block, whatever the quasiparticle basis of U, V), 0 otherwise. */
int pnamp_states_separable(pnamp_ctx *ctx);

/* Vacuum overlaps <0|Phi_q> of the states set (Det of zbareover). */
int pnamp_state_overlaps(pnamp_ctx *ctx, double *det);

/* One pair: Kernel4Triaxial(lpr, Nphi1=q0, Nphi2=q1, NLEV)
 * (Codes/PNAMP/PNAMP_Kernel3D.f90:1-761, observables norm/Z/N/Ekin/E).
 * zres may be NULL; kern may be NULL. */
int pnamp_kernel_pair(pnamp_ctx *ctx, int q0, int q1, int zval, int nval, double *zres,
                      double *kern);

/* Many pairs in one launch (replaces the pair loop of PNAMP_Main.f90:79-121).
 * zres: [npairs][n_euler][PNAMP_NOBS] complex or NULL;
 * kern: [npairs][PNAMP_NOBS][jmax+1][2jmax+1][2jmax+1] complex or NULL. */
int pnamp_kernels_batched(pnamp_ctx *ctx, int npairs, const int32_t *q0, const int32_t *q1,
                          int zval, int nval, double *zres, double *kern);

/* pnamp_kernels_batched with the valence numbers given per pair (zval[i],
 * nval[i]): one call over the pairs of several GCM instances of one model
 * space (their states set together by one pnamp_set_states, pairs indexing
 * into them).  No counterpart in the reference, which runs one nucleus at a
 * time; it lets an uneven batch of instances share one launch sequence. */
int pnamp_kernels_batched_zn(pnamp_ctx *ctx, int npairs, const int32_t *q0, const int32_t *q1, const int32_t *zval,
                             const int32_t *nval, double *zres, double *kern);

/* pnamp_kernels_batched in the PLAIN form without any device allocation.
 * No counterpart in the reference: it exists for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * pnamp_kernels_batched with the context in PNAMP_FORM_PLAIN (the form set
 * by pnamp_set_form is ignored).  Every intermediate (pair lists, ZRES,
 * status flags, projected kernels) lives in one static device array of
 * PNAMP_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: half of
 * it for the pairs, processed in chunks that fit, half for the mesh-point
 * work buffer, processed in chunks of (pair, Euler point) units.  Returns
 * PNAMP_E_SIZE if a single pair needs more than its half (per pair:
 * (8 + 4 n_euler) + 16 n_euler PNAMP_NOBS + 16 PNAMP_NOBS (jmax+1)(2jmax+1)^2
 * + 4 bytes, about 0.45 MB for 2000 Euler points and J <= 8) or a single
 * unit more than the other half (112 n^2 + 16 nphi^2 PNAMP_NOBS + 4 nphi^2
 * bytes rounded up to 256-byte blocks, 0.18 MB at n = 40, nphi = 9).  All
 * calls on one GPU share the
 * array: calls that overlap in time on the same GPU (from several threads
 * or contexts) race and must be serialised by the caller. */
#define PNAMP_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int pnamp_kernels_batched_noalloc(pnamp_ctx *ctx, int npairs, const int32_t *q0, const int32_t *q1, int zval,
                                  int nval, double *zres, double *kern);

/* E2 observables (iE2 = 1 of the reference run card,
 * Codes/PNAMP/PNAMP_Kernel3D.f90:331-360,527-710): the one-body quadrupole
 * matrices Q_{2mu}, [5][n*n] real column-major for mu = -2..2 (cME1B%Q2_2t ..
 * Q22t built by SingleParticleME, Codes/PNAMP/PNAMP_ME1B.f90:138-216).
 * Needed before pnamp_kernels_e2. */
int pnamp_set_q2(pnamp_ctx *ctx, const double *q2);

/* Orbits per species (the one-body density's dimension). */
int pnamp_norb(pnamp_ctx *ctx);

/* pnamp_kernels_batched with the E2 and one-body-density observables, in
 * the plain, factored and mixed forms (plain fold only: PNAMP_E_ARG with the
 * literal fold).  Besides zres and kern (layouts of pnamp_kernels_batched):
 *  ze2   [npairs][n_euler][2][5] complex: per Euler point the norm-weighted
 *        gauge sums of tr_s(Q_{2mu} rho) for species s (protons, neutrons),
 *        the ZRES slots 9-13 of Kernel4Triaxial (ZQ2, PNAMP_ME1B.f90:52-75);
 *  qred  [npairs][2][5][jmax+1][2jmax+5][2jmax+1] complex: reduced matrix
 *        elements <J1 K1 q0 || Q2_s || J2 K2 q1> at (s, J1 - J2 + 2, J2,
 *        K1 + jmax + 2, K2 + jmax), J1 - J2 = -2..2 (Kernel%qp / qpred,
 *        PNAMP_Kernel3D.f90:590-710): the Wigner-Eckart reduced element,
 *        <Phi0|P^{J1}_{K1M1} Q_{2mu} P^{J2}_{M2K2}|Phi1> =
 *        (-1)^(J1-M1) (J1 2 J2; -M1 mu M2) <J1 K1 || Q2 || J2 K2>;
 *  rho1b [npairs][2][norb][norb] complex: the J = 0 projected one-body
 *        transition density ZRho1BJ0 (ZRho1B_Angles / ZRho1B_Integral,
 *        Codes/PNAMP/PNAMP_Rho1B.f90:2-62), zero for orbits of different l, j.
 * Any output may be NULL. */
int pnamp_kernels_e2(pnamp_ctx *ctx, int npairs, const int32_t *q0, const int32_t *q1, int zval, int nval,
                     double *zres, double *kern, double *ze2, double *qred, double *rho1b);

/* Norm and energy of every gauge point: samples [npairs][NE][nphi][nphi][2]
 * complex = (norm, E + E0) per (pair, Euler point, phi_p, phi_n), norm as
 * summed into ZRES slot 1 before the 1/nphi^2 (PNAMP_Kernel3D.f90:294-324).
 * Factored and mixed forms only.  No counterpart in the reference: training
 * data of the learned surrogate (include/pnamp_mlp.h). */
int pnamp_mesh_samples(pnamp_ctx *ctx, int npairs, const int32_t *q0, const int32_t *q1, int zval, int nval,
                       double *samples);

/* J-projection of caller-supplied ZRES [npairs][NE][PNAMP_NOBS] complex into
 * kern (layout of pnamp_kernels_batched): the projection half of
 * Kernel4Triaxial (PNAMP_Kernel3D.f90:415-526), for kernels assembled from
 * surrogate and exact mesh points. */
int pnamp_project_zres(pnamp_ctx *ctx, int npairs, const double *zres, double *kern);

/* Hill-Wheeler solve of one J block: HNDIAG (Codes/GCM/GCM_HNDiag.f90:2-103).
 * N, H real symmetric dim x dim column-major (H norm-weighted); keeps norm
 * eigenvalues with EN(i)/EN(1) >= eps, at most nos of them (nos <= 0: all).
 * Outputs: EN[dim] norm eigenvalues (descending), E[*nkept] energies
 * (ascending), F (dim x *nkept column-major, may be NULL) amplitudes.
 * Returns 1 (IFL of the reference) if no norm eigenvalue survives. */
int pnamp_hw_solve(pnamp_ctx *ctx, int dim, const double *N, const double *H, double eps, int nos,
                   double *EN, double *E, double *F, int *nkept);

/* Natural-state selection of PreSolve_HWG (Codes/GCM/GCM_HWG.f90:2-100) for
 * one J block: with M the number of norm eigenvalues kept by HNDIAG's cutoff
 * (EN(i)/EN(1) >= eps, no NOS limit), HNDIAG is solved for every number of
 * natural states NOS = 1..M (on the device, all NOS at once); per level
 * k < kmax its energy and the variance sum_i [sum_j (H_ij - E_k N_ij) f_j]^2
 * of its amplitudes f (mean_variance, :205-254).  Outputs: EN[dim] norm
 * eigenvalues (descending), *m = M, E and var [M][kmax] (row NOS - 1; zero
 * for k >= NOS), and per level the NOS of least variance nos_best[kmax]
 * (strictly below the reference's starting value 100, else 0) with its
 * energy e_best[kmax].  Returns 1 (IFL) if M = 0. */
int pnamp_hw_presolve(pnamp_ctx *ctx, int dim, const double *N, const double *H, double eps, int kmax, double *EN,
                      int *m, double *E, double *var, int *nos_best, double *e_best);

/* Device time of the last kernel call, from HIP events on the context
 * stream: the mesh-point kernel and the whole call (ms). */
int pnamp_last_timing(pnamp_ctx *ctx, float *ms_points, float *ms_total);

/* Autotuner of the FP64 factored form (no counterpart in the reference):
 * times every candidate launch configuration on the given pairs (real
 * kernels on the context's states, each candidate captured and replayed
 * once) and keeps the fastest for later calls of this context.  Candidates
 * (species size h >= 8; smaller spaces have one): waves per SIMD of the
 * Pfaffian kernel {3, 4, 5}, of the density kernel {default (3 while
 * 3 h <= 64, else 2), 4 (h <= 16 only)}, and the pair chunk divided by
 * {1, 2, 4}; a kernel variant that spills registers to scratch memory is no
 * candidate.  choice[3] receives (Pfaffian waves, density waves, chunk
 * divisor), cand[3 i .. 3 i + 2] the same for candidate i, ms[i] its kernel
 * time (at most PNAMP_TUNE_MAX candidates), *ncand their number; any may be
 * NULL.  The form must be PNAMP_FORM_FACTORED. */
#define PNAMP_TUNE_MAX 32
int pnamp_autotune(pnamp_ctx *ctx, int npairs, const int32_t *q0, const int32_t *q1, int zval, int nval,
                   int *choice, int *cand, float *ms, int *ncand);

/* Sets the factored form's launch configuration directly (a choice of
 * pnamp_autotune kept from an earlier run): Pfaffian waves 0 (default) / 3 /
 * 4 / 5, density waves 0 (default) / 2 / 3 / 4 (4 only for h <= 16; other
 * values give the default), chunk divisor 1 / 2 / 4. */
int pnamp_set_tune(pnamp_ctx *ctx, int pf_waves, int rest_waves, int chunk_div);

/* Allows (enable = 1) or forbids (0, the default) the factored and mixed
 * forms' Pfaffian by the Schur complement of the state block,
 *   Pf([[e^{2i phi} M0, -1], [1, -Z_s]]) = Pf(-Z_s) Pf(e^{2i phi} M0 - Z_s^-1),
 * an h x h Pfaffian per mesh point instead of ZPfaffianF's 2h x 2h one
 * (Codes/PNAMP/PNAMP_Proj.f90:142-392 on the matrix of Pfaf_Over).  A pair
 * takes it when both species blocks of its bra state have
 * cond_1(Z_s) <= 1e6 (the others keep the 2h x 2h Pfaffian, both kernels
 * then run over the grid); its Pfaffians carry up to about cond(Z_s) eps of
 * relative rounding (kernels within a few 1e-9 of the 2h x 2h route), so it
 * suits the factored form's GCM-energy target, not the 1e-10 drop-in.  With
 * it allowed, pnamp_autotune also times it (Pfaffian waves 0 in its output).
 * No counterpart in the reference. */
int pnamp_set_schur_pfaffian(pnamp_ctx *ctx, int enable);

/* With the Schur-complement Pfaffian allowed: 2 if every current state
 * takes it, 1 if some do, 0 if none; 0 while it is not allowed. */
int pnamp_schur_pfaffian_active(pnamp_ctx *ctx);

/* Launch-graph counters of the factored and mixed forms: graphs captured
 * and replays launched since pnamp_create, and the kernel nodes of the last
 * captured graph (the launches one call replays).  A call of the same shape
 * and form replays the graph; pnamp_set_states with no more states than the
 * context holds keeps it (the state buffers are rewritten in place). */
int pnamp_graph_info(pnamp_ctx *ctx, int *captures, int *replays, int *nodes);

/* Number of Euler points of the context's mesh. */
int pnamp_n_euler(pnamp_ctx *ctx);

const char *pnamp_last_error(pnamp_ctx *ctx);

void pnamp_destroy(pnamp_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* PNAMP_GPU_H */
