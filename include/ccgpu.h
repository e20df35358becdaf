/*
 * ccgpu — MI355X (gfx950) CCD/(T) amplitude engine for nuclear matter on the
 * periodic momentum lattice.  C ABI: plain ints, doubles and pointers; complex
 * numbers are interleaved (re, im) doubles, layout-identical to Fortran
 * complex(c_double_complex) and hipDoubleComplex.
 *
 * Every function returns 0 on success and a nonzero status otherwise; the
 * message of the last failure is available from ccgpu_last_error().  Handles
 * are small integers; one host thread drives a handle at a time.
 *
 * Reference interfaces replaced (paths relative to deltafull-CCMkspace/ of the
 * reference):
 *   ccgpu_plan / ccgpu_create         allocate_sp_data + setup_sp_data
 *                                     (ccm_main.f90:878-1197) and
 *                                     setup_channel_structures /
 *                                     setup_ph_channel_structures
 *                                     (ccm_iter.f90:175-811, 1816-2212)
 *   ccgpu_create_from_orbitals        the same channel setup, from the host's
 *                                     own all_orbit table (ccm_modules.f90:546-614)
 *   ccgpu_generate_interaction        the V-block loops of setup_channel_structures
 *                                     (ccm_iter.f90:819-1311) plus
 *                                     normal_ordered_hamiltonian Fock/E0
 *                                     (ccm_main.f90:1586-1684), with the
 *                                     elements of chiral_NN_with_delta or
 *                                     vmom_minnesota
 *                                     (chiral_module_andreas_with_delta.f90:
 *                                     6304-6668, 8980-9111)
 *   ccgpu_set_3nf                     tnf_switch / tnf_approx and the 3NF
 *                                     regulator line of ccm_in
 *                                     (ccm_main.f90:893-900) with cE, cD;
 *                                     the normal ordering of
 *                                     normal_ordered_hamiltonian and
 *                                     setup_3nf_2bNO (ccm_main.f90:1687-2568)
 *                                     runs inside ccgpu_generate_interaction
 *   ccgpu_vnn_elements                the antisymmetrized element calls
 *                                     chiral_NN_with_delta(p,q,r,s) -
 *                                     chiral_NN_with_delta(p,q,s,r)
 *                                     (ccm_iter.f90:854, 2564)
 *   ccgpu_set_chiral                  delta_chiral_order / lambda_delta of
 *                                     ccm_in (ccm_main.f90:31) and
 *                                     init_chp_constants
 *                                     (chiral_module_andreas_with_delta.f90:
 *                                     122-808)
 *   ccgpu_set_block / ccgpu_set_fock  hand-over of the host's vnn_* blocks and
 *                                     fock_mtx/e0 (ccm_modules.f90:1167-1171, 528)
 *   ccgpu_mbpt2_start                 setup_t_amplitudes (ccm_iter.f90:2329-2405)
 *   ccgpu_residual                    t2_intermediate(0) (ccm_t2_eqn.f90:4-1304)
 *   ccgpu_update                      diis_t1t2 (ccm_diis.f90:74-249)
 *   ccgpu_energy                      ccd_energy_save (ccm_energy.f90:4-71)
 *   ccgpu_solve                       ccd_iter loop (ccm_iter.f90:1-170)
 *   ccgpu_triples                     pert_triples2 (ccm_triples.f90:528-735)
 *   ccgpu_triples_parts               its et_4 and et_5 (:706-732)
 *   ccgpu_setup_t3                    setup_t3 + the first t3_energy of the
 *                                     CCDT1 branch (ccm_iter.f90:26-40;
 *                                     ccm_t3full_channels.f90:291-667)
 *   ccgpu_t3_intermediate             t3_intermediate (ccm_t3_eqn.f90:1-192)
 *   ccgpu_t3_energy                   t3_energy (ccm_t3full_channels.f90:570-667)
 *   ccgpu_get_t2 / ccgpu_set_t2       reads/writes of t2_ccm(ch)%val
 *   ccgpu_batch_solve                 no reference counterpart: the reference
 *                                     runs one twist per process
 *                                     (ccm_main.f90:163-279)
 *   ccgpu_mlp_*                       no reference counterpart: a learned start
 *                                     that replaces setup_t_amplitudes
 *                                     (ccm_iter.f90:2329-2405) as the first
 *                                     guess; the iteration that follows is
 *                                     unchanged, so the converged energy does
 *                                     not depend on the network
 */
#ifndef CCGPU_H
#define CCGPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* matter */
#define CCGPU_PNM 0
#define CCGPU_SNM 1

/* interaction generators for ccgpu_generate_interaction */
#define CCGPU_INT_MINNESOTA 0
#define CCGPU_INT_CHIRAL 1     /* Delta-full chiral NN, order/cutoff of ccgpu_set_chiral */

/* block kinds for ccgpu_set_block (reference block names) */
#define CCGPU_BLOCK_HHHH 1
#define CCGPU_BLOCK_HHHP 2
#define CCGPU_BLOCK_HHPP 3
#define CCGPU_BLOCK_HPHP 4
#define CCGPU_BLOCK_HPPP 5
#define CCGPU_BLOCK_PPPP 6

/* precision modes of the pp/hh ladders and the ring contractions:
 * FP64  everything FP64;
 * FP32  every contraction on FP32 MFMA with FP32 accumulation (pp ladder on a
 *       complex64 copy of V_pppp);
 * V32   the pp ladder reads the complex64 copy of V_pppp (half the bytes) and
 *       computes in FP64; the rounding of V is a fixed perturbation of the
 *       problem, so the iteration and DIIS behave as in FP64
 * BF16X3 every contraction splits each operand into two bfloat16 numbers
 *       (hi = bf16(x), lo = bf16(x - hi)) and sums hi*hi + hi*lo + lo*hi on
 *       bf16 MFMA with FP32 accumulation (each real product within 3 * 2^-16
 *       of exact); the pp ladder reads the complex64 copy of V_pppp and the
 *       ring keeps its Pandya blocks in FP32 */
#define CCGPU_PREC_FP64 0
#define CCGPU_PREC_FP32 1
#define CCGPU_PREC_V32 2
#define CCGPU_PREC_BF16X3 3

typedef struct {
  int A;            /* number of nucleons (holes) */
  int matter;       /* CCGPU_PNM or CCGPU_SNM */
  double rho;       /* density in fm^-3 */
  int nmax;         /* |n_i| <= nmax per axis */
  double twist[3];  /* twist angles theta in [0, pi] per axis */
} ccgpu_lattice_params;

typedef struct {
  int nsp, nholes, nparticles;
  int nchannels;          /* (K, Tz) channels with hh and pp pairs */
  int nph_channels;       /* Pandya (particle-hole) channels */
  int max_npp, max_nhh;   /* largest unique pair counts per channel */
  long long t2_unique;    /* stored amplitudes (a<b, i<j) */
  long long t2_ordered;   /* reference-shaped count (all ordered pairs) */
  long long vpppp_unique; /* stored pppp elements (unique x unique) */
  long long ring_elems;   /* sum over ph channels of Nph * Nhp */
  double box, volume, kf;
  double pp_ladder_flops; /* 8 sum Npp^2 Nhh: real flops of one pp-ladder launch */
  double hh_ladder_flops; /* 8 sum (Npp Nhh^2 + Nhh^2 Npp): I9 and hh ladder */
  double ring_flops;      /* 8 sum (Nhp^2 Nph + Nph Nhp^2): both ring GEMMs */
} ccgpu_dims;

typedef struct {
  int subspace;      /* DIIS history length (reference: 6) */
  int diis_step;     /* DIIS extrapolation cadence (reference: 6) */
  double tol;        /* stop when |dE| <= tol (MeV); reference 1e-6 */
  int maxit;         /* iteration cap (reference: 100) */
  int precision;     /* CCGPU_PREC_* */
  int refine_every;  /* mixed precision: FP64 residual every k iterations */
  int use_graph;     /* capture one iteration in a hipGraph and replay it */
  int verbose;       /* print per-iteration lines like the reference */
  double fp32_switch;/* mixed precision: the FP32 phase ends at the first DIIS
                        extrapolation that follows an iteration with max |dE| <=
                        fp32_switch (MeV; <= 0: never there), or once every
                        FP32 |dE| <= tol; FP64 afterwards with the DIIS history
                        kept, and at least one FP64 iteration before a stop */
  int ccdt1;         /* 1: CCDT-1 energies (ccgpu_setup_t3 / t3_intermediate /
                        t3_energy inside the loop, as the CCDT1 branch of
                        ccd_iter); needs ccgpu_set_3nf and the chiral
                        interaction.  0: CCD */
  int accel;         /* CCGPU_ACCEL_REFERENCE: the update and DIIS of
                        diis_t1t2 (0.75/0.25 mixing every iteration, DIIS every
                        diis_step iterations over subspace iterates);
                        CCGPU_ACCEL_ANDERSON: Anderson/Pulay update every
                        iteration on the residual vectors f = R/D - T, with at
                        most subspace held vectors (diis_step unused); by
                        default it stops after two consecutive iterations with
                        |dE| <= tol (confirm) */
  double mix;        /* CCGPU_ACCEL_ANDERSON: T_next = sum_j c_j (T_j + mix f_j);
                        (0, 1], 0 selects 1 */
  int tail;          /* contractions after the low-precision phase:
                        CCGPU_PREC_FP64 (0, default) or CCGPU_PREC_V32 (FP64
                        arithmetic on the complex64 copy of V_pppp, a fixed
                        rounding of the Hamiltonian by at most 2^-24 relative);
                        the stopping test runs in the tail precision */
  int confirm;       /* consecutive iterations with |dE| <= tol that end a
                        solve, 1-4; 0 the default: 1 with
                        CCGPU_ACCEL_REFERENCE (the reference's test,
                        ccm_iter.f90:71-74), 2 with CCGPU_ACCEL_ANDERSON.  With
                        a low-precision phase only FP64 (or V32 tail)
                        iterations count */
} ccgpu_solve_opts;

/* convergence accelerators (ccgpu_solve_opts.accel) */
#define CCGPU_ACCEL_REFERENCE 0
#define CCGPU_ACCEL_ANDERSON 1

/* per-solve status flags (ccgpu_solve_result.status) */
#define CCGPU_STATUS_DIIS_FALLBACK 1   /* a singular DIIS/Anderson system: that
                                          solve kept its plain update for the
                                          step (the reference prints ZGESV's
                                          info and carries on,
                                          ccm_diis.f90:173-176) */
#define CCGPU_STATUS_NONFINITE 2       /* non-finite energy: the solve was
                                          stopped, the others of a batch go on */
#define CCGPU_STATUS_MAXIT 4           /* iteration cap reached before |dE| <= tol
                                          (ccm_iter.f90:71-74 stops silently) */

typedef struct {
  int iterations;
  int converged;
  double e0[2];          /* vacuum (HF) energy */
  double embpt2[2];      /* E0 + MBPT2 correlation energy */
  double eccd[2];        /* E0 + CCD correlation energy */
  double ecorr_last[2];  /* last correlation energy */
  double seconds;        /* wall time of the iteration loop */
  int fp32_iterations;   /* iterations whose contractions ran in FP32 or V32 */
  int status;            /* CCGPU_STATUS_* flags of this solve */
} ccgpu_solve_result;

/* Host-only planning of a lattice: no GPU work.  Fills dims. */
int ccgpu_plan(const ccgpu_lattice_params *p, ccgpu_dims *dims);

/* Build basis + channels for a lattice and allocate device storage. */
int ccgpu_create(const ccgpu_lattice_params *p, int *handle);

/* Same, from an explicit orbital table (the first nholes orbitals are holes).
 * k in fm^-1, e in MeV; sz and tz are +-1. */
int ccgpu_create_from_orbitals(int nsp, int nholes, const int *nx, const int *ny,
                               const int *nz, const int *sz, const int *tz,
                               const double *kx, const double *ky, const double *kz,
                               const double *e, double box, double volume, int *handle);

int ccgpu_destroy(int handle);
int ccgpu_get_dims(int handle, ccgpu_dims *dims);

/* Orbital table of a handle (arrays of length nsp). */
int ccgpu_get_orbitals(int handle, int *nx, int *ny, int *nz, int *sz, int *tz, double *e);

/* Fill every V block, the Fock matrix and E0 on the GPU. */
int ccgpu_generate_interaction(int handle, int kind);

/* Chiral order (0 LO, 2 NLO, 3 NNLO) and regulator cutoff (394, 450 or 500
 * MeV) of CCGPU_INT_CHIRAL, as the first line of ccm_in; ci = null keeps the
 * pi-N LECs of that order, else 4 values c1..c4 in GeV^-1 replace them.  The
 * default of a new handle is NNLO at 450 MeV. */
int ccgpu_set_chiral(int handle, int order, double cutoff, const double *ci);

/* NNLO three-nucleon force of CCGPU_INT_CHIRAL, normal-ordered with respect to
 * the holes: approx -1 not normal-ordered (default), 0 E0, 1 E0 and Fock, 2 E0,
 * Fock and the two-body blocks (and the <cb||kd> elements of (T)), 3 as 2 and
 * (T) with the three-body element in its numerator, as tnf_approx of ccm_in
 * (tnf_switch F is -1).  cE, cD as the deck's second line (zero below NNLO);
 * cutoff (MeV) and nexp of its regulator as the deck's last line.  The same
 * constants define the three-body element <abc||ijk> of CCDT-1, also at -1.
 * Call before ccgpu_generate_interaction. */
int ccgpu_set_3nf(int handle, int approx, double cE, double cD, double cutoff, int nexp);

/* <pq||rs> of the generated NN interaction (without the normal-ordered 3NF)
 * for n orbital quadruples (0-based),
 * complex into v[2n]; zero where momentum or isospin is not conserved. */
int ccgpu_vnn_elements(int handle, int n, const int *p, const int *q, const int *r, const int *s, double *v);

/* Hand over one reference-layout block: rows are pairs (bra_p[r], bra_q[r]),
 * columns pairs (ket_r[c], ket_s[c]), 1-based orbital indices as in the
 * reference lookup tables; val is column-major complex with leading dim ld.
 * Only canonical orderings are read; the rest follows from antisymmetry. */
int ccgpu_set_block(int handle, int kind, int nbra, const int *bra_p, const int *bra_q,
                    int nket, const int *ket_r, const int *ket_s, const double *val, int ld);

/* Fock matrix (nsp x nsp complex, column-major) and E0. */
int ccgpu_set_fock(int handle, const double *fock, double e0_re, double e0_im);

/* Derive the device-side intermediates (ring gathers, conj V) after set_block. */
int ccgpu_finalize_setup(int handle);

int ccgpu_get_e0(int handle, double *e0);     /* e0[2] */
int ccgpu_get_fock_diag(int handle, double *f); /* nsp complex */

int ccgpu_mbpt2_start(int handle, double *ecorr);                /* ecorr[2] */
int ccgpu_residual(int handle, int precision);
int ccgpu_update(int handle, int *count, int *nstep, int subspace, int diis_step);
int ccgpu_energy(int handle, double *ecorr);                     /* ecorr[2] */
int ccgpu_solve(int handle, const ccgpu_solve_opts *opts, ccgpu_solve_result *res);
/* The (T) energy the reference adds to the CCD energy: et_4 = sum |W|^2/D, or
 * at tnf_approx 3 et_5 = sum |W + W3|^2/D, over a<b<c, i<j<k with W3 =
 * <abc||ijk> of the three-nucleon force. */
int ccgpu_triples(int handle, double *et);                       /* et[2] */
/* et_4 and et_5 of the last ccgpu_triples call (et_5 is zero below
 * tnf_approx 3), the two numbers of the reference's 'CCSD(T)' line. */
int ccgpu_triples_parts(int handle, double *et4, double *et5);  /* [2] each */
/* CCDT-1 (cc_approx 'CCDT1'; needs the built-in chiral interaction and the
 * 3NF constants of ccgpu_set_3nf): the iteration of T2 is that of CCD, and
 * every energy adds sum T3 conj(W3) over a<b<c, i<j<k of the triples
 * amplitude T3 = (W3 + W(T2))/D.  ccgpu_setup_t3 returns that energy for the
 * start T3 = W3/D; ccgpu_t3_intermediate builds T3 from the current T2 (call
 * it before ccgpu_residual, as the reference does) and ccgpu_t3_energy returns
 * its energy (add it to the ccgpu_energy of the same iteration). */
int ccgpu_setup_t3(int handle, double *e3);                      /* e3[2] */
int ccgpu_t3_intermediate(int handle);
int ccgpu_t3_energy(int handle, double *e3);                     /* e3[2] */
/* Of the last ccgpu_triples call: the number of (abc, ijk) elements summed
 * (a<b<c particles, i<j<k holes, same total momentum and isospin) and its
 * wall time in seconds (kernels and the result copy). */
int ccgpu_triples_stats(int handle, long long *elements, double *seconds);

/* Amplitudes t(ab,ij) for n orbital quadruples (0-based), any ordering. */
int ccgpu_get_t2(int handle, int n, const int *a, const int *b, const int *i,
                 const int *j, double *t);
int ccgpu_set_t2(int handle, int n, const int *a, const int *b, const int *i,
                 const int *j, const double *t);
/* Whole unique amplitude vector (t2_unique complex) in storage order. */
int ccgpu_get_t2_unique(int handle, double *t);
int ccgpu_set_t2_unique(int handle, const double *t);
/* Orbital quadruples (a,b,i,j) of the storage order, 4 ints per element. */
int ccgpu_get_t2_labels(int handle, int *abij);

/* Correlation energies of the last ccgpu_solve / ccgpu_batch_solve of a handle:
 * the start (MBPT2 or learned), then one per iteration, as the reference's
 * per-iteration 'dener' lines follow them (ccm_iter.f90:88-94); up to max
 * complex values into e[2 max], the total number into count. */
int ccgpu_solve_energies(int handle, double *e, int max, int *count);

/* Solve several handles together: every kernel launch covers all of them. */
int ccgpu_batch_solve(int n, const int *handles, const ccgpu_solve_opts *opts,
                      ccgpu_solve_result *results);

/* Copy one internal device array (what: 0 R, 1 dh, 2 I7off, 3 dp, 4 I6off,
 * 5 X|I9, 6 Rt, 7 Tt, 8 I4, 9 T, 10 R0, 11 fdiag, 12 fpart, 13 W0, 14 Vt) of
 * n complex elements to the host; for tests and diagnostics. */
int ccgpu_debug_fetch(int handle, int what, double *out, long long n);

/* Per-kernel timings (ms, averaged over the recorded launches) of the last
 * solve; names are written as a ';'-separated list. */
int ccgpu_kernel_stats(char *names, int len, double *avg_ms, long long *launches, int max);

/* Diagnostics, no counterpart in the reference: the streaming-read rate of
 * this GPU's HBM (GB/s, best of reps passes of a plain 16-byte-load kernel
 * over a fresh buffer of `bytes`, timed with HIP events), the practical
 * ceiling next to the published peak for the HBM-bound kernels. */
int ccgpu_hbm_read_probe(long long bytes, int reps, double *gbs);

/* Tile shapes chosen by the autotuner so far, one entry per work list:
 * "famF/kindK: cC=MBxNB(time) ..."; returns the report length.  The tuner
 * times every candidate shape of every tile class of a batch's GEMM work lists
 * on first use and caches the choice by list signature (CCGPU_AUTOTUNE=0: the
 * default table).  No counterpart in the reference (its work partition,
 * ccm_mapping.f90:2-934, splits loops over MPI ranks). */
int ccgpu_autotune_report(char *buf, int len);

/* Warm-start network (16 -> 64 -> 64 -> 64 -> 1, GELU, BF16 MFMA with FP32
 * accumulation).  It predicts, per stored amplitude, the correction of the
 * converged CCD amplitude over its MBPT2 value from momentum, spin, density and
 * twist features:  t = t_MBPT2 + 0.01 * y * e_F / |D| with D = f_ii + f_jj -
 * f_aa - f_bb (the network predicts the numerator correction).  Parameters: a flat FP32 vector of
 * CCGPU_MLP_NPARAM values, W1[16][64] b1 W2[64][64] b2 W3[64][64] b3 w4[64] b4. */
#define CCGPU_MLP_NPARAM 9473
int ccgpu_mlp_create(unsigned long long seed, int *mlp);
int ccgpu_mlp_destroy(int mlp);
int ccgpu_mlp_get_params(int mlp, float *theta, int n);
int ccgpu_mlp_set_params(int mlp, const float *theta, int n);   /* resets Adam */
/* Append every amplitude of a converged solve to the device training set. */
int ccgpu_mlp_add_solve(int mlp, int handle, long long *nsamples);
/* Adam steps on minibatches of `batch` (multiple of 64) samples drawn with
 * replacement; loss_last = mean squared error of the last minibatch. */
int ccgpu_mlp_train(int mlp, int steps, int batch, double lr, unsigned long long seed, double *loss_last);
/* Start amplitudes of a handle from the network (instead of ccgpu_mbpt2_start);
 * returns the correlation energy of the start. */
int ccgpu_mlp_warm_start(int mlp, int handle, double *ecorr);
/* The same start for n handles in one stream, one synchronisation at the end
 * (the start of setup_t_amplitudes, ccm_iter.f90:2329-2405, for every solve of
 * a batch); the energies come with the solve. */
int ccgpu_mlp_warm_start_batch(int mlp, const int *handles, int n);
/* For checks: network output, mean-squared loss and its gradient on n host
 * samples (n x 16 features, rounded to BF16 on entry), and the features and
 * targets of a handle's amplitudes (nT x 16 and nT floats). */
int ccgpu_mlp_forward(int mlp, int n, const float *x, float *y);
int ccgpu_mlp_loss_grad(int mlp, int n, const float *x, const float *t, double *loss, float *grad);
int ccgpu_mlp_features(int mlp, int handle, float *feat, float *target);

/* Correlation energy sum t conj(r0) of n unique amplitudes (the sum of
 * ccgpu_energy) on caller-owned device buffers t, r0 (n complex each),
 * written to the device buffer ecorr (one complex), launched on `stream`
 * (a hipStream_t; null for the default stream) without synchronizing.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops).  It allocates nothing:
 * its block partials live in a static per-device array of 4096 complex values,
 * one per span of 65536 amplitudes, so n is limited to 268 435 456; larger n
 * is rejected with an error.  Calls that overlap on one GPU share that array
 * and race. */
int ccgpu_energy_noalloc(long long n, const double *t, const double *r0, double *ecorr, void *stream);

int ccgpu_last_error(char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif
