/*
 * z2dqmc.h -- C ABI of the MI355X (gfx950) Z2-gauge DQMC engine.
 *
 * Plain C types only: the reference's Fortran modules (or any FFI) bind these
 * with iso_c_binding / ctypes.  A context holds a batch of independent Markov
 * chains (one reference MPI rank each) that share the lattice size L, Ltrot and
 * Nwrap; each chain has its own couplings, seed and field.  All chains of a
 * context advance together: every entry point below acts on all of them, in
 * one batched launch per step, on the context's own HIP stream.
 *
 * Conventions: every function returns 0 on success and a negative Z2_E* code
 * on error (z2_last_error() describes it); nothing aborts.  Host buffers are
 * owned by the caller.  Matrices are N x N column-major (Fortran order) with
 * 0-based C indexing; sigma is sigma(ii, nt) column-major, ii = 1..2N fastest,
 * values +-1 as double (reference SigmaConf%sigma, src/fields.f90:10).
 * Spin-down quantities equal spin-up ones bit for bit in the reference
 * (identical operations on PropU and PropD), so one Green's function is kept.
 */
#ifndef Z2DQMC_H
#define Z2DQMC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define Z2_OK 0
#define Z2_EINVAL (-1)   /* bad argument */
#define Z2_EHIP (-2)     /* HIP runtime error */
#define Z2_ENOMEM (-3)   /* device allocation failed */
#define Z2_ESTATE (-4)   /* call out of order (e.g. sweep before pre) */

#define Z2_PREC_FP64 0   /* everything FP64 (the reference's arithmetic) */
#define Z2_PREC_MIXED 1  /* G and the delayed updates in FP32 (transformed in FP64 arithmetic),
                            FP64 propagators, QR/UDV, G recompute and ratio */

/* shared parameters: the size-defining lines of paramC_sets.txt
 * (reference src/calc_basic.f90:53-61) plus the precision mode */
typedef struct {
  int32_t Nlx, Nly, Ltrot, Nwrap;
  int32_t Nsweep;     /* sweeps per z2_sweep call (reference Nsweep) */
  int32_t precision;  /* Z2_PREC_* */
} z2_params;

/* per-chain couplings (reference RT, J, h, mu, Q, Beta) and RNG seed */
typedef struct {
  double RT, J, h, mu, Q, Beta;
  int32_t seed;
  int32_t absolute;   /* initial field mode for z2_init_fields (src/fields.f90:322-424) */
} z2_chain_desc;

typedef struct z2_ctx z2_ctx;

const char *z2_last_error(void);

/* Model_init + Propagator/WrapList make for all chains
 * (reference src/model.f90:16-36, src/process_matrix.f90:46-56,104-112). */
int z2_ctx_create(const z2_params *p, int32_t n_chains, const z2_chain_desc *chains,
                  z2_ctx **out);
int z2_ctx_destroy(z2_ctx *ctx);
int z2_ctx_info(const z2_ctx *ctx, int32_t *N, int32_t *n_chains, int32_t *precision);

/* conf_set (fresh start, confin = 0) with each chain's seed and absolute mode,
 * run on the host, then uploaded (src/fields.f90:227-309,311-427). */
int z2_init_fields(z2_ctx *ctx);
/* field, lambda and seed of one chain (confin restart / conf_out, src/fields.f90:509-557) */
int z2_set_sigma(z2_ctx *ctx, int32_t chain, const double *sigma, int32_t seed);
int z2_get_sigma(z2_ctx *ctx, int32_t chain, double *sigma);
int z2_get_seed(z2_ctx *ctx, int32_t chain, int32_t *seed);
/* lambda(N) of one chain (SigmaConf%lambda, src/fields.f90:10; drawn by
 * conf_set, lambda_random_fill + lambda_enforce_sector :559-603), kept for the
 * confin / confout files; the sweeps do not use it */
int z2_get_lambda(z2_ctx *ctx, int32_t chain, double *lambda);
int z2_set_lambda(z2_ctx *ctx, int32_t chain, const double *lambda);
/* thermalisation-size configurations (NlxTherm, NlyTherm, LtrotTherm of
 * paramC_sets.txt): conf_transfer (src/fields.f90:429-507) tiles a field of that
 * size (sigma(ii, nt), ii fastest) and its lambda onto the chain's lattice
 * (LtrotTherm <= Ltrot), sets seed; z2_init_fields_therm is the fresh start at
 * that size (conf_set there with each chain's seed and absolute mode, then
 * conf_transfer, src/fields.f90:227-309) */
int z2_set_sigma_therm(z2_ctx *ctx, int32_t chain, int32_t NlxTherm, int32_t NlyTherm, int32_t LtrotTherm,
                       const double *sigma_therm, const double *lambda_therm, int32_t seed);
int z2_init_fields_therm(z2_ctx *ctx, int32_t NlxTherm, int32_t NlyTherm, int32_t LtrotTherm);

/* LocalSweep%pre (src/local_sweep.f90:87-104) */
int z2_pre(z2_ctx *ctx);
/* LocalSweep%sweep (src/local_sweep.f90:152-195): Nsweep x (left + right pass)
 * with equal-time observables at every slice; is_beta selects the pass order
 * (1: left pass first) for all chains, or Z2_BETA_PER_CHAIN takes each chain's
 * own is_beta (reference: one is_beta per MPI rank, src/main.f90:73-77). */
#define Z2_BETA_PER_CHAIN (-1)
int z2_sweep(z2_ctx *ctx, int32_t is_beta);
/* a chain's is_beta: 1 after z2_pre, set by z2_sweep(ctx, 0 / 1), reversed by
 * an accepted Wolff trial */
int z2_set_is_beta(z2_ctx *ctx, int32_t chain, int32_t is_beta);
int z2_get_is_beta(z2_ctx *ctx, int32_t chain, int32_t *is_beta);

/* Global_sweep (src/global_update.f90:313-337, Wolff_flip :65-181,
 * ratioK_fermion and GlobalK_prop_L/R src/globalK.f90:15-151): Nglobal trials
 * per chain, each flipping one bond along a time cluster; the trial propagates
 * in the chain's own direction (Global_sweep_L when is_beta = 1, else _R) and an
 * accepted trial reverses its is_beta.  Run before z2_sweep(ctx,
 * Z2_BETA_PER_CHAIN) as the reference's main does (src/main.f90:76-77). */
int z2_global_sweep(z2_ctx *ctx, int32_t Nglobal);
/* cumulative Wolff counts: accepted, proposed, sum of cluster sizes; and the
 * running wolff_size_cluster (sum over global sweeps of the mean cluster size) */
int z2_get_global_counters(z2_ctx *ctx, int32_t chain, uint64_t out[3], double *size_cluster);
/* one LocalK_therm sweep (src/localK.f90:400-420): boson-only Metropolis over
 * every (nt, ii), nt outer (Local_sweep_therm, src/local_sweep.f90:73-85), for
 * the warm-up loop of src/main.f90:57-58 (is_warm) */
int z2_therm(z2_ctx *ctx);
/* cumulative warm-up counts: accepted, proposed */
int z2_get_therm_counters(z2_ctx *ctx, int32_t chain, uint64_t out[2]);

/* time-displaced measurement (is_tau = 1 in the reference's parameter file):
 * z2_enable_tau allocates its state (about 26 N^2 doubles per chain);
 * z2_sweep_tau(ctx, is_beta, toggle) = LocalSweep%sweep with toggle
 * (src/local_sweep.f90:152-195): with toggle, Dyn%reset + Dyn%sweep_R
 * (src/dynamics.f90:41-62) run at the start of every right pass, and the
 * Obs_tau accumulators (src/obser_tau.f90:56-133) are divided by the number of
 * such passes (Obs_tau%ave, src/obser_tau.f90:37-47).  z2_sweep(ctx, b) is
 * z2_sweep_tau(ctx, b, 0). */
int z2_enable_tau(z2_ctx *ctx);
int z2_sweep_tau(z2_ctx *ctx, int32_t is_beta, int32_t toggle);
/* Dyn%reset + Dyn%sweep_R on their own from the current state, accumulators
 * zeroed first and not normalised (unit step of the measurement) */
int z2_dyn_measure(z2_ctx *ctx);
/* time-displaced correlators of one chain, out[7 * Ltrot * N] laid out
 * [k][nt-1][d]: k = spin, pair_s, pair_d, charge_s, charge_d, single, curxx
 * (Obs_tau fields, src/obser_tau.f90:6-16), d = imj(i,j) - 1 the displacement */
int z2_get_obs_tau(z2_ctx *ctx, int32_t chain, double *out);
/* Wrap_tau drift of G(0,t), G(t,0), G(t,t): running max and running sum of the
 * per-recompute maxima (src/stabilization.f90:496-508) */
int z2_get_dyn_drift(z2_ctx *ctx, int32_t chain, double xmax[3], double xmean[3]);

/* One sweep of ctx in lockstep with an FP64 reference context of the same
 * shape (no counterpart in the reference; measures the precision deviation):
 * ctx replays ref's accept/reject decisions, and maxdiff[chain] receives
 * max |G_ctx - G_ref| over the propagated Green's functions just before each
 * recompute of the sweep. */
int z2_sweep_shadow(z2_ctx *ctx, z2_ctx *ref, int32_t is_beta, double *maxdiff);
/* the same with toggle: the time-displaced measurement (Dyn%reset + Dyn%sweep_R)
 * at the start of the right pass on both contexts, accumulators normalised as
 * in z2_sweep_tau (needs z2_enable_tau on both) */
int z2_sweep_shadow_tau(z2_ctx *ctx, z2_ctx *ref, int32_t is_beta, int32_t toggle, double *maxdiff);
/* diagnostic: one shadow sweep recording, after every step, the max over chains
 * and elements of |G_ctx - G_ref| in trace[i] and the step in code[i]
 * (kind << 16 | nt; kinds 0 Wrap_L, 1 Wrap_R, 2 Obs_equal%calc, 3 slice of the
 * left pass, 4 slice of the right pass); *nsteps <= cap steps */
int z2_sweep_shadow_trace(z2_ctx *ctx, z2_ctx *ref, int32_t is_beta, double *trace, int32_t *code, int32_t cap,
                          int32_t *nsteps);

/* fine-grained steps, all chains at slice nt (same semantics as the reference
 * routines named): */
int z2_wrap_pre(z2_ctx *ctx, int32_t nt);              /* Wrap_pre   src/stabilization.f90:374-398 */
int z2_wrap_L(z2_ctx *ctx, int32_t nt, int32_t flag);  /* Wrap_L     src/stabilization.f90:400-437 */
int z2_wrap_R(z2_ctx *ctx, int32_t nt, int32_t flag);  /* Wrap_R     src/stabilization.f90:439-476 */
int z2_slice_L(z2_ctx *ctx, int32_t nt);  /* propT_L + LocalK_prop_L  src/multiply.f90:79-93, src/localK.f90:347-373 */
int z2_slice_R(z2_ctx *ctx, int32_t nt);  /* LocalK_prop_R + propT_R  src/localK.f90:375-398, src/multiply.f90:95-108 */
int z2_obs_equal_calc(z2_ctx *ctx, int32_t nt);  /* Obs_equal%calc src/obser_equal.f90:73-149 */
int z2_obs_reset(z2_ctx *ctx);                    /* Obs_equal%reset src/obser_equal.f90:33-48 */
int z2_obs_equal_ave(z2_ctx *ctx, int32_t nobs);  /* Obs_equal%ave(Nobs) src/obser_equal.f90:50-71 */
/* Local_sweep_reset (src/local_sweep.f90:55-71): Obs_equal%reset, with toggle
 * Obs_tau%reset, and the field copied along tau for chains with |h| <= 1e-10 */
int z2_sweep_reset(z2_ctx *ctx, int32_t toggle);

/* Multiply_mod (src/multiply.f90), one procedure per call, all chains at slice nt.
 * The LocalK_prop_L/R pair below plus these compose a sweep exactly as the
 * reference's Local_sweep_L/R do (src/local_sweep.f90:106-150), with the same
 * arithmetic as z2_sweep. */
int z2_prop_K_pre(z2_ctx *ctx, int32_t nt);  /* propK_pre  src/multiply.f90:8-15: UUR <- B UUR */
int z2_prop_T_pre(z2_ctx *ctx, int32_t nt);  /* propT_pre  src/multiply.f90:65-77: UUR <- UUR e^{dtau mu} */
int z2_prop_T_L(z2_ctx *ctx, int32_t nt);    /* propT_L    src/multiply.f90:79-93: G <- G e^{+} e^{-}, UUL <- UUL e^{+} */
int z2_prop_T_R(z2_ctx *ctx, int32_t nt);    /* propT_R    src/multiply.f90:95-108: G <- G e^{+} e^{-}, UUR <- UUR e^{+} */
/* time-displaced state (after z2_enable_tau; FP64 in both precision modes) */
int z2_propgr_K_R(z2_ctx *ctx, int32_t nt);  /* propgrK_R  src/multiply.f90:43-63 */
int z2_propgr_T_R(z2_ctx *ctx, int32_t nt);  /* propgrT_R  src/multiply.f90:110-130 */
/* LocalK_mod (src/localK.f90) */
int z2_localK_prop_L(z2_ctx *ctx, int32_t nt); /* LocalK_prop_L src/localK.f90:347-373 */
int z2_localK_prop_R(z2_ctx *ctx, int32_t nt); /* LocalK_prop_R src/localK.f90:375-398 */
/* Dynamics_mod / ObserTau_mod / Wrap_tau */
int z2_dyn_reset(z2_ctx *ctx);                 /* Dyn%reset src/dynamics.f90:41-48; after a left pass */
int z2_obs_tau_reset(z2_ctx *ctx);             /* Obs_tau%reset src/obser_tau.f90:26-35 */
int z2_obs_tau_calc(z2_ctx *ctx, int32_t nt);  /* Obs_tau%calc src/obser_tau.f90:56-133 */
int z2_obs_tau_ave(z2_ctx *ctx, int32_t nobst);/* Obs_tau%ave src/obser_tau.f90:37-47 */
int z2_wrap_tau(z2_ctx *ctx, int32_t nt);      /* Wrap_tau src/stabilization.f90:478-514 */

/* results / diagnostics */
int z2_sync(z2_ctx *ctx);
int z2_get_green(z2_ctx *ctx, int32_t chain, double *G);
int z2_set_green(z2_ctx *ctx, int32_t chain, const double *G);
/* Xmaxm / Xmeanm drift of |G_recomputed - G_propagated| (src/stabilization.f90:427-430) */
int z2_get_drift(z2_ctx *ctx, int32_t chain, double *xmax, double *xmean);
/* per-bin averages after z2_sweep: flux_avg, chi_flux, pair_sc_q0, cdw_pi,
 * cdw_shift, cdw_ratio, density (src/obser_equal.f90:50-71) */
int z2_get_obs_equal(z2_ctx *ctx, int32_t chain, double out[7]);
/* cumulative local-flip counts: accepted, proposed */
int z2_get_counters(z2_ctx *ctx, int32_t chain, uint64_t out[2]);
/* delay depth k of the local-update kernel: accepted rank-2 updates are kept
 * pending (in LDS) and applied to G as one rank-2k matrix-core product per k
 * accepts (default min(16, kmax(N)), kmax set by the LDS size); the decisions
 * do not depend on k, G only through rounding */
int z2_set_delay(z2_ctx *ctx, int32_t k);
/* local-update kernel variant (no counterpart in the reference): in the mixed
 * mode at N = 256, Z2_FLIP_REGISTER (default) keeps the chain's FP32 G in the
 * registers of its workgroup for the whole slice and applies every pair of
 * accepted updates as one rank-4 matrix-core product there (no delay depth);
 * Z2_FLIP_DELAYED is the delayed-update kernel used everywhere else.  The
 * decisions are the serial ones in both; G differs by rounding. */
#define Z2_FLIP_DELAYED 0
#define Z2_FLIP_REGISTER 1
int z2_set_flip_variant(z2_ctx *ctx, int32_t variant);
/* G recompute of the stabilization (no counterpart in the reference):
 * Z2_STAB_REFERENCE (default) is stab_green (src/stabilization.f90:122-200):
 * M = (UL UR)^T + DR (VR VL) DL, column-pivoted QR, G = UL^T P V^-1 D^-1 Q^T UR^T,
 * with the reference's branch on nt > Ltrot/2.  Z2_STAB_DSPLIT is the D-split form
 * (Loh et al.): DR = DRb DRs, DL = DLb DLs with Db = max(D, 1), Ds = min(D, 1),
 * M' = DRb^-1 (UL UR)^T DLb^-1 + DRs (VR VL) DLs, G = UL^T DLb^-1 M'^-1 DRb^-1 UR^T:
 * no product of large and small scales is ever formed, which keeps G accurate at
 * large beta where the reference's form loses it; it is a deviation from the
 * reference's arithmetic (G agrees to rounding where both are accurate), for
 * opt-in use.  The time-displaced recompute is not affected. */
/* mixed mode (no counterpart in the reference): precision of the propagators
 * UUR, UUL between stabilizations (the B-products the UDV checkpoints are made
 * from): 64 (FP64, the default) or 32 (FP32, as G; about 1.6% faster at L = 16,
 * but the recomputed G then carries the products' FP32 error); set before z2_pre */
int z2_set_u_precision(z2_ctx *ctx, int32_t bits);
/* kinetic tiles (no counterpart in the reference): columns per workgroup of the
 * row-rotation kernel (8 or 16) and rows per workgroup of the column-rotation
 * kernel (8 or 32); results do not depend on them (default 8, 8) */
int z2_set_kin_tiles(z2_ctx *ctx, int32_t rows_cb, int32_t cols_rb);
#define Z2_STAB_REFERENCE 0
#define Z2_STAB_DSPLIT 1
int z2_set_stab_mode(z2_ctx *ctx, int32_t mode);
/* QR used by the stabilization (no counterpart in the reference, which calls
 * ZGEQP3): Z2_QR_BLOCKED (default) orders the columns once by decreasing norm
 * and runs a blocked Householder QR with block-reflector updates on the matrix
 * cores; Z2_QR_PIVOTED keeps ZGEQP3's column pivoting by downdated norms. Both
 * give A P = Q R with the same output convention; G agrees to rounding. */
#define Z2_QR_BLOCKED 0
#define Z2_QR_PIVOTED 1
int z2_set_qr_mode(z2_ctx *ctx, int32_t mode);
/* QDRP_decompose (src/stabilization.f90:42-68) for nmat n x n matrices that the
 * caller holds in device memory (column-major, matrix stride n*n; n <= 512),
 * enqueued on `stream` (hipStream_t as void*, NULL: the null stream):
 * A P = Q D R with the columns ordered by decreasing norm and a blocked
 * Householder QR.  On return A holds R with its rows scaled to unit diagonal
 * magnitude in the upper triangle and the Householder vectors below, D = |diag R|
 * (stride n), ipvt the 1-based column order (stride n), tau the reflector
 * scalars (stride n), and Q (optional, stride n*n) the explicit orthogonal
 * factor.  The library allocates a workspace of nmat n^2 doubles for the call
 * and releases it before returning (the call waits for the stream). */
int z2_qdrp_batched(int32_t n, int32_t nmat, double *A, double *D, int32_t *ipvt, double *tau, double *Q,
                    void *stream);

/* No counterpart in the reference: z2_qdrp_batched for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; the workspace
 * is a static device array of Z2_QDRP_SCRATCH_DOUBLES doubles (32 MiB) that
 * each GPU holds until the process ends, so nmat * n^2 must not exceed it
 * (Z2_EINVAL otherwise).  Calls nothing that allocates and does not wait for
 * the stream.  Every call on one GPU uses the same array: calls that overlap on
 * one GPU race. */
#define Z2_QDRP_SCRATCH_DOUBLES 4194304
int z2_qdrp_batched_noalloc(int32_t n, int32_t nmat, double *A, double *D, int32_t *ipvt, double *tau, double *Q,
                            void *stream);

/* HIP stream of the context (hipStream_t as void*) for timing by the caller */
int z2_get_stream(z2_ctx *ctx, void **stream);

/* ---- batches of chains of uneven size (no counterpart in the reference, which
 * runs one chain per MPI rank, src/main.f90:23-25).  Every chain has its own
 * lattice, Ltrot, Nwrap, couplings and seed.  Chains of one shape (Nlx, Nly,
 * Ltrot, Nwrap) form a group -- one context, on its own HIP stream, split into
 * near-equal parts of at most max_group chains when max_group > 0 -- and a sweep
 * of the batch advances every group: Z2_EXEC_EAGER issues the groups' launches in
 * turn on their streams, Z2_EXEC_GRAPH replays one captured HIP graph holding
 * all groups' launches as parallel branches (one graph launch per sweep; the
 * first sweep of each kind runs eagerly).  Both give the same results bit for
 * bit, and the same as one context per group.  The per-chain state and results
 * are read through the group context: z2_batch_chain_ctx maps a batch chain to
 * (context, index in it). */
typedef struct {
  int32_t Nlx, Nly, Ltrot, Nwrap;
  z2_chain_desc d;
} z2_batch_chain;
typedef struct z2_batch z2_batch;
#define Z2_EXEC_EAGER 0
#define Z2_EXEC_GRAPH 1
int z2_batch_create(const z2_batch_chain *chains, int32_t n_chains, int32_t precision, int32_t Nsweep,
                    int32_t max_group, z2_batch **out);
int z2_batch_destroy(z2_batch *b);
int z2_batch_info(z2_batch *b, int32_t *n_chains, int32_t *n_groups);
int z2_batch_chain_ctx(z2_batch *b, int32_t chain, z2_ctx **ctx, int32_t *index);
int z2_batch_group_ctx(z2_batch *b, int32_t group, z2_ctx **ctx);
int z2_batch_init_fields(z2_batch *b);
int z2_batch_enable_tau(z2_batch *b);
int z2_batch_pre(z2_batch *b);
/* LocalSweep%sweep for every chain: is_beta 0 / 1 for all, or Z2_BETA_PER_CHAIN
 * (each chain's own direction, eager); toggle: time-displaced measurement */
int z2_batch_sweep(z2_batch *b, int32_t is_beta, int32_t toggle);
int z2_batch_sync(z2_batch *b);
int z2_batch_set_exec(z2_batch *b, int32_t mode);
int z2_batch_get_exec(z2_batch *b, int32_t *mode);
/* Autotuner (before z2_batch_pre, after the fields are set): per group, on clones
 * of its chains after `warm` sweeps, the delay depth k in {4, 8, 16, 32} (up to
 * the LDS limit) and the kinetic tiles by timing `sweeps` sweeps each, then
 * Nwrap in {5, 10, 20, own} (dividing Ltrot) as the fastest whose drift
 * max|G_recomputed - G_propagated| stays <= drift_tol; then the execution mode on
 * a clone of the whole batch.  The batch's own chains are not advanced: a group
 * whose Nwrap changes is rebuilt from its chains' fields and seeds.  path (may be
 * NULL) receives the choices; z2_batch_load_tuning applies such a file (before
 * z2_batch_pre) to the groups it matches (lattice, Ltrot, precision, chain count). */
int z2_batch_autotune(z2_batch *b, int32_t warm, int32_t sweeps, double drift_tol, const char *path);
int z2_batch_load_tuning(z2_batch *b, const char *path);

/* Kernel timing (no counterpart in the reference; for roofline measurement):
 * bracket every launch of one kernel class with HIP events on the context's
 * stream; z2_profile_read sums the elapsed times since z2_profile_kernel. */
#define Z2_K_NONE 0
#define Z2_K_QRP 1
#define Z2_K_FLIP 2
#define Z2_K_KIN_ROWS 3
#define Z2_K_KIN_COLS 4
#define Z2_K_GEMM 5
#define Z2_K_ORGQR 6
#define Z2_K_TRSM 7
#define Z2_K_OBS 8
#define Z2_K_OBS_TAU 9
#define Z2_K_QT_APPLY 10  /* Q^T applied as block reflectors in the G recompute */
int z2_profile_kernel(z2_ctx *ctx, int32_t kind);
int z2_profile_read(z2_ctx *ctx, double *total_ms, int64_t *launches);
/* start and end (ms) of the first min(cap, launches) timed launches, measured
 * from the moment z2_profile_kernel last ran on ref (ctx itself or another context
 * on the same device, profiled before ctx; a later z2_profile_kernel on ref moves
 * that origin): the union over contexts shows how many launches of the class run
 * at once; *n = entries written */
int z2_profile_spans(z2_ctx *ctx, z2_ctx *ref, double *start_ms, double *end_ms, int64_t cap, int64_t *n);
/* algorithmic work of the launches timed since z2_profile_kernel: flops for
 * QR (4/3 n^3), explicit Q (4/3 n^3), Q^T application (2 n^3), GEMM (2 m n k),
 * TRSM (n^3); bytes for the kinetic and observable kernels (read + write of
 * every matrix, 2 N^2 x element size); 0 for the flip kernel, whose work
 * (4 N^2 flops per accepted update) the caller takes from z2_get_counters */
int z2_profile_work(z2_ctx *ctx, double *work);

#ifdef __cplusplus
}
#endif
#endif
