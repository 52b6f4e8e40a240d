/*
 * z2_oracle.h -- CPU restatement (FP64, real arithmetic) of the reference
 * Z2-gauge DQMC hot path.  TEST INFRASTRUCTURE ONLY: the GPU product path never
 * links, loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker.
 *
 * Every routine cites the reference file:line it follows (paths relative to the
 * reference repository root).  The model is real (real t, sigma, mu), so the
 * reference's complex(8) matrices carry exactly-zero imaginary parts and a real
 * restatement follows the same arithmetic for the element-wise parts; the two
 * spin copies go through identical operations, so one Green's function is kept
 * and the 2x2 determinant is squared (src/localK.f90:118).
 *
 * Layout: all N x N matrices are column-major, 0-based: A[i + j*N].
 * sigma is stored as sigma[(nt-1)*2N + (ii-1)] for bond ii (1..2N), slice nt.
 */
#ifndef Z2_ORACLE_H
#define Z2_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* paramC_sets.txt fields, src/calc_basic.f90:53-61 */
typedef struct {
  double RT, J, h, mu, Q;
  int32_t Nlx, Nly, Ltrot;
  double Beta;
  int32_t NlxTherm, NlyTherm, LtrotTherm;
  int32_t Nwrap, Nbin, Nsweep;
  int32_t is_tau, Nthermal;
  int32_t is_warm, Nwarm;
  int32_t is_global, Nglobal;
  int32_t absolute;
} z2o_params;

/* per-bin equal-time observables, src/obser_equal.f90:7-13 */
typedef struct {
  double flux_avg, chi_flux, pair_sc_q0, cdw_pi, cdw_shift, cdw_ratio, density;
} z2o_obs_equal;

typedef struct z2o_model z2o_model;
typedef struct z2o_chain z2o_chain;

/* ---- RNG: lib/package_of_matrix/Ran/ran_imada.f:55-68 ---- */
double z2o_ranf(int32_t *iseed);
int32_t z2o_nranf(int32_t *iseed, int32_t n); /* src/calc_basic.f90:113-121 */

/* ---- model: params + lattice + operators ---- */
z2o_model *z2o_model_new(const z2o_params *p);
void z2o_model_free(z2o_model *m);
int32_t z2o_model_ndim(const z2o_model *m);
/* lattice tables (1-based values as in the reference):
 * bond_list[2N*2] (bond-major: [b*2+k]), groups[4][N/2] concatenated with
 * group_len[4], bond_bonds[2N*7] ([b*7+k], k=0..6), imj[N*N] col-major,
 * L_bonds[N*5] ([i*5+k]), n_list[N*2] ([i*2+k]) */
void z2o_lattice_tables(const z2o_model *m, int32_t *bond_list, int32_t *groups,
                        int32_t *group_len, int32_t *bond_bonds, int32_t *imj,
                        int32_t *L_bonds, int32_t *n_list);
/* src/operator_kin.f90:61-86 (mmult_R: Mat <- B^{nflag} Mat) and 122-145
 * (mmult_L: Mat <- Mat B^{nflag}) on a caller matrix, sigma = full field */
void z2o_kin_mmult_R(const z2o_model *m, double *Mat, const double *sigma, int nt, int nflag);
void z2o_kin_mmult_L(const z2o_model *m, double *Mat, const double *sigma, int nt, int nflag);
/* src/operator_kin.f90:40-59 */
void z2o_kin_delta(const z2o_model *m, double s_old, double s_new, double D[4]);
/* src/fields.f90:40-119; sigma is the current field, sigma_new differs at (ii,nt) */
double z2o_bosonratio(const z2o_model *m, const double *sigma, const double *sigma_new, int ii, int nt);
/* forced flip of bond ii (old value s_old) applied to G in place; returns the
 * 2x2 determinant det[I + Delta (I - G)_PP] (src/localK.f90:85-165) */
double z2o_flip_update(const z2o_model *m, double *G, int ii, double s_old);
/* exp(+-dtau*mu), src/non_interact.f90:27-32 */
void z2o_mu_factors(const z2o_model *m, double *ep, double *em);

/* ---- dense linear algebra used by stabilization (LAPACK semantics) ---- */
/* column-pivoted Householder QR + D separation: src/stabilization.f90:42-68 */
void z2o_qdrp(int n, double *A, double *D, int32_t *ipvt, double *tau);
/* G = (I + UR DR VR VL DL UL)^{-1} via src/stabilization.f90:122-200 */
void z2o_stab_green(int n, int Ltrot, int nt, const double *UUR, const double *DUR,
                    const double *VUR, const double *UUL, const double *DUL,
                    const double *VUL, double *G);

/* ---- one Markov chain (one MPI rank of the reference) ---- */
/* fresh start: conf_set with seed (src/fields.f90:311-427) then conf_transfer */
z2o_chain *z2o_chain_new(const z2o_model *m, int32_t seed);
/* restart: sigma/lambda at the *therm* size, seed as read from confin */
z2o_chain *z2o_chain_new_conf(const z2o_model *m, int32_t seed, const double *sigma_therm,
                              const double *lambda_therm);
void z2o_chain_free(z2o_chain *c);

void z2o_therm(z2o_chain *c);                 /* src/local_sweep.f90:73-85 */
void z2o_pre(z2o_chain *c);                   /* src/local_sweep.f90:87-104 */
void z2o_sweep(z2o_chain *c, int toggle);     /* src/local_sweep.f90:152-195 */
void z2o_global_sweep(z2o_chain *c);          /* src/global_update.f90:313-337 */
/* the first nslices steps of a left pass (src/local_sweep.f90:111-120); a timing
 * sample for the CPU baseline */
void z2o_sweep_L_partial(z2o_chain *c, int nslices);
/* the left pass (src/local_sweep.f90:106-124), the right pass (:126-150), and
 * Dyn%reset + Dyn%sweep_R alone on the current state (src/dynamics.f90:41-62) */
void z2o_pass_L(z2o_chain *c);
void z2o_pass_R(z2o_chain *c, int toggle);
void z2o_dyn_measure(z2o_chain *c);
/* single-slice pieces for fine-grained parity (src/localK.f90:347-398) */
void z2o_slice_L(z2o_chain *c, int nt);
void z2o_slice_R(z2o_chain *c, int nt);

/* state access */
int32_t z2o_get_seed(const z2o_chain *c);
void z2o_set_seed(z2o_chain *c, int32_t seed);
void z2o_get_sigma(const z2o_chain *c, double *sigma);
void z2o_set_sigma(z2o_chain *c, const double *sigma);
void z2o_get_lambda(const z2o_chain *c, double *lambda);
void z2o_get_green(const z2o_chain *c, double *G);
void z2o_set_green(z2o_chain *c, const double *G);
/* which: 0 UUR 1 VUR 2 UUL 3 VUL (N*N); D: 0 DUR 1 DUL (N) */
void z2o_get_prop_mat(const z2o_chain *c, int which, double *out);
void z2o_get_prop_d(const z2o_chain *c, int which, double *out);
int z2o_get_is_beta(const z2o_chain *c);
void z2o_set_is_beta(z2o_chain *c, int is_beta);
/* drift: Xmaxm, Xmeanm of the propagator (src/process_matrix.f90:10) */
void z2o_get_drift(const z2o_chain *c, double *xmax, double *xmean);
/* counters: [0] local accepts, [1] local proposals, [2] global accepts,
 * [3] global proposals, [4] therm accepts, [5] therm proposals,
 * [6] sum of Wolff cluster sizes (as double) */
void z2o_get_counters(const z2o_chain *c, double *out);
/* running acceptance sums (AccCounter%acc): [0] Acc_Kl [1] Acc_Kg [2] Acc_Kt
 * [3] Wolff%size_cluster */
void z2o_get_acc(const z2o_chain *c, double *out);
void z2o_get_obs_equal(const z2o_chain *c, z2o_obs_equal *out);
/* the per-slice sums Obs_equal_calc adds for slice nt on the current G
 * (src/obser_equal.f90:73-149): flux_tau, pair_sum, cdw_sum, cdw_shift_sum,
 * density_slice; the accumulators are not touched */
void z2o_obs_equal_terms(const z2o_chain *c, int nt, double out[5]);
/* time-displaced arrays (real parts), 7 x N x Ltrot in the order
 * spin, pair_s, pair_d, charge_s, charge_d, single, curxx, col-major (imj,nt) */
void z2o_get_obs_tau(const z2o_chain *c, double *out);
/* time-displaced drift maxima/means (3 each): PropGreen%Xmaxm/Xmeanm */
void z2o_get_dyn_drift(const z2o_chain *c, double *xmax3, double *xmean3);
/* smallest | |ratio| - r | met by any local proposal since the last reset */
double z2o_get_min_gap(const z2o_chain *c);
void z2o_reset_min_gap(z2o_chain *c);
double z2o_get_gmin_gap(const z2o_chain *c);
/* optional per-proposal log: (bond, slice, accept) packed as
 * bond | slice<<16 | accept<<31 into a caller buffer; cap entries */
void z2o_set_decision_log(z2o_chain *c, uint32_t *buf, int64_t cap);
int64_t z2o_get_decision_count(const z2o_chain *c);

/* whole program, src/main.f90:1-137 for one rank: init (fresh or conf),
 * optional warm-up, pre, Nbin bins of [global sweep] + local sweep.
 * obs_out: Nbin records; may be NULL. */
void z2o_run_program(z2o_chain *c, z2o_obs_equal *obs_out);

#ifdef __cplusplus
}
#endif
#endif
