/*
 * z2_oracle.c -- CPU restatement of the reference Z2-gauge DQMC hot path.
 *
 * TEST INFRASTRUCTURE.  This file is the checker: only tests/, the smoke()
 * entry and the cpu_baseline leg of bench.py use it.  The GPU library never
 * links or calls it.
 *
 * Arithmetic follows the reference statement by statement (same operand order,
 * no FMA contraction: build with -ffp-contract=off), in real FP64.  The
 * reference keeps complex(8) matrices whose imaginary parts stay exactly zero
 * for this real model, so each complex product/sum reduces to the real one
 * written here.  Dense linear algebra follows the netlib LAPACK/BLAS
 * algorithms the reference calls (ZGEQP3/ZLAQP2, ZUNGQR/ZUNG2R, ZUNMQR/ZUNM2R,
 * ZTRMM, ZTRSM, ZGEMM, ZLAPMR/ZLAPMT, LINPACK ZGEFA/ZGEDI) restated for real
 * data; the system LAPACK the reference links is unpinned, so those parts match
 * it to rounding, not bit for bit.
 */
#include "z2_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ZERO_H 1.0e-10   /* src/calc_basic.f90:5 */
#define UPBOUND 1.0e+200 /* src/calc_basic.f90:7 */
#define RATIO_EPS 1.0e-300 /* src/globalK.f90:9 */

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "z2_oracle: out of memory\n");
    abort();
  }
  return p;
}

/* ======================================================================
 * RNG  (lib/package_of_matrix/Ran/ran_imada.f:55-68)
 * iq <- iq*48828125 with 32-bit wrap; if negative, (iq + 2147483647) + 1.
 * ====================================================================== */
double z2o_ranf(int32_t *iq) {
  uint32_t x = (uint32_t)(*iq) * 48828125u;
  int32_t s = (int32_t)x;
  if (s < 0) s = (int32_t)((uint32_t)s + 2147483648u);
  *iq = s;
  return (double)s / 2147483648.0;
}

/* src/calc_basic.f90:113-121: nint(ranf*N + 0.5), clamped to [1,N] */
int32_t z2o_nranf(int32_t *iq, int32_t n) {
  double x = z2o_ranf(iq) * (double)n + 0.5;
  int32_t r = (int32_t)lround(x); /* nint: half away from zero */
  if (r < 1) r = 1;
  if (r > n) r = n;
  return r;
}

/* src/calc_basic.f90:124-130 */
static int npbc(int nr, int L) {
  if (nr < 1) return nr + L;
  if (nr > L) return nr - L;
  return nr;
}

/* ======================================================================
 * Model: parameters, lattice, operators
 * ====================================================================== */
struct z2o_model {
  z2o_params p;
  int N, N2, T, Nst;
  int NT, N2T; /* therm sizes */
  double Dtau;
  /* lattice (1-based values, 0-based storage) */
  int *n_list;        /* N x 2: [i*2+k] */
  int *inv_n_list;    /* Nlx x Nly: [(nx-1) + (ny-1)*Nlx] */
  int *bond_list;     /* 2N x 2: [b*2+k] */
  int *inv_bond_list; /* N x N: [(i-1) + (j-1)*N] */
  int *group[4];
  int glen[4];
  int *bond_bonds; /* 2N x 7: [b*7+k] */
  int *imj;        /* N x N col-major */
  int *L_bonds;    /* N x 5 */
  /* operators */
  double alpha, alpha_2;   /* src/operator_kin.f90:22-29 */
  double expMU_P, expMU_M; /* src/non_interact.f90:27-32 */
};

/* src/lattice.f90:25-221 */
static void lattice_make(z2o_model *m) {
  const int Nlx = m->p.Nlx, Nly = m->p.Nly, N = m->N;
  m->n_list = xcalloc((size_t)N * 2, sizeof(int));
  m->inv_n_list = xcalloc((size_t)Nlx * Nly, sizeof(int));
  int nc = 0;
  for (int ny = 1; ny <= Nly; ++ny)
    for (int nx = 1; nx <= Nlx; ++nx) {
      nc++;
      m->n_list[(nc - 1) * 2 + 0] = nx;
      m->n_list[(nc - 1) * 2 + 1] = ny;
      m->inv_n_list[(nx - 1) + (ny - 1) * Nlx] = nc;
    }
#define INVN(x, y) m->inv_n_list[((x)-1) + ((y)-1) * Nlx]
#define NL(i, k) m->n_list[((i)-1) * 2 + (k)-1]
  m->bond_list = xcalloc((size_t)2 * N * 2, sizeof(int));
  m->inv_bond_list = xcalloc((size_t)N * N, sizeof(int));
  for (int k = 0; k < N * N; ++k) m->inv_bond_list[k] = -1;
#define IBL(i, j) m->inv_bond_list[((i)-1) + ((j)-1) * N]
  nc = 0;
  for (int ny = 1; ny <= Nly; ++ny)
    for (int nx = 1; nx <= Nlx; ++nx) {
      nc++;
      int ii = INVN(nx, ny);
      int ix = INVN(npbc(nx + 1, Nlx), ny);
      m->bond_list[(nc - 1) * 2 + 0] = ii;
      m->bond_list[(nc - 1) * 2 + 1] = ix;
      IBL(ii, ix) = nc;
      nc++;
      int iy = INVN(nx, npbc(ny + 1, Nly));
      m->bond_list[(nc - 1) * 2 + 0] = ii;
      m->bond_list[(nc - 1) * 2 + 1] = iy;
      IBL(ii, iy) = nc;
    }
  /* checkerboard groups, src/lattice.f90:72-102 */
  int cap = 2 * N;
  for (int g = 0; g < 4; ++g) {
    m->group[g] = xcalloc((size_t)cap, sizeof(int));
    m->glen[g] = 0;
  }
  for (nc = 1; nc <= 2 * N; ++nc) {
    int ii = m->bond_list[(nc - 1) * 2 + 0], jj = m->bond_list[(nc - 1) * 2 + 1];
    int ix = NL(ii, 1), iy = NL(ii, 2), jx = NL(jj, 1), jy = NL(jj, 2);
    if (npbc(ix + 1, Nlx) == jx) {
      if (ix % 2 == 0) m->group[0][m->glen[0]++] = nc;
      else m->group[1][m->glen[1]++] = nc;
    } else if (npbc(iy + 1, Nly) == jy) {
      if (iy % 2 == 0) m->group[2][m->glen[2]++] = nc;
      else m->group[3][m->glen[3]++] = nc;
    }
  }
  /* plaquette neighbours, src/lattice.f90:104-141 */
  m->bond_bonds = xcalloc((size_t)2 * N * 7, sizeof(int));
#define BB(b, k) m->bond_bonds[((b)-1) * 7 + (k)]
  for (nc = 1; nc <= 2 * N; ++nc) {
    int ii = m->bond_list[(nc - 1) * 2 + 0], jj = m->bond_list[(nc - 1) * 2 + 1];
    int ix = NL(ii, 1), iy = NL(ii, 2), jx = NL(jj, 1), jy = NL(jj, 2);
    BB(nc, 0) = nc;
    if (npbc(ix + 1, Nlx) == jx) {
      int nx = INVN(ix, npbc(iy + 1, Nly));
      BB(nc, 1) = IBL(ii, nx);
      int ny = INVN(jx, npbc(jy + 1, Nly));
      BB(nc, 2) = IBL(nx, ny);
      BB(nc, 3) = IBL(jj, ny);
      nx = INVN(ix, npbc(iy - 1, Nly));
      BB(nc, 4) = IBL(nx, ii);
      ny = INVN(jx, npbc(jy - 1, Nly));
      BB(nc, 5) = IBL(nx, ny);
      BB(nc, 6) = IBL(ny, jj);
    } else if (npbc(iy + 1, Nly) == jy) {
      int nx = INVN(npbc(jx - 1, Nlx), jy);
      BB(nc, 1) = IBL(nx, jj);
      int ny = INVN(npbc(ix - 1, Nlx), iy);
      BB(nc, 2) = IBL(ny, nx);
      BB(nc, 3) = IBL(ny, ii);
      nx = INVN(npbc(jx + 1, Nlx), jy);
      BB(nc, 4) = IBL(jj, nx);
      ny = INVN(npbc(ix + 1, Nlx), iy);
      BB(nc, 5) = IBL(ny, nx);
      BB(nc, 6) = IBL(ii, ny);
    }
  }
  /* imj, src/lattice.f90:144-155 */
  m->imj = xcalloc((size_t)N * N, sizeof(int));
  for (int jj = 1; jj <= N; ++jj)
    for (int ii = 1; ii <= N; ++ii) {
      int imjx = npbc(NL(ii, 1) - NL(jj, 1), Nlx);
      int imjy = npbc(NL(ii, 2) - NL(jj, 2), Nly);
      m->imj[(ii - 1) + (jj - 1) * N] = INVN(imjx, imjy);
    }
  /* L_bonds, src/lattice.f90:184-195 */
  m->L_bonds = xcalloc((size_t)N * 5, sizeof(int));
  for (int iy = 1; iy <= Nly; ++iy)
    for (int ix = 1; ix <= Nlx; ++ix) {
      int ii = INVN(ix, iy);
      m->L_bonds[(ii - 1) * 5 + 0] = ii;
      m->L_bonds[(ii - 1) * 5 + 1] = INVN(npbc(ix + 1, Nlx), iy);
      m->L_bonds[(ii - 1) * 5 + 2] = INVN(ix, npbc(iy + 1, Nly));
      m->L_bonds[(ii - 1) * 5 + 3] = INVN(npbc(ix - 1, Nlx), iy);
      m->L_bonds[(ii - 1) * 5 + 4] = INVN(ix, npbc(iy - 1, Nly));
    }
#undef BB
#undef IBL
#undef NL
#undef INVN
}

z2o_model *z2o_model_new(const z2o_params *p) {
  z2o_model *m = xcalloc(1, sizeof(z2o_model));
  m->p = *p;
  m->N = p->Nlx * p->Nly; /* src/calc_basic.f90:91-102 */
  m->N2 = 2 * m->N;
  m->T = p->Ltrot;
  m->NT = p->NlxTherm * p->NlyTherm;
  m->N2T = 2 * m->NT;
  m->Dtau = p->Beta / (double)p->Ltrot;
  if (p->Nwrap <= 0 || p->Ltrot % p->Nwrap != 0) {
    fprintf(stderr, "z2_oracle: Ltrot must be a multiple of Nwrap\n");
    free(m);
    return NULL;
  }
  m->Nst = p->Ltrot / p->Nwrap;
  lattice_make(m);
  m->alpha = 0.5 * m->Dtau * p->RT;
  m->alpha_2 = m->Dtau * p->RT;
  m->expMU_P = exp(m->Dtau * p->mu);
  m->expMU_M = exp(-m->Dtau * p->mu);
  return m;
}

void z2o_model_free(z2o_model *m) {
  if (!m) return;
  free(m->n_list);
  free(m->inv_n_list);
  free(m->bond_list);
  free(m->inv_bond_list);
  for (int g = 0; g < 4; ++g) free(m->group[g]);
  free(m->bond_bonds);
  free(m->imj);
  free(m->L_bonds);
  free(m);
}

int32_t z2o_model_ndim(const z2o_model *m) { return m->N; }

void z2o_mu_factors(const z2o_model *m, double *ep, double *em) {
  *ep = m->expMU_P;
  *em = m->expMU_M;
}

void z2o_lattice_tables(const z2o_model *m, int32_t *bond_list, int32_t *groups,
                        int32_t *group_len, int32_t *bond_bonds, int32_t *imj,
                        int32_t *L_bonds, int32_t *n_list) {
  const int N = m->N;
  if (bond_list)
    for (int k = 0; k < 4 * N; ++k) bond_list[k] = m->bond_list[k];
  if (groups) {
    int off = 0;
    for (int g = 0; g < 4; ++g)
      for (int k = 0; k < m->glen[g]; ++k) groups[off++] = m->group[g][k];
  }
  if (group_len)
    for (int g = 0; g < 4; ++g) group_len[g] = m->glen[g];
  if (bond_bonds)
    for (int k = 0; k < 14 * N; ++k) bond_bonds[k] = m->bond_bonds[k];
  if (imj)
    for (int k = 0; k < N * N; ++k) imj[k] = m->imj[k];
  if (L_bonds)
    for (int k = 0; k < 5 * N; ++k) L_bonds[k] = m->L_bonds[k];
  if (n_list)
    for (int k = 0; k < 2 * N; ++k) n_list[k] = m->n_list[k];
}

/* ---- kinetic checkerboard: src/operator_kin.f90:61-179 ---- */
#define SIG(sig, ii, nt, N2) (sig)[((size_t)(nt)-1) * (N2) + (ii)-1]

/* rows P1,P2 of Mat <- [C S; S C] rows, src/operator_kin.f90:88-120 */
static void process_group_R(const z2o_model *m, double *Mat, int g, const double *sigma, int nt,
                            int nflag) {
  const int N = m->N;
  for (int k = 0; k < m->glen[g]; ++k) {
    int no = m->group[g][k];
    double vec = SIG(sigma, no, nt, m->N2);
    double arg = ((double)nflag * m->alpha) * vec;
    double C = cosh(arg), S = sinh(arg);
    int p1 = m->bond_list[(no - 1) * 2 + 0] - 1, p2 = m->bond_list[(no - 1) * 2 + 1] - 1;
    for (int jj = 0; jj < N; ++jj) {
      double a = Mat[p1 + (size_t)jj * N], b = Mat[p2 + (size_t)jj * N];
      double v1 = C * a + S * b;
      double v2 = S * a + C * b;
      Mat[p1 + (size_t)jj * N] = v1;
      Mat[p2 + (size_t)jj * N] = v2;
    }
  }
}

/* columns P1,P2 of Mat <- columns [C S; S C], src/operator_kin.f90:147-179 */
static void process_group_L(const z2o_model *m, double *Mat, int g, const double *sigma, int nt,
                            int nflag) {
  const int N = m->N;
  for (int k = 0; k < m->glen[g]; ++k) {
    int no = m->group[g][k];
    double vec = SIG(sigma, no, nt, m->N2);
    double arg = ((double)nflag * m->alpha) * vec;
    double C = cosh(arg), S = sinh(arg);
    int p1 = m->bond_list[(no - 1) * 2 + 0] - 1, p2 = m->bond_list[(no - 1) * 2 + 1] - 1;
    double *c1 = Mat + (size_t)p1 * N, *c2 = Mat + (size_t)p2 * N;
    for (int jj = 0; jj < N; ++jj) {
      double a = c1[jj], b = c2[jj];
      double u1 = a * C + b * S;
      double u2 = a * S + b * C;
      c1[jj] = u1;
      c2[jj] = u2;
    }
  }
}

static const int kin_order[8] = {0, 1, 2, 3, 3, 2, 1, 0};

void z2o_kin_mmult_R(const z2o_model *m, double *Mat, const double *sigma, int nt, int nflag) {
  for (int s = 0; s < 8; ++s) process_group_R(m, Mat, kin_order[s], sigma, nt, nflag);
}

void z2o_kin_mmult_L(const z2o_model *m, double *Mat, const double *sigma, int nt, int nflag) {
  for (int s = 0; s < 8; ++s) process_group_L(m, Mat, kin_order[s], sigma, nt, nflag);
}

/* src/operator_kin.f90:40-59 */
void z2o_kin_delta(const z2o_model *m, double s_old, double s_new, double D[4]) {
  double C_new = cosh(m->alpha_2 * s_new);
  double C_old = cosh(m->alpha_2 * s_old);
  double S_new = sinh(m->alpha_2 * s_new);
  double S_old = sinh(m->alpha_2 * s_old);
  double d11 = C_new * C_old - S_new * S_old - 1.0;
  double d12 = S_new * C_old - C_new * S_old;
  /* column-major 2x2: D[0]=D11 D[1]=D21 D[2]=D12 D[3]=D22 */
  D[0] = d11;
  D[1] = d12;
  D[2] = d12;
  D[3] = d11;
}

/* src/fields.f90:40-119.  sigma is the stored field; the proposed value at
 * (ii,nt) is s_new, all neighbours are read from sigma. */
static double bosonratio_local(const z2o_model *m, const double *sigma, double s_new, int ii,
                               int nt) {
  const int N2 = m->N2, T = m->T;
  double alpha = m->Dtau * m->p.J, gamma;
  int use_temporal;
  if (fabs(m->p.h) > ZERO_H) {
    double tanh_arg = tanh(m->Dtau * m->p.h);
    if (tanh_arg < 1.0e-300) tanh_arg = 1.0e-300;
    gamma = -0.5 * log(tanh_arg);
    use_temporal = 1;
  } else {
    gamma = 0.0;
    use_temporal = 0;
  }
  double log_upbound = log(UPBOUND);
  double s_old = SIG(sigma, ii, nt, N2);
  double tmp_new = 0.0, tmp_old = 0.0, tmp;
  const int *bb = m->bond_bonds + (size_t)(ii - 1) * 7;
  tmp = s_new;
  for (int nf = 1; nf <= 3; ++nf) tmp = tmp * SIG(sigma, bb[nf], nt, N2);
  tmp_new = tmp_new + tmp;
  tmp = s_new;
  for (int nf = 4; nf <= 6; ++nf) tmp = tmp * SIG(sigma, bb[nf], nt, N2);
  tmp_new = tmp_new + tmp;
  tmp = s_old;
  for (int nf = 1; nf <= 3; ++nf) tmp = tmp * SIG(sigma, bb[nf], nt, N2);
  tmp_old = tmp_old + tmp;
  tmp = s_old;
  for (int nf = 4; nf <= 6; ++nf) tmp = tmp * SIG(sigma, bb[nf], nt, N2);
  tmp_old = tmp_old + tmp;
  double expo = alpha * (tmp_new - tmp_old);
  if (expo > log_upbound) expo = log_upbound;
  if (expo < -log_upbound) expo = -log_upbound;
  double ratio = exp(expo);
  if (use_temporal) {
    tmp_new = 0.0;
    tmp_old = 0.0;
    for (int nf = -1; nf <= 1; nf += 2) {
      double nb = SIG(sigma, ii, npbc(nt + nf, T), N2);
      tmp_new = tmp_new + s_new * nb;
      tmp_old = tmp_old + s_old * nb;
    }
    expo = gamma * (tmp_new - tmp_old);
    if (expo > log_upbound) expo = log_upbound;
    if (expo < -log_upbound) expo = -log_upbound;
    ratio = ratio * exp(expo);
  }
  return ratio;
}

double z2o_bosonratio(const z2o_model *m, const double *sigma, const double *sigma_new, int ii,
                      int nt) {
  return bosonratio_local(m, sigma, SIG(sigma_new, ii, nt, m->N2), ii, nt);
}

/* ======================================================================
 * Dense linear algebra (netlib algorithms restated for real data)
 * ====================================================================== */
/* classic scaled 2-norm (dnrm2 / dznrm2 with zero imaginary parts) */
static double dnrm2(int n, const double *x, int incx) {
  if (n < 1) return 0.0;
  if (n == 1) return fabs(x[0]);
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < n; ++i) {
    double v = x[(size_t)i * incx];
    if (v != 0.0) {
      double a = fabs(v);
      if (scale < a) {
        double r = scale / a;
        ssq = 1.0 + ssq * r * r;
        scale = a;
      } else {
        double r = a / scale;
        ssq = ssq + r * r;
      }
    }
  }
  return scale * sqrt(ssq);
}

/* dlapy2 == dlapy3(x,0,z) for real data */
static double dlapy2(double x, double y) {
  double xa = fabs(x), ya = fabs(y);
  double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
  if (z == 0.0) return w;
  double q = z / w;
  return w * sqrt(1.0 + q * q);
}

/* dlarfg / zlarfg for real data: H*(alpha;x) = (beta;0) */
static void dlarfg(int n, double *alpha, double *x, int incx, double *tau) {
  if (n <= 1) {
    *tau = 0.0;
    return;
  }
  double xnorm = dnrm2(n - 1, x, incx);
  if (xnorm == 0.0) {
    *tau = 0.0;
    return;
  }
  double beta = -copysign(dlapy2(*alpha, xnorm), *alpha);
  const double safmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;
  const double rsafmn = 1.0 / safmin;
  int knt = 0;
  if (fabs(beta) < safmin) {
    do {
      knt++;
      for (int i = 0; i < n - 1; ++i) x[(size_t)i * incx] *= rsafmn;
      beta *= rsafmn;
      *alpha *= rsafmn;
    } while (fabs(beta) < safmin && knt < 20);
    xnorm = dnrm2(n - 1, x, incx);
    beta = -copysign(dlapy2(*alpha, xnorm), *alpha);
  }
  *tau = (beta - *alpha) / beta;
  double sc = 1.0 / (*alpha - beta);
  for (int i = 0; i < n - 1; ++i) x[(size_t)i * incx] *= sc;
  for (int j = 0; j < knt; ++j) beta *= safmin;
  *alpha = beta;
}

/* C(m x n) <- (I - tau v v^T) C   (zlarf 'Left': zgemv 'C' + zgerc) */
static void dlarf_left(int m, int n, const double *v, double tau, double *C, int ldc,
                       double *work) {
  if (tau == 0.0) return;
  for (int j = 0; j < n; ++j) {
    double t = 0.0;
    const double *c = C + (size_t)j * ldc;
    for (int i = 0; i < m; ++i) t = t + c[i] * v[i];
    work[j] = t;
  }
  for (int j = 0; j < n; ++j) {
    double t = -tau * work[j];
    if (work[j] != 0.0) {
      double *c = C + (size_t)j * ldc;
      for (int i = 0; i < m; ++i) c[i] = c[i] + v[i] * t;
    }
  }
}

/* C(m x n) <- C (I - tau v v^T)   (zlarf 'Right': zgemv 'N' + zgerc) */
static void dlarf_right(int m, int n, const double *v, double tau, double *C, int ldc,
                        double *work) {
  if (tau == 0.0) return;
  for (int i = 0; i < m; ++i) work[i] = 0.0;
  for (int j = 0; j < n; ++j) {
    double t = v[j];
    if (t != 0.0) {
      const double *c = C + (size_t)j * ldc;
      for (int i = 0; i < m; ++i) work[i] = work[i] + t * c[i];
    }
  }
  for (int j = 0; j < n; ++j) {
    if (v[j] != 0.0) {
      double t = -tau * v[j];
      double *c = C + (size_t)j * ldc;
      for (int i = 0; i < m; ++i) c[i] = c[i] + work[i] * t;
    }
  }
}

/* ZGEQP3 with all columns free (unblocked ZLAQP2 form, modern norm
 * downdating with tol3z = sqrt(eps)).  jpvt is 1-based on output. */
static void geqp3(int n, double *A, int32_t *jpvt, double *tau, double *work) {
  double *vn1 = work, *vn2 = work + n, *w = work + 2 * n;
  const double tol3z = sqrt(1.1102230246251565e-16);
  for (int j = 0; j < n; ++j) {
    jpvt[j] = j + 1;
    vn1[j] = dnrm2(n, A + (size_t)j * n, 1);
    vn2[j] = vn1[j];
  }
  for (int i = 0; i < n; ++i) {
    /* pivot: first index of max vn1[i..n) (idamax) */
    int pvt = i;
    double best = fabs(vn1[i]);
    for (int j = i + 1; j < n; ++j)
      if (fabs(vn1[j]) > best) {
        best = fabs(vn1[j]);
        pvt = j;
      }
    if (pvt != i) {
      double *a = A + (size_t)pvt * n, *b = A + (size_t)i * n;
      for (int k = 0; k < n; ++k) {
        double t = a[k];
        a[k] = b[k];
        b[k] = t;
      }
      int32_t it = jpvt[pvt];
      jpvt[pvt] = jpvt[i];
      jpvt[i] = it;
      vn1[pvt] = vn1[i];
      vn2[pvt] = vn2[i];
    }
    double *aii = A + i + (size_t)i * n;
    if (i < n - 1) dlarfg(n - i, aii, aii + 1, 1, &tau[i]);
    else dlarfg(1, aii, aii, 1, &tau[i]);
    if (i < n - 1) {
      double save = *aii;
      *aii = 1.0;
      dlarf_left(n - i, n - i - 1, aii, tau[i], A + i + (size_t)(i + 1) * n, n, w);
      *aii = save;
    }
    for (int j = i + 1; j < n; ++j) {
      if (vn1[j] != 0.0) {
        double r = fabs(A[i + (size_t)j * n]) / vn1[j];
        double temp = 1.0 - r * r;
        if (temp < 0.0) temp = 0.0;
        double q = vn1[j] / vn2[j];
        double temp2 = temp * q * q;
        if (temp2 <= tol3z) {
          if (i < n - 1) {
            vn1[j] = dnrm2(n - i - 1, A + (i + 1) + (size_t)j * n, 1);
            vn2[j] = vn1[j];
          } else {
            vn1[j] = 0.0;
            vn2[j] = 0.0;
          }
        } else {
          vn1[j] = vn1[j] * sqrt(temp);
        }
      }
    }
  }
}

/* ZUNG2R: form Q (n x n) from the reflectors in A */
static void orgqr(int n, double *A, const double *tau, double *work) {
  for (int i = n - 1; i >= 0; --i) {
    double *aii = A + i + (size_t)i * n;
    if (i < n - 1) {
      *aii = 1.0;
      dlarf_left(n - i, n - i - 1, aii, tau[i], A + i + (size_t)(i + 1) * n, n, work);
      for (int k = 1; k < n - i; ++k) aii[k] = -tau[i] * aii[k];
    }
    *aii = 1.0 - tau[i];
    for (int l = 0; l < i; ++l) A[l + (size_t)i * n] = 0.0;
  }
}

/* ZUNM2R side='L', trans='C': C <- Q^T C (H(1) applied first) */
static void ormqr_LT(int n, double *A, const double *tau, double *C, double *work) {
  for (int i = 0; i < n; ++i) {
    double *aii = A + i + (size_t)i * n;
    double save = *aii;
    *aii = 1.0;
    dlarf_left(n - i, n, aii, tau[i], C + i, n, work);
    *aii = save;
  }
}

/* ZUNM2R side='R', trans='N': C <- C Q (H(1) applied first) */
static void ormqr_RN(int n, double *A, const double *tau, double *C, double *work) {
  for (int i = 0; i < n; ++i) {
    double *aii = A + i + (size_t)i * n;
    double save = *aii;
    *aii = 1.0;
    dlarf_right(n, n - i, aii, tau[i], C + (size_t)i * n, n, work);
    *aii = save;
  }
}

/* ZTRMM('L','U','N','N'): B <- R B */
static void trmm_LUNN(int n, const double *R, double *B) {
  for (int j = 0; j < n; ++j) {
    double *b = B + (size_t)j * n;
    for (int k = 0; k < n; ++k) {
      if (b[k] != 0.0) {
        double temp = b[k];
        const double *r = R + (size_t)k * n;
        for (int i = 0; i < k; ++i) b[i] = b[i] + temp * r[i];
        temp = temp * r[k];
        b[k] = temp;
      }
    }
  }
}

/* ZTRMM('R','U','C','N'): B <- B R^T */
static void trmm_RUTN(int n, const double *R, double *B) {
  for (int k = 0; k < n; ++k) {
    const double *r = R + (size_t)k * n;
    double *bk = B + (size_t)k * n;
    for (int j = 0; j < k; ++j) {
      if (r[j] != 0.0) {
        double temp = r[j];
        double *bj = B + (size_t)j * n;
        for (int i = 0; i < n; ++i) bj[i] = bj[i] + temp * bk[i];
      }
    }
    double temp = r[k];
    if (temp != 1.0)
      for (int i = 0; i < n; ++i) bk[i] = temp * bk[i];
  }
}

/* ZTRSM('L','U','N','N'): B <- R^{-1} B */
static void trsm_LUNN(int n, const double *R, double *B) {
  for (int j = 0; j < n; ++j) {
    double *b = B + (size_t)j * n;
    for (int k = n - 1; k >= 0; --k) {
      if (b[k] != 0.0) {
        const double *r = R + (size_t)k * n;
        b[k] = b[k] / r[k];
        for (int i = 0; i < k; ++i) b[i] = b[i] - b[k] * r[i];
      }
    }
  }
}

/* ZTRSM('R','U','C','N'): B <- B R^{-T} */
static void trsm_RUTN(int n, const double *R, double *B) {
  for (int k = n - 1; k >= 0; --k) {
    const double *r = R + (size_t)k * n;
    double *bk = B + (size_t)k * n;
    double temp = 1.0 / r[k];
    for (int i = 0; i < n; ++i) bk[i] = temp * bk[i];
    for (int j = 0; j < k; ++j) {
      if (r[j] != 0.0) {
        double t = r[j];
        double *bj = B + (size_t)j * n;
        for (int i = 0; i < n; ++i) bj[i] = bj[i] - t * bk[i];
      }
    }
  }
}

/* ZGEMM with op in {N, T}: C <- op(A) op(B) (beta = 0), square n */
static void gemm(char ta, char tb, int n, const double *A, const double *B, double *C) {
  if (ta == 'N' && tb == 'N') {
    for (int j = 0; j < n; ++j) {
      double *c = C + (size_t)j * n;
      for (int i = 0; i < n; ++i) c[i] = 0.0;
      for (int l = 0; l < n; ++l) {
        double temp = B[l + (size_t)j * n];
        if (temp != 0.0) {
          const double *a = A + (size_t)l * n;
          for (int i = 0; i < n; ++i) c[i] = c[i] + temp * a[i];
        }
      }
    }
  } else if (ta == 'T' && tb == 'T') {
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double temp = 0.0;
        const double *a = A + (size_t)i * n;
        for (int l = 0; l < n; ++l) temp = temp + a[l] * B[j + (size_t)l * n];
        C[i + (size_t)j * n] = temp;
      }
  } else if (ta == 'T' && tb == 'N') {
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double temp = 0.0;
        const double *a = A + (size_t)i * n, *b = B + (size_t)j * n;
        for (int l = 0; l < n; ++l) temp = temp + a[l] * b[l];
        C[i + (size_t)j * n] = temp;
      }
  } else { /* 'N','T' */
    for (int j = 0; j < n; ++j) {
      double *c = C + (size_t)j * n;
      for (int i = 0; i < n; ++i) c[i] = 0.0;
      for (int l = 0; l < n; ++l) {
        double temp = B[j + (size_t)l * n];
        if (temp != 0.0) {
          const double *a = A + (size_t)l * n;
          for (int i = 0; i < n; ++i) c[i] = c[i] + temp * a[i];
        }
      }
    }
  }
}

static void transpose(int n, const double *A, double *B) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) B[j + (size_t)i * n] = A[i + (size_t)j * n];
}

/* ZLAPMR(forward): new row i = old row k(i);  backward: new row k(i) = old row i */
static void lapmr(int forward, int n, double *X, const int32_t *k, double *tmp) {
  memcpy(tmp, X, sizeof(double) * (size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      if (forward) X[i + (size_t)j * n] = tmp[(k[i] - 1) + (size_t)j * n];
      else X[(k[i] - 1) + (size_t)j * n] = tmp[i + (size_t)j * n];
    }
}

/* ZLAPMT(forward): new col i = old col k(i);  backward: new col k(i) = old col i */
static void lapmt(int forward, int n, double *X, const int32_t *k, double *tmp) {
  memcpy(tmp, X, sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n; ++i) {
    if (forward) memcpy(X + (size_t)i * n, tmp + (size_t)(k[i] - 1) * n, sizeof(double) * n);
    else memcpy(X + (size_t)(k[i] - 1) * n, tmp + (size_t)i * n, sizeof(double) * n);
  }
}

/* QDRP_decompose, src/stabilization.f90:42-68 */
void z2o_qdrp(int n, double *A, double *D, int32_t *ipvt, double *tau) {
  double *work = xcalloc((size_t)3 * n + 8, sizeof(double));
  geqp3(n, A, ipvt, tau, work);
  for (int i = 0; i < n; ++i) {
    double X = fabs(A[i + (size_t)i * n]);
    D[i] = X;
    for (int j = i; j < n; ++j) A[i + (size_t)j * n] = A[i + (size_t)j * n] / X;
  }
  free(work);
}

/* stab_green, src/stabilization.f90:122-200 */
void z2o_stab_green(int n, int Ltrot, int nt, const double *UUR, const double *DUR,
                    const double *VUR, const double *UUL, const double *DUL,
                    const double *VUL, double *G) {
  size_t nn = (size_t)n * n;
  double *VRVL = xcalloc(nn, sizeof(double)), *inv = xcalloc(nn, sizeof(double));
  double *temp = xcalloc(nn, sizeof(double)), *mat = xcalloc(nn, sizeof(double));
  double *DUP = xcalloc((size_t)n, sizeof(double)), *tau = xcalloc((size_t)n, sizeof(double));
  double *work = xcalloc((size_t)3 * n + 8, sizeof(double));
  int32_t *ipvt = xcalloc((size_t)n, sizeof(int32_t));
  gemm('N', 'N', n, VUR, VUL, VRVL);
  gemm('T', 'T', n, UUR, UUL, inv);
  /* temp(:,nr) = VRVL(:,nr)*DUL(nr); VRVL(nl,:) = DUR(nl)*temp(nl,:) (zaxpy on zero) */
  for (int nr = 0; nr < n; ++nr)
    for (int i = 0; i < n; ++i) temp[i + (size_t)nr * n] = 0.0 + DUL[nr] * VRVL[i + (size_t)nr * n];
  for (int nl = 0; nl < n; ++nl)
    for (int j = 0; j < n; ++j) VRVL[nl + (size_t)j * n] = 0.0 + DUR[nl] * temp[nl + (size_t)j * n];
  for (size_t k = 0; k < nn; ++k) mat[k] = inv[k] + VRVL[k];
  if (nt > Ltrot / 2) {
    transpose(n, mat, temp);
    memcpy(mat, temp, nn * sizeof(double));
  }
  z2o_qdrp(n, mat, DUP, ipvt, tau);
  if (nt < Ltrot / 2 + 1) {
    transpose(n, UUR, temp);
    ormqr_LT(n, mat, tau, temp, work);
    for (int nl = 0; nl < n; ++nl)
      for (int j = 0; j < n; ++j) temp[nl + (size_t)j * n] = temp[nl + (size_t)j * n] / DUP[nl];
    trsm_LUNN(n, mat, temp);
    lapmr(0, n, temp, ipvt, inv);
    gemm('T', 'N', n, UUL, temp, G);
  } else {
    transpose(n, UUL, temp);
    ormqr_RN(n, mat, tau, temp, work);
    for (int nr = 0; nr < n; ++nr)
      for (int i = 0; i < n; ++i) temp[i + (size_t)nr * n] = temp[i + (size_t)nr * n] / DUP[nr];
    trsm_RUTN(n, mat, temp);
    lapmt(0, n, temp, ipvt, inv);
    gemm('N', 'T', n, temp, UUR, G);
  }
  free(VRVL);
  free(inv);
  free(temp);
  free(mat);
  free(DUP);
  free(tau);
  free(work);
  free(ipvt);
}

/* LINPACK ZGEFA + ZGEDI(job=1) restated: Ainv <- A^{-1}
 * (lib/package_of_matrix/Modules/mat_mod.f90:296-336) */
static void linpack_inv(int n, const double *A, double *Ainv, double *work, int32_t *ipvt) {
  memcpy(Ainv, A, sizeof(double) * (size_t)n * n);
  double *a = Ainv;
#define AA(i, j) a[(i) + (size_t)(j) * n]
  for (int k = 0; k < n - 1; ++k) {
    int l = k;
    double best = fabs(AA(k, k));
    for (int i = k + 1; i < n; ++i)
      if (fabs(AA(i, k)) > best) {
        best = fabs(AA(i, k));
        l = i;
      }
    ipvt[k] = l;
    if (AA(l, k) == 0.0) continue;
    if (l != k) {
      double t = AA(l, k);
      AA(l, k) = AA(k, k);
      AA(k, k) = t;
    }
    double t = -(1.0 / AA(k, k));
    for (int i = k + 1; i < n; ++i) AA(i, k) = t * AA(i, k);
    for (int j = k + 1; j < n; ++j) {
      t = AA(l, j);
      if (l != k) {
        AA(l, j) = AA(k, j);
        AA(k, j) = t;
      }
      for (int i = k + 1; i < n; ++i) AA(i, j) = AA(i, j) + t * AA(i, k);
    }
  }
  ipvt[n - 1] = n - 1;
  for (int k = 0; k < n; ++k) {
    AA(k, k) = 1.0 / AA(k, k);
    double t = -AA(k, k);
    for (int i = 0; i < k; ++i) AA(i, k) = t * AA(i, k);
    for (int j = k + 1; j < n; ++j) {
      t = AA(k, j);
      AA(k, j) = 0.0;
      for (int i = 0; i <= k; ++i) AA(i, j) = AA(i, j) + t * AA(i, k);
    }
  }
  for (int k = n - 2; k >= 0; --k) {
    for (int i = k + 1; i < n; ++i) {
      work[i] = AA(i, k);
      AA(i, k) = 0.0;
    }
    for (int j = k + 1; j < n; ++j) {
      double t = work[j];
      for (int i = 0; i < n; ++i) AA(i, k) = AA(i, k) + t * AA(i, j);
    }
    int l = ipvt[k];
    if (l != k)
      for (int i = 0; i < n; ++i) {
        double t = AA(i, k);
        AA(i, k) = AA(i, l);
        AA(i, l) = t;
      }
  }
#undef AA
}

/* UDV_C (lib/package_of_matrix/Modules/mat_mod.f90:604-711) restated with an
 * unpivoted Householder QR: A = U D V, V unit-diagonal-magnitude upper
 * triangular, sign fix so that det(R) > 0 as in the reference. */
static void udv_plain(int n, const double *A, double *U, double *D, double *V, double *work) {
  size_t nn = (size_t)n * n;
  double *tmp = work, *tau = work + nn, *w = work + nn + n;
  memcpy(tmp, A, nn * sizeof(double));
  for (int i = 0; i < n; ++i) {
    double *aii = tmp + i + (size_t)i * n;
    if (i < n - 1) dlarfg(n - i, aii, aii + 1, 1, &tau[i]);
    else tau[i] = 0.0;
    if (i < n - 1) {
      double save = *aii;
      *aii = 1.0;
      dlarf_left(n - i, n - i - 1, aii, tau[i], tmp + i + (size_t)(i + 1) * n, n, w);
      *aii = save;
    }
  }
  for (size_t k = 0; k < nn; ++k) V[k] = 0.0;
  double detv = 1.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) V[i + (size_t)j * n] = tmp[i + (size_t)j * n];
  for (int i = 0; i < n; ++i) detv = detv * tmp[i + (size_t)i * n];
  orgqr(n, tmp, tau, w);
  memcpy(U, tmp, nn * sizeof(double));
  if (detv < 0.0) {
    for (int i = 0; i < n; ++i) U[i] = -U[i];
    for (int j = 0; j < n; ++j) V[(size_t)j * n] = -V[(size_t)j * n];
  }
  for (int i = 0; i < n; ++i) D[i] = fabs(V[i + (size_t)i * n]);
  for (int i = 0; i < n; ++i) {
    double z = 1.0 / D[i];
    for (int j = i; j < n; ++j) V[i + (size_t)j * n] = V[i + (size_t)j * n] * z;
  }
}

/* ======================================================================
 * Chain state
 * ====================================================================== */
typedef struct {
  double *UUR, *VUR, *UUL, *VUL, *DUR, *DUL, *Gr;
  double Xmaxm, Xmeanm;
} prop_t; /* src/process_matrix.f90:6-15 */

typedef struct {
  double *URlist, *VRlist, *ULlist, *VLlist, *DRlist, *DLlist;
} wrlist_t; /* src/process_matrix.f90:26-33 */

typedef struct {
  double *Gr00, *Grtt, *Grt0, *Gr0t;
  double Xmaxm[3], Xmeanm[3];
} propgr_t; /* src/process_matrix.f90:17-24 */

typedef struct {
  double NC, ACC, acc;
} acc_t; /* src/process_matrix.f90:35-43 */

typedef struct {
  double flux_sum, flux_sq, flux_avg, chi_flux, pair_sc_acc, pair_sc_q0;
  double cdw_pi_acc, cdw_shift_acc, cdw_pi, cdw_shift, cdw_ratio, density;
} obseq_t;

struct z2o_chain {
  const z2o_model *m;
  int N, N2, T, Nst;
  double *sigma, *sigma_new, *lambda;
  double *gsig_new; /* GlobalUpdate_mod private sigma_new */
  int32_t seed;
  int is_beta;
  prop_t prop, gprop, dprop;
  wrlist_t wr, gwr;
  propgr_t pg, pg_tmp;
  acc_t Kl, Kt, Kg;
  double wolff_size_cluster;
  double n_local_acc, n_local_prop, n_glob_acc, n_glob_prop, n_therm_acc, n_therm_prop,
      cluster_sum;
  obseq_t oe;
  double *otau; /* 7 x N x T */
  int have_dyn;
  /* scratch */
  double *Gtmp, *work;
  int *tau_mask, *queue;
  double min_gap;
  double gmin_gap; /* smallest |min(total,0) - log(r)| of the global decisions */
  uint32_t *log_buf;
  int64_t log_cap, log_n;
};

static void prop_make(prop_t *p, int N) {
  size_t nn = (size_t)N * N;
  p->UUR = xcalloc(nn, sizeof(double));
  p->VUR = xcalloc(nn, sizeof(double));
  p->UUL = xcalloc(nn, sizeof(double));
  p->VUL = xcalloc(nn, sizeof(double));
  p->Gr = xcalloc(nn, sizeof(double));
  p->DUR = xcalloc((size_t)N, sizeof(double));
  p->DUL = xcalloc((size_t)N, sizeof(double));
  for (int i = 0; i < N; ++i) {
    p->UUR[i + (size_t)i * N] = 1.0;
    p->VUR[i + (size_t)i * N] = 1.0;
    p->UUL[i + (size_t)i * N] = 1.0;
    p->VUL[i + (size_t)i * N] = 1.0;
    p->Gr[i + (size_t)i * N] = 1.0;
    p->DUR[i] = 1.0;
    p->DUL[i] = 1.0;
  }
  p->Xmaxm = 0.0;
  p->Xmeanm = 0.0;
}

static void prop_free(prop_t *p) {
  free(p->UUR);
  free(p->VUR);
  free(p->UUL);
  free(p->VUL);
  free(p->Gr);
  free(p->DUR);
  free(p->DUL);
}

/* Prop_assign, src/process_matrix.f90:58-66 */
static void prop_assign(prop_t *dst, const prop_t *src, int N) {
  size_t nn = (size_t)N * N * sizeof(double);
  memcpy(dst->UUL, src->UUL, nn);
  memcpy(dst->VUL, src->VUL, nn);
  memcpy(dst->DUL, src->DUL, N * sizeof(double));
  memcpy(dst->UUR, src->UUR, nn);
  memcpy(dst->VUR, src->VUR, nn);
  memcpy(dst->DUR, src->DUR, N * sizeof(double));
  memcpy(dst->Gr, src->Gr, nn);
  dst->Xmaxm = src->Xmaxm;
  dst->Xmeanm = src->Xmeanm;
}

static void wr_make(wrlist_t *w, int N, int Nst) {
  size_t sz = (size_t)N * N * (Nst + 1);
  w->URlist = xcalloc(sz, sizeof(double));
  w->VRlist = xcalloc(sz, sizeof(double));
  w->ULlist = xcalloc(sz, sizeof(double));
  w->VLlist = xcalloc(sz, sizeof(double));
  w->DRlist = xcalloc((size_t)N * (Nst + 1), sizeof(double));
  w->DLlist = xcalloc((size_t)N * (Nst + 1), sizeof(double));
}

static void wr_free(wrlist_t *w) {
  free(w->URlist);
  free(w->VRlist);
  free(w->ULlist);
  free(w->VLlist);
  free(w->DRlist);
  free(w->DLlist);
}

static void wr_assign(wrlist_t *dst, const wrlist_t *src, int N, int Nst) {
  size_t sz = (size_t)N * N * (Nst + 1) * sizeof(double);
  memcpy(dst->URlist, src->URlist, sz);
  memcpy(dst->VRlist, src->VRlist, sz);
  memcpy(dst->ULlist, src->ULlist, sz);
  memcpy(dst->VLlist, src->VLlist, sz);
  memcpy(dst->DRlist, src->DRlist, (size_t)N * (Nst + 1) * sizeof(double));
  memcpy(dst->DLlist, src->DLlist, (size_t)N * (Nst + 1) * sizeof(double));
}

static void propgr_make(propgr_t *g, int N) {
  size_t nn = (size_t)N * N;
  g->Gr00 = xcalloc(nn, sizeof(double));
  g->Grtt = xcalloc(nn, sizeof(double));
  g->Grt0 = xcalloc(nn, sizeof(double));
  g->Gr0t = xcalloc(nn, sizeof(double));
  for (int i = 0; i < 3; ++i) g->Xmaxm[i] = g->Xmeanm[i] = 0.0;
}

static void propgr_free(propgr_t *g) {
  free(g->Gr00);
  free(g->Grtt);
  free(g->Grt0);
  free(g->Gr0t);
}

/* ---- conf_set (src/fields.f90:311-427) at the therm size ---- */
static void lambda_enforce_sector(double *lam, int n, double Q) {
  double prod = 1.0;
  for (int i = 0; i < n; ++i) {
    lam[i] = lam[i] >= 0.0 ? 1.0 : -1.0;
    prod = prod * lam[i];
  }
  double target = Q >= 0.0 ? 1.0 : -1.0;
  if (prod * target < 0.0) lam[n - 1] = -lam[n - 1];
}

static void lambda_random_fill(double *lam, int n, int32_t *seed, double Q) {
  for (int i = 0; i < n; ++i) {
    double X = z2o_ranf(seed);
    lam[i] = X >= 0.5 ? 1.0 : -1.0;
  }
  lambda_enforce_sector(lam, n, Q);
}

static void conf_set(const z2o_model *m, double *sig, double *lam, int32_t *itmp) {
  const z2o_params *p = &m->p;
  const int N2T = m->N2T, LT = p->LtrotTherm;
  if (p->absolute == 1) {
    for (int nt = 1; nt <= LT; ++nt)
      for (int ii = 1; ii <= N2T; ++ii) {
        double X = z2o_ranf(itmp);
        SIG(sig, ii, nt, N2T) = X >= 0.5 ? 1.0 : -1.0;
      }
  } else if (p->absolute == 2) {
    const double PI = acos(-1.0);
    for (int nt = 1; nt <= LT; ++nt)
      for (int ii = 1; ii <= N2T; ++ii) {
        double X1 = z2o_ranf(itmp), X2 = z2o_ranf(itmp);
        double X = sqrt(-2.0 * log(X1)) * cos(2.0 * PI * X2);
        SIG(sig, ii, nt, N2T) = X >= 0.0 ? 1.0 : -1.0;
      }
  } else if (p->absolute == 3) {
    for (int k = 0; k < N2T * LT; ++k) sig[k] = 1.0;
  } else if (p->absolute == 4) {
    const int Lx = p->NlxTherm;
    for (int nt = 1; nt <= LT; ++nt)
      for (int nc = 1; nc <= N2T; ++nc) {
        int ii = (nc - 1) / 2 + 1; /* bond_list_therm(nc,1) */
        int nx = (ii - 1) % Lx + 1, ny = (ii - 1) / Lx + 1;
        SIG(sig, nc, nt, N2T) = ((nx + ny) % 2 == 0) ? 1.0 : -1.0;
      }
  } else if (p->absolute == 5) {
    for (int ii = 1; ii <= N2T; ++ii) {
      double X = z2o_ranf(itmp);
      double v = X >= 0.5 ? 1.0 : -1.0;
      for (int nt = 1; nt <= LT; ++nt) SIG(sig, ii, nt, N2T) = v;
    }
  } else {
    fprintf(stderr, "incorrect absolute input %d\n", p->absolute);
  }
  lambda_random_fill(lam, m->NT, itmp, p->Q);
}

/* conf_transfer, src/fields.f90:429-507 */
static void conf_transfer(const z2o_model *m, double *sigma, double *lambda, const double *sig_t,
                          const double *lam_t) {
  const z2o_params *p = &m->p;
  const int N2 = m->N2, N2T = m->N2T, LxT = p->NlxTherm, LyT = p->NlyTherm;
  for (int nt = 1; nt <= p->LtrotTherm; ++nt)
    for (int b = 1; b <= N2; ++b) {
      int ii = m->bond_list[(b - 1) * 2], jj = m->bond_list[(b - 1) * 2 + 1];
      int nx = m->n_list[(ii - 1) * 2], ny = m->n_list[(ii - 1) * 2 + 1];
      int jx = m->n_list[(jj - 1) * 2], jy = m->n_list[(jj - 1) * 2 + 1];
      int nxt = (nx - 1) % LxT + 1, nyt = (ny - 1) % LyT + 1;
      int bt = 0;
      if (npbc(nx + 1, p->Nlx) == jx) bt = (nyt - 1) * LxT * 2 + (nxt - 1) * 2 + 1;
      else if (npbc(ny + 1, p->Nly) == jy) bt = (nyt - 1) * LxT * 2 + (nxt - 1) * 2 + 2;
      SIG(sigma, b, nt, N2) = SIG(sig_t, bt, nt, N2T);
    }
  for (int nt = p->LtrotTherm + 1; nt <= p->Ltrot; ++nt) {
    int ntt = nt - p->LtrotTherm;
    for (int b = 1; b <= N2; ++b) SIG(sigma, b, nt, N2) = SIG(sigma, b, ntt, N2);
  }
  for (int ii = 1; ii <= m->N; ++ii) {
    int nx = m->n_list[(ii - 1) * 2], ny = m->n_list[(ii - 1) * 2 + 1];
    int nxt = (nx - 1) % LxT + 1, nyt = (ny - 1) % LyT + 1;
    lambda[ii - 1] = lam_t[(nxt - 1) + (nyt - 1) * LxT];
  }
  lambda_enforce_sector(lambda, m->N, p->Q);
}

static z2o_chain *chain_alloc(const z2o_model *m) {
  z2o_chain *c = xcalloc(1, sizeof(z2o_chain));
  c->m = m;
  c->N = m->N;
  c->N2 = m->N2;
  c->T = m->T;
  c->Nst = m->Nst;
  size_t ns = (size_t)c->N2 * c->T;
  c->sigma = xcalloc(ns, sizeof(double));
  c->sigma_new = xcalloc(ns, sizeof(double));
  c->gsig_new = xcalloc(ns, sizeof(double));
  c->lambda = xcalloc((size_t)c->N, sizeof(double));
  prop_make(&c->prop, c->N);
  wr_make(&c->wr, c->N, c->Nst);
  c->is_beta = 1;
  c->Gtmp = xcalloc((size_t)c->N * c->N, sizeof(double));
  c->work = xcalloc((size_t)c->N * c->N + 8 * (size_t)c->N + 64, sizeof(double));
  c->tau_mask = xcalloc((size_t)c->T, sizeof(int));
  c->queue = xcalloc((size_t)c->T, sizeof(int));
  c->min_gap = INFINITY;
  c->gmin_gap = INFINITY;
  if (m->p.is_global) {
    prop_make(&c->gprop, c->N);
    wr_make(&c->gwr, c->N, c->Nst);
  }
  if (m->p.is_tau) {
    prop_make(&c->dprop, c->N);
    propgr_make(&c->pg, c->N);
    propgr_make(&c->pg_tmp, c->N);
    c->otau = xcalloc((size_t)7 * c->N * c->T, sizeof(double));
    c->have_dyn = 1;
  }
  return c;
}

z2o_chain *z2o_chain_new(const z2o_model *m, int32_t seed) {
  z2o_chain *c = chain_alloc(m);
  double *sig_t = xcalloc((size_t)m->N2T * m->p.LtrotTherm, sizeof(double));
  double *lam_t = xcalloc((size_t)m->NT, sizeof(double));
  c->seed = seed;
  conf_set(m, sig_t, lam_t, &c->seed);
  conf_transfer(m, c->sigma, c->lambda, sig_t, lam_t);
  free(sig_t);
  free(lam_t);
  memcpy(c->sigma_new, c->sigma, sizeof(double) * (size_t)c->N2 * c->T);
  return c;
}

z2o_chain *z2o_chain_new_conf(const z2o_model *m, int32_t seed, const double *sigma_therm,
                              const double *lambda_therm) {
  z2o_chain *c = chain_alloc(m);
  double *lam_t = xcalloc((size_t)m->NT, sizeof(double));
  c->seed = seed;
  if (lambda_therm) memcpy(lam_t, lambda_therm, sizeof(double) * m->NT);
  else lambda_random_fill(lam_t, m->NT, &c->seed, m->p.Q); /* src/fields.f90:273-279 */
  conf_transfer(m, c->sigma, c->lambda, sigma_therm, lam_t);
  free(lam_t);
  memcpy(c->sigma_new, c->sigma, sizeof(double) * (size_t)c->N2 * c->T);
  return c;
}

void z2o_chain_free(z2o_chain *c) {
  if (!c) return;
  free(c->sigma);
  free(c->sigma_new);
  free(c->gsig_new);
  free(c->lambda);
  prop_free(&c->prop);
  wr_free(&c->wr);
  if (c->m->p.is_global) {
    prop_free(&c->gprop);
    wr_free(&c->gwr);
  }
  if (c->have_dyn) {
    prop_free(&c->dprop);
    propgr_free(&c->pg);
    propgr_free(&c->pg_tmp);
    free(c->otau);
  }
  free(c->Gtmp);
  free(c->work);
  free(c->tau_mask);
  free(c->queue);
  free(c);
}

/* ======================================================================
 * Propagation (src/multiply.f90, src/non_interact.f90)
 * ====================================================================== */
static void scale_all(double *A, size_t n, double f) {
  for (size_t k = 0; k < n; ++k) A[k] = A[k] * f;
}

/* ======================================================================
 * Stabilization (src/stabilization.f90)
 * ====================================================================== */
/* stab_UR, src/stabilization.f90:70-94 */
static void stab_UR(z2o_chain *c, prop_t *P) {
  const int n = c->N;
  size_t nn = (size_t)n * n;
  double *mat = c->Gtmp, *tmp = c->work, *tau = c->work + nn;
  int32_t ipvt[n];
  memcpy(mat, P->UUR, nn * sizeof(double));
  for (int i = 0; i < n; ++i)
    for (int r = 0; r < n; ++r) mat[r + (size_t)i * n] = mat[r + (size_t)i * n] * P->DUR[i];
  z2o_qdrp(n, mat, P->DUR, ipvt, tau);
  lapmr(1, n, P->VUR, ipvt, tmp);
  trmm_LUNN(n, mat, P->VUR);
  double *w = c->work + nn + n;
  orgqr(n, mat, tau, w);
  memcpy(P->UUR, mat, nn * sizeof(double));
}

/* stab_UL, src/stabilization.f90:96-120 */
static void stab_UL(z2o_chain *c, prop_t *P) {
  const int n = c->N;
  size_t nn = (size_t)n * n;
  double *mat = c->Gtmp, *tmp = c->work, *tau = c->work + nn;
  int32_t ipvt[n];
  transpose(n, P->UUL, mat);
  for (int i = 0; i < n; ++i)
    for (int r = 0; r < n; ++r) mat[r + (size_t)i * n] = mat[r + (size_t)i * n] * P->DUL[i];
  z2o_qdrp(n, mat, P->DUL, ipvt, tau);
  lapmt(1, n, P->VUL, ipvt, tmp);
  trmm_RUTN(n, mat, P->VUL);
  double *w = c->work + nn + n;
  orgqr(n, mat, tau, w);
  transpose(n, mat, P->UUL);
}

static void stab_green_prop(z2o_chain *c, const prop_t *P, int nt, double *G) {
  z2o_stab_green(c->N, c->T, nt, P->UUR, P->DUR, P->VUR, P->UUL, P->DUL, P->VUL, G);
}

/* compare_mat, src/stabilization.f90:365-372 */
static double compare_mat(const double *A, const double *B, size_t n) {
  double d = 0.0;
  for (size_t k = 0; k < n; ++k) {
    double x = fabs(A[k] - B[k]);
    if (x > d) d = x;
  }
  return d;
}

#define LST(lst, k, N) ((lst) + (size_t)(k) * (N) * (N))
#define DST(lst, k, N) ((lst) + (size_t)(k) * (N))

/* Wrap_pre, src/stabilization.f90:374-398 */
static void wrap_pre(z2o_chain *c, prop_t *P, wrlist_t *W, int nt) {
  const int n = c->N;
  size_t nn = (size_t)n * n;
  int st = nt / c->m->p.Nwrap;
  if (nt != 0) stab_UR(c, P);
  memcpy(LST(W->URlist, st, n), P->UUR, nn * sizeof(double));
  memcpy(LST(W->VRlist, st, n), P->VUR, nn * sizeof(double));
  memcpy(DST(W->DRlist, st, n), P->DUR, n * sizeof(double));
  if (nt == c->T) {
    double *G = xcalloc(nn, sizeof(double));
    stab_green_prop(c, P, nt, G);
    memcpy(P->Gr, G, nn * sizeof(double));
    free(G);
  }
}

/* Wrap_L, src/stabilization.f90:400-437 */
static void wrap_L(z2o_chain *c, prop_t *P, wrlist_t *W, int nt, int flag) {
  const int n = c->N;
  size_t nn = (size_t)n * n;
  int st = nt / c->m->p.Nwrap;
  memcpy(P->UUR, LST(W->URlist, st, n), nn * sizeof(double));
  memcpy(P->VUR, LST(W->VRlist, st, n), nn * sizeof(double));
  memcpy(P->DUR, DST(W->DRlist, st, n), n * sizeof(double));
  if (nt == 0) {
    memset(W->URlist, 0, nn * (c->Nst + 1) * sizeof(double));
    memset(W->VRlist, 0, nn * (c->Nst + 1) * sizeof(double));
    memset(W->DRlist, 0, (size_t)n * (c->Nst + 1) * sizeof(double));
  }
  if (nt != c->T) {
    double *G = xcalloc(nn, sizeof(double));
    stab_UL(c, P);
    stab_green_prop(c, P, nt, G);
    double dif = compare_mat(G, P->Gr, nn);
    if (dif > P->Xmaxm) P->Xmaxm = dif;
    if (flag) P->Xmeanm = P->Xmeanm + dif;
    memcpy(P->Gr, G, nn * sizeof(double));
    free(G);
  }
  memcpy(LST(W->ULlist, st, n), P->UUL, nn * sizeof(double));
  memcpy(LST(W->VLlist, st, n), P->VUL, nn * sizeof(double));
  memcpy(DST(W->DLlist, st, n), P->DUL, n * sizeof(double));
}

/* Wrap_R, src/stabilization.f90:439-476 */
static void wrap_R(z2o_chain *c, prop_t *P, wrlist_t *W, int nt, int flag) {
  const int n = c->N;
  size_t nn = (size_t)n * n;
  int st = nt / c->m->p.Nwrap;
  memcpy(P->UUL, LST(W->ULlist, st, n), nn * sizeof(double));
  memcpy(P->VUL, LST(W->VLlist, st, n), nn * sizeof(double));
  memcpy(P->DUL, DST(W->DLlist, st, n), n * sizeof(double));
  if (nt == c->T) {
    memset(W->ULlist, 0, nn * (c->Nst + 1) * sizeof(double));
    memset(W->VLlist, 0, nn * (c->Nst + 1) * sizeof(double));
    memset(W->DLlist, 0, (size_t)n * (c->Nst + 1) * sizeof(double));
  }
  if (nt != 0) {
    double *G = xcalloc(nn, sizeof(double));
    stab_UR(c, P);
    stab_green_prop(c, P, nt, G);
    double dif = compare_mat(G, P->Gr, nn);
    if (dif > P->Xmaxm) P->Xmaxm = dif;
    if (flag) P->Xmeanm = P->Xmeanm + dif;
    memcpy(P->Gr, G, nn * sizeof(double));
    free(G);
  }
  memcpy(LST(W->URlist, st, n), P->UUR, nn * sizeof(double));
  memcpy(LST(W->VRlist, st, n), P->VUR, nn * sizeof(double));
  memcpy(DST(W->DRlist, st, n), P->DUR, n * sizeof(double));
}

/* stab_green_big, src/stabilization.f90:202-264 -> pg_tmp */
static void stab_green_big(z2o_chain *c, const prop_t *P) {
  const int n = c->N, n2 = 2 * n;
  size_t nn = (size_t)n * n, NN = (size_t)n2 * n2;
  double *ULUR = xcalloc(nn, sizeof(double)), *VRVL = xcalloc(nn, sizeof(double));
  double *invULUR = xcalloc(nn, sizeof(double)), *invVRVL = xcalloc(nn, sizeof(double));
  double *big = xcalloc(NN, sizeof(double)), *bigU = xcalloc(NN, sizeof(double));
  double *bigV = xcalloc(NN, sizeof(double)), *bigD = xcalloc((size_t)n2, sizeof(double));
  double *ibU = xcalloc(NN, sizeof(double)), *ibV = xcalloc(NN, sizeof(double));
  double *left = xcalloc(NN, sizeof(double)), *right = xcalloc(NN, sizeof(double));
  double *tot = xcalloc(NN, sizeof(double));
  double *iUR = xcalloc(nn, sizeof(double)), *iUL = xcalloc(nn, sizeof(double));
  double *iVR = xcalloc(nn, sizeof(double)), *iVL = xcalloc(nn, sizeof(double));
  double *work = xcalloc(NN + 4 * (size_t)n2 + 8, sizeof(double));
  int32_t *ipvt = xcalloc((size_t)n2, sizeof(int32_t));
  gemm('N', 'N', n, P->UUL, P->UUR, ULUR);
  gemm('N', 'N', n, P->VUR, P->VUL, VRVL);
  linpack_inv(n, ULUR, invULUR, work, ipvt);
  linpack_inv(n, VRVL, invVRVL, work, ipvt);
#define BG(M, i, j) (M)[(i) + (size_t)(j) * n2]
  for (int nl = 0; nl < n; ++nl)
    for (int nr = 0; nr < n; ++nr) {
      BG(big, nl, nr) = invVRVL[nl + (size_t)nr * n];
      BG(big, nl + n, nr + n) = invULUR[nl + (size_t)nr * n];
    }
  for (int nr = 0; nr < n; ++nr) {
    BG(big, nr, nr + n) = P->DUL[nr];
    BG(big, nr + n, nr) = -P->DUR[nr];
  }
  udv_plain(n2, big, bigU, bigD, bigV, work);
  linpack_inv(n2, bigU, ibU, work, ipvt);
  linpack_inv(n2, bigV, ibV, work, ipvt);
  linpack_inv(n, P->UUR, iUR, work, ipvt);
  linpack_inv(n, P->UUL, iUL, work, ipvt);
  linpack_inv(n, P->VUR, iVR, work, ipvt);
  linpack_inv(n, P->VUL, iVL, work, ipvt);
  memset(big, 0, NN * sizeof(double));
  for (int nl = 0; nl < n; ++nl)
    for (int nr = 0; nr < n; ++nr) {
      BG(big, nl, nr) = iVR[nl + (size_t)nr * n];
      BG(big, nl + n, nr + n) = iUL[nl + (size_t)nr * n];
    }
  gemm('N', 'N', n2, big, ibV, left);
  memset(big, 0, NN * sizeof(double));
  for (int nl = 0; nl < n; ++nl)
    for (int nr = 0; nr < n; ++nr) {
      BG(big, nl, nr) = iVL[nl + (size_t)nr * n];
      BG(big, nl + n, nr + n) = iUR[nl + (size_t)nr * n];
    }
  gemm('N', 'N', n2, ibU, big, right);
  for (int nr = 0; nr < n2; ++nr)
    for (int nl = 0; nl < n2; ++nl) BG(left, nl, nr) = BG(left, nl, nr) / bigD[nr];
  gemm('N', 'N', n2, left, right, tot);
  for (int nl = 0; nl < n; ++nl)
    for (int nr = 0; nr < n; ++nr) {
      c->pg_tmp.Gr00[nl + (size_t)nr * n] = BG(tot, nl, nr);
      c->pg_tmp.Gr0t[nl + (size_t)nr * n] = BG(tot, nl, nr + n);
      c->pg_tmp.Grt0[nl + (size_t)nr * n] = BG(tot, nl + n, nr);
      c->pg_tmp.Grtt[nl + (size_t)nr * n] = BG(tot, nl + n, nr + n);
    }
#undef BG
  free(ULUR);
  free(VRVL);
  free(invULUR);
  free(invVRVL);
  free(big);
  free(bigU);
  free(bigV);
  free(bigD);
  free(ibU);
  free(ibV);
  free(left);
  free(right);
  free(tot);
  free(iUR);
  free(iUL);
  free(iVR);
  free(iVL);
  free(work);
  free(ipvt);
}

/* Wrap_tau, src/stabilization.f90:478-514 */
static void wrap_tau(z2o_chain *c, int nt) {
  const int n = c->N;
  size_t nn = (size_t)n * n;
  int st = nt / c->m->p.Nwrap;
  prop_t *P = &c->dprop;
  memcpy(P->UUL, LST(c->wr.ULlist, st, n), nn * sizeof(double));
  memcpy(P->VUL, LST(c->wr.VLlist, st, n), nn * sizeof(double));
  memcpy(P->DUL, DST(c->wr.DLlist, st, n), n * sizeof(double));
  stab_UR(c, P);
  stab_green_big(c, P);
  double dif;
  dif = compare_mat(c->pg_tmp.Gr0t, c->pg.Gr0t, nn);
  if (dif > c->pg.Xmaxm[0]) c->pg.Xmaxm[0] = dif;
  c->pg.Xmeanm[0] += dif;
  dif = compare_mat(c->pg_tmp.Grt0, c->pg.Grt0, nn);
  if (dif > c->pg.Xmaxm[1]) c->pg.Xmaxm[1] = dif;
  c->pg.Xmeanm[1] += dif;
  dif = compare_mat(c->pg_tmp.Grtt, c->pg.Grtt, nn);
  if (dif > c->pg.Xmaxm[2]) c->pg.Xmaxm[2] = dif;
  c->pg.Xmeanm[2] += dif;
  memcpy(c->pg.Gr00, c->pg_tmp.Gr00, nn * sizeof(double));
  memcpy(c->pg.Gr0t, c->pg_tmp.Gr0t, nn * sizeof(double));
  memcpy(c->pg.Grt0, c->pg_tmp.Grt0, nn * sizeof(double));
  memcpy(c->pg.Grtt, c->pg_tmp.Grtt, nn * sizeof(double));
}

/* ======================================================================
 * Local updates (src/localK.f90)
 * ====================================================================== */
static void acc_count(acc_t *a, int toggle) {
  a->NC = a->NC + 1;
  if (toggle) a->ACC = a->ACC + 1;
}
static void acc_ratio(acc_t *a) {
  if (a->NC > 0.0) a->acc = a->acc + a->ACC / a->NC;
}
static void acc_reset(acc_t *a) {
  a->NC = 0.0;
  a->ACC = 0.0;
}

/* rank-2 update of G after an accepted/forced flip, src/localK.f90:127-165 */
static void rank2_update_n(int N, double *work, double *G, const double D[4], double P11,
                           double P12, double P21, double P22, double det, int p1, int p2) {
  double Pi11 = P22 / det, Pi12 = (-P12) / det, Pi21 = (-P21) / det, Pi22 = P11 / det;
  double *U = work, *V = work + 2 * (size_t)N; /* U: N x 2, V: 2 x N (row-major rows) */
  double *temp = work + 4 * (size_t)N;
  for (int no = 0; no < 2; ++no) {
    double d1 = D[no + 0], d2 = D[no + 2]; /* Delta(no,1), Delta(no,2) */
    int pn = no == 0 ? p1 : p2;
    for (int j = 0; j < N; ++j) {
      U[j + (size_t)no * N] = G[j + (size_t)pn * N];
      V[no * (size_t)N + j] = -(d1 * G[p1 + (size_t)j * N]) - d2 * G[p2 + (size_t)j * N];
    }
    V[no * (size_t)N + p1] = V[no * (size_t)N + p1] + d1;
    V[no * (size_t)N + p2] = V[no * (size_t)N + p2] + d2;
  }
  for (int j = 0; j < N; ++j) {
    temp[j] = U[j] * Pi11 + U[j + N] * Pi21;
    temp[j + N] = U[j] * Pi12 + U[j + N] * Pi22;
  }
  for (int l = 0; l < N; ++l) {
    double v1 = V[l], v2 = V[N + l];
    double *g = G + (size_t)l * N;
    for (int j = 0; j < N; ++j) {
      double diff = temp[j] * v1 + temp[j + N] * v2;
      g[j] = g[j] - diff;
    }
  }
}

static void rank2_update(z2o_chain *c, double *G, const double D[4], double P11, double P12,
                         double P21, double P22, double det, int p1, int p2) {
  rank2_update_n(c->N, c->work, G, D, P11, P12, P21, P22, det, p1, p2);
}

/* forced flip of bond ii on a caller G (the accept branch of LocalK_metro,
 * src/localK.f90:85-165); returns the 2x2 determinant */
double z2o_flip_update(const z2o_model *m, double *G, int ii, double s_old) {
  const int N = m->N;
  int p1 = m->bond_list[(ii - 1) * 2] - 1, p2 = m->bond_list[(ii - 1) * 2 + 1] - 1;
  double D[4];
  z2o_kin_delta(m, s_old, -s_old, D);
  double A11 = 1.0 - G[p1 + (size_t)p1 * N], A21 = 0.0 - G[p2 + (size_t)p1 * N];
  double A12 = 0.0 - G[p1 + (size_t)p2 * N], A22 = 1.0 - G[p2 + (size_t)p2 * N];
  double P11 = 1.0 + (D[0] * A11 + D[2] * A21);
  double P21 = 0.0 + (D[1] * A11 + D[3] * A21);
  double P12 = 0.0 + (D[0] * A12 + D[2] * A22);
  double P22 = 1.0 + (D[1] * A12 + D[3] * A22);
  double det = P11 * P22 - P12 * P21;
  double *work = xcalloc((size_t)6 * N + 8, sizeof(double));
  rank2_update_n(N, work, G, D, P11, P12, P21, P22, det, p1, p2);
  free(work);
  return det;
}

/* LocalK_metro, src/localK.f90:45-173 */
static void localK_metro(z2o_chain *c, int ii, int nt) {
  const z2o_model *m = c->m;
  const int N = c->N, N2 = c->N2;
  double *G = c->prop.Gr;
  double s_old = SIG(c->sigma, ii, nt, N2);
  SIG(c->sigma_new, ii, nt, N2) = -s_old;
  double s_new = SIG(c->sigma_new, ii, nt, N2);
  int p1 = m->bond_list[(ii - 1) * 2] - 1, p2 = m->bond_list[(ii - 1) * 2 + 1] - 1;
  double D[4];
  z2o_kin_delta(m, s_old, s_new, D);
  /* A = I - G_PP */
  double A11 = 1.0 - G[p1 + (size_t)p1 * N], A21 = 0.0 - G[p2 + (size_t)p1 * N];
  double A12 = 0.0 - G[p1 + (size_t)p2 * N], A22 = 1.0 - G[p2 + (size_t)p2 * N];
  /* matmul(Delta, A) then + I */
  double P11 = 1.0 + (D[0] * A11 + D[2] * A21);
  double P21 = 0.0 + (D[1] * A11 + D[3] * A21);
  double P12 = 0.0 + (D[0] * A12 + D[2] * A22);
  double P22 = 1.0 + (D[1] * A12 + D[3] * A22);
  double det = P11 * P22 - P12 * P21;
  double ratio_fermion = det * det;
  double ratio_boson = bosonratio_local(m, c->sigma, s_new, ii, nt);
  double ratio_re = ratio_fermion * ratio_boson;
  double ratio_abs = fabs(ratio_re);
  double r = z2o_ranf(&c->seed);
  double gap = fabs(ratio_abs - r);
  if (gap < c->min_gap) c->min_gap = gap;
  int accept = ratio_abs > r;
  c->n_local_prop += 1;
  if (c->log_buf && c->log_n < c->log_cap)
    c->log_buf[c->log_n] = (uint32_t)ii | ((uint32_t)nt << 16) | ((uint32_t)accept << 31);
  c->log_n++;
  if (accept) {
    acc_count(&c->Kl, 1);
    c->n_local_acc += 1;
    rank2_update(c, G, D, P11, P12, P21, P22, det, p1, p2);
    SIG(c->sigma, ii, nt, N2) = SIG(c->sigma_new, ii, nt, N2);
  } else {
    acc_count(&c->Kl, 0);
    SIG(c->sigma_new, ii, nt, N2) = SIG(c->sigma, ii, nt, N2);
  }
}

/* LocalK_prop_L, src/localK.f90:347-373 */
static void localK_prop_L(z2o_chain *c, int nt) {
  for (int ii = c->N2; ii >= 1; --ii) localK_metro(c, ii, nt);
  z2o_kin_mmult_L(c->m, c->prop.Gr, c->sigma, nt, 1);
  z2o_kin_mmult_R(c->m, c->prop.Gr, c->sigma, nt, -1);
  z2o_kin_mmult_L(c->m, c->prop.UUL, c->sigma, nt, 1);
}

/* LocalK_prop_R, src/localK.f90:375-398 */
static void localK_prop_R(z2o_chain *c, int nt) {
  z2o_kin_mmult_R(c->m, c->prop.Gr, c->sigma, nt, 1);
  z2o_kin_mmult_L(c->m, c->prop.Gr, c->sigma, nt, -1);
  for (int ii = 1; ii <= c->N2; ++ii) localK_metro(c, ii, nt);
  z2o_kin_mmult_R(c->m, c->prop.UUR, c->sigma, nt, 1);
}

/* LocalK_therm, src/localK.f90:400-420 */
static void localK_therm(z2o_chain *c, int ii, int nt) {
  const int N2 = c->N2;
  SIG(c->sigma_new, ii, nt, N2) = -SIG(c->sigma, ii, nt, N2);
  double rb = bosonratio_local(c->m, c->sigma, SIG(c->sigma_new, ii, nt, N2), ii, nt);
  double r = z2o_ranf(&c->seed);
  c->n_therm_prop += 1;
  if (rb > r) {
    acc_count(&c->Kt, 1);
    c->n_therm_acc += 1;
    SIG(c->sigma, ii, nt, N2) = SIG(c->sigma_new, ii, nt, N2);
  } else {
    acc_count(&c->Kt, 0);
    SIG(c->sigma_new, ii, nt, N2) = SIG(c->sigma, ii, nt, N2);
  }
}

/* propT_L / propT_R on the main propagator, src/multiply.f90:79-108 */
static void propT_L(z2o_chain *c, prop_t *P) {
  size_t nn = (size_t)c->N * c->N;
  const z2o_model *m = c->m;
  scale_all(P->Gr, nn, m->expMU_P);
  scale_all(P->Gr, nn, m->expMU_M);
  scale_all(P->UUL, nn, m->expMU_P);
}
static void propT_R(z2o_chain *c, prop_t *P) {
  size_t nn = (size_t)c->N * c->N;
  const z2o_model *m = c->m;
  scale_all(P->Gr, nn, m->expMU_P);
  scale_all(P->Gr, nn, m->expMU_M);
  scale_all(P->UUR, nn, m->expMU_P);
}

void z2o_slice_L(z2o_chain *c, int nt) {
  propT_L(c, &c->prop);
  localK_prop_L(c, nt);
}
void z2o_slice_R(z2o_chain *c, int nt) {
  localK_prop_R(c, nt);
  propT_R(c, &c->prop);
}

/* ======================================================================
 * Observables
 * ====================================================================== */
static void obs_equal_reset(obseq_t *o) { memset(o, 0, sizeof(*o)); }

/* the per-slice sums of Obs_equal_calc, src/obser_equal.f90:73-149:
 * out = {flux_tau, pair_sum, cdw_sum, cdw_shift_sum, density_slice} */
static void obs_equal_terms(const z2o_chain *c, int nt, double out[5]) {
  const z2o_model *m = c->m;
  const int N = c->N, N2 = c->N2;
  const double *G = c->prop.Gr;
  const double PI = acos(-1.0);
  double qshift_x = PI - 2.0 * PI / (double)m->p.Nlx;
  double flux_tau = 0.0;
  for (int site = 1; site <= N; ++site) {
    int right_site = m->L_bonds[(site - 1) * 5 + 1];
    int up_site = m->L_bonds[(site - 1) * 5 + 2];
    int diag_site = m->L_bonds[(up_site - 1) * 5 + 1];
    int bb = m->inv_bond_list[(site - 1) + (size_t)(right_site - 1) * N];
    int bl = m->inv_bond_list[(site - 1) + (size_t)(up_site - 1) * N];
    int bt = m->inv_bond_list[(up_site - 1) + (size_t)(diag_site - 1) * N];
    int up_of_right = m->L_bonds[(right_site - 1) * 5 + 2];
    int br = m->inv_bond_list[(right_site - 1) + (size_t)(up_of_right - 1) * N];
    flux_tau = flux_tau + SIG(c->sigma, bb, nt, N2) * SIG(c->sigma, br, nt, N2) *
                              SIG(c->sigma, bt, nt, N2) * SIG(c->sigma, bl, nt, N2);
  }
  flux_tau = flux_tau / (double)N;
  double pair_sum = 0.0, cdw_sum = 0.0, cdw_shift_sum = 0.0, density_slice = 0.0;
  for (int ii = 0; ii < N; ++ii) {
    double n_i = 1.0 - G[ii + (size_t)ii * N];
    density_slice = density_slice + n_i + n_i;
    int xi = m->n_list[ii * 2], yi = m->n_list[ii * 2 + 1];
    double stag_i = ((xi + yi) % 2 == 0) ? 1.0 : -1.0;
    for (int jj = 0; jj < N; ++jj) {
      double n_j = 1.0 - G[jj + (size_t)jj * N];
      double delta = ii == jj ? 1.0 : 0.0;
      double gij = G[ii + (size_t)jj * N], gji = G[jj + (size_t)ii * N];
      pair_sum = pair_sum + (delta - gij) * (delta - gij);
      double dc = n_i * n_j - (delta - gji) * gij;
      dc = dc + n_i * n_j - (delta - gji) * gij;
      dc = dc + n_i * n_j + n_i * n_j;
      int xj = m->n_list[jj * 2], yj = m->n_list[jj * 2 + 1];
      double stag_j = ((xj + yj) % 2 == 0) ? 1.0 : -1.0;
      double phase_pi = stag_i * stag_j;
      double dx = (double)(xi - xj), dy = (double)(yi - yj);
      double phase_shift = cos(qshift_x * dx + PI * dy);
      cdw_sum = cdw_sum + phase_pi * dc;
      cdw_shift_sum = cdw_shift_sum + phase_shift * dc;
    }
  }
  out[0] = flux_tau;
  out[1] = pair_sum;
  out[2] = cdw_sum;
  out[3] = cdw_shift_sum;
  out[4] = density_slice;
}

void z2o_obs_equal_terms(const z2o_chain *c, int nt, double out[5]) { obs_equal_terms(c, nt, out); }

/* Obs_equal_calc, src/obser_equal.f90:73-149 */
static void obs_equal_calc(z2o_chain *c, int nt) {
  obseq_t *o = &c->oe;
  double t[5];
  obs_equal_terms(c, nt, t);
  o->flux_sum = o->flux_sum + t[0];
  o->flux_sq = o->flux_sq + t[0] * t[0];
  o->pair_sc_acc = o->pair_sc_acc + t[1];
  o->cdw_pi_acc = o->cdw_pi_acc + t[2];
  o->cdw_shift_acc = o->cdw_shift_acc + t[3];
  o->density = o->density + t[4] / (double)c->N;
}

/* Obs_equal_ave, src/obser_equal.f90:50-71 */
static void obs_equal_ave(z2o_chain *c, int Nobs) {
  obseq_t *o = &c->oe;
  if (Nobs <= 0) return;
  double znorm = 1.0 / (double)Nobs;
  double fm = o->flux_sum * znorm, fsq = o->flux_sq * znorm;
  o->flux_avg = fm;
  double v = fsq - fm * fm;
  o->chi_flux = c->m->p.Beta * (double)c->N * (v > 0.0 ? v : 0.0);
  o->pair_sc_q0 = (o->pair_sc_acc * znorm) / (double)c->N;
  o->cdw_pi = (o->cdw_pi_acc * znorm) / (double)c->N;
  o->cdw_shift = (o->cdw_shift_acc * znorm) / (double)c->N;
  if (o->cdw_pi > 1.0e-12) o->cdw_ratio = 1.0 - o->cdw_shift / o->cdw_pi;
  else o->cdw_ratio = 0.0;
  o->density = o->density * znorm;
}

/* Obs_tau_calc, src/obser_tau.f90:56-133 (real parts; both spins equal) */
static void obs_tau_calc(z2o_chain *c, int nt) {
  const z2o_model *m = c->m;
  const int N = c->N, N2 = c->N2, T = c->T;
  const double *Gtt = c->pg.Grtt, *G00 = c->pg.Gr00, *Gt0 = c->pg.Grt0, *G0t = c->pg.Gr0t;
  double *spin = c->otau, *pair_s = spin + (size_t)N * T, *pair_d = pair_s + (size_t)N * T;
  double *charge_s = pair_d + (size_t)N * T, *charge_d = charge_s + (size_t)N * T;
  double *single = charge_d + (size_t)N * T, *curxx = single + (size_t)N * T;
  double *tt = c->work, *t0 = c->work + N;
  for (int i = 0; i < N; ++i) {
    double x = 1.0 - Gtt[i + (size_t)i * N]; /* ZKRON - transpose(Grtt) diag */
    tt[i] = x + x;
    double y = 1.0 - G00[i + (size_t)i * N];
    t0[i] = y + y;
  }
  const double RT = m->p.RT;
  for (int jj = 0; jj < N; ++jj)
    for (int ii = 0; ii < N; ++ii) {
      size_t idx = (size_t)(m->imj[ii + (size_t)jj * N] - 1) + (size_t)(nt - 1) * N;
      spin[idx] = spin[idx] + SIG(c->sigma, ii + 1, nt, N2) * SIG(c->sigma, jj + 1, 1, N2);
      double g0t = G0t[jj + (size_t)ii * N];
      double diag = g0t * g0t;
      pair_d[idx] = pair_d[idx] + (diag + diag) * 4.0;
      pair_s[idx] = pair_s[idx] + (diag + diag) * 4.0;
      double a = g0t * Gt0[ii + (size_t)jj * N];
      double cd = (tt[ii] - tt[ii]) * (t0[jj] - t0[jj]);
      double cs = (tt[ii] + tt[ii]) * (t0[jj] + t0[jj]);
      charge_d[idx] = charge_d[idx] - a - a - a - a + cd;
      charge_s[idx] = charge_s[idx] - a - a - a - a + cs;
      double term = Gt0[ii + (size_t)jj * N] + Gt0[ii + (size_t)jj * N];
      single[idx] = single[idx] + term + term;
    }
  for (int jj = 0; jj < N; ++jj)
    for (int ii = 0; ii < N; ++ii) {
      size_t idx = (size_t)(m->imj[ii + (size_t)jj * N] - 1) + (size_t)(nt - 1) * N;
      int ip = m->L_bonds[ii * 5 + 1] - 1, jp = m->L_bonds[jj * 5 + 1] - 1;
      double term = -(G0t[jj + (size_t)ip * N] * Gt0[ii + (size_t)jp * N]) -
                    G0t[jp + (size_t)ii * N] * Gt0[ip + (size_t)jj * N] +
                    G0t[jj + (size_t)ii * N] * Gt0[ip + (size_t)jp * N] +
                    G0t[jp + (size_t)ip * N] * Gt0[ii + (size_t)jj * N];
      curxx[idx] = curxx[idx] - RT * RT * term - RT * RT * term;
      curxx[idx] = curxx[idx] - RT * RT * term - RT * RT * term;
    }
}

/* ======================================================================
 * Sweeps (src/local_sweep.f90, src/dynamics.f90)
 * ====================================================================== */
/* Local_sweep_reset, src/local_sweep.f90:56-71 */
static void sweep_reset(z2o_chain *c, int toggle) {
  acc_reset(&c->Kl);
  acc_reset(&c->Kt);
  memcpy(c->sigma_new, c->sigma, sizeof(double) * (size_t)c->N2 * c->T);
  obs_equal_reset(&c->oe);
  if (toggle && c->have_dyn) memset(c->otau, 0, sizeof(double) * 7 * (size_t)c->N * c->T);
  if (fabs(c->m->p.h) <= ZERO_H) {
    for (int ii = 1; ii <= c->N2; ++ii) {
      double sval = SIG(c->sigma, ii, 1, c->N2);
      for (int nt = 1; nt <= c->T; ++nt) SIG(c->sigma, ii, nt, c->N2) = sval;
    }
  }
}

void z2o_therm(z2o_chain *c) {
  sweep_reset(c, 0);
  for (int nt = 1; nt <= c->T; ++nt)
    for (int ii = 1; ii <= c->N2; ++ii) localK_therm(c, ii, nt);
  acc_ratio(&c->Kt);
}

void z2o_pre(z2o_chain *c) {
  const z2o_model *m = c->m;
  sweep_reset(c, 0);
  wrap_pre(c, &c->prop, &c->wr, 0);
  for (int nt = 1; nt <= c->T; ++nt) {
    if (m->p.RT > ZERO_H) z2o_kin_mmult_R(m, c->prop.UUR, c->sigma, nt, 1);
    scale_all(c->prop.UUR, (size_t)c->N * c->N, m->expMU_P);
    if (nt % m->p.Nwrap == 0) wrap_pre(c, &c->prop, &c->wr, nt);
  }
}

static void sweep_L(z2o_chain *c, int *Nobs) {
  const int w = c->m->p.Nwrap;
  for (int nt = c->T; nt >= 1; --nt) {
    if (nt % w == 0) wrap_L(c, &c->prop, &c->wr, nt, 1);
    obs_equal_calc(c, nt);
    (*Nobs)++;
    propT_L(c, &c->prop);
    localK_prop_L(c, nt);
  }
  wrap_L(c, &c->prop, &c->wr, 0, 1);
}

/* the first nslices slices of a left pass (timing sample for the CPU baseline) */
void z2o_sweep_L_partial(z2o_chain *c, int nslices) {
  const int w = c->m->p.Nwrap;
  int Nobs = 0;
  for (int nt = c->T; nt >= 1 && nt > c->T - nslices; --nt) {
    if (nt % w == 0) wrap_L(c, &c->prop, &c->wr, nt, 1);
    obs_equal_calc(c, nt);
    Nobs++;
    propT_L(c, &c->prop);
    localK_prop_L(c, nt);
  }
}

/* Dyn_reset + Dyn_sweep_R, src/dynamics.f90:41-62 */
static void dyn_sweep(z2o_chain *c) {
  const z2o_model *m = c->m;
  const int N = c->N, w = m->p.Nwrap;
  size_t nn = (size_t)N * N;
  prop_assign(&c->dprop, &c->prop, N);
  memcpy(c->pg.Gr00, c->prop.Gr, nn * sizeof(double));
  memcpy(c->pg.Grtt, c->prop.Gr, nn * sizeof(double));
  memcpy(c->pg.Grt0, c->prop.Gr, nn * sizeof(double));
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i)
      c->pg.Gr0t[i + (size_t)j * N] = -((i == j ? 1.0 : 0.0) - c->prop.Gr[i + (size_t)j * N]);
  for (int nt = 1; nt <= c->T; ++nt) {
    obs_tau_calc(c, nt);
    if (m->p.RT > ZERO_H) {
      z2o_kin_mmult_R(m, c->pg.Grt0, c->sigma, nt, 1);
      z2o_kin_mmult_L(m, c->pg.Gr0t, c->sigma, nt, -1);
      z2o_kin_mmult_R(m, c->pg.Grtt, c->sigma, nt, 1);
      z2o_kin_mmult_L(m, c->pg.Grtt, c->sigma, nt, -1);
      z2o_kin_mmult_R(m, c->dprop.UUR, c->sigma, nt, 1);
    }
    scale_all(c->pg.Grt0, nn, m->expMU_P);
    scale_all(c->pg.Gr0t, nn, m->expMU_M);
    scale_all(c->pg.Grtt, nn, m->expMU_P);
    scale_all(c->pg.Grtt, nn, m->expMU_M);
    scale_all(c->dprop.UUR, nn, m->expMU_P);
    if (nt % w == 0) wrap_tau(c, nt);
  }
}

static void sweep_R(z2o_chain *c, int toggle, int *Nobs, int *Nobst) {
  const int w = c->m->p.Nwrap;
  if (toggle && c->have_dyn) {
    dyn_sweep(c);
    (*Nobst)++;
  }
  wrap_R(c, &c->prop, &c->wr, 0, 1);
  for (int nt = 1; nt <= c->T; ++nt) {
    localK_prop_R(c, nt);
    propT_R(c, &c->prop);
    if (nt % w == 0) wrap_R(c, &c->prop, &c->wr, nt, 1);
    obs_equal_calc(c, nt);
    (*Nobs)++;
  }
}

/* single passes and the time-displaced measurement on their own (test hooks) */
void z2o_pass_L(z2o_chain *c) {
  int Nobs = 0;
  sweep_L(c, &Nobs);
}

void z2o_pass_R(z2o_chain *c, int toggle) {
  int Nobs = 0, Nobst = 0;
  sweep_R(c, toggle, &Nobs, &Nobst);
}

void z2o_dyn_measure(z2o_chain *c) {
  if (!c->have_dyn) return;
  memset(c->otau, 0, sizeof(double) * 7 * (size_t)c->N * c->T);
  dyn_sweep(c);
}

void z2o_sweep(z2o_chain *c, int toggle) {
  int Nobs = 0, Nobst = 0;
  sweep_reset(c, toggle);
  for (int nsw = 1; nsw <= c->m->p.Nsweep; ++nsw) {
    if (c->is_beta) {
      sweep_L(c, &Nobs);
      sweep_R(c, toggle, &Nobs, &Nobst);
    } else {
      sweep_R(c, toggle, &Nobs, &Nobst);
      sweep_L(c, &Nobs);
    }
  }
  obs_equal_ave(c, Nobs);
  if (toggle && c->have_dyn && Nobst > 0) {
    double z = 1.0 / (double)Nobst;
    for (size_t k = 0; k < 7 * (size_t)c->N * c->T; ++k) c->otau[k] = z * c->otau[k];
  }
  acc_ratio(&c->Kl);
}

/* ======================================================================
 * Global (Wolff) update: src/global_update.f90, src/globalK.f90:15-151
 * ====================================================================== */
static double ratioK_fermion(z2o_chain *c, double *G, int ii, int nt) {
  const z2o_model *m = c->m;
  const int N = c->N, N2 = c->N2;
  double s_old = SIG(c->sigma, ii, nt, N2), s_new = SIG(c->gsig_new, ii, nt, N2);
  if (s_new == s_old) return 1.0;
  int p1 = m->bond_list[(ii - 1) * 2] - 1, p2 = m->bond_list[(ii - 1) * 2 + 1] - 1;
  double D[4];
  z2o_kin_delta(m, s_old, s_new, D);
  double A11 = 1.0 - G[p1 + (size_t)p1 * N], A21 = 0.0 - G[p2 + (size_t)p1 * N];
  double A12 = 0.0 - G[p1 + (size_t)p2 * N], A22 = 1.0 - G[p2 + (size_t)p2 * N];
  double P11 = 1.0 + (D[0] * A11 + D[2] * A21);
  double P21 = 0.0 + (D[1] * A11 + D[3] * A21);
  double P12 = 0.0 + (D[0] * A12 + D[2] * A22);
  double P22 = 1.0 + (D[1] * A12 + D[3] * A22);
  double det = P11 * P22 - P12 * P21;
  double r = fabs(det * det);
  if (r < RATIO_EPS) r = RATIO_EPS;
  rank2_update(c, G, D, P11, P12, P21, P22, det, p1, p2);
  return r;
}

static double compute_time_bond_prob(const z2o_model *m) {
  double p;
  if (fabs(m->p.h) > ZERO_H) {
    double ta = tanh(m->Dtau * fabs(m->p.h));
    if (ta < RATIO_EPS) ta = RATIO_EPS;
    double gamma = -0.5 * log(ta);
    p = 1.0 - exp(-2.0 * gamma);
  } else {
    p = 0.0;
  }
  if (p > 1.0) p = 1.0;
  if (p < 0.0) p = 0.0;
  return p;
}

/* Wolff_flip, src/global_update.f90:65-181 */
static void wolff_flip(z2o_chain *c, int *size_cluster, double *log_ratio_space) {
  const z2o_model *m = c->m;
  const int T = c->T, N2 = c->N2;
  int bond = z2o_nranf(&c->seed, N2);
  if (bond < 1) bond = 1;
  double p_time = compute_time_bond_prob(m);
  int *mask = c->tau_mask, *queue = c->queue;
  for (int t = 0; t < T; ++t) mask[t] = 0;
  for (int t = 0; t < T; ++t) queue[t] = 0;
  queue[0] = z2o_nranf(&c->seed, T);
  if (queue[0] < 1) queue[0] = 1;
  int head = 0, tail = 0;
  mask[queue[0] - 1] = 1;
  while (head <= tail) {
    int cur = queue[head++];
    for (int off = -1; off <= 1; off += 2) {
      int nb = npbc(cur + off, T);
      if (mask[nb - 1]) continue;
      if (SIG(c->gsig_new, bond, nb, N2) == SIG(c->gsig_new, bond, cur, N2)) {
        if (p_time > 0.0 && z2o_ranf(&c->seed) < p_time) {
          queue[++tail] = nb;
          mask[nb - 1] = 1;
        }
      }
    }
  }
  int cnt = 0;
  for (int t = 0; t < T; ++t) cnt += mask[t];
  *size_cluster = cnt;
  if (cnt == 0) {
    *log_ratio_space = 0.0;
    return;
  }
  double lr = 0.0, alpha = m->Dtau * m->p.J;
  if (alpha != 0.0) {
    for (int t = 1; t <= T; ++t) {
      if (!mask[t - 1]) continue;
      for (int start = 1; start <= 4; start += 3) {
        double prod = SIG(c->gsig_new, bond, t, N2);
        int bad = 0;
        for (int k = 0; k <= 2; ++k) {
          int nbb = m->bond_bonds[(bond - 1) * 7 + start + k];
          if (nbb <= 0) {
            bad = 1;
            break;
          }
          prod = prod * SIG(c->gsig_new, nbb, t, N2);
        }
        lr = lr + (bad ? 0.0 : -2.0 * alpha * prod);
      }
    }
  }
  *log_ratio_space = lr;
  for (int t = 1; t <= T; ++t)
    if (mask[t - 1]) SIG(c->gsig_new, bond, t, N2) = -SIG(c->gsig_new, bond, t, N2);
}

static void globalK_prop_L(z2o_chain *c, double *lrf, int nt) {
  prop_t *P = &c->gprop;
  for (int ii = c->N2; ii >= 1; --ii) {
    double r = ratioK_fermion(c, P->Gr, ii, nt);
    *lrf = *lrf + log(r > RATIO_EPS ? r : RATIO_EPS);
  }
  z2o_kin_mmult_L(c->m, P->Gr, c->gsig_new, nt, 1);
  z2o_kin_mmult_R(c->m, P->Gr, c->gsig_new, nt, -1);
  z2o_kin_mmult_L(c->m, P->UUL, c->gsig_new, nt, 1);
}

static void globalK_prop_R(z2o_chain *c, double *lrf, int nt) {
  prop_t *P = &c->gprop;
  z2o_kin_mmult_R(c->m, P->Gr, c->sigma, nt, 1);
  z2o_kin_mmult_L(c->m, P->Gr, c->sigma, nt, -1);
  for (int ii = 1; ii <= c->N2; ++ii) {
    double r = ratioK_fermion(c, P->Gr, ii, nt);
    *lrf = *lrf + log(r > RATIO_EPS ? r : RATIO_EPS);
  }
  z2o_kin_mmult_R(c->m, P->UUR, c->gsig_new, nt, 1);
}

static void global_accept_reject(z2o_chain *c, double total, int beta_if_accept) {
  double r = z2o_ranf(&c->seed);
  double lhs = total < 0.0 ? total : 0.0;
  double rr = r > RATIO_EPS ? r : RATIO_EPS;
  double gap = fabs(lhs - log(rr));
  if (gap < c->gmin_gap) c->gmin_gap = gap;
  c->n_glob_prop += 1;
  size_t ns = sizeof(double) * (size_t)c->N2 * c->T;
  if (lhs >= log(rr)) {
    acc_count(&c->Kg, 1);
    c->n_glob_acc += 1;
    memcpy(c->sigma, c->gsig_new, ns);
    prop_assign(&c->prop, &c->gprop, c->N);
    wr_assign(&c->wr, &c->gwr, c->N, c->Nst);
    c->is_beta = beta_if_accept;
  } else {
    acc_count(&c->Kg, 0);
    memcpy(c->gsig_new, c->sigma, ns);
    prop_assign(&c->gprop, &c->prop, c->N);
    wr_assign(&c->gwr, &c->wr, c->N, c->Nst);
    c->is_beta = !beta_if_accept;
  }
}

void z2o_global_sweep(z2o_chain *c) {
  const z2o_model *m = c->m;
  if (!m->p.is_global) return;
  const int w = m->p.Nwrap;
  prop_assign(&c->gprop, &c->prop, c->N);
  wr_assign(&c->gwr, &c->wr, c->N, c->Nst);
  memcpy(c->gsig_new, c->sigma, sizeof(double) * (size_t)c->N2 * c->T);
  acc_reset(&c->Kg);
  double size_cluster = 0.0;
  for (int ngl = 1; ngl <= m->p.Nglobal; ++ngl) {
    int sc;
    double lrs, lrf = 0.0;
    if (c->is_beta) {
      wolff_flip(c, &sc, &lrs);
      for (int nt = c->T; nt >= 1; --nt) {
        if (nt % w == 0) wrap_L(c, &c->gprop, &c->gwr, nt, 0);
        propT_L(c, &c->gprop);
        globalK_prop_L(c, &lrf, nt);
      }
      wrap_L(c, &c->gprop, &c->gwr, 0, 0);
      global_accept_reject(c, lrf + lrs, 0);
    } else {
      wolff_flip(c, &sc, &lrs);
      wrap_R(c, &c->gprop, &c->gwr, 0, 0);
      for (int nt = 1; nt <= c->T; ++nt) {
        globalK_prop_R(c, &lrf, nt);
        propT_R(c, &c->gprop);
        if (nt % w == 0) wrap_R(c, &c->gprop, &c->gwr, nt, 0);
      }
      global_accept_reject(c, lrf + lrs, 1);
    }
    size_cluster = size_cluster + (double)sc;
    c->cluster_sum += sc;
  }
  acc_ratio(&c->Kg);
  size_cluster = size_cluster / (double)m->p.Nglobal;
  c->wolff_size_cluster = c->wolff_size_cluster + size_cluster;
}

/* main program for one rank, src/main.f90:39-103 */
void z2o_run_program(z2o_chain *c, z2o_obs_equal *obs_out) {
  const z2o_params *p = &c->m->p;
  if (p->is_warm)
    for (int nth = 1; nth <= p->Nwarm; ++nth) z2o_therm(c);
  z2o_pre(c);
  c->is_beta = 1;
  int istau = 0;
  for (int nbc = 1; nbc <= p->Nbin; ++nbc) {
    if (nbc > p->Nthermal) istau = p->is_tau;
    if (p->is_global) z2o_global_sweep(c);
    z2o_sweep(c, istau);
    if (obs_out) z2o_get_obs_equal(c, &obs_out[nbc - 1]);
  }
}

/* ======================================================================
 * accessors
 * ====================================================================== */
int32_t z2o_get_seed(const z2o_chain *c) { return c->seed; }
void z2o_set_seed(z2o_chain *c, int32_t seed) { c->seed = seed; }
void z2o_get_sigma(const z2o_chain *c, double *s) {
  memcpy(s, c->sigma, sizeof(double) * (size_t)c->N2 * c->T);
}
void z2o_set_sigma(z2o_chain *c, const double *s) {
  memcpy(c->sigma, s, sizeof(double) * (size_t)c->N2 * c->T);
  memcpy(c->sigma_new, s, sizeof(double) * (size_t)c->N2 * c->T);
}
void z2o_get_lambda(const z2o_chain *c, double *l) { memcpy(l, c->lambda, sizeof(double) * c->N); }
void z2o_get_green(const z2o_chain *c, double *G) {
  memcpy(G, c->prop.Gr, sizeof(double) * (size_t)c->N * c->N);
}
void z2o_set_green(z2o_chain *c, const double *G) {
  memcpy(c->prop.Gr, G, sizeof(double) * (size_t)c->N * c->N);
}
void z2o_get_prop_mat(const z2o_chain *c, int which, double *out) {
  const double *src = which == 0 ? c->prop.UUR : which == 1 ? c->prop.VUR
                    : which == 2 ? c->prop.UUL : c->prop.VUL;
  memcpy(out, src, sizeof(double) * (size_t)c->N * c->N);
}
void z2o_get_prop_d(const z2o_chain *c, int which, double *out) {
  memcpy(out, which == 0 ? c->prop.DUR : c->prop.DUL, sizeof(double) * c->N);
}
int z2o_get_is_beta(const z2o_chain *c) { return c->is_beta; }
void z2o_set_is_beta(z2o_chain *c, int b) { c->is_beta = b; }
void z2o_get_drift(const z2o_chain *c, double *xmax, double *xmean) {
  *xmax = c->prop.Xmaxm;
  *xmean = c->prop.Xmeanm;
}
void z2o_get_counters(const z2o_chain *c, double *o) {
  o[0] = c->n_local_acc;
  o[1] = c->n_local_prop;
  o[2] = c->n_glob_acc;
  o[3] = c->n_glob_prop;
  o[4] = c->n_therm_acc;
  o[5] = c->n_therm_prop;
  o[6] = c->cluster_sum;
}
void z2o_get_acc(const z2o_chain *c, double *o) {
  o[0] = c->Kl.acc;
  o[1] = c->Kg.acc;
  o[2] = c->Kt.acc;
  o[3] = c->wolff_size_cluster;
}
void z2o_get_obs_equal(const z2o_chain *c, z2o_obs_equal *out) {
  out->flux_avg = c->oe.flux_avg;
  out->chi_flux = c->oe.chi_flux;
  out->pair_sc_q0 = c->oe.pair_sc_q0;
  out->cdw_pi = c->oe.cdw_pi;
  out->cdw_shift = c->oe.cdw_shift;
  out->cdw_ratio = c->oe.cdw_ratio;
  out->density = c->oe.density;
}
void z2o_get_obs_tau(const z2o_chain *c, double *out) {
  if (c->have_dyn) memcpy(out, c->otau, sizeof(double) * 7 * (size_t)c->N * c->T);
}
void z2o_get_dyn_drift(const z2o_chain *c, double *xmax3, double *xmean3) {
  for (int i = 0; i < 3; ++i) {
    xmax3[i] = c->have_dyn ? c->pg.Xmaxm[i] : 0.0;
    xmean3[i] = c->have_dyn ? c->pg.Xmeanm[i] : 0.0;
  }
}
double z2o_get_min_gap(const z2o_chain *c) { return c->min_gap; }
void z2o_reset_min_gap(z2o_chain *c) { c->min_gap = INFINITY; }
double z2o_get_gmin_gap(const z2o_chain *c) { return c->gmin_gap; }
void z2o_set_decision_log(z2o_chain *c, uint32_t *buf, int64_t cap) {
  c->log_buf = buf;
  c->log_cap = cap;
  c->log_n = 0;
}
int64_t z2o_get_decision_count(const z2o_chain *c) { return c->log_n; }
