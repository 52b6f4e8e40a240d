// Host side of the gfx950 Z2-gauge DQMC engine: context, C ABI (include/z2dqmc.h)
// and the per-sweep orchestration of the batched kernels.  The order of steps
// follows the reference drivers LocalSweep_mod / Stabilize_mod / LocalK_mod
// (src/local_sweep.f90, src/stabilization.f90, src/localK.f90) exactly; every
// step acts on all chains of the context in one launch.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/z2dqmc.h"
#include "flip_kernels.hpp"
#include "flip_rg.hpp"
#include "global_kernels.hpp"
#include "kin_kernels.hpp"
#include "linalg_kernels.hpp"
#include "obs_kernels.hpp"
#include "qr_blocked.hpp"
#include "tau_kernels.hpp"

static thread_local std::string g_err;

static int set_err(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPC(x)                                                                         \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      return set_err(Z2_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
  } while (0)

#define CHK(x)             \
  do {                     \
    int r_ = (x);          \
    if (r_ != Z2_OK) return r_; \
  } while (0)

extern "C" const char *z2_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------- RNG (host)
// ran_imada.f:55-68
static double ranf_host(int32_t *iq) {
  uint32_t x = (uint32_t)(*iq) * 48828125u;
  int32_t s = (int32_t)x;
  if (s < 0) s = (int32_t)((uint32_t)s + 2147483648u);
  *iq = s;
  return (double)s / 2147483648.0;
}

// One per-chain device array of a context: `bytes` per chain, `slots` slot-major
// copies (slot stride = all chains).  shift() moves the array's pointer by a
// number of chains, so that the launch code sees a contiguous range of chains as
// a whole context.
enum : int { CF_PERSIST = 1, CF_TRIAL = 2 };  // moves with its chain; changed by a Wolff trial
struct ChainField {
  std::function<char *()> get;
  std::function<void(long)> shift;
  size_t bytes = 0;
  int slots = 1;
  int flags = 0;
  char *backup = nullptr;  // Wolff trial backup (CF_TRIAL), all chains
};

struct z2_ctx {
  z2_params p;
  // nch: chains the launch code acts on (all chains, or the range of a ChainRange);
  // nch_all: chains of the context
  int N = 0, N2 = 0, T = 0, w = 0, Nst = 0, nch = 0, nch_all = 0, Lx = 0, Ly = 0;
  hipStream_t stream = nullptr;
  // propagator wraps deferred to the next stabilization: ud_n slices nt0, nt0 + step,
  // ... of UUL (ud_cols = 1, left pass) or UUR (0, right pass) not applied yet
  int ud_n = 0, ud_nt0 = 0, ud_step = 0, ud_cols = 0;
  std::vector<z2_chain_desc> desc;
  std::vector<ChainConst> hcc;
  ChainConst *dcc = nullptr;
  LatticeDev lat{};
  int *d_int = nullptr;  // all integer lattice tables
  double *d_cosT = nullptr;
  uint32_t *d_apow = nullptr;
  // chain state (FP64)
  double *G = nullptr, *UUR = nullptr, *VUR = nullptr, *UUL = nullptr, *VUL = nullptr;
  double *DUR = nullptr, *DUL = nullptr;
  double *URl = nullptr, *VRl = nullptr, *ULl = nullptr, *VLl = nullptr, *DRl = nullptr,
         *DLl = nullptr;  // slot-major: [st][chain][.]
  int8_t *sig = nullptr;
  int32_t *seed = nullptr;
  double *Tbuf = nullptr, *Vbuf = nullptr;
  long upd_stride = 0;
  int *nacc = nullptr, *mu_done = nullptr;
  int kdelay = 16;  // delay depth of the flip kernel
  int flip_rg = 1;  // FP32 at N = 256: the flip kernel with G in registers (k_flip_rg)
  int kin_cb = 8, kin_rb = 8;  // kinetic tiles: columns per row-rotation block, rows per column-rotation block
  int kdelay_max = 16;
  size_t lds_max = 65536;
  unsigned long long *accc = nullptr;
  double *xmax = nullptr, *xmean = nullptr;
  double *obs_acc = nullptr, *obs_res = nullptr;
  double *opart = nullptr;  // per chain and slot: tile partial sums of the equal-time observables
  double *odiag = nullptr, *oflux = nullptr;  // per chain and slot: 1 - G(i, i), plaquette products
  int opart_stride = 0;     // tiles per chain
  int obs_pend = 0;         // slices measured whose sums k_obs_finish has not yet added (slots 0 ..)
  // scratch
  double *M = nullptr, *Q = nullptr, *X = nullptr, *Y = nullptr, *Gn = nullptr;
  double *tau = nullptr, *Dtmp = nullptr;
  int *ipvt = nullptr;
  double *qrw = nullptr, *qrw2 = nullptr;  // workspaces of the blocked QR (N x N, 2N x 2N)
  double *ortw = nullptr;  // panel T factors of the explicit Q (ceil(N/16) x 32 x 32 per chain)
  const double *ortw_src = nullptr;  // the matrix whose split QR wrote the T factors now in ortw
  size_t ortw_per = 0;
  int qr_mode = Z2_QR_BLOCKED;
  int stab_mode = Z2_STAB_REFERENCE;
  double *dsc = nullptr;  // D-split scales [chain][4][N]
  // mixed precision: propagated G, UUR, UUL and the delayed updates in FP32
  bool mixed = false;
  bool u64 = false;  // mixed mode: the right/left propagators UUR, UUL (the B-products) kept in FP64
  float *Gf = nullptr, *UURf = nullptr, *UULf = nullptr, *Tbf = nullptr, *Vbf = nullptr;
  // decisions of the last flip launch (bit k = proposal k), and decisions to replay
  uint32_t *rec = nullptr;
  const uint32_t *force = nullptr;
  int dwords = 0;
  double *cmp = nullptr;  // per-chain max |G - G_other| (shadow runs)
  // time-displaced measurement (Dyn%sweep_R): its own right propagator d*, the four
  // Green's functions G(0,0), G(0,t), G(t,0), G(t,t), the 2N x 2N recompute scratch
  bool tau_on = false;
  double *dUUR = nullptr, *dVUR = nullptr, *dDUR = nullptr, *dUUL = nullptr, *dVUL = nullptr, *dDUL = nullptr;
  double *G00 = nullptr, *G0t = nullptr, *Gt0 = nullptr, *Gtt = nullptr;
  double *iVR = nullptr, *iVL = nullptr;
  double *bigA = nullptr, *bigQ = nullptr, *bigZ = nullptr, *bigR = nullptr, *tau2 = nullptr, *D2 = nullptr;
  int *ipvt2 = nullptr;
  double *otau = nullptr;       // [chain][7][T][N]
  double *dyn_drift = nullptr;  // [chain][6]: max of G0t, Gt0, Gtt drifts, then their sums
  // Wolff global update: trial bond, cluster mask and queue, log ratios, accept
  // flag, counters; boson warm-up counters
  int *gbond = nullptr, *gmask = nullptr, *gqueue = nullptr, *gacc = nullptr;
  double *glr = nullptr, *gsz = nullptr;
  unsigned long long *gcnt = nullptr, *tcnt = nullptr;
  bool gmode = false;      // the flip kernel runs in forced (Wolff trial) mode
  bool gbackup = false;    // trial backups allocated
  // chains are stored grouped by pass direction (is_beta 1 first); phys maps a
  // chain's index to its storage slot, isb is the direction per slot
  std::vector<int> phys, isb;
  std::vector<double> lam;  // lambda(N) per storage slot (host only: the sweep does not use it)
  std::vector<ChainField> fields;
  int *d_perm = nullptr;
  char *reorder_tmp = nullptr;
  std::vector<void *> allocs;
  bool have_fields = false, have_pre = false;
  // 1: every chain has just finished a left pass (the place of Dyn%reset, the start
  // of a right pass); 0: not; -1: unknown (chains in different directions)
  int left_done = 0;
  // kernel timing hooks: HIP events around every launch of one kernel class
  int prof_kind = 0;
  hipEvent_t prof_ref = nullptr;  // recorded by z2_profile_kernel: time zero of z2_profile_spans
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev;
  size_t prof_used = 0;
  double prof_work = 0.0;
};

static int prof_begin(z2_ctx *c, int kind) {
  if (c->prof_kind != kind) return Z2_OK;
  if (c->prof_used == c->prof_ev.size()) {
    hipEvent_t a, b;
    HIPC(hipEventCreate(&a));
    HIPC(hipEventCreate(&b));
    c->prof_ev.push_back({a, b});
  }
  HIPC(hipEventRecord(c->prof_ev[c->prof_used].first, c->stream));
  return Z2_OK;
}

// work: the launch's algorithmic work (flops, or bytes for the memory-bound classes)
static int prof_end(z2_ctx *c, int kind, double work) {
  if (c->prof_kind != kind) return Z2_OK;
  HIPC(hipEventRecord(c->prof_ev[c->prof_used].second, c->stream));
  c->prof_used++;
  c->prof_work += work;
  return Z2_OK;
}

template <typename T>
static int dalloc(z2_ctx *c, T **ptr, size_t n) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(T) * (n ? n : 1));
  if (e != hipSuccess) return set_err(Z2_ENOMEM, "hipMalloc(%zu bytes): %s", sizeof(T) * n, hipGetErrorString(e));
  c->allocs.push_back(p);
  *ptr = reinterpret_cast<T *>(p);
  return Z2_OK;
}

static void free_ctx(z2_ctx *c) {
  for (auto &e : c->prof_ev) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  if (c->prof_ref) (void)hipEventDestroy(c->prof_ref);
  for (void *p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

template <typename T>
static void reg_field(z2_ctx *c, T *&ptr, size_t per_chain, int slots, int flags) {
  ChainField f;
  f.get = [&ptr]() { return reinterpret_cast<char *>(ptr); };
  f.shift = [&ptr, per_chain](long dc) {
    if (ptr) ptr += dc * (long)per_chain;
  };
  f.bytes = per_chain * sizeof(T);
  f.slots = slots;
  f.flags = flags;
  c->fields.push_back(f);
}

// every per-chain device array with its per-chain size (arrays allocated later,
// e.g. by z2_enable_tau, are registered while still null)
static void register_fields(z2_ctx *c) {
  const size_t N = c->N, nn = N * N, T = c->T, N2 = c->N2, us = (size_t)c->upd_stride;
  const int L = c->Nst + 1, P = CF_PERSIST, PT = CF_PERSIST | CF_TRIAL;
  reg_field(c, c->G, nn, 1, PT);
  reg_field(c, c->UUR, nn, 1, PT);
  reg_field(c, c->VUR, nn, 1, PT);
  reg_field(c, c->UUL, nn, 1, PT);
  reg_field(c, c->VUL, nn, 1, PT);
  reg_field(c, c->DUR, N, 1, PT);
  reg_field(c, c->DUL, N, 1, PT);
  reg_field(c, c->URl, nn, L, PT);
  reg_field(c, c->VRl, nn, L, PT);
  reg_field(c, c->ULl, nn, L, PT);
  reg_field(c, c->VLl, nn, L, PT);
  reg_field(c, c->DRl, N, L, PT);
  reg_field(c, c->DLl, N, L, PT);
  reg_field(c, c->sig, T * N2, 1, PT);
  reg_field(c, c->seed, 1, 1, P);
  reg_field(c, c->Tbuf, us, 1, 0);
  reg_field(c, c->Vbuf, us, 1, 0);
  reg_field(c, c->nacc, 1, 1, 0);
  reg_field(c, c->mu_done, 1, 1, 0);
  reg_field(c, c->accc, 2, 1, P);
  reg_field(c, c->xmax, 1, 1, PT);
  reg_field(c, c->xmean, 1, 1, PT);
  reg_field(c, c->obs_acc, OBS_NACC, 1, P);
  reg_field(c, c->obs_res, RES_N, 1, P);
  reg_field(c, c->opart, OBS_SLOTS * 3 * (size_t)c->opart_stride, 1, 0);
  reg_field(c, c->odiag, OBS_SLOTS * N, 1, 0);
  reg_field(c, c->oflux, OBS_SLOTS * N, 1, 0);
  for (double **m : {&c->M, &c->Q, &c->X, &c->Y, &c->Gn, &c->qrw}) reg_field(c, *m, nn, 1, 0);
  reg_field(c, c->tau, N, 1, 0);
  reg_field(c, c->Dtmp, N, 1, 0);
  reg_field(c, c->dsc, 4 * N, 1, 0);
  reg_field(c, c->ortw, c->ortw_per, 1, 0);
  reg_field(c, c->ipvt, N, 1, 0);
  reg_field(c, c->rec, (size_t)c->dwords, 1, 0);
  reg_field(c, c->cmp, 1, 1, 0);
  reg_field(c, c->Gf, nn, 1, PT);
  reg_field(c, c->UURf, nn, 1, PT);
  reg_field(c, c->UULf, nn, 1, PT);
  reg_field(c, c->Tbf, us, 1, 0);
  reg_field(c, c->Vbf, us, 1, 0);
  reg_field(c, c->dcc, 1, 1, P);
  for (double **m : {&c->dUUR, &c->dVUR, &c->dUUL, &c->dVUL, &c->G00, &c->G0t, &c->Gt0, &c->Gtt, &c->iVR, &c->iVL})
    reg_field(c, *m, nn, 1, 0);
  for (double **m : {&c->bigA, &c->bigQ, &c->bigZ, &c->bigR, &c->qrw2}) reg_field(c, *m, 4 * nn, 1, 0);
  reg_field(c, c->dDUR, N, 1, 0);
  reg_field(c, c->dDUL, N, 1, 0);
  reg_field(c, c->tau2, 2 * N, 1, 0);
  reg_field(c, c->D2, 2 * N, 1, 0);
  reg_field(c, c->ipvt2, 2 * N, 1, 0);
  reg_field(c, c->otau, (size_t)TAU_NOBS * T * N, 1, P);
  reg_field(c, c->dyn_drift, 6, 1, P);
  reg_field(c, c->gbond, 1, 1, 0);
  reg_field(c, c->gmask, T, 1, 0);
  reg_field(c, c->gqueue, T, 1, 0);
  reg_field(c, c->glr, 2, 1, 0);
  reg_field(c, c->gacc, 1, 1, 0);
  reg_field(c, c->gcnt, 3, 1, P);
  reg_field(c, c->gsz, 2, 1, P);
  reg_field(c, c->tcnt, 2, 1, P);
}

// A contiguous range [c0, c0 + cnt) of storage slots seen by the launch code as
// the whole context (every per-chain pointer moved by c0 chains), on `stream`.
static int u_flush(z2_ctx *c);
static int obs_flush(z2_ctx *c);

struct ChainRange {
  z2_ctx *c;
  int c0, nsave;
  hipStream_t ssave;
  ChainRange(z2_ctx *ctx, int first, int cnt, hipStream_t st) : c(ctx), c0(first), nsave(ctx->nch), ssave(ctx->stream) {
    for (auto &f : c->fields) f.shift(c0);
    c->nch = cnt;
    c->stream = st;
  }
  ~ChainRange() {
    for (auto &f : c->fields) f.shift(-c0);
    c->nch = nsave;
    c->stream = ssave;
  }
};

// per-chain constants with the reference's libm calls
static ChainConst make_const(const z2_chain_desc &d, int Ltrot) {
  ChainConst k;
  double Dtau = d.Beta / (double)Ltrot;
  double alpha = 0.5 * Dtau * d.RT;  // src/operator_kin.f90:22-29
  double alpha_2 = Dtau * d.RT;
  k.Cp = cosh((1.0 * alpha) * 1.0);
  k.Cm = cosh((-1.0 * alpha) * 1.0);
  k.Sp = sinh((1.0 * alpha) * 1.0);
  k.Sm = sinh((-1.0 * alpha) * 1.0);
  for (int q = 0; q < 2; ++q) {  // src/operator_kin.f90:40-59
    double s_old = q == 0 ? 1.0 : -1.0, s_new = -s_old;
    double C_new = cosh(alpha_2 * s_new), C_old = cosh(alpha_2 * s_old);
    double S_new = sinh(alpha_2 * s_new), S_old = sinh(alpha_2 * s_old);
    double d11 = C_new * C_old - S_new * S_old - 1.0;
    double d12 = S_new * C_old - C_new * S_old;
    k.delta[q][0] = d11;
    k.delta[q][1] = d12;
    k.delta[q][2] = d12;
    k.delta[q][3] = d11;
  }
  // src/fields.f90:52-117
  double ab = Dtau * d.J, gamma = 0.0;
  bool temporal = fabs(d.h) > 1.0e-10;
  if (temporal) {
    double ta = tanh(Dtau * d.h);
    if (ta < 1.0e-300) ta = 1.0e-300;
    gamma = -0.5 * log(ta);
  }
  double lub = log(1.0e200);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double dS = 4.0 * (i - 1), dT = 4.0 * (j - 1);
      double e = ab * dS;
      if (e > lub) e = lub;
      if (e < -lub) e = -lub;
      double r = exp(e);
      if (temporal) {
        double e2 = gamma * dT;
        if (e2 > lub) e2 = lub;
        if (e2 < -lub) e2 = -lub;
        r = r * exp(e2);
      }
      k.boson[i * 3 + j] = r;
    }
  k.ep = exp(Dtau * d.mu);
  k.em = exp(-Dtau * d.mu);
  k.beta = d.Beta;
  k.rt = d.RT;
  // compute_time_bond_prob (src/global_update.f90:88-100) and the plaquette
  // coupling of plaquette_log_ratio (:151-172)
  double pt = 0.0;
  if (fabs(d.h) > 1.0e-10) {
    double ta = tanh(Dtau * fabs(d.h));
    if (ta < 1.0e-300) ta = 1.0e-300;
    double g = -0.5 * log(ta);
    pt = 1.0 - exp(-2.0 * g);
  }
  if (pt > 1.0) pt = 1.0;
  if (pt < 0.0) pt = 0.0;
  k.ptime = pt;
  k.walpha = Dtau * d.J;
  return k;
}

static int npbc(int nr, int L) { return nr < 1 ? nr + L : (nr > L ? nr - L : nr); }

extern "C" int z2_ctx_create(const z2_params *p, int32_t n_chains, const z2_chain_desc *chains,
                             z2_ctx **out) {
  if (!p || !chains || !out || n_chains <= 0) return set_err(Z2_EINVAL, "null argument or n_chains <= 0");
  if (p->Nlx < 2 || p->Nly < 2 || (p->Nlx % 2) || (p->Nly % 2))
    return set_err(Z2_EINVAL, "Nlx, Nly must be even and >= 2 (checkerboard groups are perfect matchings only for even L)");
  if (p->Ltrot < 1 || p->Nwrap < 1 || p->Ltrot % p->Nwrap)
    return set_err(Z2_EINVAL, "Ltrot must be a positive multiple of Nwrap");
  if (p->precision != Z2_PREC_FP64 && p->precision != Z2_PREC_MIXED)
    return set_err(Z2_EINVAL, "unknown precision mode %d", p->precision);
  if (p->Nlx * p->Nly > 1024) return set_err(Z2_EINVAL, "N = Nlx*Nly must be <= 1024");
  z2_ctx *c = new z2_ctx();
  c->p = *p;
  c->Lx = p->Nlx;
  c->Ly = p->Nly;
  c->N = p->Nlx * p->Nly;
  c->N2 = 2 * c->N;
  c->T = p->Ltrot;
  c->w = p->Nwrap;
  c->Nst = p->Ltrot / p->Nwrap;
  c->nch = n_chains;
  c->nch_all = n_chains;
  for (int i = 0; i < n_chains; ++i) c->phys.push_back(i);
  c->isb.assign(n_chains, 1);
  c->lam.assign((size_t)n_chains * c->N, 1.0);
  c->desc.assign(chains, chains + n_chains);
  if (c->p.Nsweep < 1) c->p.Nsweep = 1;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    int r = set_err(Z2_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    delete c;
    return r;
  }
  int rc = Z2_OK;
  do {
    const int N = c->N, N2 = c->N2, Lx = c->Lx, Ly = c->Ly;
    // lattice, src/lattice.f90:25-155 (0-based)
    auto site = [&](int x, int y) { return (npbc(y, Ly) - 1) * Lx + (npbc(x, Lx) - 1); };
    std::vector<int> bond_p(2 * N2), grp, bb(6 * N2), stag(N), xy(2 * N), plaq(4 * N);
    for (int y = 1; y <= Ly; ++y)
      for (int x = 1; x <= Lx; ++x) {
        int s = site(x, y);
        bond_p[2 * (2 * s)] = s;
        bond_p[2 * (2 * s) + 1] = site(x + 1, y);
        bond_p[2 * (2 * s + 1)] = s;
        bond_p[2 * (2 * s + 1) + 1] = site(x, y + 1);
        xy[2 * s] = x;
        xy[2 * s + 1] = y;
        stag[s] = ((x + y) % 2 == 0) ? 1 : -1;
      }
    auto bx = [&](int x, int y) { return 2 * site(x, y); };      // +x bond from (x,y)
    auto by = [&](int x, int y) { return 2 * site(x, y) + 1; };  // +y bond from (x,y)
    std::vector<int> g4[4];
    for (int b = 0; b < N2; ++b) {
      int s = b / 2, x = xy[2 * s], y = xy[2 * s + 1];
      if (b % 2 == 0) g4[(x % 2 == 0) ? 0 : 1].push_back(b);
      else g4[(y % 2 == 0) ? 2 : 3].push_back(b);
      int *q = &bb[6 * b];
      if (b % 2 == 0) {  // horizontal, src/lattice.f90:280-292
        q[0] = by(x, y); q[1] = bx(x, y + 1); q[2] = by(x + 1, y);
        q[3] = by(x, y - 1); q[4] = bx(x, y - 1); q[5] = by(x + 1, y - 1);
      } else {           // vertical, src/lattice.f90:293-305
        q[0] = bx(x - 1, y + 1); q[1] = by(x - 1, y); q[2] = bx(x - 1, y);
        q[3] = bx(x, y + 1); q[4] = by(x + 1, y); q[5] = bx(x, y);
      }
      if (b % 2 == 0) {  // plaquette with this bond at the bottom: bottom,right,top,left
        int pl = 4 * s;
        plaq[pl] = bx(x, y); plaq[pl + 1] = by(x + 1, y); plaq[pl + 2] = bx(x, y + 1); plaq[pl + 3] = by(x, y);
      }
    }
    for (int g = 0; g < 4; ++g) {
      if ((int)g4[g].size() != N / 2) { rc = set_err(Z2_EINVAL, "checkerboard group size"); break; }
      grp.insert(grp.end(), g4[g].begin(), g4[g].end());
    }
    if (rc) break;
    std::vector<int> all;
    size_t off_bond = all.size(); all.insert(all.end(), bond_p.begin(), bond_p.end());
    size_t off_grp = all.size(); all.insert(all.end(), grp.begin(), grp.end());
    size_t off_bb = all.size(); all.insert(all.end(), bb.begin(), bb.end());
    size_t off_stag = all.size(); all.insert(all.end(), stag.begin(), stag.end());
    size_t off_xy = all.size(); all.insert(all.end(), xy.begin(), xy.end());
    size_t off_plaq = all.size(); all.insert(all.end(), plaq.begin(), plaq.end());
    if ((rc = dalloc(c, &c->d_int, all.size()))) break;
    if (hipMemcpy(c->d_int, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice) != hipSuccess) { rc = set_err(Z2_EHIP, "copy tables"); break; }
    c->lat.N = N;
    c->lat.N2 = N2;
    c->lat.Lx = Lx;
    c->lat.Ly = Ly;
    c->lat.bond_p = c->d_int + off_bond;
    c->lat.grp = c->d_int + off_grp;
    c->lat.bond_bonds = c->d_int + off_bb;
    c->lat.stag = c->d_int + off_stag;
    c->lat.xy = c->d_int + off_xy;
    c->lat.plaq = c->d_int + off_plaq;
    // cos table, src/obser_equal.f90:92,139
    const double PI = acos(-1.0);
    double qx = PI - 2.0 * PI / (double)Lx;
    std::vector<double> cosT((2 * Lx - 1) * (2 * Ly - 1));
    for (int dy = -(Ly - 1); dy <= Ly - 1; ++dy)
      for (int dx = -(Lx - 1); dx <= Lx - 1; ++dx)
        cosT[(dx + Lx - 1) + (2 * Lx - 1) * (dy + Ly - 1)] = cos(qx * (double)dx + PI * (double)dy);
    if ((rc = dalloc(c, &c->d_cosT, cosT.size()))) break;
    if (hipMemcpy(c->d_cosT, cosT.data(), sizeof(double) * cosT.size(), hipMemcpyHostToDevice) != hipSuccess) { rc = set_err(Z2_EHIP, "copy cos"); break; }
    std::vector<uint32_t> apow(N2 + 1);
    uint64_t a = 1;
    for (int k = 0; k <= N2; ++k) { apow[k] = (uint32_t)a; a = (a * 48828125ull) & 0x7fffffffull; }
    if ((rc = dalloc(c, &c->d_apow, apow.size()))) break;
    if (hipMemcpy(c->d_apow, apow.data(), sizeof(uint32_t) * apow.size(), hipMemcpyHostToDevice) != hipSuccess) { rc = set_err(Z2_EHIP, "copy apow"); break; }
    for (int i = 0; i < n_chains; ++i) c->hcc.push_back(make_const(chains[i], c->T));
    if ((rc = dalloc(c, &c->dcc, n_chains))) break;
    if (hipMemcpy(c->dcc, c->hcc.data(), sizeof(ChainConst) * n_chains, hipMemcpyHostToDevice) != hipSuccess) { rc = set_err(Z2_EHIP, "copy consts"); break; }
    // state
    const size_t nn = (size_t)N * N, B = (size_t)n_chains;
    const size_t L = (size_t)(c->Nst + 1);
    if ((rc = dalloc(c, &c->G, B * nn))) break;
    if ((rc = dalloc(c, &c->UUR, B * nn))) break;
    if ((rc = dalloc(c, &c->VUR, B * nn))) break;
    if ((rc = dalloc(c, &c->UUL, B * nn))) break;
    if ((rc = dalloc(c, &c->VUL, B * nn))) break;
    if ((rc = dalloc(c, &c->DUR, B * N))) break;
    if ((rc = dalloc(c, &c->DUL, B * N))) break;
    if ((rc = dalloc(c, &c->URl, L * B * nn))) break;
    if ((rc = dalloc(c, &c->VRl, L * B * nn))) break;
    if ((rc = dalloc(c, &c->ULl, L * B * nn))) break;
    if ((rc = dalloc(c, &c->VLl, L * B * nn))) break;
    if ((rc = dalloc(c, &c->DRl, L * B * N))) break;
    if ((rc = dalloc(c, &c->DLl, L * B * N))) break;
    if ((rc = dalloc(c, &c->sig, B * c->T * N2))) break;
    if ((rc = dalloc(c, &c->seed, B))) break;
    c->upd_stride = (long)4 * N * N2;  // up to 2N accepts x (2 vectors of N)
    if (p->precision == Z2_PREC_FP64) {
      if ((rc = dalloc(c, &c->Tbuf, B * c->upd_stride))) break;
      if ((rc = dalloc(c, &c->Vbuf, B * c->upd_stride))) break;
    }
    if ((rc = dalloc(c, &c->nacc, B))) break;
    if ((rc = dalloc(c, &c->mu_done, B))) break;
    if ((rc = dalloc(c, &c->accc, 2 * B))) break;
    if ((rc = dalloc(c, &c->xmax, B))) break;
    if ((rc = dalloc(c, &c->xmean, B))) break;
    if ((rc = dalloc(c, &c->obs_acc, B * OBS_NACC))) break;
    if ((rc = dalloc(c, &c->obs_res, B * RES_N))) break;
    {
      const int ntb = (N + OBS_TB - 1) / OBS_TB;
      c->opart_stride = ntb * ntb;
      if ((rc = dalloc(c, &c->opart, B * OBS_SLOTS * 3 * (size_t)c->opart_stride))) break;
      if ((rc = dalloc(c, &c->odiag, B * OBS_SLOTS * (size_t)N))) break;
      if ((rc = dalloc(c, &c->oflux, B * OBS_SLOTS * (size_t)N))) break;
    }
    if ((rc = dalloc(c, &c->M, B * nn))) break;
    if ((rc = dalloc(c, &c->Q, B * nn))) break;
    if ((rc = dalloc(c, &c->X, B * nn))) break;
    if ((rc = dalloc(c, &c->Y, B * nn))) break;
    if ((rc = dalloc(c, &c->Gn, B * nn))) break;
    if ((rc = dalloc(c, &c->tau, B * N))) break;
    if ((rc = dalloc(c, &c->Dtmp, B * N))) break;
    if ((rc = dalloc(c, &c->ipvt, B * N))) break;
    if ((rc = dalloc(c, &c->qrw, B * nn))) break;
    if ((rc = dalloc(c, &c->dsc, B * 4 * (size_t)N))) break;
    c->ortw_per = (size_t)((N + 15) / 16) * 32 * 32;
    if ((rc = dalloc(c, &c->ortw, B * c->ortw_per))) break;
    c->dwords = (N2 + 31) / 32;
    {
      int dev = 0, lds = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
        c->lds_max = (size_t)lds;
      // flip kernel LDS: cached blocks and tables plus 4N doubles per pending update
      // (512 bytes kept for its static shared variables)
      int cap = 0;
      while (cap < N2 && flip_lds_bytes(N, cap + 1) + 512 <= c->lds_max) ++cap;
      if (cap < 1) { rc = set_err(Z2_EINVAL, "N too large for the flip kernel's LDS"); break; }
      c->kdelay_max = cap;
      c->kdelay = c->kdelay_max < 16 ? c->kdelay_max : 16;
    }
    if ((rc = dalloc(c, &c->rec, B * c->dwords))) break;
    if ((rc = dalloc(c, &c->cmp, B))) break;
    c->mixed = p->precision == Z2_PREC_MIXED;
    c->u64 = c->mixed;  // the B-products in FP64 by default (z2_set_u_precision)
    if (c->mixed) {
      if ((rc = dalloc(c, &c->Gf, B * nn))) break;
      if ((rc = dalloc(c, &c->UURf, B * nn))) break;
      if ((rc = dalloc(c, &c->UULf, B * nn))) break;
      if ((rc = dalloc(c, &c->Tbf, B * c->upd_stride))) break;
      if ((rc = dalloc(c, &c->Vbf, B * c->upd_stride))) break;
    }
    // Propagator%make / WrapList%make: identities, zero lists
    std::vector<double> I(B * nn, 0.0), ones(B * N, 1.0);
    for (size_t b = 0; b < B; ++b)
      for (int i = 0; i < N; ++i) I[b * nn + i + (size_t)i * N] = 1.0;
    for (double *m : {c->G, c->UUR, c->VUR, c->UUL, c->VUL})
      if (hipMemcpy(m, I.data(), sizeof(double) * B * nn, hipMemcpyHostToDevice) != hipSuccess) rc = Z2_EHIP;
    for (double *d : {c->DUR, c->DUL})
      if (hipMemcpy(d, ones.data(), sizeof(double) * B * N, hipMemcpyHostToDevice) != hipSuccess) rc = Z2_EHIP;
    if (rc) { rc = set_err(Z2_EHIP, "init copies"); break; }
    if (hipMemset(c->URl, 0, sizeof(double) * L * B * nn) || hipMemset(c->VRl, 0, sizeof(double) * L * B * nn) ||
        hipMemset(c->ULl, 0, sizeof(double) * L * B * nn) || hipMemset(c->VLl, 0, sizeof(double) * L * B * nn) ||
        hipMemset(c->DRl, 0, sizeof(double) * L * B * N) || hipMemset(c->DLl, 0, sizeof(double) * L * B * N) ||
        hipMemset(c->accc, 0, sizeof(unsigned long long) * 2 * B) || hipMemset(c->xmax, 0, sizeof(double) * B) ||
        hipMemset(c->xmean, 0, sizeof(double) * B) || hipMemset(c->obs_acc, 0, sizeof(double) * B * OBS_NACC) ||
        hipMemset(c->obs_res, 0, sizeof(double) * B * RES_N) || hipMemset(c->nacc, 0, sizeof(int) * B) ||
        hipMemset(c->mu_done, 0, sizeof(int) * B)) {
      rc = set_err(Z2_EHIP, "memset");
      break;
    }
    if (c->mixed) {
      std::vector<float> If(I.begin(), I.end());
      for (float *m : {c->Gf, c->UURf, c->UULf})
        if (hipMemcpy(m, If.data(), sizeof(float) * B * nn, hipMemcpyHostToDevice) != hipSuccess) rc = Z2_EHIP;
      if (rc) { rc = set_err(Z2_EHIP, "init copies (fp32)"); break; }
    }
    // Wolff and warm-up state
    if ((rc = dalloc(c, &c->gbond, B)) || (rc = dalloc(c, &c->gmask, B * c->T)) ||
        (rc = dalloc(c, &c->gqueue, B * c->T)) || (rc = dalloc(c, &c->glr, 2 * B)) || (rc = dalloc(c, &c->gacc, B)) ||
        (rc = dalloc(c, &c->gcnt, 3 * B)) || (rc = dalloc(c, &c->gsz, 2 * B)) || (rc = dalloc(c, &c->tcnt, 2 * B)) ||
        (rc = dalloc(c, &c->d_perm, B)))
      break;
    if (hipMemset(c->gcnt, 0, sizeof(unsigned long long) * 3 * B) || hipMemset(c->gsz, 0, sizeof(double) * 2 * B) ||
        hipMemset(c->tcnt, 0, sizeof(unsigned long long) * 2 * B) || hipMemset(c->gacc, 0, sizeof(int) * B)) {
      rc = set_err(Z2_EHIP, "memset");
      break;
    }
    register_fields(c);
    std::vector<int32_t> seeds(B);
    for (size_t b = 0; b < B; ++b) seeds[b] = chains[b].seed;
    if (hipMemcpy(c->seed, seeds.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice) != hipSuccess) { rc = set_err(Z2_EHIP, "seed"); break; }
  } while (0);
  if (rc) {
    free_ctx(c);
    return rc;
  }
  *out = c;
  return Z2_OK;
}

extern "C" int z2_ctx_destroy(z2_ctx *c) {
  if (!c) return Z2_OK;
  (void)hipStreamSynchronize(c->stream);
  free_ctx(c);
  return Z2_OK;
}

extern "C" int z2_ctx_info(const z2_ctx *c, int32_t *N, int32_t *nch, int32_t *prec) {
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (N) *N = c->N;
  if (nch) *nch = c->nch;
  if (prec) *prec = c->p.precision;
  return Z2_OK;
}

extern "C" int z2_set_delay(z2_ctx *c, int32_t k) {
  if (!c || k < 1 || k > c->kdelay_max)
    return set_err(Z2_EINVAL, "delay depth must be in [1, %d] for N = %d", c ? c->kdelay_max : 0, c ? c->N : 0);
  c->kdelay = k;
  return Z2_OK;
}

extern "C" int z2_set_flip_variant(z2_ctx *c, int32_t v) {
  if (!c || (v != Z2_FLIP_DELAYED && v != Z2_FLIP_REGISTER)) return set_err(Z2_EINVAL, "bad flip kernel variant");
  c->flip_rg = v == Z2_FLIP_REGISTER;
  return Z2_OK;
}

extern "C" int z2_set_u_precision(z2_ctx *c, int32_t bits) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c || (bits != 32 && bits != 64)) return set_err(Z2_EINVAL, "propagator precision must be 32 or 64 bits");
  if (!c->mixed) return bits == 64 ? Z2_OK : set_err(Z2_EINVAL, "FP64 contexts propagate in FP64");
  if (c->have_pre) return set_err(Z2_ESTATE, "set the propagator precision before z2_pre");
  c->u64 = bits == 64;
  return Z2_OK;
}

extern "C" int z2_set_kin_tiles(z2_ctx *c, int32_t rows_cb, int32_t cols_rb) {
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if ((rows_cb != 8 && rows_cb != 16) || (cols_rb != 8 && cols_rb != 32))
    return set_err(Z2_EINVAL, "kinetic tiles: rows kernel 8 or 16 columns, columns kernel 8 or 32 rows");
  // FP64 tiles (the time-displaced path runs the kinetic kernels in FP64 in both modes)
  const size_t tab = sizeof(int) * 2 * (size_t)c->N;
  if (8 * (size_t)std::max(rows_cb, cols_rb) * c->N + tab > c->lds_max)
    return set_err(Z2_EINVAL, "kinetic tile exceeds the LDS at N = %d (FP64 kernels)", c->N);
  c->kin_cb = rows_cb;
  c->kin_rb = cols_rb;
  return Z2_OK;
}

extern "C" int z2_set_stab_mode(z2_ctx *c, int32_t mode) {
  if (!c || (mode != Z2_STAB_REFERENCE && mode != Z2_STAB_DSPLIT)) return set_err(Z2_EINVAL, "bad stabilization mode");
  c->stab_mode = mode;
  return Z2_OK;
}

extern "C" int z2_set_qr_mode(z2_ctx *c, int32_t mode) {
  if (!c || (mode != Z2_QR_BLOCKED && mode != Z2_QR_PIVOTED)) return set_err(Z2_EINVAL, "bad QR mode");
  c->qr_mode = mode;
  return Z2_OK;
}

extern "C" int z2_get_stream(z2_ctx *c, void **s) {
  if (!c || !s) return set_err(Z2_EINVAL, "null argument");
  *s = (void *)c->stream;
  return Z2_OK;
}

extern "C" int z2_sync(z2_ctx *c) {
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  HIPC(hipStreamSynchronize(c->stream));
  return Z2_OK;
}

// ---------------------------------------------------------------- fields
// conf_set, src/fields.f90:311-427 (therm size == run size: conf_transfer is the identity)
static void conf_set_host(int Lx, int Ly, int T, int absolute, double Q, int32_t *seed, std::vector<double> &sig,
                          double *lam) {
  const int N = Lx * Ly, N2 = 2 * N;
  sig.assign((size_t)N2 * T, 1.0);
  if (absolute == 1) {
    for (int nt = 0; nt < T; ++nt)
      for (int ii = 0; ii < N2; ++ii) sig[(size_t)nt * N2 + ii] = ranf_host(seed) >= 0.5 ? 1.0 : -1.0;
  } else if (absolute == 2) {
    const double PI = acos(-1.0);
    for (int nt = 0; nt < T; ++nt)
      for (int ii = 0; ii < N2; ++ii) {
        double X1 = ranf_host(seed), X2 = ranf_host(seed);
        double X = sqrt(-2.0 * log(X1)) * cos(2.0 * PI * X2);
        sig[(size_t)nt * N2 + ii] = X >= 0.0 ? 1.0 : -1.0;
      }
  } else if (absolute == 4) {
    for (int nt = 0; nt < T; ++nt)
      for (int nc = 0; nc < N2; ++nc) {
        int s = nc / 2, x = s % Lx + 1, y = s / Lx + 1;
        sig[(size_t)nt * N2 + nc] = ((x + y) % 2 == 0) ? 1.0 : -1.0;
      }
  } else if (absolute == 5) {
    for (int ii = 0; ii < N2; ++ii) {
      double v = ranf_host(seed) >= 0.5 ? 1.0 : -1.0;
      for (int nt = 0; nt < T; ++nt) sig[(size_t)nt * N2 + ii] = v;
    }
  }
  // lambda_random_fill + lambda_enforce_sector (src/fields.f90:559-603)
  double prod = 1.0;
  for (int i = 0; i < N; ++i) {
    lam[i] = ranf_host(seed) >= 0.5 ? 1.0 : -1.0;
    prod = prod * lam[i];
  }
  if (prod * (Q >= 0.0 ? 1.0 : -1.0) < 0.0) lam[N - 1] = -lam[N - 1];
}

static int upload_sigma(z2_ctx *c, int chain, const double *sigma) {
  std::vector<int8_t> s((size_t)c->T * c->N2);
  for (size_t k = 0; k < s.size(); ++k) {
    if (sigma[k] != 1.0 && sigma[k] != -1.0) return set_err(Z2_EINVAL, "sigma values must be +-1");
    s[k] = sigma[k] > 0 ? 1 : -1;
  }
  HIPC(hipMemcpyAsync(c->sig + (size_t)chain * s.size(), s.data(), s.size(), hipMemcpyHostToDevice, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return Z2_OK;
}

extern "C" int z2_init_fields(z2_ctx *c) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  std::vector<double> sig;
  std::vector<int32_t> seeds(c->nch);
  for (int b = 0; b < c->nch; ++b) {
    int32_t s = c->desc[b].seed;
    int ab = c->desc[b].absolute;
    if (ab < 1 || ab > 5) return set_err(Z2_EINVAL, "absolute must be 1..5");
    conf_set_host(c->Lx, c->Ly, c->T, ab, c->desc[b].Q, &s, sig, c->lam.data() + (size_t)b * c->N);
    CHK(upload_sigma(c, b, sig.data()));
    seeds[b] = s;
  }
  HIPC(hipMemcpy(c->seed, seeds.data(), sizeof(int32_t) * c->nch, hipMemcpyHostToDevice));
  c->have_fields = true;
  return Z2_OK;
}

extern "C" int z2_set_sigma(z2_ctx *c, int32_t chain, const double *sigma, int32_t seed) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c || !sigma || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];  // storage slot
  CHK(upload_sigma(c, chain, sigma));
  HIPC(hipMemcpy(c->seed + chain, &seed, sizeof(int32_t), hipMemcpyHostToDevice));
  c->have_fields = true;
  return Z2_OK;
}

extern "C" int z2_get_sigma(z2_ctx *c, int32_t chain, double *sigma) {
  if (!c || !sigma || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];  // storage slot
  std::vector<int8_t> s((size_t)c->T * c->N2);
  HIPC(hipMemcpyAsync(s.data(), c->sig + (size_t)chain * s.size(), s.size(), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  for (size_t k = 0; k < s.size(); ++k) sigma[k] = (double)s[k];
  return Z2_OK;
}

extern "C" int z2_get_lambda(z2_ctx *c, int32_t chain, double *lambda) {
  if (!c || !lambda || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  const double *l = c->lam.data() + (size_t)c->phys[chain] * c->N;
  std::copy(l, l + c->N, lambda);
  return Z2_OK;
}

extern "C" int z2_set_lambda(z2_ctx *c, int32_t chain, const double *lambda) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c || !lambda || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  for (int i = 0; i < c->N; ++i)
    if (lambda[i] != 1.0 && lambda[i] != -1.0) return set_err(Z2_EINVAL, "lambda values must be +-1");
  std::copy(lambda, lambda + c->N, c->lam.data() + (size_t)c->phys[chain] * c->N);
  return Z2_OK;
}

// conf_transfer, src/fields.f90:429-507: the field of a thermalisation-size lattice
// (LxT x LyT sites, TT slices) tiled periodically onto the run's lattice (bond of
// site (x, y) in direction d <- bond of site ((x-1) % LxT + 1, (y-1) % LyT + 1) in
// direction d; slice nt <- slice nt - TT beyond TT), lambda likewise by site, then
// the sector fix (lambda_enforce_sector, :580-603)
static int conf_transfer_host(const z2_ctx *c, int LxT, int LyT, int TT, double Q, const double *sig_t,
                              const double *lam_t, std::vector<double> &sig, double *lam) {
  if (LxT < 1 || LyT < 1 || TT < 1 || TT > c->T)
    return set_err(Z2_EINVAL, "thermalisation size %d x %d x %d does not tile %d x %d x %d", LxT, LyT, TT, c->Lx,
                   c->Ly, c->T);
  const int N2 = c->N2, N2T = 2 * LxT * LyT;
  sig.assign((size_t)N2 * c->T, 0.0);
  for (int nt = 0; nt < TT; ++nt)
    for (int b = 0; b < N2; ++b) {
      const int site = b / 2, d = b % 2, x = site % c->Lx, y = site / c->Lx;
      const int bt = 2 * ((y % LyT) * LxT + (x % LxT)) + d;
      sig[(size_t)nt * N2 + b] = sig_t[(size_t)nt * N2T + bt];
    }
  for (int nt = TT; nt < c->T; ++nt)
    for (int b = 0; b < N2; ++b) sig[(size_t)nt * N2 + b] = sig[(size_t)(nt - TT) * N2 + b];
  double prod = 1.0;
  for (int i = 0; i < c->N; ++i) {
    const int x = i % c->Lx, y = i / c->Lx;
    lam[i] = lam_t[(y % LyT) * LxT + (x % LxT)] >= 0.0 ? 1.0 : -1.0;
    prod = prod * lam[i];
  }
  if (prod * (Q >= 0.0 ? 1.0 : -1.0) < 0.0) lam[c->N - 1] = -lam[c->N - 1];
  return Z2_OK;
}

extern "C" int z2_set_sigma_therm(z2_ctx *c, int32_t chain, int32_t NlxTherm, int32_t NlyTherm, int32_t LtrotTherm,
                                  const double *sigma_therm, const double *lambda_therm, int32_t seed) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c || !sigma_therm || !lambda_therm || chain < 0 || chain >= c->nch_all)
    return set_err(Z2_EINVAL, "bad argument");
  std::vector<double> sig, lam(c->N);
  CHK(conf_transfer_host(c, NlxTherm, NlyTherm, LtrotTherm, c->desc[c->phys[chain]].Q, sigma_therm, lambda_therm,
                         sig, lam.data()));
  CHK(z2_set_sigma(c, chain, sig.data(), seed));
  return z2_set_lambda(c, chain, lam.data());
}

extern "C" int z2_init_fields_therm(z2_ctx *c, int32_t NlxTherm, int32_t NlyTherm, int32_t LtrotTherm) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (NlxTherm < 1 || NlyTherm < 1 || LtrotTherm < 1) return set_err(Z2_EINVAL, "bad thermalisation size");
  std::vector<double> sig_t, lam_t((size_t)NlxTherm * NlyTherm);
  for (int k = 0; k < c->nch_all; ++k) {
    const z2_chain_desc &d = c->desc[c->phys[k]];
    if (d.absolute < 1 || d.absolute > 5) return set_err(Z2_EINVAL, "absolute must be 1..5");
    int32_t s = d.seed;
    conf_set_host(NlxTherm, NlyTherm, LtrotTherm, d.absolute, d.Q, &s, sig_t, lam_t.data());
    CHK(z2_set_sigma_therm(c, k, NlxTherm, NlyTherm, LtrotTherm, sig_t.data(), lam_t.data(), s));
  }
  return Z2_OK;
}

extern "C" int z2_get_seed(z2_ctx *c, int32_t chain, int32_t *seed) {
  if (!c || !seed || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];  // storage slot
  HIPC(hipMemcpyAsync(seed, c->seed + chain, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return Z2_OK;
}

static int get_mat(z2_ctx *c, const double *base, int chain, double *out) {
  if (!c || !out || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];  // storage slot
  size_t nn = (size_t)c->N * c->N;
  HIPC(hipMemcpyAsync(out, base + chain * nn, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return Z2_OK;
}

extern "C" int z2_get_green(z2_ctx *c, int32_t chain, double *G) {
  if (c && c->mixed) {
    if (!G || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
    chain = c->phys[chain];  // storage slot
    size_t nn = (size_t)c->N * c->N;
    std::vector<float> f(nn);
    HIPC(hipMemcpyAsync(f.data(), c->Gf + chain * nn, sizeof(float) * nn, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < nn; ++k) G[k] = (double)f[k];
    return Z2_OK;
  }
  return get_mat(c, c ? c->G : nullptr, chain, G);
}

extern "C" int z2_set_green(z2_ctx *c, int32_t chain, const double *G) {
  if (!c || !G || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];  // storage slot
  size_t nn = (size_t)c->N * c->N;
  if (c->mixed) {
    std::vector<float> f(G, G + nn);
    HIPC(hipMemcpyAsync(c->Gf + chain * nn, f.data(), sizeof(float) * nn, hipMemcpyHostToDevice, c->stream));
  } else {
    HIPC(hipMemcpyAsync(c->G + chain * nn, G, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  }
  HIPC(hipStreamSynchronize(c->stream));
  return Z2_OK;
}

extern "C" int z2_get_drift(z2_ctx *c, int32_t chain, double *xmax, double *xmean) {
  if (!c || !xmax || !xmean || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];  // storage slot
  HIPC(hipMemcpyAsync(xmax, c->xmax + chain, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipMemcpyAsync(xmean, c->xmean + chain, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return Z2_OK;
}

extern "C" int z2_get_obs_equal(z2_ctx *c, int32_t chain, double out[7]) {
  if (!c || !out || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];  // storage slot
  HIPC(hipMemcpyAsync(out, c->obs_res + (size_t)chain * RES_N, sizeof(double) * RES_N, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return Z2_OK;
}

extern "C" int z2_get_counters(z2_ctx *c, int32_t chain, uint64_t out[2]) {
  if (!c || !out || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];  // storage slot
  unsigned long long v[2];
  HIPC(hipMemcpyAsync(v, c->accc + 2 * chain, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  out[0] = v[0];
  out[1] = v[1];
  return Z2_OK;
}

// ---------------------------------------------------------------- launch helpers
#define LAUNCH_CHECK()                                                            \
  do {                                                                            \
    hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess) return set_err(Z2_EHIP, "launch %s:%d: %s", __FILE__, __LINE__, hipGetErrorString(e_)); \
  } while (0)

// buffers of the propagated quantities in precision R
template <typename R> struct PB;
template <> struct PB<double> {
  static double *G(z2_ctx *c) { return c->G; }
  static double *UUR(z2_ctx *c) { return c->UUR; }
  static double *UUL(z2_ctx *c) { return c->UUL; }
  static double *Tb(z2_ctx *c) { return c->Tbuf; }
  static double *Vb(z2_ctx *c) { return c->Vbuf; }
};
template <> struct PB<float> {
  static float *G(z2_ctx *c) { return c->Gf; }
  static float *UUR(z2_ctx *c) { return c->UURf; }
  static float *UUL(z2_ctx *c) { return c->UULf; }
  static float *Tb(z2_ctx *c) { return c->Tbf; }
  static float *Vb(z2_ctx *c) { return c->Vbf; }
};

template <typename R>
static KinArgs<R> kin_args(z2_ctx *c, int nt) {
  KinArgs<R> a;
  memset(&a, 0, sizeof(a));
  a.sig = c->sig;
  a.T = c->T;
  a.nt = nt;
  a.cc = c->dcc;
  a.Tbuf = PB<R>::Tb(c);
  a.Vbuf = PB<R>::Vb(c);
  a.nacc = c->nacc;
  a.mu_done = c->mu_done;
  a.upd_stride = c->upd_stride;
  a.lat = c->lat;
  return a;
}

template <typename R>
static KinMat<R> kmat(R *m, int nflag, int flush = 0, int pre1 = 0, int pre2 = 0, int mid1 = 0,
                      int mid2 = 0, int post = 0) {
  KinMat<R> k;
  k.mat = m;
  k.nflag = nflag;
  k.flush = flush;
  k.pre1 = pre1;
  k.pre2 = pre2;
  k.mid1 = mid1;
  k.mid2 = mid2;
  k.post = post;
  k.post2 = 0;
  k.pre_cond = 0;
  return k;
}

static constexpr int FLIP_THREADS = 512;  // must not exceed k_flip's __launch_bounds__

// Kinetic tiles: columns per block of the row-rotation kernel (kin_cb) and rows per
// block of the column-rotation kernel (kin_rb), chosen per context from the
// compiled sizes (z2_set_kin_tiles, the autotuner).  Every element sees the same
// operations in the same order whatever the tile, so G does not depend on it.
template <typename R, int CB, int NC>
static int launch_rows_t(z2_ctx *c, KinArgs<R> &a, int nmat) {
  dim3 grid((c->N + CB - 1) / CB, c->nch, nmat);
  size_t sh = sizeof(double) * CB * c->N + sizeof(int) * 2 * c->N;  // FP64 tile + rotation table
  if (sh > 65536)
    HIPC(hipFuncSetAttribute((const void *)k_kin_rows<R, CB, NC>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  hipLaunchKernelGGL((k_kin_rows<R, CB, NC>), grid, dim3(KIN_NT), sh, c->stream, a);
  LAUNCH_CHECK();
  return Z2_OK;
}

template <typename R, int RB, int NC>
static int launch_cols_t(z2_ctx *c, KinArgs<R> &a, int nmat) {
  dim3 grid((c->N + RB - 1) / RB, c->nch, nmat);
  size_t sh = sizeof(double) * RB * c->N + sizeof(int) * 2 * c->N;  // FP64 tile + rotation table
  if (sh > 65536)
    HIPC(hipFuncSetAttribute((const void *)k_kin_cols<R, RB, NC>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  hipLaunchKernelGGL((k_kin_cols<R, RB, NC>), grid, dim3(KIN_NT), sh, c->stream, a);
  LAUNCH_CHECK();
  return Z2_OK;
}

template <typename R, int NC>
static int launch_rows_n(z2_ctx *c, KinArgs<R> &a, int nmat) {
  return c->kin_cb == 16 ? launch_rows_t<R, 16, NC>(c, a, nmat) : launch_rows_t<R, 8, NC>(c, a, nmat);
}

template <typename R, int NC>
static int launch_cols_n(z2_ctx *c, KinArgs<R> &a, int nmat) {
  return c->kin_rb == 32 ? launch_cols_t<R, 32, NC>(c, a, nmat) : launch_cols_t<R, 8, NC>(c, a, nmat);
}

// the kinetic kernels compiled for N = 64, 144, 256 (L = 8, 12, 16), else for any N
template <typename R>
static int launch_rows(z2_ctx *c, KinArgs<R> &a, int nmat) {
  CHK(prof_begin(c, Z2_K_KIN_ROWS));
  switch (c->N) {
    case 64: CHK((launch_rows_n<R, 64>(c, a, nmat))); break;
    case 144: CHK((launch_rows_n<R, 144>(c, a, nmat))); break;
    case 256: CHK((launch_rows_n<R, 256>(c, a, nmat))); break;
    default: CHK((launch_rows_n<R, 0>(c, a, nmat)));
  }
  CHK(prof_end(c, Z2_K_KIN_ROWS, 2.0 * sizeof(R) * c->N * c->N * (double)nmat * c->nch));
  return Z2_OK;
}

template <typename R>
static int launch_cols(z2_ctx *c, KinArgs<R> &a, int nmat) {
  CHK(prof_begin(c, Z2_K_KIN_COLS));
  switch (c->N) {
    case 64: CHK((launch_cols_n<R, 64>(c, a, nmat))); break;
    case 144: CHK((launch_cols_n<R, 144>(c, a, nmat))); break;
    case 256: CHK((launch_cols_n<R, 256>(c, a, nmat))); break;
    default: CHK((launch_cols_n<R, 0>(c, a, nmat)));
  }
  CHK(prof_end(c, Z2_K_KIN_COLS, 2.0 * sizeof(R) * c->N * c->N * (double)nmat * c->nch));
  return Z2_OK;
}

template <typename R>
static int launch_elem_kin(z2_ctx *c, KinArgs<R> &a, int nmat) {
  int nb = (c->N * c->N + 255) / 256;
  if (nb > 64) nb = 64;
  dim3 grid(nb, c->nch, nmat);
  hipLaunchKernelGGL((k_kin_elem<R>), grid, dim3(256), 0, c->stream, a);
  LAUNCH_CHECK();
  return Z2_OK;
}

template <typename R>
static int launch_flip(z2_ctx *c, int nt, int descending, int mu_on_read, int final_flush = 0) {
  FlipArgs<R> f;
  f.G = PB<R>::G(c);
  f.sig = c->sig;
  f.seed = c->seed;
  f.Tbuf = PB<R>::Tb(c);
  f.Vbuf = PB<R>::Vb(c);
  f.nacc = c->nacc;
  f.acc_count = c->accc;
  f.cc = c->dcc;
  f.apow = c->d_apow;
  f.upd_stride = c->upd_stride;
  f.T = c->T;
  f.nt = nt;
  f.descending = descending;
  f.mu_on_read = mu_on_read;
  f.kdelay = c->kdelay;
  f.mu_done = c->mu_done;
  f.final_flush = final_flush;
  f.force = c->force;
  f.record = c->rec;
  f.dwords = c->dwords;
  f.lat = c->lat;
  f.gbond = c->gmode ? c->gbond : nullptr;
  f.gmask = c->gmode ? c->gmask : nullptr;
  f.glr = c->gmode ? c->glr : nullptr;
  static_assert(FLIP_THREADS == 512, "k_flip is compiled for 512 threads per block");
  if constexpr (sizeof(R) == 4) {
    if (c->N == 256 && c->flip_rg) {
      // G of the chain in the registers of its workgroup (k_flip_rg); no update
      // is left pending, the delay depth does not apply
      const size_t shr = flip_rg_lds_bytes(c->N);
      if (shr > 65536)
        HIPC(hipFuncSetAttribute((const void *)k_flip_rg<256>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shr));
      CHK(prof_begin(c, Z2_K_FLIP));
      hipLaunchKernelGGL((k_flip_rg<256>), dim3(c->nch), dim3(FLIP_THREADS), shr, c->stream, f);
      LAUNCH_CHECK();
      CHK(prof_end(c, Z2_K_FLIP, 0.0));
      return Z2_OK;
    }
  }
  size_t sh = flip_lds_bytes(c->N, c->kdelay);
  const int mpt = (c->N2 + FLIP_THREADS - 1) / FLIP_THREADS;
  // FP32 at N = 64, 144, 256 (L = 8, 12, 16): kernels compiled for that size
  // (constant distances between the pending vectors and between G's columns;
  // measured 14% faster at N = 256; the FP64 kernel compiled this way was slower)
  const int nc = (sizeof(R) == 4 && mpt == 1 && (c->N == 64 || c->N == 144 || c->N == 256)) ? c->N : 0;
  constexpr bool f32 = sizeof(R) == 4;
  const void *fn = nc == 64    ? (const void *)k_flip<R, 1, f32 ? 64 : 0>
                   : nc == 144 ? (const void *)k_flip<R, 1, f32 ? 144 : 0>
                   : nc == 256 ? (const void *)k_flip<R, 1, f32 ? 256 : 0>
                   : (mpt == 1 ? (const void *)k_flip<R, 1> : (mpt == 2 ? (const void *)k_flip<R, 2> : (const void *)k_flip<R, 4>));
  if (mpt > 4) return set_err(Z2_EINVAL, "flip kernel supports 2N <= %d", 4 * FLIP_THREADS);
  if (sh > 65536) HIPC(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  CHK(prof_begin(c, Z2_K_FLIP));
  if (nc == 64) hipLaunchKernelGGL((k_flip<R, 1, f32 ? 64 : 0>), dim3(c->nch), dim3(FLIP_THREADS), sh, c->stream, f);
  else if (nc == 144) hipLaunchKernelGGL((k_flip<R, 1, f32 ? 144 : 0>), dim3(c->nch), dim3(FLIP_THREADS), sh, c->stream, f);
  else if (nc == 256) hipLaunchKernelGGL((k_flip<R, 1, f32 ? 256 : 0>), dim3(c->nch), dim3(FLIP_THREADS), sh, c->stream, f);
  else if (mpt == 1) hipLaunchKernelGGL((k_flip<R, 1>), dim3(c->nch), dim3(FLIP_THREADS), sh, c->stream, f);
  else if (mpt == 2) hipLaunchKernelGGL((k_flip<R, 2>), dim3(c->nch), dim3(FLIP_THREADS), sh, c->stream, f);
  else hipLaunchKernelGGL((k_flip<R, 4>), dim3(c->nch), dim3(FLIP_THREADS), sh, c->stream, f);
  LAUNCH_CHECK();
  CHK(prof_end(c, Z2_K_FLIP, 0.0));  // 4 N^2 flops per accepted update: counted by the caller
  return Z2_OK;
}

// fill with a value (kernels, not hipMemsetAsync: every step of a sweep stays a
// kernel or a copy, so that a sweep can be captured in a HIP graph)
__global__ void __launch_bounds__(256) k_fill(double *x, size_t n, double v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = v;
}

static int dzero(z2_ctx *c, double *x, size_t n) {
  if (!n) return Z2_OK;
  size_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)(nb > 1024 ? 1024 : nb)), dim3(256), 0, c->stream, x, n, 0.0);
  LAUNCH_CHECK();
  return Z2_OK;
}

template <typename Tin, typename Tout>
__global__ void __launch_bounds__(256) k_cvt(const Tin *in, Tout *out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (Tout)in[i];
}

template <typename Tin, typename Tout>
static int cvt(z2_ctx *c, const Tin *in, Tout *out) {
  size_t n = (size_t)c->nch * c->N * c->N;
  int nb = (int)((n + 255) / 256);
  if (nb > 2048) nb = 2048;
  hipLaunchKernelGGL((k_cvt<Tin, Tout>), dim3(nb), dim3(256), 0, c->stream, in, out, n);
  LAUNCH_CHECK();
  return Z2_OK;
}

// per-chain max |A - B| into out (max with the value already there)
template <typename Ra, typename Rb>
__global__ void __launch_bounds__(256) k_maxdiff(const Ra *A, const Rb *B, double *out, int N) {
  const int chain = blockIdx.x;
  const Ra *a = A + (size_t)chain * N * N;
  const Rb *b = B + (size_t)chain * N * N;
  double m = 0.0;
  for (int i = threadIdx.x; i < N * N; i += blockDim.x) {
    double d = fabs((double)a[i] - (double)b[i]);
    m = d > m ? d : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = out[chain];
    for (int w = 0; w < 4; ++w) mm = red[w] > mm ? red[w] : mm;
    out[chain] = mm;
  }
}

static int elem(z2_ctx *c, const double *in, double *out, int mode, const double *d = nullptr,
                const int *p = nullptr, int dstride = 1);
// elem with a scale vector of chain stride dstride * N
static int elem_d(z2_ctx *c, const double *in, double *out, int mode, const double *d, int dstride) {
  return elem(c, in, out, mode, d, nullptr, dstride);
}
static int elem(z2_ctx *c, const double *in, double *out, int mode, const double *d, const int *p, int dstride) {
  ElemArgs e;
  e.in = in;
  e.out = out;
  e.d = d;
  e.p = p;
  e.N = c->N;
  e.mode = mode;
  e.dstride = dstride;
  int nb = (c->N * c->N + 255) / 256;
  if (nb > 64) nb = 64;
  hipLaunchKernelGGL(k_elem, dim3(nb, c->nch), dim3(256), 0, c->stream, e);
  LAUNCH_CHECK();
  return Z2_OK;
}

// general batched GEMM: C(M x Nc) = alpha op(A) op(B) + beta C, chain strides sA/sB/sC
static GemmArgs gemm_args(int M, int Nc, int K, const double *A, int lda, long sA, int ta, const double *B, int ldb,
                          long sB, int tb, double *C, int ldc, long sC, double alpha, double beta, int triA, int triB) {
  GemmArgs g;
  g.A = A;
  g.B = B;
  g.C = C;
  g.sA = sA;
  g.sB = sB;
  g.sC = sC;
  g.M = M;
  g.Nc = Nc;
  g.K = K;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = ldc;
  g.ta = ta;
  g.tb = tb;
  g.triA = triA;
  g.triB = triB;
  g.alpha = alpha;
  g.beta = beta;
  return g;
}

static int gemm_g(z2_ctx *c, int M, int Nc, int K, const double *A, int lda, long sA, int ta,
                  const double *B, int ldb, long sB, int tb, double *C, int ldc, long sC, double alpha = 1.0,
                  double beta = 0.0, int triA = 0, int triB = 0) {
  const GemmArgs g = gemm_args(M, Nc, K, A, lda, sA, ta, B, ldb, sB, tb, C, ldc, sC, alpha, beta, triA, triB);
  dim3 grid((M + GT - 1) / GT, (Nc + GT - 1) / GT, c->nch);
  CHK(prof_begin(c, Z2_K_GEMM));
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, c->stream, g, g, c->nch);
  LAUNCH_CHECK();
  CHK(prof_end(c, Z2_K_GEMM, 2.0 * M * Nc * (double)K * c->nch));
  return Z2_OK;
}

// two independent square N x N GEMMs on per-chain matrices in one launch
static int gemm2(z2_ctx *c, const double *A0, int ta0, const double *B0, int tb0, double *C0, const double *A1, int ta1,
                 const double *B1, int tb1, double *C1) {
  const int N = c->N;
  const long nn = (long)N * N;
  const GemmArgs g0 = gemm_args(N, N, N, A0, N, nn, ta0, B0, N, nn, tb0, C0, N, nn, 1.0, 0.0, 0, 0);
  const GemmArgs g1 = gemm_args(N, N, N, A1, N, nn, ta1, B1, N, nn, tb1, C1, N, nn, 1.0, 0.0, 0, 0);
  dim3 grid((N + GT - 1) / GT, (N + GT - 1) / GT, 2 * c->nch);
  CHK(prof_begin(c, Z2_K_GEMM));
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, c->stream, g0, g1, c->nch);
  LAUNCH_CHECK();
  CHK(prof_end(c, Z2_K_GEMM, 4.0 * N * N * (double)N * c->nch));
  return Z2_OK;
}

// square N x N GEMM on per-chain matrices
static int gemm(z2_ctx *c, const double *A, int ta, const double *B, int tb, double *C, int triA = 0,
                int triB = 0) {
  const int N = c->N;
  const long nn = (long)N * N;
  return gemm_g(c, N, N, N, A, N, nn, ta, B, N, nn, tb, C, N, nn, 1.0, 0.0, triA, triB);
}

// register-blocked QR kernels for N <= 1024 (RPL = ceil(N/64) rows per lane,
// CPW columns per wavefront pass, sized to stay within 256 VGPRs), plain ones above
#define QR_CASES(X)                                                                              \
  X(1, 8) X(2, 8) X(3, 8) X(4, 8) X(5, 4) X(6, 4) X(7, 4) X(8, 4) X(9, 2) X(10, 2) X(11, 2) X(12, 2) \
  X(13, 2) X(14, 2) X(15, 2) X(16, 2)

static int launch_qrp(int N, dim3 grid, size_t sh, hipStream_t st, QrpArgs q) {
  const int rpl = (N + 63) / 64;
  switch (rpl) {
#define QCASE(R, W) case R: hipLaunchKernelGGL((k_qrp_reg<R, W>), grid, dim3(QR_REG_THREADS), sh, st, q); break;
    QR_CASES(QCASE)
#undef QCASE
    default: hipLaunchKernelGGL(k_qrp, grid, dim3(1024), sh, st, q);
  }
  LAUNCH_CHECK();
  return Z2_OK;
}

static int launch_orgqr(int N, dim3 grid, hipStream_t st, const double *A, const double *tau, double *Q) {
  const int rpl = (N + 63) / 64;
  switch (rpl) {
#define OCASE(R, W) case R: hipLaunchKernelGGL((k_orgqr_reg<R, W>), grid, dim3(QR_REG_THREADS), 0, st, A, tau, Q, N); break;
    QR_CASES(OCASE)
#undef OCASE
    default: hipLaunchKernelGGL(k_orgqr, grid, dim3(1024), 0, st, A, tau, Q, N);
  }
  LAUNCH_CHECK();
  return Z2_OK;
}

// QR of the chains' n x n matrices A (column-pivoted unless nopiv): R with rows
// scaled to unit diagonal magnitude in the upper triangle, D = |diag R|
// panel width and trailing tile of the blocked QR that fit the LDS for order n (0: none)
static bool qrb_config(size_t lds_max, int n, int *nb, int *tc) {
  static const int cand[3][2] = {{32, 32}, {32, 16}, {16, 16}};
  if (n > 512) return false;  // register panel: at most 8 rows per lane
  for (auto &k : cand)
    if (qrb_lds_bytes(n, k[0], k[1]) <= lds_max) {
      *nb = k[0];
      *tc = k[1];
      return true;
    }
  return false;
}

// one launch of the blocked QR (orgqr = 0) or of its explicit Q (orgqr = 1)
template <int NB, int RPL, int ORG>
static int launch_qrb_t(int nmat, size_t sh, hipStream_t st, const QrbArgs &q) {
  const void *fn = ORG ? (const void *)k_orgqr_blk<NB, RPL> : (const void *)k_qr_blk<NB, RPL>;
  if (sh > 65536) HIPC(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  if (ORG) hipLaunchKernelGGL((k_orgqr_blk<NB, RPL>), dim3(nmat), dim3(512), sh, st, q);
  else hipLaunchKernelGGL((k_qr_blk<NB, RPL>), dim3(nmat), dim3(512), sh, st, q);
  LAUNCH_CHECK();
  return Z2_OK;
}

template <int NB, int ORG>
static int launch_qrb_nb(int rpl, int nmat, size_t sh, hipStream_t st, const QrbArgs &q) {
  switch (rpl) {
    case 1: return launch_qrb_t<NB, 1, ORG>(nmat, sh, st, q);
    case 2: return launch_qrb_t<NB, 2, ORG>(nmat, sh, st, q);
    case 3: return launch_qrb_t<NB, 3, ORG>(nmat, sh, st, q);
    case 4: return launch_qrb_t<NB, 4, ORG>(nmat, sh, st, q);
    case 5: return launch_qrb_t<NB, 5, ORG>(nmat, sh, st, q);
    case 6: return launch_qrb_t<NB, 6, ORG>(nmat, sh, st, q);
    case 7: return launch_qrb_t<NB, 7, ORG>(nmat, sh, st, q);
    case 8: return launch_qrb_t<NB, 8, ORG>(nmat, sh, st, q);
  }
  return set_err(Z2_EINVAL, "blocked QR supports N <= 512");
}

static int launch_qrb(int orgqr, int nb, int nmat, size_t sh, hipStream_t st, const QrbArgs &q) {
  const int rpl = (q.N + 63) / 64;
  if (orgqr) return nb == 32 ? launch_qrb_nb<32, 1>(rpl, nmat, sh, st, q) : launch_qrb_nb<16, 1>(rpl, nmat, sh, st, q);
  return nb == 32 ? launch_qrb_nb<32, 0>(rpl, nmat, sh, st, q) : launch_qrb_nb<16, 0>(rpl, nmat, sh, st, q);
}

// explicit Q over many workgroups per matrix (panel T factors, then one launch per panel)
template <int NB, int RPL>
static int launch_orgqr_split_t(int n, int tc, int nmat, size_t sh, hipStream_t st, const QrbArgs &q, double *Tws,
                                int apply = 0, int have_t = 0) {
  const int npan = (n + NB - 1) / NB;
  if (sh > 65536) {
    HIPC(hipFuncSetAttribute((const void *)k_orgqr_t<NB, RPL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    HIPC(hipFuncSetAttribute((const void *)k_orgqr_tiles<NB, RPL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    HIPC(hipFuncSetAttribute((const void *)k_qt_apply<NB, RPL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  }
  if (!have_t) {  // the factorization's own T factors are not at hand: rebuild them
    hipLaunchKernelGGL((k_orgqr_t<NB, RPL>), dim3(npan, nmat), dim3(512), sh, st, q, Tws);
    LAUNCH_CHECK();
  }
  if (apply)
    hipLaunchKernelGGL((k_qt_apply<NB, RPL>), dim3((n + tc - 1) / tc, nmat), dim3(512), sh, st, q,
                       (const double *)Tws);
  else
    hipLaunchKernelGGL((k_orgqr_tiles<NB, RPL>), dim3((n + tc - 1) / tc, nmat), dim3(512), sh, st, q,
                       (const double *)Tws);
  LAUNCH_CHECK();
  return Z2_OK;
}

template <int NB>
static int launch_orgqr_split_nb(int rpl, int n, int tc, int nmat, size_t sh, hipStream_t st, const QrbArgs &q,
                                 double *Tws, int apply = 0, int have_t = 0) {
  switch (rpl) {
    case 1: return launch_orgqr_split_t<NB, 1>(n, tc, nmat, sh, st, q, Tws, apply, have_t);
    case 2: return launch_orgqr_split_t<NB, 2>(n, tc, nmat, sh, st, q, Tws, apply, have_t);
    case 3: return launch_orgqr_split_t<NB, 3>(n, tc, nmat, sh, st, q, Tws, apply, have_t);
    case 4: return launch_orgqr_split_t<NB, 4>(n, tc, nmat, sh, st, q, Tws, apply, have_t);
    case 5: return launch_orgqr_split_t<NB, 5>(n, tc, nmat, sh, st, q, Tws, apply, have_t);
    case 6: return launch_orgqr_split_t<NB, 6>(n, tc, nmat, sh, st, q, Tws, apply, have_t);
    case 7: return launch_orgqr_split_t<NB, 7>(n, tc, nmat, sh, st, q, Tws, apply, have_t);
    case 8: return launch_orgqr_split_t<NB, 8>(n, tc, nmat, sh, st, q, Tws, apply, have_t);
  }
  return set_err(Z2_EINVAL, "blocked QR supports N <= 512");
}

// workgroups per chain that copy the ordered columns beside panel 0
constexpr int QR_GATHER_WG = 8;

// the QR over many workgroups per matrix (order; panel 0 with the copy of the
// ordered columns to W; then per panel: trailing update, next panel)
template <int NB, int RPL>
static int launch_qr_split_t(z2_ctx *c, size_t sh, const QrbArgs &q) {
  const int n = q.N, tc = q.tc, npan = (n + NB - 1) / NB;
  if (sh > 65536) {
    HIPC(hipFuncSetAttribute((const void *)k_qr_panel<NB, RPL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    HIPC(hipFuncSetAttribute((const void *)k_qr_panel0<NB, RPL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    HIPC(hipFuncSetAttribute((const void *)k_qr_trail<NB, RPL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  }
  hipLaunchKernelGGL((k_qr_order<RPL>), dim3(c->nch), dim3(512), sizeof(double) * n, c->stream, q);
  LAUNCH_CHECK();
  const int lookahead = tc == NB;  // tile 0 holds exactly the next panel's columns
  QrbArgs qw = q;
  if (n > NB) {
    // panel 0 into W with W = A P beyond it, one launch; the first trail launch
    // copies panel 0 to A
    qw.p0w = 1;
    hipLaunchKernelGGL((k_qr_panel0<NB, RPL>), dim3(1 + QR_GATHER_WG, c->nch), dim3(512), sh, c->stream, qw, c->ortw);
    LAUNCH_CHECK();
  } else {
    CHK(elem(c, q.A, q.W, 5, nullptr, q.ipvt));  // W = A P
    hipLaunchKernelGGL((k_qr_panel<NB, RPL>), dim3(c->nch), dim3(512), sh, c->stream, q, c->ortw, 0);
    LAUNCH_CHECK();
  }
  for (int p = 0; p < npan; ++p) {
    const int j0 = p * NB, w = n - j0 < NB ? n - j0 : NB;
    if (j0 + w >= n) break;
    const int ntile = (n - j0 - w + tc - 1) / tc;
    hipLaunchKernelGGL((k_qr_trail<NB, RPL>), dim3(ntile, c->nch), dim3(512), sh, c->stream, p == 0 ? qw : q, c->ortw,
                       p, lookahead);
    LAUNCH_CHECK();
    if (!lookahead) {
      hipLaunchKernelGGL((k_qr_panel<NB, RPL>), dim3(c->nch), dim3(512), sh, c->stream, q, c->ortw, p + 1);
      LAUNCH_CHECK();
    }
  }
  return Z2_OK;
}

template <int NB>
static int launch_qr_split_nb(z2_ctx *c, int rpl, size_t sh, const QrbArgs &q) {
  switch (rpl) {
    case 1: return launch_qr_split_t<NB, 1>(c, sh, q);
    case 2: return launch_qr_split_t<NB, 2>(c, sh, q);
    case 3: return launch_qr_split_t<NB, 3>(c, sh, q);
    case 4: return launch_qr_split_t<NB, 4>(c, sh, q);
    case 5: return launch_qr_split_t<NB, 5>(c, sh, q);
    case 6: return launch_qr_split_t<NB, 6>(c, sh, q);
    case 7: return launch_qr_split_t<NB, 7>(c, sh, q);
    case 8: return launch_qr_split_t<NB, 8>(c, sh, q);
  }
  return set_err(Z2_EINVAL, "blocked QR supports N <= 512");
}

static double *qr_workspace(z2_ctx *c, int n) { return n == c->N ? c->qrw : (n == 2 * c->N ? c->qrw2 : nullptr); }

static int qrp_n(z2_ctx *c, double *A, double *D, double *tau, int *ipvt, int n, int nopiv) {
  int nb = 0, tc = 0;
  double *ws = qr_workspace(c, n);
  if (c->qr_mode == Z2_QR_BLOCKED && ws && qrb_config(c->lds_max, n, &nb, &tc)) {
    QrbArgs q;
    memset(&q, 0, sizeof(q));
    q.A = A;
    q.W = ws;
    q.tau = tau;
    q.ipvt = ipvt;
    q.D = D;
    q.N = n;
    q.tc = tc;
    q.nopiv = nopiv;
    CHK(prof_begin(c, Z2_K_QRP));
    if (n == c->N && c->ortw) {
      const int rpl = (n + 63) / 64;
      const size_t sh = qrb_lds_bytes(n, nb, tc);
      if (nb == 32) CHK(launch_qr_split_nb<32>(c, rpl, sh, q));
      else CHK(launch_qr_split_nb<16>(c, rpl, sh, q));
      c->ortw_src = A;  // every panel's T factor of this factorization is now in ortw
    } else {
      CHK(launch_qrb(0, nb, c->nch, qrb_lds_bytes(n, nb, tc), c->stream, q));
    }
    CHK(prof_end(c, Z2_K_QRP, 4.0 / 3.0 * n * n * (double)n * c->nch));
    return Z2_OK;
  }
  QrpArgs q;
  q.A = A;
  q.tau = tau;
  q.ipvt = ipvt;
  q.D = D;
  q.N = n;
  q.nopiv = nopiv;
  size_t sh = sizeof(double) * (2 * n + 32) + sizeof(int) * 32;
  CHK(prof_begin(c, Z2_K_QRP));
  CHK(launch_qrp(n, dim3(c->nch), sh, c->stream, q));
  CHK(prof_end(c, Z2_K_QRP, 4.0 / 3.0 * n * n * (double)n * c->nch));
  return Z2_OK;
}

static int qrp(z2_ctx *c, double *A, double *D) { return qrp_n(c, A, D, c->tau, c->ipvt, c->N, 0); }

static int orgqr_n(z2_ctx *c, const double *A, const double *tau, double *Qo, int n) {
  int nb = 0, tc = 0;
  if (c->qr_mode == Z2_QR_BLOCKED && qr_workspace(c, n) && qrb_config(c->lds_max, n, &nb, &tc)) {
    QrbArgs q;
    memset(&q, 0, sizeof(q));
    q.Ain = A;
    q.tin = tau;
    q.Q = Qo;
    q.N = n;
    q.tc = tc;
    CHK(prof_begin(c, Z2_K_ORGQR));
    if (n == c->N && c->ortw) {
      const int rpl = (n + 63) / 64;
      const size_t sh = qrb_lds_bytes(n, nb, tc);
      const int have_t = c->ortw_src == A;
      if (nb == 32) CHK(launch_orgqr_split_nb<32>(rpl, n, tc, c->nch, sh, c->stream, q, c->ortw, 0, have_t));
      else CHK(launch_orgqr_split_nb<16>(rpl, n, tc, c->nch, sh, c->stream, q, c->ortw, 0, have_t));
      c->ortw_src = A;  // rebuilt or kept: ortw holds A's T factors
    } else {
      CHK(launch_qrb(1, nb, c->nch, qrb_lds_bytes(n, nb, tc), c->stream, q));
    }
    CHK(prof_end(c, Z2_K_ORGQR, 4.0 / 3.0 * n * n * (double)n * c->nch));
    return Z2_OK;
  }
  CHK(prof_begin(c, Z2_K_ORGQR));
  CHK(launch_orgqr(n, dim3(c->nch), c->stream, A, tau, Qo));
  CHK(prof_end(c, Z2_K_ORGQR, 4.0 / 3.0 * n * n * (double)n * c->nch));
  return Z2_OK;
}

static int orgqr(z2_ctx *c, const double *A, double *Qo) { return orgqr_n(c, A, c->tau, Qo, c->N); }

// X = Q^T B^T (xtrans = 0, B read transposed) or X = B^T Q = (Q^T B)^T (xtrans = 1),
// Q from the last blocked QR of A (reflectors in A, tau): the block reflectors are
// applied to the tiles of B directly; without the split blocked path, Q is formed
// and multiplied
static int apply_qt(z2_ctx *c, const double *A, const double *B, double *X, int xtrans) {
  const int n = c->N;
  int nb = 0, tc = 0;
  if (c->qr_mode == Z2_QR_BLOCKED && c->ortw && qr_workspace(c, n) && qrb_config(c->lds_max, n, &nb, &tc) &&
      tc <= 32) {
    QrbArgs q;
    memset(&q, 0, sizeof(q));
    q.Ain = A;
    q.tin = c->tau;
    q.Bq = B;
    q.Xq = X;
    q.btrans = xtrans ? 0 : 1;
    q.xtrans = xtrans;
    q.N = n;
    q.tc = tc;
    CHK(prof_begin(c, Z2_K_QT_APPLY));
    const int rpl = (n + 63) / 64;
    const size_t sh = qrb_lds_bytes(n, nb, tc);
    const int have_t = c->ortw_src == A;
    if (nb == 32) CHK(launch_orgqr_split_nb<32>(rpl, n, tc, c->nch, sh, c->stream, q, c->ortw, 1, have_t));
    else CHK(launch_orgqr_split_nb<16>(rpl, n, tc, c->nch, sh, c->stream, q, c->ortw, 1, have_t));
    c->ortw_src = A;
    // Q^T applied to an n x n right-hand side: sum_j 4 (n - j) n = 2 n^3 flops
    CHK(prof_end(c, Z2_K_QT_APPLY, 2.0 * n * n * (double)n * c->nch));
    return Z2_OK;
  }
  CHK(orgqr(c, A, c->Q));
  if (xtrans) return gemm(c, B, 1, c->Q, 0, X, 0, 0);
  return gemm(c, c->Q, 1, B, 1, X, 0, 0);
}

// n x n triangular solves for nb matrices; the right-hand-side block shrinks so that
// (cbt + 1) n doubles fit the LDS of one workgroup
template <int MJ>
static int launch_trsm_reg(int n, int nb, hipStream_t st, const double *R, double *X, int right,
                           const double *dsc = nullptr, const int *perm = nullptr, double *Y = nullptr) {
  const size_t sh = sizeof(double) * 16 * (size_t)n;
  if (sh > 65536)
    HIPC(hipFuncSetAttribute((const void *)k_trsm_reg<MJ>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  dim3 grid((n + 15) / 16, nb);
  if (right) hipLaunchKernelGGL(k_trsm_reg<MJ>, grid, dim3(256), sh, st, R, X, n, n, 1, 1, dsc, perm, Y);
  else hipLaunchKernelGGL(k_trsm_reg<MJ>, grid, dim3(256), sh, st, R, X, n, 1, n, 0, dsc, perm, Y);
  LAUNCH_CHECK();
  return Z2_OK;
}

static int launch_trsm(int n, int nb, hipStream_t st, size_t lds_max, const double *R, double *X, int right) {
  // register kernel: X column in 16 lanes, up to 64 rows per lane
  if (sizeof(double) * 16 * (size_t)n <= lds_max) {
    if (n <= 64) return launch_trsm_reg<4>(n, nb, st, R, X, right);
    if (n <= 128) return launch_trsm_reg<8>(n, nb, st, R, X, right);
    if (n <= 256) return launch_trsm_reg<16>(n, nb, st, R, X, right);
    if (n <= 512) return launch_trsm_reg<32>(n, nb, st, R, X, right);
    if (n <= 1024) return launch_trsm_reg<64>(n, nb, st, R, X, right);
  }
  int cbt = CBT;
  while (cbt > 1 && sizeof(double) * (size_t)(cbt + 1) * n > lds_max) --cbt;
  size_t sh = sizeof(double) * (size_t)(cbt + 1) * n;
  if (sh > lds_max) return set_err(Z2_EINVAL, "triangular solve of order %d exceeds the LDS", n);
  if (sh > 65536) HIPC(hipFuncSetAttribute((const void *)k_trsm, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  dim3 grid((n + cbt - 1) / cbt, nb);
  if (right) hipLaunchKernelGGL(k_trsm, grid, dim3(256), sh, st, R, X, n, n, 1, 1, cbt);
  else hipLaunchKernelGGL(k_trsm, grid, dim3(256), sh, st, R, X, n, 1, n, 0, cbt);
  LAUNCH_CHECK();
  return Z2_OK;
}

static int trsm(z2_ctx *c, const double *R, double *X, int right, int n = 0) {
  CHK(prof_begin(c, Z2_K_TRSM));
  const int m = n ? n : c->N;
  CHK(launch_trsm(m, c->nch, c->stream, c->lds_max, R, X, right));
  CHK(prof_end(c, Z2_K_TRSM, (double)m * m * (double)m * c->nch));
  return Z2_OK;
}

// Y = P R^{-1} D^{-1} X (right = 0) or Y = X D^{-1} R^{-1} P^T (right = 1), N x N:
// the row (column) division by D, the triangular solve and the row (column)
// permutation, in one launch where the register solve applies (X is left
// unchanged then), else the three launches it fuses
static int trsm_dp(z2_ctx *c, const double *R, double *X, int right, const double *D, const int *perm, double *Y) {
  const int n = c->N;
  if (sizeof(double) * 16 * (size_t)n <= c->lds_max && n <= 1024) {
    CHK(prof_begin(c, Z2_K_TRSM));
    if (n <= 64) CHK(launch_trsm_reg<4>(n, c->nch, c->stream, R, X, right, D, perm, Y));
    else if (n <= 128) CHK(launch_trsm_reg<8>(n, c->nch, c->stream, R, X, right, D, perm, Y));
    else if (n <= 256) CHK(launch_trsm_reg<16>(n, c->nch, c->stream, R, X, right, D, perm, Y));
    else if (n <= 512) CHK(launch_trsm_reg<32>(n, c->nch, c->stream, R, X, right, D, perm, Y));
    else CHK(launch_trsm_reg<64>(n, c->nch, c->stream, R, X, right, D, perm, Y));
    CHK(prof_end(c, Z2_K_TRSM, (double)n * n * (double)n * c->nch));
    return Z2_OK;
  }
  CHK(elem(c, X, X, right ? 9 : 8, D));
  CHK(trsm(c, R, X, right));
  return elem(c, X, Y, right ? 7 : 6, nullptr, perm);
}

// device copies as a kernel (every step of a sweep stays a kernel launch, so that
// a sweep captures into a HIP graph of kernel nodes only)
__global__ void __launch_bounds__(256) k_dcopy(double *__restrict__ dst, const double *__restrict__ src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

static int dcopy(z2_ctx *c, double *dst, const double *src, size_t n) {
  if (!n) return Z2_OK;
  size_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_dcopy, dim3((unsigned)(nb > 2048 ? 2048 : nb)), dim3(256), 0, c->stream, dst, src, n);
  LAUNCH_CHECK();
  return Z2_OK;
}

// three copies in one launch (a UDV triple to or from its checkpoint slot): index
// i of the concatenation of the three ranges
struct CopySeg {
  double *dst;
  const double *src;
  size_t n;
};

__global__ void __launch_bounds__(256) k_dcopy3(CopySeg a, CopySeg b, CopySeg d) {
  const size_t ab = a.n + b.n, total = ab + d.n;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    if (i < a.n) a.dst[i] = a.src[i];
    else if (i < ab) b.dst[i - a.n] = b.src[i - a.n];
    else d.dst[i - ab] = d.src[i - ab];
  }
}

static int dcopy3(z2_ctx *c, CopySeg a, CopySeg b, CopySeg d) {
  const size_t n = a.n + b.n + d.n;
  if (!n) return Z2_OK;
  size_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_dcopy3, dim3((unsigned)(nb > 2048 ? 2048 : nb)), dim3(256), 0, c->stream, a, b, d);
  LAUNCH_CHECK();
  return Z2_OK;
}

// ---------------------------------------------------------------- stabilization
// stab_UR, src/stabilization.f90:70-94, on the right propagator (U, D, V)
static int stab_UR_buf(z2_ctx *c, double *U, double *D, double *V) {
  CHK(elem(c, U, c->M, 2, D));
  CHK(qrp(c, c->M, D));
  CHK(elem(c, V, c->X, 4, nullptr, c->ipvt));
  CHK(gemm(c, c->M, 0, c->X, 0, V, 1, 0));
  CHK(orgqr(c, c->M, U));
  return Z2_OK;
}

static int stab_UR(z2_ctx *c) { return stab_UR_buf(c, c->UUR, c->DUR, c->VUR); }

// stab_UL, src/stabilization.f90:96-120
static int stab_UL(z2_ctx *c) {
  CHK(elem(c, c->UUL, c->M, 3, c->DUL));
  CHK(qrp(c, c->M, c->DUL));
  CHK(elem(c, c->VUL, c->X, 5, nullptr, c->ipvt));
  CHK(gemm(c, c->X, 0, c->M, 1, c->VUL, 0, 1));
  CHK(orgqr(c, c->M, c->Q));
  CHK(elem(c, c->Q, c->UUL, 1));
  return Z2_OK;
}

// stab_green, src/stabilization.f90:122-200 -> c->Gn
// D-split recompute (no counterpart in the reference; opt-in, Z2_STAB_DSPLIT):
// G = UL^T DLb^{-1} M'^{-1} DRb^{-1} UR^T with M' = DRb^{-1} (UL UR)^T DLb^{-1} + DRs VR VL DLs
// and M' P = Q D V as in the reference path, so M'^{-1} = P V^{-1} D^{-1} Q^T.
static int stab_green_split(z2_ctx *c) {
  CHK(gemm(c, c->VUR, 0, c->VUL, 0, c->Y));
  CHK(gemm(c, c->UUR, 1, c->UUL, 1, c->X));
  hipLaunchKernelGGL(k_dsplit, dim3(c->nch), dim3(256), 0, c->stream, (const double *)c->DUR,
                     (const double *)c->DUL, c->dsc, c->N);
  LAUNCH_CHECK();
  {
    const int nb = c->N < 64 ? c->N : 64;
    hipLaunchKernelGGL(k_build_m_split, dim3(nb, c->nch), dim3(256), 0, c->stream, (const double *)c->X,
                       (const double *)c->Y, (const double *)c->dsc, c->M, c->N);
    LAUNCH_CHECK();
  }
  CHK(qrp(c, c->M, c->Dtmp));
  CHK(orgqr(c, c->M, c->Q));
  // DRb^{-1} UR^T = (UR DRb^{-1})^T: divide the columns of UR by DRb
  CHK(elem_d(c, c->UUR, c->Gn, 9, c->dsc, 4));
  CHK(gemm(c, c->Q, 1, c->Gn, 1, c->X));           // Q^T DRb^{-1} UR^T
  CHK(trsm_dp(c, c->M, c->X, 0, c->Dtmp, c->ipvt, c->Y));  // P V^{-1} D^{-1} ...
  CHK(elem_d(c, c->Y, c->Y, 8, c->dsc + 2 * (size_t)c->N, 4));  // DLb^{-1} ...
  CHK(gemm(c, c->UUL, 1, c->Y, 0, c->Gn));         // UL^T ...
  return Z2_OK;
}

static int stab_green(z2_ctx *c, int nt) {
  if (c->stab_mode == Z2_STAB_DSPLIT) return stab_green_split(c);
  CHK(gemm2(c, c->VUR, 0, c->VUL, 0, c->Y, c->UUR, 1, c->UUL, 1, c->X));  // Y = VR VL, X = UR^T UL^T
  const int trans = nt > c->T / 2;
  {
    int nb = (c->N * c->N + 255) / 256;
    if (nb > 64) nb = 64;
    hipLaunchKernelGGL(k_build_m, dim3(nb, c->nch), dim3(256), 0, c->stream, (const double *)c->X,
                       (const double *)c->Y, (const double *)c->DUR, (const double *)c->DUL, c->M, c->N, trans);
    LAUNCH_CHECK();
  }
  CHK(qrp(c, c->M, c->Dtmp));
  if (nt < c->T / 2 + 1) {
    CHK(apply_qt(c, c->M, c->UUR, c->X, 0));  // X = Q^T UUR^T
    CHK(trsm_dp(c, c->M, c->X, 0, c->Dtmp, c->ipvt, c->Y));
    CHK(gemm(c, c->UUL, 1, c->Y, 0, c->Gn));
  } else {
    CHK(apply_qt(c, c->M, c->UUL, c->X, 1));  // X = UUL^T Q
    CHK(trsm_dp(c, c->M, c->X, 1, c->Dtmp, c->ipvt, c->Y));
    CHK(gemm(c, c->Y, 0, c->UUR, 1, c->Gn));
  }
  return Z2_OK;
}

static int drift_replace(z2_ctx *c, int accumulate) {
  // partial maxima in Dtmp (N per chain), free once the recomputed G is in Gn
  const int nb = c->N < 16 ? c->N : 16;
  if (c->mixed)
    hipLaunchKernelGGL((k_drift_replace<float>), dim3(nb, c->nch), dim3(256), 0, c->stream, (const double *)c->Gn,
                       c->Gf, c->Dtmp, c->N, c->N);
  else
    hipLaunchKernelGGL((k_drift_replace<double>), dim3(nb, c->nch), dim3(256), 0, c->stream, (const double *)c->Gn,
                       c->G, c->Dtmp, c->N, c->N);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_drift_finish, dim3((c->nch + 63) / 64), dim3(64), 0, c->stream, (const double *)c->Dtmp,
                     c->N, nb, c->xmax, c->xmean, c->nch, accumulate);
  LAUNCH_CHECK();
  return Z2_OK;
}

static inline size_t slot(const z2_ctx *c, int st) { return (size_t)st * c->nch_all * c->N * c->N; }
static inline size_t slotd(const z2_ctx *c, int st) { return (size_t)st * c->nch_all * c->N; }

extern "C" int z2_wrap_pre(z2_ctx *c, int32_t nt) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c || nt < 0 || nt > c->T || nt % c->w) return set_err(Z2_EINVAL, "incorrect preortho time slice %d", nt);
  if (!c->have_fields) return set_err(Z2_ESTATE, "fields not initialised (z2_init_fields / z2_set_sigma)");
  const int st = nt / c->w;
  const size_t nn = (size_t)c->nch * c->N * c->N, nd = (size_t)c->nch * c->N;
  if (nt != 0) {
    if (c->mixed && !c->u64) CHK(cvt(c, (const float *)c->UURf, c->UUR));
    CHK(stab_UR(c));
  }
  CHK(dcopy3(c, {c->URl + slot(c, st), c->UUR, nn}, {c->VRl + slot(c, st), c->VUR, nn},
             {c->DRl + slotd(c, st), c->DUR, nd}));
  if (nt == c->T) {
    CHK(stab_green(c, nt));
    if (c->mixed) CHK(cvt(c, (const double *)c->Gn, c->Gf));
    else CHK(dcopy(c, c->G, c->Gn, nn));
    // the end of Local_sweep_pre: G(beta) = G(0) is in place
    c->isb.assign(c->nch_all, 1);  // the reference's main starts every chain with is_beta = .true.
    c->have_pre = true;
    c->left_done = 0;
  }
  if (c->mixed && !c->u64) CHK(cvt(c, (const double *)c->UUR, c->UURf));
  return Z2_OK;
}

extern "C" int z2_wrap_L(z2_ctx *c, int32_t nt, int32_t flag) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c || nt < 0 || nt > c->T || nt % c->w) return set_err(Z2_EINVAL, "incorrect ortholeft time slice %d", nt);
  const int st = nt / c->w;
  const size_t nn = (size_t)c->nch * c->N * c->N, nd = (size_t)c->nch * c->N;
  CHK(dcopy3(c, {c->UUR, c->URl + slot(c, st), nn}, {c->VUR, c->VRl + slot(c, st), nn},
             {c->DUR, c->DRl + slotd(c, st), nd}));
  if (nt != c->T) {
    if (c->mixed && !c->u64) CHK(cvt(c, (const float *)c->UULf, c->UUL));
    CHK(stab_UL(c));
    CHK(stab_green(c, nt));
    CHK(drift_replace(c, flag));
  }
  CHK(dcopy3(c, {c->ULl + slot(c, st), c->UUL, nn}, {c->VLl + slot(c, st), c->VUL, nn},
             {c->DLl + slotd(c, st), c->DUL, nd}));
  if (c->mixed && !c->u64) CHK(cvt(c, (const double *)c->UUL, c->UULf));
  if (nt == 0) c->left_done = c->nch == c->nch_all ? 1 : -1;
  return Z2_OK;
}

extern "C" int z2_wrap_R(z2_ctx *c, int32_t nt, int32_t flag) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c || nt < 0 || nt > c->T || nt % c->w) return set_err(Z2_EINVAL, "incorrect orthoright time slice %d", nt);
  const int st = nt / c->w;
  const size_t nn = (size_t)c->nch * c->N * c->N, nd = (size_t)c->nch * c->N;
  CHK(dcopy3(c, {c->UUL, c->ULl + slot(c, st), nn}, {c->VUL, c->VLl + slot(c, st), nn},
             {c->DUL, c->DLl + slotd(c, st), nd}));
  if (nt != 0) {
    if (c->mixed && !c->u64) CHK(cvt(c, (const float *)c->UURf, c->UUR));
    CHK(stab_UR(c));
    CHK(stab_green(c, nt));
    CHK(drift_replace(c, flag));
  }
  CHK(dcopy3(c, {c->URl + slot(c, st), c->UUR, nn}, {c->VRl + slot(c, st), c->VUR, nn},
             {c->DRl + slotd(c, st), c->DUR, nd}));
  if (c->mixed && !c->u64) CHK(cvt(c, (const double *)c->UUR, c->UURf));
  if (nt == c->T) c->left_done = c->nch == c->nch_all ? 0 : -1;
  return Z2_OK;
}

// ---------------------------------------------------------------- slices
// propT_L + LocalK_prop_L (src/multiply.f90:79-93, src/localK.f90:347-373)
// the propagators UUR, UUL in the precision of G (one launch with G), or in FP64
// while G is FP32 (mixed mode with FP64 B-products)
template <typename R>
static bool u_with_g(const z2_ctx *c) {
  return sizeof(R) == 8 || !c->u64;
}

// U <- U B^{nflag} (cols) or B^{nflag} U (rows) in FP64, with scalar factors
static int kin_u64(z2_ctx *c, int nt, double *U, int cols, int nflag, int pre, int post) {
  KinArgs<double> d = kin_args<double>(c, nt);
  d.m[0] = kmat<double>(U, nflag, 0, pre, 0, 0, 0, post);
  return cols ? launch_cols<double>(c, d, 1) : launch_rows<double>(c, d, 1);
}

// ns slices' wraps of one propagator in one launch (k_kin_cols_multi / k_kin_rows_multi)
template <typename R, int NC>
static int launch_u_multi_n(z2_ctx *c, KinArgs<R> &a, int cols, int nt0, int ns, int step) {
  const size_t tab = sizeof(int) * 2 * (size_t)c->N;
  if (cols) {
    const int RB = c->kin_rb;
    const size_t sh = sizeof(double) * RB * c->N + tab;
    const void *fn = RB == 32 ? (const void *)k_kin_cols_multi<R, 32, NC> : (const void *)k_kin_cols_multi<R, 8, NC>;
    if (sh > 65536) HIPC(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    dim3 grid((c->N + RB - 1) / RB, c->nch, 1);
    if (RB == 32) hipLaunchKernelGGL((k_kin_cols_multi<R, 32, NC>), grid, dim3(KIN_NT), sh, c->stream, a, nt0, ns, step);
    else hipLaunchKernelGGL((k_kin_cols_multi<R, 8, NC>), grid, dim3(KIN_NT), sh, c->stream, a, nt0, ns, step);
  } else {
    const int CB = c->kin_cb;
    const size_t sh = sizeof(double) * CB * c->N + tab;
    const void *fn = CB == 16 ? (const void *)k_kin_rows_multi<R, 16, NC> : (const void *)k_kin_rows_multi<R, 8, NC>;
    if (sh > 65536) HIPC(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    dim3 grid((c->N + CB - 1) / CB, c->nch, 1);
    if (CB == 16) hipLaunchKernelGGL((k_kin_rows_multi<R, 16, NC>), grid, dim3(KIN_NT), sh, c->stream, a, nt0, ns, step);
    else hipLaunchKernelGGL((k_kin_rows_multi<R, 8, NC>), grid, dim3(KIN_NT), sh, c->stream, a, nt0, ns, step);
  }
  LAUNCH_CHECK();
  return Z2_OK;
}

template <typename R>
static int launch_u_multi(z2_ctx *c, R *U, int cols, int nt0, int ns, int step) {
  KinArgs<R> a = kin_args<R>(c, nt0);
  // UUL <- (UUL*ep) B per left-pass slice, UUR <- (B UUR)*ep per right-pass slice
  a.m[0] = cols ? kmat<R>(U, +1, 0, FAC_EP) : kmat<R>(U, +1, 0, 0, 0, 0, 0, FAC_EP);
  const int kind = cols ? Z2_K_KIN_COLS : Z2_K_KIN_ROWS;
  CHK(prof_begin(c, kind));
  switch (c->N) {
    case 64: CHK((launch_u_multi_n<R, 64>(c, a, cols, nt0, ns, step))); break;
    case 144: CHK((launch_u_multi_n<R, 144>(c, a, cols, nt0, ns, step))); break;
    case 256: CHK((launch_u_multi_n<R, 256>(c, a, cols, nt0, ns, step))); break;
    default: CHK((launch_u_multi_n<R, 0>(c, a, cols, nt0, ns, step)));
  }
  CHK(prof_end(c, kind, 2.0 * sizeof(R) * c->N * c->N * (double)c->nch));
  return Z2_OK;
}

// apply the deferred propagator wraps (before anything reads or writes the
// propagators or changes a field the deferred slices were made from)
static int u_flush(z2_ctx *c) {
  CHK(obs_flush(c));  // the observables' pending sums too: every caller is a point where they may be read
  if (!c->ud_n) return Z2_OK;
  const int ns = c->ud_n;
  c->ud_n = 0;
  if (!c->mixed || c->u64)
    return launch_u_multi<double>(c, c->ud_cols ? c->UUL : c->UUR, c->ud_cols, c->ud_nt0, ns, c->ud_step);
  return launch_u_multi<float>(c, c->ud_cols ? c->UULf : c->UURf, c->ud_cols, c->ud_nt0, ns, c->ud_step);
}

// the propagator wrap of slice nt (left pass: cols = 1, slices in decreasing order;
// right pass: rows, increasing), deferred: between two stabilizations nothing reads
// UUL or UUR, so an interval's wraps run as one launch at its end
static int u_defer(z2_ctx *c, int nt, int cols) {
  const int step = cols ? -1 : 1;
  if (c->ud_n && (c->ud_cols != cols || nt != c->ud_nt0 + c->ud_n * c->ud_step)) CHK(u_flush(c));
  if (!c->ud_n) {
    c->ud_nt0 = nt;
    c->ud_cols = cols;
    c->ud_step = step;
  }
  ++c->ud_n;
  return Z2_OK;
}

template <typename R>
static int slice_L_impl(z2_ctx *c, int nt) {
  CHK(launch_flip<R>(c, nt, 1, 1));
  KinArgs<R> a = kin_args<R>(c, nt);
  a.m[0] = kmat<R>(PB<R>::G(c), +1, 1, FAC_EP, FAC_EM);  // G <- (flush((G*ep)*em)) B
  a.m[0].pre_cond = 1;                                    // mu already in G if the flip kernel flushed
  CHK(launch_cols<R>(c, a, 1));
  CHK(u_defer(c, nt, 1));                                 // UUL <- (UUL*ep) B
  a.m[0] = kmat<R>(PB<R>::G(c), -1);                     // G <- B^-1 G
  CHK(launch_rows<R>(c, a, 1));
  return Z2_OK;
}

// LocalK_prop_R + propT_R (src/localK.f90:375-398, src/multiply.f90:95-108)
template <typename R>
static int slice_R_impl(z2_ctx *c, int nt) {
  KinArgs<R> a = kin_args<R>(c, nt);
  a.m[0] = kmat<R>(PB<R>::G(c), +1);                         // G <- B G
  CHK(launch_rows<R>(c, a, 1));
  a.m[0] = kmat<R>(PB<R>::G(c), -1);                         // G <- G B^-1
  CHK(launch_cols<R>(c, a, 1));
  // G <- (flush(G)*ep)*em: in FP32 with in-kernel flushes on, by the flip kernel's own
  // matrix-core flush; in FP64 (the FP64 matrix rate is the vector rate, so the
  // elementwise kernel over many workgroups per chain is faster) or with the delay
  // depth covering the slice (updates applied one by one, in acceptance order), by
  // k_kin_elem
  const int in_kernel = sizeof(R) == 4 && c->kdelay < c->N2;
  CHK(launch_flip<R>(c, nt, 0, 0, in_kernel));
  if (!in_kernel) {
    a.m[0] = kmat<R>(PB<R>::G(c), 0, 1, 0, 0, FAC_EP, FAC_EM);
    CHK(launch_elem_kin<R>(c, a, 1));
  }
  return u_defer(c, nt, 0);  // UUR <- (B UUR)*ep
}

extern "C" int z2_slice_L(z2_ctx *c, int32_t nt) {
  if (!c || nt < 1 || nt > c->T) return set_err(Z2_EINVAL, "bad slice %d", nt);
  return c->mixed ? slice_L_impl<float>(c, nt) : slice_L_impl<double>(c, nt);
}

extern "C" int z2_slice_R(z2_ctx *c, int32_t nt) {
  if (!c || nt < 1 || nt > c->T) return set_err(Z2_EINVAL, "bad slice %d", nt);
  return c->mixed ? slice_R_impl<float>(c, nt) : slice_R_impl<double>(c, nt);
}

// propK_pre + propT_pre for all chains at slice nt (src/multiply.f90:8-15,65-77)
static int pre_prop(z2_ctx *c, int nt) {
  if (c->mixed && !c->u64) {
    KinArgs<float> a = kin_args<float>(c, nt);
    a.m[0] = kmat<float>(c->UURf, +1, 0, 0, 0, 0, 0, FAC_EP);
    return launch_rows<float>(c, a, 1);
  }
  KinArgs<double> a = kin_args<double>(c, nt);
  a.m[0] = kmat<double>(c->UUR, +1, 0, 0, 0, 0, 0, FAC_EP);
  return launch_rows<double>(c, a, 1);
}

// ---------------------------------------------------------------- observables
// The equal-time observables of a slice are measured when the sweep reaches it
// (k_obs_tiles: the pair sums by tiles, the diagonal and the plaquette products
// into a slot) and added to the bin later, by one k_obs_finish launch for all
// slices pending (at the next stabilization, the end of the sweep, or before
// anything reads or resets the bin), slice by slice in the order measured.
template <typename R>
static ObsArgs<R> obs_args(z2_ctx *c, const R *G, int nt) {
  ObsArgs<R> o;
  o.G = G;
  o.sig = c->sig;
  o.T = c->T;
  o.nt = nt;
  o.acc = c->obs_acc;
  o.cosT = c->d_cosT;
  o.Lx = c->Lx;
  o.Ly = c->Ly;
  o.lat = c->lat;
  o.part = c->opart;
  o.part_stride = c->opart_stride;
  o.diag = c->odiag;
  o.flux = c->oflux;
  o.slot = 0;
  return o;
}

static int obs_flush(z2_ctx *c) {
  if (!c->obs_pend) return Z2_OK;
  const int ns = c->obs_pend;
  c->obs_pend = 0;
  hipLaunchKernelGGL(k_obs_finish, dim3(c->nch), dim3(256), 0, c->stream, obs_args<double>(c, nullptr, 0), ns);
  LAUNCH_CHECK();
  return Z2_OK;
}

extern "C" int z2_obs_reset(z2_ctx *c) {
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  CHK(obs_flush(c));
  CHK(dzero(c, c->obs_acc, (size_t)c->nch * OBS_NACC));
  return Z2_OK;
}

template <typename R>
static int obs_tiles_launch(z2_ctx *c, const R *G, int nt) {
  if (c->obs_pend == OBS_SLOTS) CHK(obs_flush(c));
  ObsArgs<R> o = obs_args<R>(c, G, nt);
  o.slot = c->obs_pend;
  const int ntb = (c->N + OBS_TB - 1) / OBS_TB;
  hipLaunchKernelGGL((k_obs_tiles<R>), dim3(ntb * ntb, c->nch), dim3(256), 0, c->stream, o);
  LAUNCH_CHECK();
  c->obs_pend++;
  return Z2_OK;
}

// slice nt measured; its sums are added to the bin by the next obs_flush
static int obs_equal_defer(z2_ctx *c, int nt) {
  CHK(prof_begin(c, Z2_K_OBS));
  if (c->mixed) CHK(obs_tiles_launch<float>(c, c->Gf, nt));
  else CHK(obs_tiles_launch<double>(c, c->G, nt));
  CHK(prof_end(c, Z2_K_OBS, 2.0 * (c->mixed ? 4.0 : 8.0) * c->N * (double)c->N * c->nch));
  return Z2_OK;
}

extern "C" int z2_obs_equal_calc(z2_ctx *c, int32_t nt) {
  if (!c || nt < 1 || nt > c->T) return set_err(Z2_EINVAL, "bad slice %d", nt);
  CHK(obs_equal_defer(c, nt));
  return obs_flush(c);
}

// Obs_equal%ave (src/obser_equal.f90:50-71); nobs > 0 is the caller's sample
// count, else the count of z2_obs_equal_calc calls since the reset
static int obs_ave(z2_ctx *c, int nobs = 0) {
  CHK(obs_flush(c));
  hipLaunchKernelGGL(k_obs_ave, dim3((c->nch + 63) / 64), dim3(64), 0, c->stream, (const double *)c->obs_acc,
                     c->obs_res, (const ChainConst *)c->dcc, c->N, c->nch, nobs);
  LAUNCH_CHECK();
  return Z2_OK;
}

// ---------------------------------------------------------------- time-displaced
extern "C" int z2_enable_tau(z2_ctx *c) {
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (c->tau_on) return Z2_OK;
  const size_t nn = (size_t)c->N * c->N, B = (size_t)c->nch_all, n2n2 = 4 * nn;
  double **mats[] = {&c->dUUR, &c->dVUR, &c->dUUL, &c->dVUL, &c->G00, &c->G0t, &c->Gt0, &c->Gtt, &c->iVR, &c->iVL};
  for (double **m : mats) CHK(dalloc(c, m, B * nn));
  for (double **m : {&c->bigA, &c->bigQ, &c->bigZ, &c->bigR}) CHK(dalloc(c, m, B * n2n2));
  for (double **d : {&c->dDUR, &c->dDUL}) CHK(dalloc(c, d, B * c->N));
  CHK(dalloc(c, &c->tau2, B * 2 * c->N));
  CHK(dalloc(c, &c->qrw2, B * n2n2));
  CHK(dalloc(c, &c->D2, B * 2 * c->N));
  CHK(dalloc(c, &c->ipvt2, B * 2 * c->N));
  CHK(dalloc(c, &c->otau, B * TAU_NOBS * c->T * c->N));
  CHK(dalloc(c, &c->dyn_drift, B * 6));
  HIPC(hipMemsetAsync(c->otau, 0, sizeof(double) * B * TAU_NOBS * c->T * c->N, c->stream));
  HIPC(hipMemsetAsync(c->dyn_drift, 0, sizeof(double) * B * 6, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  c->tau_on = true;
  return Z2_OK;
}

static dim3 ew_grid(const z2_ctx *c, size_t n) {
  size_t nb = (n + 255) / 256;
  return dim3((unsigned)(nb > 64 ? 64 : nb), c->nch);
}

// Ainv <- A^{-1} for each chain's N x N A: A P = Q D R (pivoted QR, R unit-diagonal
// magnitude), so A^{-1} = P R^{-1} D^{-1} Q^T.  Scratch: M, Q, X, Dtmp, tau, ipvt.
static int ginv(z2_ctx *c, const double *A, double *Ainv) {
  CHK(elem(c, A, c->M, 0));
  CHK(qrp(c, c->M, c->Dtmp));
  CHK(orgqr(c, c->M, c->Q));
  CHK(elem(c, c->Q, c->X, 1));
  CHK(trsm_dp(c, c->M, c->X, 0, c->Dtmp, c->ipvt, Ainv));
  return Z2_OK;
}

// stab_green_big, src/stabilization.f90:202-264, on the d* propagators -> bigZ (2N x 2N)
//   big = [[ (V_R V_L)^{-1}, D_L ], [ -D_R, (U_L U_R)^{-1} ]] = Qb Db Rb (plain QR)
//   G_big = diag(V_R^{-1}, U_L^{-1}) Rb^{-1} Db^{-1} . Qb^T diag(V_L^{-1}, U_R^{-1})
// U_R, U_L are orthogonal (explicit Householder Q), so their inverses are transposes.
static int stab_green_big(z2_ctx *c) {
  const int N = c->N, N2 = 2 * N;
  const long nn = (long)N * N, n4 = 4 * nn;
  CHK(ginv(c, c->dVUR, c->iVR));
  CHK(ginv(c, c->dVUL, c->iVL));
  CHK(gemm(c, c->iVL, 0, c->iVR, 0, c->Y));     // (V_R V_L)^{-1}
  CHK(gemm(c, c->dUUR, 1, c->dUUL, 1, c->X));   // (U_L U_R)^{-1}
  {
    dim3 g = ew_grid(c, (size_t)n4);
    hipLaunchKernelGGL(k_assemble_big, g, dim3(256), 0, c->stream, (const double *)c->Y, (const double *)c->X,
                       (const double *)c->dDUL, (const double *)c->dDUR, c->bigA, N);
    LAUNCH_CHECK();
  }
  CHK(qrp_n(c, c->bigA, c->D2, c->tau2, c->ipvt2, N2, 1));
  CHK(orgqr_n(c, c->bigA, c->tau2, c->bigQ, N2));
  // right = Qb^T diag(V_L^{-1}, U_R^T)
  CHK(gemm_g(c, N2, N, N, c->bigQ, N2, n4, 1, c->iVL, N, nn, 0, c->bigR, N2, n4));
  CHK(gemm_g(c, N2, N, N, c->bigQ + N, N2, n4, 1, c->dUUR, N, nn, 1, c->bigR + (size_t)N * N2, N2, n4));
  // Z = Rb^{-1}
  {
    dim3 g = ew_grid(c, (size_t)n4);
    hipLaunchKernelGGL(k_eye, g, dim3(256), 0, c->stream, c->bigZ, N2);
    LAUNCH_CHECK();
  }
  CHK(trsm(c, c->bigA, c->bigZ, 0, N2));
  // left = diag(V_R^{-1}, U_L^T) Z Db^{-1}  (into bigQ)
  CHK(gemm_g(c, N, N2, N, c->iVR, N, nn, 0, c->bigZ, N2, n4, 0, c->bigQ, N2, n4));
  CHK(gemm_g(c, N, N2, N, c->dUUL, N, nn, 1, c->bigZ + N, N2, n4, 0, c->bigQ + N, N2, n4));
  {
    dim3 g = ew_grid(c, (size_t)n4);
    hipLaunchKernelGGL(k_div_cols, g, dim3(256), 0, c->stream, c->bigQ, (const double *)c->D2, N2);
    LAUNCH_CHECK();
  }
  CHK(gemm_g(c, N2, N2, N2, c->bigQ, N2, n4, 0, c->bigR, N2, n4, 0, c->bigZ, N2, n4));
  return Z2_OK;
}

// Wrap_tau, src/stabilization.f90:478-514
static int wrap_tau(z2_ctx *c, int nt) {
  CHK(u_flush(c));
  const int st = nt / c->w;
  const size_t nn = (size_t)c->nch * c->N * c->N, nd = (size_t)c->nch * c->N;
  CHK(dcopy(c, c->dUUL, c->ULl + slot(c, st), nn));
  CHK(dcopy(c, c->dVUL, c->VLl + slot(c, st), nn));
  CHK(dcopy(c, c->dDUL, c->DLl + slotd(c, st), nd));
  CHK(stab_UR_buf(c, c->dUUR, c->dDUR, c->dVUR));
  CHK(stab_green_big(c));
  hipLaunchKernelGGL(k_tau_replace, dim3(c->nch), dim3(256), 0, c->stream, (const double *)c->bigZ, c->G00,
                     c->G0t, c->Gt0, c->Gtt, c->dyn_drift, c->N);
  LAUNCH_CHECK();
  return Z2_OK;
}

static int obs_tau(z2_ctx *c, int nt) {
  TauObsArgs o;
  o.G00 = c->G00; o.Gtt = c->Gtt; o.Gt0 = c->Gt0; o.G0t = c->G0t;
  o.sig = c->sig; o.acc = c->otau; o.cc = c->dcc;
  o.T = c->T; o.nt = nt; o.Lx = c->Lx; o.Ly = c->Ly; o.lat = c->lat;
  const int th = ((c->N + 63) / 64) * 64;
  CHK(prof_begin(c, Z2_K_OBS_TAU));
  hipLaunchKernelGGL(k_obs_tau, dim3(c->nch), dim3(th), 0, c->stream, o);
  LAUNCH_CHECK();
  CHK(prof_end(c, Z2_K_OBS_TAU, 0.0));
  return Z2_OK;
}

// Dyn_reset, src/dynamics.f90:41-48: the second right propagator from the current
// one, G(0,0) = G(t,t) = G(t,0) = G, G(0,t) = -(I - G)
static int dyn_reset(z2_ctx *c) {
  CHK(u_flush(c));
  const size_t nn = (size_t)c->nch * c->N * c->N, nd = (size_t)c->nch * c->N;
  CHK(dcopy(c, c->dUUR, c->UUR, nn));
  CHK(dcopy(c, c->dVUR, c->VUR, nn));
  CHK(dcopy(c, c->dDUR, c->DUR, nd));
  dim3 g = ew_grid(c, (size_t)c->N * c->N);
  if (c->mixed)
    hipLaunchKernelGGL((k_dyn_reset<float>), g, dim3(256), 0, c->stream, (const float *)c->Gf, c->G00, c->Gtt,
                       c->Gt0, c->G0t, c->N);
  else
    hipLaunchKernelGGL((k_dyn_reset<double>), g, dim3(256), 0, c->stream, (const double *)c->G, c->G00, c->Gtt,
                       c->Gt0, c->G0t, c->N);
  LAUNCH_CHECK();
  return Z2_OK;
}

// Dyn_reset + Dyn%sweep_R, src/dynamics.f90:41-62: starts from the current G and
// right propagator, steps G(t,0), G(0,t), G(t,t) and U_R through all slices.
static int dyn_sweep(z2_ctx *c) {
  CHK(u_flush(c));
  CHK(dyn_reset(c));
  for (int nt = 1; nt <= c->T; ++nt) {
    CHK(obs_tau(c, nt));
    KinArgs<double> a = kin_args<double>(c, nt);
    a.m[0] = kmat<double>(c->Gt0, +1, 0, 0, 0, 0, 0, FAC_EP);   // G(t,0) <- (B G(t,0)) ep
    a.m[1] = kmat<double>(c->Gtt, +1);                          // G(t,t) <- B G(t,t)
    a.m[2] = kmat<double>(c->dUUR, +1, 0, 0, 0, 0, 0, FAC_EP);  // U_R <- (B U_R) ep
    CHK(launch_rows<double>(c, a, 3));
    a.m[0] = kmat<double>(c->G0t, -1, 0, 0, 0, 0, 0, FAC_EM);   // G(0,t) <- (G(0,t) B^-1) em
    a.m[1] = kmat<double>(c->Gtt, -1, 0, 0, 0, 0, 0, FAC_EP);   // G(t,t) <- ((G(t,t) B^-1) ep) em
    a.m[1].post2 = FAC_EM;
    CHK(launch_cols<double>(c, a, 2));
    if (nt % c->w == 0) CHK(wrap_tau(c, nt));
  }
  return Z2_OK;
}

extern "C" int z2_dyn_measure(z2_ctx *c) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (!c->tau_on) return set_err(Z2_ESTATE, "time-displaced measurement not enabled (z2_enable_tau)");
  if (!c->have_pre) return set_err(Z2_ESTATE, "z2_pre must run before z2_dyn_measure");
  if (c->left_done == 0)
    return set_err(Z2_ESTATE, "the time-displaced measurement starts a right pass: run it after a left pass");
  CHK(dzero(c, c->otau, (size_t)c->nch * TAU_NOBS * c->T * c->N));
  return dyn_sweep(c);
}

static int h_zero_sync(z2_ctx *c);

// ---------------------------------------------------------------- fine-grained steps
// The reference's Multiply_mod / LocalK_mod / ObserEqual / ObserTau / Dynamics /
// Wrap_tau procedures one by one, for callers that sequence a sweep themselves.
// Each acts on all chains of the context; composing them in the reference's
// order (src/local_sweep.f90:106-150) gives the same arithmetic as z2_sweep.
template <typename R>
static int mu_elem(z2_ctx *c, int nt, R *m0, int f0a, int f0b, R *m1, int f1) {
  KinArgs<R> a = kin_args<R>(c, nt);
  a.m[0] = kmat<R>(m0, 0, 0, f0a, f0b);
  int nmat = 1;
  if (m1) {
    a.m[1] = kmat<R>(m1, 0, 0, f1);
    nmat = 2;
  }
  return launch_elem_kin<R>(c, a, nmat);
}

static int check_slice(z2_ctx *c, int nt) {
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (nt < 1 || nt > c->T) return set_err(Z2_EINVAL, "bad slice %d", nt);
  return Z2_OK;
}

// propK_pre, src/multiply.f90:8-15: UUR <- B(nt) UUR
extern "C" int z2_prop_K_pre(z2_ctx *c, int32_t nt) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  CHK(check_slice(c, nt));
  if (c->mixed && !c->u64) {
    KinArgs<float> a = kin_args<float>(c, nt);
    a.m[0] = kmat<float>(c->UURf, +1);
    return launch_rows<float>(c, a, 1);
  }
  KinArgs<double> a = kin_args<double>(c, nt);
  a.m[0] = kmat<double>(c->UUR, +1);
  return launch_rows<double>(c, a, 1);
}

// propT_pre, src/multiply.f90:65-77: UUR <- UUR exp(dtau mu)
extern "C" int z2_prop_T_pre(z2_ctx *c, int32_t nt) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  CHK(check_slice(c, nt));
  if (c->mixed && !c->u64) return mu_elem<float>(c, nt, c->UURf, FAC_EP, FAC_NONE, (float *)nullptr, 0);
  return mu_elem<double>(c, nt, c->UUR, FAC_EP, FAC_NONE, (double *)nullptr, 0);
}

// propT_L, src/multiply.f90:79-93: G <- (G e^{+}) e^{-}, UUL <- UUL e^{+}
extern "C" int z2_prop_T_L(z2_ctx *c, int32_t nt) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  CHK(check_slice(c, nt));
  if (c->mixed && c->u64) {
    CHK(mu_elem<float>(c, nt, c->Gf, FAC_EP, FAC_EM, (float *)nullptr, 0));
    return mu_elem<double>(c, nt, c->UUL, FAC_EP, FAC_NONE, (double *)nullptr, 0);
  }
  if (c->mixed) return mu_elem<float>(c, nt, c->Gf, FAC_EP, FAC_EM, c->UULf, FAC_EP);
  return mu_elem<double>(c, nt, c->G, FAC_EP, FAC_EM, c->UUL, FAC_EP);
}

// propT_R, src/multiply.f90:95-108: G <- (G e^{+}) e^{-}, UUR <- UUR e^{+}
extern "C" int z2_prop_T_R(z2_ctx *c, int32_t nt) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  CHK(check_slice(c, nt));
  if (c->mixed && c->u64) {
    CHK(mu_elem<float>(c, nt, c->Gf, FAC_EP, FAC_EM, (float *)nullptr, 0));
    return mu_elem<double>(c, nt, c->UUR, FAC_EP, FAC_NONE, (double *)nullptr, 0);
  }
  if (c->mixed) return mu_elem<float>(c, nt, c->Gf, FAC_EP, FAC_EM, c->UURf, FAC_EP);
  return mu_elem<double>(c, nt, c->G, FAC_EP, FAC_EM, c->UUR, FAC_EP);
}

// LocalK_prop_L, src/localK.f90:347-373: flips of bonds 2N..1, then
// G <- B^{-1} (G B) and UUL <- UUL B (the pending updates flushed first)
template <typename R>
static int localK_L_impl(z2_ctx *c, int nt) {
  CHK(launch_flip<R>(c, nt, 1, 0));
  KinArgs<R> a = kin_args<R>(c, nt);
  a.m[0] = kmat<R>(PB<R>::G(c), +1, 1);
  if (u_with_g<R>(c)) {
    a.m[1] = kmat<R>(PB<R>::UUL(c), +1);
    CHK(launch_cols<R>(c, a, 2));
  } else {
    CHK(launch_cols<R>(c, a, 1));
    CHK(kin_u64(c, nt, c->UUL, 1, +1, 0, 0));
  }
  a.m[0] = kmat<R>(PB<R>::G(c), -1);
  return launch_rows<R>(c, a, 1);
}

// LocalK_prop_R, src/localK.f90:375-398: G <- (B G) B^{-1}, flips of bonds
// 1..2N, the pending updates flushed, then UUR <- B UUR
template <typename R>
static int localK_R_impl(z2_ctx *c, int nt) {
  KinArgs<R> a = kin_args<R>(c, nt);
  a.m[0] = kmat<R>(PB<R>::G(c), +1);
  CHK(launch_rows<R>(c, a, 1));
  a.m[0] = kmat<R>(PB<R>::G(c), -1);
  CHK(launch_cols<R>(c, a, 1));
  const int in_kernel = sizeof(R) == 4 && c->kdelay < c->N2;
  CHK(launch_flip<R>(c, nt, 0, 0, in_kernel ? 2 : 0));  // 2: final flush without the mu factors
  if (!in_kernel) {
    a.m[0] = kmat<R>(PB<R>::G(c), 0, 1);
    CHK(launch_elem_kin<R>(c, a, 1));
  }
  if (!u_with_g<R>(c)) return kin_u64(c, nt, c->UUR, 0, +1, 0, 0);
  a.m[0] = kmat<R>(PB<R>::UUR(c), +1);
  return launch_rows<R>(c, a, 1);
}

extern "C" int z2_localK_prop_L(z2_ctx *c, int32_t nt) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  CHK(check_slice(c, nt));
  return c->mixed ? localK_L_impl<float>(c, nt) : localK_L_impl<double>(c, nt);
}

extern "C" int z2_localK_prop_R(z2_ctx *c, int32_t nt) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  CHK(check_slice(c, nt));
  return c->mixed ? localK_R_impl<float>(c, nt) : localK_R_impl<double>(c, nt);
}

static int need_tau(z2_ctx *c) {
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (!c->tau_on) return set_err(Z2_ESTATE, "time-displaced measurement not enabled (z2_enable_tau)");
  return Z2_OK;
}

// propgrK_R, src/multiply.f90:43-63: G(t,0) <- B G(t,0), G(0,t) <- G(0,t) B^{-1},
// G(t,t) <- B G(t,t) B^{-1}, U_R <- B U_R on the time-displaced state (FP64)
extern "C" int z2_propgr_K_R(z2_ctx *c, int32_t nt) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  CHK(need_tau(c));
  CHK(check_slice(c, nt));
  KinArgs<double> a = kin_args<double>(c, nt);
  a.m[0] = kmat<double>(c->Gt0, +1);
  a.m[1] = kmat<double>(c->Gtt, +1);
  a.m[2] = kmat<double>(c->dUUR, +1);
  CHK(launch_rows<double>(c, a, 3));
  a.m[0] = kmat<double>(c->G0t, -1);
  a.m[1] = kmat<double>(c->Gtt, -1);
  return launch_cols<double>(c, a, 2);
}

// propgrT_R, src/multiply.f90:110-130: G(t,0) e^{+}, G(0,t) e^{-}, (G(t,t) e^{+}) e^{-}, U_R e^{+}
extern "C" int z2_propgr_T_R(z2_ctx *c, int32_t nt) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  CHK(need_tau(c));
  CHK(check_slice(c, nt));
  KinArgs<double> a = kin_args<double>(c, nt);
  a.m[0] = kmat<double>(c->Gt0, 0, 0, FAC_EP);
  a.m[1] = kmat<double>(c->G0t, 0, 0, FAC_EM);
  a.m[2] = kmat<double>(c->Gtt, 0, 0, FAC_EP, FAC_EM);
  a.m[3] = kmat<double>(c->dUUR, 0, 0, FAC_EP);
  return launch_elem_kin<double>(c, a, 4);
}

// Dyn%reset, src/dynamics.f90:41-48 (Propagator%asgn, Propgr_reset src/process_matrix.f90:88-96)
extern "C" int z2_dyn_reset(z2_ctx *c) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  CHK(need_tau(c));
  if (!c->have_pre) return set_err(Z2_ESTATE, "z2_pre must run before z2_dyn_reset");
  if (c->left_done == 0)
    return set_err(Z2_ESTATE, "Dyn%%reset belongs at the start of a right pass (after a left pass)");
  return dyn_reset(c);
}

// Obs_tau%calc, src/obser_tau.f90:56-133
extern "C" int z2_obs_tau_calc(z2_ctx *c, int32_t nt) {
  CHK(need_tau(c));
  CHK(check_slice(c, nt));
  return obs_tau(c, nt);
}

// Obs_tau%reset (src/obser_tau.f90:26-35) and Obs_tau%ave (:37-47)
extern "C" int z2_obs_tau_reset(z2_ctx *c) {
  CHK(need_tau(c));
  return dzero(c, c->otau, (size_t)c->nch * TAU_NOBS * c->T * c->N);
}

extern "C" int z2_obs_tau_ave(z2_ctx *c, int32_t nobst) {
  CHK(need_tau(c));
  if (nobst < 1) return set_err(Z2_EINVAL, "Nobst must be positive");
  hipLaunchKernelGGL(k_scale, dim3(256), dim3(256), 0, c->stream, c->otau,
                     (size_t)c->nch * TAU_NOBS * c->T * c->N, 1.0 / (double)nobst);
  LAUNCH_CHECK();
  return Z2_OK;
}

// Wrap_tau, src/stabilization.f90:478-514
extern "C" int z2_wrap_tau(z2_ctx *c, int32_t nt) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  CHK(need_tau(c));
  if (nt < 1 || nt > c->T || nt % c->w) return set_err(Z2_EINVAL, "incorrect orthobig time slice %d", nt);
  return wrap_tau(c, nt);
}

// Obs_equal%ave(Nobs), src/obser_equal.f90:50-71
extern "C" int z2_obs_equal_ave(z2_ctx *c, int32_t nobs) {
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (nobs < 1) return set_err(Z2_EINVAL, "Nobs must be positive");
  return obs_ave(c, nobs);
}

// Local_sweep_reset, src/local_sweep.f90:55-71: observables (and with toggle the
// time-displaced accumulators) zeroed, the field synchronised along tau where h ~ 0
extern "C" int z2_sweep_reset(z2_ctx *c, int32_t toggle) {
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (toggle && !c->tau_on) return set_err(Z2_ESTATE, "time-displaced measurement not enabled (z2_enable_tau)");
  CHK(h_zero_sync(c));
  CHK(z2_obs_reset(c));
  if (toggle) CHK(dzero(c, c->otau, (size_t)c->nch * TAU_NOBS * c->T * c->N));
  return Z2_OK;
}

extern "C" int z2_get_obs_tau(z2_ctx *c, int32_t chain, double *out) {
  if (!c || !out || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];  // storage slot
  if (!c->tau_on) return set_err(Z2_ESTATE, "time-displaced measurement not enabled (z2_enable_tau)");
  const size_t n = (size_t)TAU_NOBS * c->T * c->N;
  HIPC(hipMemcpyAsync(out, c->otau + chain * n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return Z2_OK;
}

extern "C" int z2_get_dyn_drift(z2_ctx *c, int32_t chain, double xmax[3], double xmean[3]) {
  if (!c || !xmax || !xmean || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];  // storage slot
  if (!c->tau_on) return set_err(Z2_ESTATE, "time-displaced measurement not enabled (z2_enable_tau)");
  double v[6];
  HIPC(hipMemcpyAsync(v, c->dyn_drift + 6 * chain, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  for (int k = 0; k < 3; ++k) {
    xmax[k] = v[k];
    xmean[k] = v[3 + k];
  }
  return Z2_OK;
}

// ---------------------------------------------------------------- drivers
// Local_sweep_reset's h ~ 0 synchronisation (src/local_sweep.f90:63-69)
static int h_zero_sync(z2_ctx *c) {
  for (int b = 0; b < c->nch_all; ++b) {
    if (fabs(c->desc[c->phys[b]].h) > 1.0e-10) continue;
    std::vector<double> s((size_t)c->T * c->N2);
    CHK(z2_get_sigma(c, b, s.data()));
    for (int nt = 1; nt < c->T; ++nt)
      for (int ii = 0; ii < c->N2; ++ii) s[(size_t)nt * c->N2 + ii] = s[ii];
    int32_t sd;
    CHK(z2_get_seed(c, b, &sd));
    CHK(z2_set_sigma(c, b, s.data(), sd));
  }
  return Z2_OK;
}

extern "C" int z2_pre(z2_ctx *c) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (!c->have_fields) return set_err(Z2_ESTATE, "fields not initialised (z2_init_fields / z2_set_sigma)");
  CHK(h_zero_sync(c));
  CHK(z2_obs_reset(c));
  CHK(z2_wrap_pre(c, 0));
  for (int nt = 1; nt <= c->T; ++nt) {
    CHK(pre_prop(c, nt));
    if (nt % c->w == 0) CHK(z2_wrap_pre(c, nt));
  }
  return Z2_OK;
}

// One sweep as a list of steps (src/local_sweep.f90:106-150), so that two
// contexts can be advanced in lockstep.
enum StepKind { ST_WRAP_L, ST_WRAP_R, ST_OBS, ST_SLICE_L, ST_SLICE_R, ST_DYN };
struct Step {
  StepKind kind;
  int nt;
  int flag = 1;  // wrap steps: 1 adds the drift to its mean (Wrap_L/R flag)
};

static void steps_L(const z2_ctx *c, std::vector<Step> &v) {
  for (int nt = c->T; nt >= 1; --nt) {
    if (nt % c->w == 0) v.push_back({ST_WRAP_L, nt});
    v.push_back({ST_OBS, nt});
    v.push_back({ST_SLICE_L, nt});
  }
  v.push_back({ST_WRAP_L, 0});
}

static void steps_R(const z2_ctx *c, std::vector<Step> &v, int toggle) {
  if (toggle) v.push_back({ST_DYN, 0});
  v.push_back({ST_WRAP_R, 0});
  for (int nt = 1; nt <= c->T; ++nt) {
    v.push_back({ST_SLICE_R, nt});
    if (nt % c->w == 0) v.push_back({ST_WRAP_R, nt});
    v.push_back({ST_OBS, nt});
  }
}

static void sweep_steps(const z2_ctx *c, int is_beta, int toggle, std::vector<Step> &v) {
  for (int s = 0; s < c->p.Nsweep; ++s) {
    if (is_beta) {
      steps_L(c, v);
      steps_R(c, v, toggle);
    } else {
      steps_R(c, v, toggle);
      steps_L(c, v);
    }
  }
}

static int run_step(z2_ctx *c, const Step &st) {
  switch (st.kind) {
    case ST_WRAP_L: return z2_wrap_L(c, st.nt, st.flag);
    case ST_WRAP_R: return z2_wrap_R(c, st.nt, st.flag);
    case ST_OBS:
      if (st.nt < 1 || st.nt > c->T) return set_err(Z2_EINVAL, "bad slice %d", st.nt);
      return obs_equal_defer(c, st.nt);
    case ST_SLICE_L: return z2_slice_L(c, st.nt);
    case ST_SLICE_R: return z2_slice_R(c, st.nt);
    case ST_DYN: return dyn_sweep(c);
  }
  return set_err(Z2_EINVAL, "bad step");
}

// true when the step replaces G by a recomputed one (Wrap_L nt != Ltrot, Wrap_R nt != 0)
static bool step_recomputes(const z2_ctx *c, const Step &st) {
  return (st.kind == ST_WRAP_L && st.nt != c->T) || (st.kind == ST_WRAP_R && st.nt != 0);
}

// Local_sweep with toggle (src/local_sweep.f90:73-150): with toggle the
// time-displaced observables are reset, measured at the start of every right
// pass and divided by the number of such passes.
// ---------------------------------------------------------------- chain order
// Chains are stored grouped by pass direction so that each group is a
// contiguous range the batched kernels can run on.
static int copy_chains(z2_ctx *c, const char *src, char *dst, const ChainField &f, int c0, int cnt,
                       const int *perm, const int *skip) {
  if (cnt <= 0) return Z2_OK;
  const size_t off = (size_t)c0 * f.bytes, sstride = (size_t)c->nch_all * f.bytes;
  const size_t blocks = (f.bytes / 4 + 255) / 256;
  dim3 grid((unsigned)(blocks > 256 ? 256 : (blocks ? blocks : 1)), cnt, f.slots);
  if (f.bytes % 16 == 0)
    hipLaunchKernelGGL(k_copy_chains<uint4>, grid, dim3(256), 0, c->stream, (const uint4 *)(src + off),
                       (uint4 *)(dst + off), f.bytes / 16, sstride / 16, perm, skip);
  else if (f.bytes % 8 == 0)
    hipLaunchKernelGGL(k_copy_chains<uint2>, grid, dim3(256), 0, c->stream, (const uint2 *)(src + off),
                       (uint2 *)(dst + off), f.bytes / 8, sstride / 8, perm, skip);
  else if (f.bytes % 4 == 0)
    hipLaunchKernelGGL(k_copy_chains<uint32_t>, grid, dim3(256), 0, c->stream, (const uint32_t *)(src + off),
                       (uint32_t *)(dst + off), f.bytes / 4, sstride / 4, perm, skip);
  else
    return set_err(Z2_EINVAL, "per-chain array size %zu is not a multiple of 4 bytes", f.bytes);
  LAUNCH_CHECK();
  return Z2_OK;
}

// storage slot j <- slot perm[j] for every array that moves with its chain
static int reorder(z2_ctx *c, const std::vector<int> &perm) {
  const int B = c->nch_all;
  size_t need = 0;
  for (auto &f : c->fields)  // including arrays not allocated yet, so that the scratch fits them later
    if (f.flags & CF_PERSIST) need = std::max(need, f.bytes * f.slots * (size_t)B);
  if (!c->reorder_tmp) CHK(dalloc(c, &c->reorder_tmp, need));
  HIPC(hipMemcpyAsync(c->d_perm, perm.data(), sizeof(int) * B, hipMemcpyHostToDevice, c->stream));
  for (auto &f : c->fields) {
    char *a = f.get();
    if (!(f.flags & CF_PERSIST) || !a) continue;
    CHK(copy_chains(c, a, c->reorder_tmp, f, 0, B, c->d_perm, nullptr));
    HIPC(hipMemcpyAsync(a, c->reorder_tmp, f.bytes * f.slots * (size_t)B, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPC(hipStreamSynchronize(c->stream));  // perm's host buffer is released on return
  std::vector<int> inv(B);
  for (int j = 0; j < B; ++j) inv[perm[j]] = j;
  std::vector<ChainConst> hcc(B);
  std::vector<z2_chain_desc> desc(B);
  std::vector<int> isb(B);
  std::vector<double> lam(c->lam.size());
  for (int j = 0; j < B; ++j) {
    hcc[j] = c->hcc[perm[j]];
    desc[j] = c->desc[perm[j]];
    isb[j] = c->isb[perm[j]];
    std::copy(c->lam.begin() + (size_t)perm[j] * c->N, c->lam.begin() + (size_t)(perm[j] + 1) * c->N,
              lam.begin() + (size_t)j * c->N);
  }
  c->hcc.swap(hcc);
  c->desc.swap(desc);
  c->isb.swap(isb);
  c->lam.swap(lam);
  for (int k = 0; k < B; ++k) c->phys[k] = inv[c->phys[k]];
  return Z2_OK;
}

// stores the chains with is_beta = 1 first; *n1 = their number
static int partition(z2_ctx *c, int *n1) {
  CHK(u_flush(c));
  std::vector<int> perm;
  for (int j = 0; j < c->nch_all; ++j)
    if (c->isb[j]) perm.push_back(j);
  *n1 = (int)perm.size();
  for (int j = 0; j < c->nch_all; ++j)
    if (!c->isb[j]) perm.push_back(j);
  bool ident = true;
  for (int j = 0; j < c->nch_all; ++j) ident = ident && perm[j] == j;
  return ident ? Z2_OK : reorder(c, perm);
}

// the steps of a sweep for the chain range of one direction
static int run_steps(z2_ctx *c, const std::vector<Step> &v, int *nobst) {
  for (const Step &st : v) {
    CHK(run_step(c, st));
    if (nobst) *nobst += st.kind == ST_DYN;
  }
  return Z2_OK;
}

// the device work of one LocalSweep%sweep with every chain in direction is_beta:
// observables reset, the steps, the averages -- kernel launches only (no host
// synchronisation, no memset), so it can be captured in a HIP graph
static int sweep_prologue(z2_ctx *c, int toggle) {
  CHK(z2_obs_reset(c));
  if (toggle) CHK(dzero(c, c->otau, (size_t)c->nch * TAU_NOBS * c->T * c->N));
  return Z2_OK;
}

static int sweep_epilogue(z2_ctx *c, int toggle, int nobst) {
  CHK(u_flush(c));
  CHK(obs_ave(c));
  if (toggle && nobst > 0) {
    hipLaunchKernelGGL(k_scale, dim3(256), dim3(256), 0, c->stream, c->otau, (size_t)c->nch * TAU_NOBS * c->T * c->N,
                       1.0 / (double)nobst);
    LAUNCH_CHECK();
  }
  return Z2_OK;
}

static int sweep_core(z2_ctx *c, int is_beta, int toggle) {
  CHK(sweep_prologue(c, toggle));
  std::vector<Step> v;
  sweep_steps(c, is_beta, toggle, v);
  int nobst = 0;
  CHK(run_steps(c, v, &nobst));
  return sweep_epilogue(c, toggle, nobst);
}

static int sweep_check(z2_ctx *c, int is_beta, int toggle) {
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (!c->have_pre) return set_err(Z2_ESTATE, "z2_pre must run before z2_sweep");
  if (toggle && !c->tau_on) return set_err(Z2_ESTATE, "time-displaced measurement not enabled (z2_enable_tau)");
  if (is_beta != 0 && is_beta != 1 && is_beta != Z2_BETA_PER_CHAIN)
    return set_err(Z2_EINVAL, "is_beta must be 0, 1 or Z2_BETA_PER_CHAIN");
  return Z2_OK;
}

extern "C" int z2_sweep_tau(z2_ctx *c, int32_t is_beta, int32_t toggle) {
  CHK(sweep_check(c, is_beta, toggle));
  CHK(h_zero_sync(c));
  if (is_beta != Z2_BETA_PER_CHAIN) {
    c->isb.assign(c->nch_all, is_beta);
    return sweep_core(c, is_beta, toggle);
  }
  // each chain in its own order (after Wolff updates the directions differ)
  CHK(z2_obs_reset(c));
  const size_t ntau = (size_t)c->nch * TAU_NOBS * c->T * c->N;
  if (toggle) CHK(dzero(c, c->otau, ntau));
  int n1;
  CHK(partition(c, &n1));
  int nob[2] = {0, 0};
  for (int g = 0; g < 2; ++g) {
    const int c0 = g == 0 ? 0 : n1, cnt = g == 0 ? n1 : c->nch_all - n1;
    if (!cnt) continue;
    std::vector<Step> v;
    sweep_steps(c, g == 0 ? 1 : 0, toggle, v);
    ChainRange r(c, c0, cnt, c->stream);
    CHK(run_steps(c, v, &nob[g]));
    CHK(u_flush(c));  // in this chain range
  }
  const int nobst = nob[0] > nob[1] ? nob[0] : nob[1];
  CHK(obs_ave(c));
  if (toggle && nobst > 0) {
    hipLaunchKernelGGL(k_scale, dim3(256), dim3(256), 0, c->stream, c->otau, ntau, 1.0 / (double)nobst);
    LAUNCH_CHECK();
  }
  return Z2_OK;
}

extern "C" int z2_sweep(z2_ctx *c, int32_t is_beta) { return z2_sweep_tau(c, is_beta, 0); }

extern "C" int z2_set_is_beta(z2_ctx *c, int32_t chain, int32_t is_beta) {
  if (!c || chain < 0 || chain >= c->nch_all || (is_beta != 0 && is_beta != 1))
    return set_err(Z2_EINVAL, "bad argument");
  c->isb[c->phys[chain]] = is_beta;
  return Z2_OK;
}

extern "C" int z2_get_is_beta(z2_ctx *c, int32_t chain, int32_t *is_beta) {
  if (!c || !is_beta || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  *is_beta = c->isb[c->phys[chain]];
  return Z2_OK;
}

// ---------------------------------------------------------------- Wolff global update
static GlobArgs glob_args(z2_ctx *c) {
  GlobArgs g;
  g.sig = c->sig;
  g.seed = c->seed;
  g.cc = c->dcc;
  g.lat = c->lat;
  g.T = c->T;
  g.nch = c->nch;
  g.bond = c->gbond;
  g.mask = c->gmask;
  g.queue = c->gqueue;
  g.lr = c->glr;
  g.acc = c->gacc;
  g.cnt = c->gcnt;
  g.sz = c->gsz;
  return g;
}

// copy of everything a trial pass changes (restore = 0), or its restoration for
// the chains of the range whose trial was rejected (restore = 1)
static int trial_state(z2_ctx *c, int c0, int cnt, int restore) {
  if (!c->gbackup) {
    for (auto &f : c->fields)
      if ((f.flags & CF_TRIAL) && f.get()) CHK(dalloc(c, &f.backup, f.bytes * f.slots * (size_t)c->nch_all));
    c->gbackup = true;
  }
  for (auto &f : c->fields) {
    char *a = f.get();
    if (!(f.flags & CF_TRIAL) || !a) continue;
    if (restore) {
      // skip flag = accepted
      CHK(copy_chains(c, f.backup, a, f, c0, cnt, nullptr, c->gacc + c0));
    } else {
      CHK(copy_chains(c, a, f.backup, f, c0, cnt, nullptr, nullptr));
    }
  }
  return Z2_OK;
}

// Global_sweep_L / Global_sweep_R (src/global_update.f90:229-311) as steps: the
// propagation of a left or right pass with the flips of the cluster forced and
// the checkpoints stored without adding to the drift mean (Wrap_L/R flag 0)
static void global_steps(const z2_ctx *c, int is_beta, std::vector<Step> &v) {
  if (is_beta) {
    for (int nt = c->T; nt >= 1; --nt) {
      if (nt % c->w == 0) v.push_back({ST_WRAP_L, nt, 0});
      v.push_back({ST_SLICE_L, nt, 0});
    }
    v.push_back({ST_WRAP_L, 0, 0});
  } else {
    v.push_back({ST_WRAP_R, 0, 0});
    for (int nt = 1; nt <= c->T; ++nt) {
      v.push_back({ST_SLICE_R, nt, 0});
      if (nt % c->w == 0) v.push_back({ST_WRAP_R, nt, 0});
    }
  }
}

// Global_sweep (src/global_update.f90:313-337) for every chain, each in its own
// direction; an accepted trial reverses the chain's is_beta.
extern "C" int z2_global_sweep(z2_ctx *c, int32_t nglobal) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c || nglobal < 1) return set_err(Z2_EINVAL, "null ctx or Nglobal < 1");
  if (!c->have_pre) return set_err(Z2_ESTATE, "z2_pre must run before z2_global_sweep");
  if (c->force) return set_err(Z2_ESTATE, "global sweep during a replay");
  const dim3 g1((c->nch_all + 63) / 64), b1(64);
  for (int ngl = 0; ngl < nglobal; ++ngl) {
    int n1;
    CHK(partition(c, &n1));
    for (int g = 0; g < 2; ++g) {
      const int c0 = g == 0 ? 0 : n1, cnt = g == 0 ? n1 : c->nch_all - n1;
      if (!cnt) continue;
      CHK(trial_state(c, c0, cnt, 0));
      {
        ChainRange r(c, c0, cnt, c->stream);
        const dim3 gr((cnt + 63) / 64);
        hipLaunchKernelGGL(k_wolff_cluster, gr, b1, 0, c->stream, glob_args(c));
        LAUNCH_CHECK();
        std::vector<Step> v;
        global_steps(c, g == 0 ? 1 : 0, v);
        c->gmode = true;
        int rc = run_steps(c, v, nullptr);
        c->gmode = false;
        CHK(rc);
        CHK(u_flush(c));  // in this chain range
        hipLaunchKernelGGL(k_wolff_accept, gr, b1, 0, c->stream, glob_args(c));
        LAUNCH_CHECK();
      }
      CHK(trial_state(c, c0, cnt, 1));
    }
    std::vector<int> acc(c->nch_all);
    HIPC(hipMemcpyAsync(acc.data(), c->gacc, sizeof(int) * c->nch_all, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    for (int j = 0; j < c->nch_all; ++j)
      if (acc[j]) c->isb[j] = !c->isb[j];
  }
  c->left_done = -1;  // trial passes of either direction, some restored
  hipLaunchKernelGGL(k_wolff_finish, g1, b1, 0, c->stream, glob_args(c), (int)nglobal);
  LAUNCH_CHECK();
  return Z2_OK;
}

extern "C" int z2_get_global_counters(z2_ctx *c, int32_t chain, uint64_t out[3], double *size_cluster) {
  if (!c || !out || !size_cluster || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];
  unsigned long long v[3];
  double z[2];
  HIPC(hipMemcpyAsync(v, c->gcnt + 3 * chain, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipMemcpyAsync(z, c->gsz + 2 * chain, sizeof(z), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  for (int k = 0; k < 3; ++k) out[k] = v[k];
  *size_cluster = z[1];
  return Z2_OK;
}

// LocalK_therm sweep (src/localK.f90:400-420), the boson-only warm-up
extern "C" int z2_therm(z2_ctx *c) {
  if (c) CHK(u_flush(c));  // deferred propagator wraps first
  if (!c) return set_err(Z2_EINVAL, "null ctx");
  if (!c->have_fields) return set_err(Z2_ESTATE, "fields not initialised (z2_init_fields / z2_set_sigma)");
  CHK(h_zero_sync(c));  // Local_sweep_reset (src/local_sweep.f90:63-69)
  ThermArgs a;
  a.sig = c->sig;
  a.seed = c->seed;
  a.cc = c->dcc;
  a.lat = c->lat;
  a.T = c->T;
  a.cnt = c->tcnt;
  const size_t sh = ((3 * (size_t)c->N2 + 15) & ~(size_t)15) + 16 * sizeof(double);
  hipLaunchKernelGGL(k_therm, dim3(c->nch_all), dim3(256), sh, c->stream, a);
  LAUNCH_CHECK();
  return Z2_OK;
}

extern "C" int z2_get_therm_counters(z2_ctx *c, int32_t chain, uint64_t out[2]) {
  if (!c || !out || chain < 0 || chain >= c->nch_all) return set_err(Z2_EINVAL, "bad argument");
  chain = c->phys[chain];
  unsigned long long v[2];
  HIPC(hipMemcpyAsync(v, c->tcnt + 2 * chain, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  out[0] = v[0];
  out[1] = v[1];
  return Z2_OK;
}

template <typename Rb>
static int maxdiff(z2_ctx *ref, z2_ctx *c, const Rb *Gc) {
  hipLaunchKernelGGL((k_maxdiff<double, Rb>), dim3(c->nch), dim3(256), 0, c->stream, (const double *)ref->G,
                     Gc, c->cmp, c->N);
  LAUNCH_CHECK();
  return Z2_OK;
}

extern "C" int z2_sweep_shadow_tau(z2_ctx *c, z2_ctx *ref, int32_t is_beta, int32_t toggle, double *maxdiff_out) {
  if (!c || !ref || !maxdiff_out) return set_err(Z2_EINVAL, "null argument");
  if (ref->mixed || c->N != ref->N || c->T != ref->T || c->w != ref->w || c->nch != ref->nch ||
      c->p.Nsweep != ref->p.Nsweep)
    return set_err(Z2_EINVAL, "shadow needs an FP64 reference context of the same shape");
  if (!c->have_pre || !ref->have_pre) return set_err(Z2_ESTATE, "z2_pre must run on both contexts");
  if (c->phys != ref->phys) return set_err(Z2_ESTATE, "shadow needs both contexts' chains in the same storage order");
  if (is_beta != 0 && is_beta != 1) return set_err(Z2_EINVAL, "shadow sweeps take is_beta 0 or 1");
  if (toggle && (!c->tau_on || !ref->tau_on))
    return set_err(Z2_ESTATE, "time-displaced measurement not enabled on both contexts (z2_enable_tau)");
  for (z2_ctx *x : {c, ref}) CHK(z2_sweep_reset(x, toggle));
  CHK(dzero(c, c->cmp, (size_t)c->nch));
  std::vector<Step> v;
  sweep_steps(ref, is_beta, toggle, v);
  c->force = ref->rec;
  int rc = Z2_OK, nobst = 0;
  for (const Step &st : v) {
    if (step_recomputes(ref, st)) {  // propagated G of both, before the recompute
      HIPC(hipStreamSynchronize(ref->stream));
      rc = c->mixed ? maxdiff<float>(ref, c, c->Gf) : maxdiff<double>(ref, c, c->G);
      if (rc) break;
    }
    nobst += st.kind == ST_DYN;
    if ((rc = run_step(ref, st))) break;
    HIPC(hipStreamSynchronize(ref->stream));
    if ((rc = run_step(c, st))) break;
    HIPC(hipStreamSynchronize(c->stream));
  }
  c->force = nullptr;
  CHK(rc);
  for (z2_ctx *x : {c, ref}) {
    CHK(obs_ave(x));
    if (nobst > 0) CHK(z2_obs_tau_ave(x, nobst));
  }
  // cmp is indexed by storage slot; the caller's index is the chain's own
  std::vector<double> slotv(c->nch_all);
  HIPC(hipMemcpyAsync(slotv.data(), c->cmp, sizeof(double) * c->nch_all, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->nch_all; ++k) maxdiff_out[k] = slotv[c->phys[k]];
  return Z2_OK;
}

extern "C" int z2_sweep_shadow(z2_ctx *c, z2_ctx *ref, int32_t is_beta, double *maxdiff_out) {
  return z2_sweep_shadow_tau(c, ref, is_beta, 0, maxdiff_out);
}

// The deviation's course through one shadow sweep: after every step, max over the
// chains and elements of |G_ctx - G_ref| (and the step: kind << 16 | nt, kinds
// 0 Wrap_L, 1 Wrap_R, 2 Obs_equal, 3 slice L, 4 slice R); a diagnostic with a host
// synchronisation per step.
extern "C" int z2_sweep_shadow_trace(z2_ctx *c, z2_ctx *ref, int32_t is_beta, double *trace, int32_t *code,
                                     int32_t cap, int32_t *nsteps) {
  if (!c || !ref || !trace || !code || !nsteps || cap < 1) return set_err(Z2_EINVAL, "null argument");
  if (ref->mixed || c->N != ref->N || c->T != ref->T || c->w != ref->w || c->nch != ref->nch ||
      c->p.Nsweep != ref->p.Nsweep)
    return set_err(Z2_EINVAL, "shadow needs an FP64 reference context of the same shape");
  if (!c->have_pre || !ref->have_pre) return set_err(Z2_ESTATE, "z2_pre must run on both contexts");
  if (c->phys != ref->phys) return set_err(Z2_ESTATE, "shadow needs both contexts' chains in the same storage order");
  if (is_beta != 0 && is_beta != 1) return set_err(Z2_EINVAL, "shadow sweeps take is_beta 0 or 1");
  for (z2_ctx *x : {c, ref}) CHK(z2_sweep_reset(x, 0));
  std::vector<Step> v;
  sweep_steps(ref, is_beta, 0, v);
  if ((int)v.size() > cap) return set_err(Z2_EINVAL, "trace holds %d steps, the sweep has %zu", cap, v.size());
  c->force = ref->rec;
  int rc = Z2_OK;
  std::vector<double> m(c->nch);
  int i = 0;
  for (const Step &st : v) {
    if ((rc = run_step(ref, st))) break;
    HIPC(hipStreamSynchronize(ref->stream));
    if ((rc = run_step(c, st))) break;
    if ((rc = dzero(c, c->cmp, (size_t)c->nch))) break;
    rc = c->mixed ? maxdiff<float>(ref, c, c->Gf) : maxdiff<double>(ref, c, c->G);
    if (rc) break;
    HIPC(hipMemcpyAsync(m.data(), c->cmp, sizeof(double) * c->nch, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    trace[i] = *std::max_element(m.begin(), m.end());
    code[i] = ((int)st.kind << 16) | st.nt;
    ++i;
  }
  c->force = nullptr;
  CHK(rc);
  for (z2_ctx *x : {c, ref}) CHK(obs_ave(x));
  *nsteps = i;
  return Z2_OK;
}

// ---------------------------------------------------------------- batched QDRP
// QDRP_decompose (src/stabilization.f90:42-68) on caller matrices in device
// memory: A P = Q D R by the blocked Householder QR with the columns ordered by
// decreasing norm, one workgroup per matrix.  W is the QR's staging copy of the
// partially reduced columns (nmat n x n doubles); every element of W is written
// before it is read, so it needs no clearing.
static int qdrp_batched(int n, int nmat, double *A, double *D, int32_t *ipvt, double *tau, double *Q,
                        hipStream_t st, double *W) {
  int dev = 0, lds = 65536;
  HIPC(hipGetDevice(&dev));
  HIPC(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  int nb = 0, tc = 0;
  if (!qrb_config((size_t)lds, n, &nb, &tc)) return set_err(Z2_EINVAL, "order %d too large for the blocked QR", n);
  QrbArgs q;
  memset(&q, 0, sizeof(q));
  q.A = A; q.W = W; q.tau = tau; q.ipvt = ipvt; q.D = D; q.N = n; q.tc = tc; q.nopiv = 0;
  q.Ain = A; q.tin = tau; q.Q = Q;
  const size_t sh = qrb_lds_bytes(n, nb, tc);
  CHK(launch_qrb(0, nb, nmat, sh, st, q));
  if (Q) CHK(launch_qrb(1, nb, nmat, sh, st, q));
  return Z2_OK;
}

static int qdrp_args(int32_t n, int32_t nmat, const double *A, const double *D, const int32_t *ipvt,
                     const double *tau) {
  if (n < 1 || nmat < 1 || !A || !D || !ipvt || !tau) return set_err(Z2_EINVAL, "bad argument");
  return Z2_OK;
}

extern "C" int z2_qdrp_batched(int32_t n, int32_t nmat, double *A, double *D, int32_t *ipvt, double *tau,
                               double *Q, void *stream) {
  CHK(qdrp_args(n, nmat, A, D, ipvt, tau));
  hipStream_t st = (hipStream_t)stream;
  void *W = nullptr;
  hipError_t e = hipMalloc(&W, sizeof(double) * (size_t)nmat * n * n);
  if (e != hipSuccess) return set_err(Z2_ENOMEM, "hipMalloc(workspace): %s", hipGetErrorString(e));
  int rc = qdrp_batched(n, nmat, A, D, ipvt, tau, Q, st, (double *)W);
  hipError_t es = hipStreamSynchronize(st);  // W is released on return
  (void)hipFree(W);
  if (rc) return rc;
  if (es != hipSuccess) return set_err(Z2_EHIP, "hipStreamSynchronize: %s", hipGetErrorString(es));
  return Z2_OK;
}

// The same without allocation: the QR's staging copy W lives in `scratch`, a
// file-scope device array (every GPU the library runs on holds its own copy
// until the process ends).  In host code `scratch` names a host-side stand-in,
// not a device address: its device address is taken once per device with
// hipGetSymbolAddress and only that address is handed to kernels.  All calls on
// one GPU share the array, so calls that overlap on one GPU (two host threads, or
// two streams without an ordering between them) race on it.
static __device__ __align__(256) double scratch[Z2_QDRP_SCRATCH_DOUBLES];

static constexpr int SCRATCH_MAX_DEVICES = 64;
static double *g_scratch_dev[SCRATCH_MAX_DEVICES];
static std::mutex g_scratch_mu;

// device address of this device's `scratch` (the device must be current)
static int scratch_ptr(double **out) {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return set_err(Z2_EINVAL, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return set_err(Z2_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[dev] = (double *)p;
  }
  *out = g_scratch_dev[dev];
  return Z2_OK;
}

extern "C" int z2_qdrp_batched_noalloc(int32_t n, int32_t nmat, double *A, double *D, int32_t *ipvt, double *tau,
                                       double *Q, void *stream) {
  CHK(qdrp_args(n, nmat, A, D, ipvt, tau));
  if ((size_t)nmat * n * n > (size_t)Z2_QDRP_SCRATCH_DOUBLES)
    return set_err(Z2_EINVAL, "nmat * n^2 = %zu exceeds the %d doubles of the static workspace", (size_t)nmat * n * n,
                   (int)Z2_QDRP_SCRATCH_DOUBLES);
  double *W = nullptr;
  CHK(scratch_ptr(&W));
  return qdrp_batched(n, nmat, A, D, ipvt, tau, Q, (hipStream_t)stream, W);
}

// ---------------------------------------------------------------- kernel test hooks
// (include/z2dqmc_test.h) single-matrix runs of the dense kernels for parity tests.
namespace {
struct TmpBufs {
  std::vector<void *> v;
  ~TmpBufs() {
    for (void *p : v) (void)hipFree(p);
  }
  template <typename T>
  T *get(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (n ? n : 1)) != hipSuccess) return nullptr;
    v.push_back(p);
    return (T *)p;
  }
};
}  // namespace

extern "C" int z2_test_gemm(int32_t n, const double *A, int32_t ta, const double *B, int32_t tb,
                            int32_t triA, int32_t triB, double *C) {
  if (n < 1 || !A || !B || !C) return set_err(Z2_EINVAL, "bad argument");
  TmpBufs t;
  size_t nn = (size_t)n * n;
  double *dA = t.get<double>(nn), *dB = t.get<double>(nn), *dC = t.get<double>(nn);
  if (!dA || !dB || !dC) return set_err(Z2_ENOMEM, "alloc");
  HIPC(hipMemcpy(dA, A, nn * 8, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(dB, B, nn * 8, hipMemcpyHostToDevice));
  GemmArgs g;
  g.A = dA; g.B = dB; g.C = dC; g.sA = g.sB = g.sC = (long)nn; g.M = g.Nc = g.K = n;
  g.lda = g.ldb = g.ldc = n;
  g.ta = ta; g.tb = tb; g.triA = triA; g.triB = triB; g.alpha = 1.0; g.beta = 0.0;
  dim3 grid((n + GT - 1) / GT, (n + GT - 1) / GT, 1);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, 0, g, g, 1);
  LAUNCH_CHECK();
  HIPC(hipDeviceSynchronize());
  HIPC(hipMemcpy(C, dC, nn * 8, hipMemcpyDeviceToHost));
  return Z2_OK;
}

extern "C" int z2_test_qrp(int32_t n, double *A, double *D, int32_t *ipvt, double *tau, double *Q) {
  if (n < 1 || !A || !D || !ipvt || !tau) return set_err(Z2_EINVAL, "bad argument");
  TmpBufs t;
  size_t nn = (size_t)n * n;
  double *dA = t.get<double>(nn), *dD = t.get<double>(n), *dt = t.get<double>(n), *dQ = t.get<double>(nn);
  int *dp = t.get<int>(n);
  if (!dA || !dD || !dt || !dp || !dQ) return set_err(Z2_ENOMEM, "alloc");
  HIPC(hipMemcpy(dA, A, nn * 8, hipMemcpyHostToDevice));
  QrpArgs q;
  q.A = dA; q.tau = dt; q.ipvt = dp; q.D = dD; q.N = n; q.nopiv = 0;
  size_t sh = sizeof(double) * (2 * n + 32) + sizeof(int) * 32;
  CHK(launch_qrp(n, dim3(1), sh, 0, q));
  if (Q) CHK(launch_orgqr(n, dim3(1), 0, (const double *)dA, (const double *)dt, dQ));
  HIPC(hipDeviceSynchronize());
  HIPC(hipMemcpy(A, dA, nn * 8, hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(D, dD, n * 8, hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(tau, dt, n * 8, hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(ipvt, dp, n * 4, hipMemcpyDeviceToHost));
  if (Q) HIPC(hipMemcpy(Q, dQ, nn * 8, hipMemcpyDeviceToHost));
  return Z2_OK;
}

extern "C" int z2_test_qr_blocked(int32_t n, double *A, double *D, int32_t *ipvt, double *tau, double *Q) {
  if (n < 1 || !A || !D || !ipvt || !tau) return set_err(Z2_EINVAL, "bad argument");
  int dev = 0, lds = 65536;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lds <= 0)
    lds = 65536;
  int nb = 0, tc = 0;
  if (!qrb_config((size_t)lds, n, &nb, &tc)) return set_err(Z2_EINVAL, "order %d too large for the blocked QR", n);
  TmpBufs t;
  size_t nn = (size_t)n * n;
  double *dA = t.get<double>(nn), *dW = t.get<double>(nn), *dD = t.get<double>(n), *dt = t.get<double>(n);
  double *dQ = t.get<double>(nn);
  int *dp = t.get<int>(n);
  if (!dA || !dW || !dD || !dt || !dp || !dQ) return set_err(Z2_ENOMEM, "alloc");
  HIPC(hipMemcpy(dA, A, nn * 8, hipMemcpyHostToDevice));
  QrbArgs q;
  memset(&q, 0, sizeof(q));
  q.A = dA; q.W = dW; q.tau = dt; q.ipvt = dp; q.D = dD; q.N = n; q.tc = tc; q.nopiv = 0;
  q.Ain = dA; q.tin = dt; q.Q = dQ;
  const size_t sh = qrb_lds_bytes(n, nb, tc);
  CHK(launch_qrb(0, nb, 1, sh, 0, q));
  if (Q) CHK(launch_qrb(1, nb, 1, sh, 0, q));
  HIPC(hipDeviceSynchronize());
  HIPC(hipMemcpy(A, dA, nn * 8, hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(D, dD, n * 8, hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(tau, dt, n * 8, hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(ipvt, dp, n * 4, hipMemcpyDeviceToHost));
  if (Q) HIPC(hipMemcpy(Q, dQ, nn * 8, hipMemcpyDeviceToHost));
  return Z2_OK;
}

// plain streaming copy (calibrates hardware byte counters against a known traffic)
__global__ void __launch_bounds__(256) k_copy(const double *__restrict__ in, double *__restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

extern "C" int z2_test_copy(int64_t n, int32_t reps) {
  if (n < 1 || reps < 1) return set_err(Z2_EINVAL, "bad argument");
  TmpBufs t;
  double *a = t.get<double>((size_t)n), *b = t.get<double>((size_t)n);
  if (!a || !b) return set_err(Z2_ENOMEM, "alloc");
  HIPC(hipMemset(a, 0, sizeof(double) * (size_t)n));
  HIPC(hipDeviceSynchronize());
  for (int r = 0; r < reps; ++r) {
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, (const double *)a, b, (size_t)n);
    LAUNCH_CHECK();
  }
  HIPC(hipDeviceSynchronize());
  return Z2_OK;
}

extern "C" int z2_test_trsm(int32_t n, const double *R, double *X, int32_t right) {
  if (n < 1 || !R || !X) return set_err(Z2_EINVAL, "bad argument");
  TmpBufs t;
  size_t nn = (size_t)n * n;
  double *dR = t.get<double>(nn), *dX = t.get<double>(nn);
  if (!dR || !dX) return set_err(Z2_ENOMEM, "alloc");
  HIPC(hipMemcpy(dR, R, nn * 8, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(dX, X, nn * 8, hipMemcpyHostToDevice));
  int dev = 0, lds = 65536;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lds <= 0)
    lds = 65536;
  CHK(launch_trsm(n, 1, 0, (size_t)lds, dR, dX, right));
  HIPC(hipDeviceSynchronize());
  HIPC(hipMemcpy(X, dX, nn * 8, hipMemcpyDeviceToHost));
  return Z2_OK;
}

// ---------------------------------------------------------------- kernel timing hooks
extern "C" int z2_profile_kernel(z2_ctx *c, int32_t kind) {
  if (!c || kind < 0 || kind > Z2_K_QT_APPLY) return set_err(Z2_EINVAL, "bad kernel class %d", kind);
  HIPC(hipStreamSynchronize(c->stream));
  if (!c->prof_ref) HIPC(hipEventCreate(&c->prof_ref));
  HIPC(hipEventRecord(c->prof_ref, c->stream));
  c->prof_kind = kind;
  c->prof_used = 0;
  c->prof_work = 0.0;
  return Z2_OK;
}

extern "C" int z2_profile_spans(z2_ctx *c, z2_ctx *ref, double *start_ms, double *end_ms, int64_t cap,
                                int64_t *n) {
  if (!c || !ref || !start_ms || !end_ms || !n || cap < 0) return set_err(Z2_EINVAL, "bad argument");
  if (!ref->prof_ref) return set_err(Z2_ESTATE, "z2_profile_kernel has not run on the reference context");
  HIPC(hipStreamSynchronize(c->stream));
  HIPC(hipStreamSynchronize(ref->stream));
  const int64_t m = std::min<int64_t>((int64_t)c->prof_used, cap);
  for (int64_t k = 0; k < m; ++k) {
    float a = 0.f, b = 0.f;
    HIPC(hipEventElapsedTime(&a, ref->prof_ref, c->prof_ev[k].first));
    HIPC(hipEventElapsedTime(&b, ref->prof_ref, c->prof_ev[k].second));
    start_ms[k] = a;
    end_ms[k] = b;
  }
  *n = m;
  return Z2_OK;
}

extern "C" int z2_profile_read(z2_ctx *c, double *total_ms, int64_t *launches) {
  if (!c || !total_ms || !launches) return set_err(Z2_EINVAL, "null argument");
  HIPC(hipStreamSynchronize(c->stream));
  double t = 0.0;
  for (size_t k = 0; k < c->prof_used; ++k) {
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, c->prof_ev[k].first, c->prof_ev[k].second));
    t += ms;
  }
  *total_ms = t;
  *launches = (int64_t)c->prof_used;
  return Z2_OK;
}

extern "C" int z2_profile_work(z2_ctx *c, double *work) {
  if (!c || !work) return set_err(Z2_EINVAL, "null argument");
  *work = c->prof_work;
  return Z2_OK;
}

#include "batch.hpp"
