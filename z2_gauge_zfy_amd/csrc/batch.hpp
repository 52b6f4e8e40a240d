// Batches of independent chains of uneven lattice size, Ltrot, Nwrap and
// couplings (the reference runs one chain per MPI rank, src/main.f90:23-25,
// src/fields.f90:243-304).  Chains are grouped by shape (Nlx, Nly, Ltrot,
// Nwrap); each group is one context, whose batched kernels are compiled for its
// size, on its own HIP stream.  A sweep of the batch issues every group's steps in
// turn on the groups' streams (eager), or replays one captured HIP graph that
// holds all groups' launches as parallel branches (one graph launch per sweep).
// The autotuner picks, per group, the delay depth, the kinetic tiles and Nwrap,
// then the execution mode, by timing sweeps of clones of the batch's chains.
// Included at the end of z2dqmc.hip (host code on the context internals).
#pragma once
#include <chrono>
#include <memory>

struct z2_batch {
  std::vector<z2_ctx *> grp;               // one context per group
  std::vector<std::vector<int>> members;   // group -> batch chain indices, in group order
  std::vector<std::pair<int, int>> where;  // batch chain -> (group, index in the group)
  std::vector<z2_batch_chain> chains;
  int precision = Z2_PREC_FP64, Nsweep = 1, max_group = 0;
  int exec = Z2_EXEC_EAGER;
  bool have_pre = false;
  // the captured sweep: valid for (is_beta, toggle) and the groups' settings in `stamp`
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_isb = -1, g_toggle = -1;
  std::vector<long> stamp;
  std::vector<int> swept;  // eager sweeps run per (is_beta, toggle) key: capture after the first
  hipEvent_t ev_fork = nullptr, ev_done = nullptr;
  std::vector<hipEvent_t> ev_join;
};

static void batch_drop_graph(z2_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->g_isb = b->g_toggle = -1;
}

static void batch_free(z2_batch *b) {
  batch_drop_graph(b);
  for (z2_ctx *c : b->grp)
    if (c) (void)z2_ctx_destroy(c);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_done) (void)hipEventDestroy(b->ev_done);
  for (hipEvent_t e : b->ev_join)
    if (e) (void)hipEventDestroy(e);
  delete b;
}

// a context for `idx` chains of the batch (all of one shape) with Nwrap = nwrap
static int make_group_ctx(const z2_batch *b, const std::vector<int> &idx, int nwrap, z2_ctx **out) {
  const z2_batch_chain &h = b->chains[idx[0]];
  z2_params p;
  p.Nlx = h.Nlx;
  p.Nly = h.Nly;
  p.Ltrot = h.Ltrot;
  p.Nwrap = nwrap;
  p.Nsweep = b->Nsweep;
  p.precision = b->precision;
  std::vector<z2_chain_desc> d;
  for (int k : idx) d.push_back(b->chains[k].d);
  return z2_ctx_create(&p, (int32_t)d.size(), d.data(), out);
}

// the events a sweep synchronises the groups' streams with
static int batch_events(z2_batch *b) {
  b->ev_join.assign(b->grp.size(), nullptr);
  hipError_t e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming);
  for (size_t g = 0; g < b->grp.size() && e == hipSuccess; ++g)
    e = hipEventCreateWithFlags(&b->ev_join[g], hipEventDisableTiming);
  if (e != hipSuccess) return set_err(Z2_EHIP, "hipEventCreate: %s", hipGetErrorString(e));
  b->swept.assign(4, 0);
  return Z2_OK;
}

extern "C" int z2_batch_create(const z2_batch_chain *chains, int32_t n, int32_t precision, int32_t Nsweep,
                               int32_t max_group, z2_batch **out) {
  if (!chains || n < 1 || !out || Nsweep < 1 || max_group < 0 ||
      (precision != Z2_PREC_FP64 && precision != Z2_PREC_MIXED))
    return set_err(Z2_EINVAL, "bad argument");
  *out = nullptr;
  z2_batch *b = new z2_batch();
  b->chains.assign(chains, chains + n);
  b->precision = precision;
  b->Nsweep = Nsweep;
  b->max_group = max_group;
  // groups of one shape, in order of first appearance; split into near-equal parts
  // of at most max_group chains, dealt round-robin (chain j of the shape to part
  // j mod parts), so that every part gets a share of each run of similar chains
  // (chains next to each other often share couplings, and a launch lasts as long
  // as its slowest chain)
  std::vector<std::vector<int>> shapes;
  for (int k = 0; k < n; ++k) {
    const z2_batch_chain &c = chains[k];
    bool found = false;
    for (auto &g : shapes) {
      const z2_batch_chain &h = chains[g[0]];
      if (h.Nlx == c.Nlx && h.Nly == c.Nly && h.Ltrot == c.Ltrot && h.Nwrap == c.Nwrap) {
        g.push_back(k);
        found = true;
        break;
      }
    }
    if (!found) shapes.push_back({k});
  }
  for (auto &g : shapes) {
    const int parts = max_group > 0 ? ((int)g.size() + max_group - 1) / max_group : 1;
    for (int q = 0; q < parts; ++q) {
      std::vector<int> part;
      for (size_t j = q; j < g.size(); j += parts) part.push_back(g[j]);
      b->members.push_back(part);
    }
  }
  b->where.resize(n);
  for (size_t gi = 0; gi < b->members.size(); ++gi) {
    z2_ctx *c = nullptr;
    int rc = make_group_ctx(b, b->members[gi], chains[b->members[gi][0]].Nwrap, &c);
    if (rc) {
      batch_free(b);
      return rc;
    }
    b->grp.push_back(c);
    for (size_t j = 0; j < b->members[gi].size(); ++j) b->where[b->members[gi][j]] = {(int)gi, (int)j};
  }
  int rc = batch_events(b);
  if (rc) {
    batch_free(b);
    return rc;
  }
  *out = b;
  return Z2_OK;
}

extern "C" int z2_batch_destroy(z2_batch *b) {
  if (!b) return set_err(Z2_EINVAL, "null batch");
  for (z2_ctx *c : b->grp) (void)hipStreamSynchronize(c->stream);
  batch_free(b);
  return Z2_OK;
}

extern "C" int z2_batch_info(z2_batch *b, int32_t *n_chains, int32_t *n_groups) {
  if (!b || !n_chains || !n_groups) return set_err(Z2_EINVAL, "null argument");
  *n_chains = (int32_t)b->chains.size();
  *n_groups = (int32_t)b->grp.size();
  return Z2_OK;
}

extern "C" int z2_batch_chain_ctx(z2_batch *b, int32_t chain, z2_ctx **ctx, int32_t *index) {
  if (!b || !ctx || !index || chain < 0 || chain >= (int32_t)b->chains.size())
    return set_err(Z2_EINVAL, "bad argument");
  *ctx = b->grp[b->where[chain].first];
  *index = b->where[chain].second;
  return Z2_OK;
}

extern "C" int z2_batch_group_ctx(z2_batch *b, int32_t group, z2_ctx **ctx) {
  if (!b || !ctx || group < 0 || group >= (int32_t)b->grp.size()) return set_err(Z2_EINVAL, "bad argument");
  *ctx = b->grp[group];
  return Z2_OK;
}

extern "C" int z2_batch_init_fields(z2_batch *b) {
  if (!b) return set_err(Z2_EINVAL, "null batch");
  for (z2_ctx *c : b->grp) CHK(z2_init_fields(c));
  return Z2_OK;
}

extern "C" int z2_batch_enable_tau(z2_batch *b) {
  if (!b) return set_err(Z2_EINVAL, "null batch");
  for (z2_ctx *c : b->grp) CHK(z2_enable_tau(c));
  batch_drop_graph(b);
  return Z2_OK;
}

extern "C" int z2_batch_pre(z2_batch *b) {
  if (!b) return set_err(Z2_EINVAL, "null batch");
  for (z2_ctx *c : b->grp) CHK(z2_pre(c));
  b->have_pre = true;
  b->swept.assign(4, 0);
  return Z2_OK;
}

extern "C" int z2_batch_sync(z2_batch *b) {
  if (!b) return set_err(Z2_EINVAL, "null batch");
  for (z2_ctx *c : b->grp) HIPC(hipStreamSynchronize(c->stream));
  return Z2_OK;
}

extern "C" int z2_batch_set_exec(z2_batch *b, int32_t mode) {
  if (!b || (mode != Z2_EXEC_EAGER && mode != Z2_EXEC_GRAPH)) return set_err(Z2_EINVAL, "bad argument");
  b->exec = mode;
  return Z2_OK;
}

extern "C" int z2_batch_get_exec(z2_batch *b, int32_t *mode) {
  if (!b || !mode) return set_err(Z2_EINVAL, "bad argument");
  *mode = b->exec;
  return Z2_OK;
}

// the host-side settings a captured sweep bakes into its launches
static std::vector<long> batch_stamp(const z2_batch *b) {
  std::vector<long> v;
  for (const z2_ctx *c : b->grp) {
    v.push_back(c->kdelay);
    v.push_back(c->kin_cb);
    v.push_back(c->kin_rb);
    v.push_back(c->qr_mode);
    v.push_back(c->stab_mode);
    v.push_back(c->tau_on);
    v.push_back(c->prof_kind);
  }
  return v;
}

// every group's steps of one sweep, issued in turn on the groups' streams so that
// each stream's queue fills while the others run
static int batch_issue(z2_batch *b, int is_beta, int toggle) {
  const size_t G = b->grp.size();
  std::vector<std::vector<Step>> steps(G);
  std::vector<int> nobst(G, 0);
  size_t most = 0;
  for (size_t g = 0; g < G; ++g) {
    CHK(sweep_prologue(b->grp[g], toggle));
    sweep_steps(b->grp[g], is_beta, toggle, steps[g]);
    most = std::max(most, steps[g].size());
  }
  for (size_t i = 0; i < most; ++i)
    for (size_t g = 0; g < G; ++g)
      if (i < steps[g].size()) {
        CHK(run_step(b->grp[g], steps[g][i]));
        nobst[g] += steps[g][i].kind == ST_DYN;
      }
  for (size_t g = 0; g < G; ++g) CHK(sweep_epilogue(b->grp[g], toggle, nobst[g]));
  return Z2_OK;
}

// one sweep captured as a graph: the first group's stream forks to the others by
// an event, every group's launches are recorded on its own stream, all join back
static int batch_capture(z2_batch *b, int is_beta, int toggle) {
  batch_drop_graph(b);
  hipStream_t s0 = b->grp[0]->stream;
  HIPC(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
  int rc = Z2_OK;
  hipError_t e = hipEventRecord(b->ev_fork, s0);
  for (size_t g = 1; g < b->grp.size() && e == hipSuccess; ++g) e = hipStreamWaitEvent(b->grp[g]->stream, b->ev_fork, 0);
  if (e == hipSuccess) rc = batch_issue(b, is_beta, toggle);
  for (size_t g = 1; g < b->grp.size() && e == hipSuccess && !rc; ++g) {
    e = hipEventRecord(b->ev_join[g], b->grp[g]->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s0, b->ev_join[g], 0);
  }
  hipGraph_t graph = nullptr;
  hipError_t ee = hipStreamEndCapture(s0, &graph);
  if (rc || e != hipSuccess || ee != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;
    return set_err(Z2_EHIP, "sweep capture: %s", hipGetErrorString(e != hipSuccess ? e : ee));
  }
  b->graph = graph;
  e = hipGraphInstantiateWithFlags(&b->gexec, graph, 0);
  if (e != hipSuccess) {
    batch_drop_graph(b);
    return set_err(Z2_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  b->g_isb = is_beta;
  b->g_toggle = toggle;
  b->stamp = batch_stamp(b);
  return Z2_OK;
}

extern "C" int z2_batch_sweep(z2_batch *b, int32_t is_beta, int32_t toggle) {
  if (!b) return set_err(Z2_EINVAL, "null batch");
  if (!b->have_pre) return set_err(Z2_ESTATE, "z2_batch_pre must run before z2_batch_sweep");
  if (toggle != 0 && toggle != 1) return set_err(Z2_EINVAL, "toggle must be 0 or 1");
  if (is_beta == Z2_BETA_PER_CHAIN) {  // directions differ after Wolff updates: per group, eager
    for (z2_ctx *c : b->grp) CHK(z2_sweep_tau(c, is_beta, toggle));
    return Z2_OK;
  }
  for (z2_ctx *c : b->grp) {
    CHK(sweep_check(c, is_beta, toggle));
    CHK(h_zero_sync(c));
    c->isb.assign(c->nch_all, is_beta);
  }
  const int key = 2 * (is_beta ? 1 : 0) + toggle;
  bool profiling = false;
  for (z2_ctx *c : b->grp) profiling = profiling || c->prof_kind != 0;
  // graph mode: the first sweep of a kind runs eagerly (the kernels' launch
  // attributes are set up), the next ones replay the captured graph
  if (b->exec != Z2_EXEC_GRAPH || profiling || !b->swept[key]) {
    CHK(batch_issue(b, is_beta, toggle));
    b->swept[key] = 1;
    return Z2_OK;
  }
  if (!b->gexec || b->g_isb != is_beta || b->g_toggle != toggle || b->stamp != batch_stamp(b))
    CHK(batch_capture(b, is_beta, toggle));
  hipStream_t s0 = b->grp[0]->stream;
  // the graph runs on the first group's stream after everything already queued on
  // any group's stream, and every group's stream waits for it
  for (size_t g = 1; g < b->grp.size(); ++g) {
    HIPC(hipEventRecord(b->ev_join[g], b->grp[g]->stream));
    HIPC(hipStreamWaitEvent(s0, b->ev_join[g], 0));
  }
  HIPC(hipGraphLaunch(b->gexec, s0));
  HIPC(hipEventRecord(b->ev_done, s0));
  for (size_t g = 1; g < b->grp.size(); ++g) HIPC(hipStreamWaitEvent(b->grp[g]->stream, b->ev_done, 0));
  return Z2_OK;
}

// ---------------------------------------------------------------- autotuner
// A group's tuned settings.
struct GroupTune {
  int nwrap, kdelay, cb, rb;
};

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// field, seed and lambda of chain k of context c (host copies)
struct ChainState {
  std::vector<double> sig, lam;
  int32_t seed = 0;
};

static int save_state(z2_ctx *c, std::vector<ChainState> &st) {
  st.resize(c->nch_all);
  for (int k = 0; k < c->nch_all; ++k) {
    st[k].sig.resize((size_t)c->T * c->N2);
    st[k].lam.resize(c->N);
    CHK(z2_get_sigma(c, k, st[k].sig.data()));
    CHK(z2_get_lambda(c, k, st[k].lam.data()));
    CHK(z2_get_seed(c, k, &st[k].seed));
  }
  return Z2_OK;
}

static int load_state(z2_ctx *c, const std::vector<ChainState> &st) {
  for (int k = 0; k < c->nch_all; ++k) {
    CHK(z2_set_sigma(c, k, st[k].sig.data(), st[k].seed));
    CHK(z2_set_lambda(c, k, st[k].lam.data()));
  }
  return Z2_OK;
}

static int timed_sweeps(z2_ctx *c, int n, double *t) {
  CHK(z2_sync(c));
  const double t0 = now_s();
  for (int i = 0; i < n; ++i) CHK(z2_sweep(c, 1));
  CHK(z2_sync(c));
  *t = (now_s() - t0) / n;
  return Z2_OK;
}

// RAII holder of a clone context
struct CtxHolder {
  z2_ctx *c = nullptr;
  ~CtxHolder() {
    if (c) (void)z2_ctx_destroy(c);
  }
};

static int tune_group(z2_batch *b, int g, int warm, int sweeps, double drift_tol, GroupTune *out,
                      std::vector<ChainState> &warmed) {
  z2_ctx *orig = b->grp[g];
  std::vector<ChainState> st0;
  CHK(save_state(orig, st0));
  GroupTune best{orig->w, orig->kdelay, orig->kin_cb, orig->kin_rb};
  // delay depth and tiles on a clone with the group's own Nwrap, after warm-up sweeps
  CtxHolder h;
  CHK(make_group_ctx(b, b->members[g], orig->w, &h.c));
  CHK(load_state(h.c, st0));
  CHK(z2_pre(h.c));
  for (int i = 0; i < warm; ++i) CHK(z2_sweep(h.c, 1));
  double tbest = 1e300, t = 0.0;
  for (int k : {4, 8, 16, 32}) {
    if (k > h.c->kdelay_max) continue;
    CHK(z2_set_delay(h.c, k));
    CHK(timed_sweeps(h.c, sweeps, &t));
    if (t < tbest) {
      tbest = t;
      best.kdelay = k;
    }
  }
  CHK(z2_set_delay(h.c, best.kdelay));
  tbest = 1e300;
  for (int cb : {8, 16})
    for (int rb : {8, 32}) {
      if (z2_set_kin_tiles(h.c, cb, rb) != Z2_OK) continue;
      CHK(timed_sweeps(h.c, sweeps, &t));
      if (t < tbest) {
        tbest = t;
        best.cb = cb;
        best.rb = rb;
      }
    }
  CHK(save_state(h.c, warmed));
  // Nwrap: the fastest whose drift (max |G_recomputed - G_propagated| over the timed
  // sweeps and all chains) stays within drift_tol, all from the same warmed state
  double tw = 1e300;
  int nw_best = orig->w;
  for (int nw : {5, 10, 20, orig->w}) {
    if (orig->T % nw) continue;
    CtxHolder hw;
    CHK(make_group_ctx(b, b->members[g], nw, &hw.c));
    CHK(load_state(hw.c, warmed));
    CHK(z2_set_delay(hw.c, std::min(best.kdelay, hw.c->kdelay_max)));
    CHK(z2_set_kin_tiles(hw.c, best.cb, best.rb));
    CHK(z2_pre(hw.c));
    CHK(timed_sweeps(hw.c, sweeps, &t));
    double dmax = 0.0;
    for (int k = 0; k < hw.c->nch_all; ++k) {
      double xm = 0.0, xa = 0.0;
      CHK(z2_get_drift(hw.c, k, &xm, &xa));
      dmax = std::max(dmax, xm);
    }
    if (dmax <= drift_tol && t < tw) {
      tw = t;
      nw_best = nw;
    }
  }
  best.nwrap = nw_best;
  *out = best;
  return Z2_OK;
}

// group g of the batch rebuilt with Nwrap = nwrap from the chains' current state
static int regroup(z2_batch *b, int g, int nwrap) {
  z2_ctx *orig = b->grp[g];
  if (orig->w == nwrap) return Z2_OK;
  std::vector<ChainState> st;
  CHK(save_state(orig, st));
  z2_ctx *c = nullptr;
  CHK(make_group_ctx(b, b->members[g], nwrap, &c));
  int rc = load_state(c, st);
  if (rc) {
    (void)z2_ctx_destroy(c);
    return rc;
  }
  for (int k : b->members[g]) b->chains[k].Nwrap = nwrap;
  (void)z2_ctx_destroy(orig);
  b->grp[g] = c;
  batch_drop_graph(b);
  return Z2_OK;
}

static int apply_tune(z2_batch *b, int g, const GroupTune &t) {
  CHK(regroup(b, g, t.nwrap));
  CHK(z2_set_delay(b->grp[g], std::min(t.kdelay, b->grp[g]->kdelay_max)));
  return z2_set_kin_tiles(b->grp[g], t.cb, t.rb);
}

static void tune_line(const z2_batch *b, int g, char *buf, size_t n) {
  const z2_ctx *c = b->grp[g];
  snprintf(buf, n, "group %d %d %d %d %d %d %d %d %d\n", c->Lx, c->Ly, c->T, b->precision, c->nch_all, c->w,
           c->kdelay, c->kin_cb, c->kin_rb);
}

static int write_tuning(const z2_batch *b, const char *path) {
  FILE *f = fopen(path, "w");
  if (!f) return set_err(Z2_EINVAL, "cannot write %s", path);
  fprintf(f, "# z2dqmc batch tuning: group Nlx Nly Ltrot precision n_chains Nwrap delay rows_cb cols_rb\n");
  char line[256];
  for (size_t g = 0; g < b->grp.size(); ++g) {
    tune_line(b, (int)g, line, sizeof(line));
    fputs(line, f);
  }
  fprintf(f, "exec %d\n", b->exec);
  fclose(f);
  return Z2_OK;
}

extern "C" int z2_batch_autotune(z2_batch *b, int32_t warm, int32_t sweeps, double drift_tol, const char *path) {
  if (!b || warm < 0 || sweeps < 1 || !(drift_tol > 0.0)) return set_err(Z2_EINVAL, "bad argument");
  if (b->have_pre) return set_err(Z2_ESTATE, "z2_batch_autotune runs before z2_batch_pre");
  for (z2_ctx *c : b->grp)
    if (!c->have_fields) return set_err(Z2_ESTATE, "fields not initialised (z2_batch_init_fields)");
  const size_t G = b->grp.size();
  std::vector<GroupTune> tune(G);
  std::vector<std::vector<ChainState>> warmed(G);
  for (size_t g = 0; g < G; ++g) CHK(tune_group(b, (int)g, warm, sweeps, drift_tol, &tune[g], warmed[g]));
  // execution mode: a clone of the whole batch (same groups) from the warmed states,
  // eager then graph
  {
    std::unique_ptr<z2_batch, void (*)(z2_batch *)> cl(new z2_batch(), [](z2_batch *x) { batch_free(x); });
    cl->chains = b->chains;
    cl->precision = b->precision;
    cl->Nsweep = b->Nsweep;
    cl->members = b->members;
    cl->where = b->where;
    for (size_t g = 0; g < G; ++g) {
      z2_ctx *c = nullptr;
      CHK(make_group_ctx(b, b->members[g], tune[g].nwrap, &c));
      cl->grp.push_back(c);
      CHK(load_state(c, warmed[g]));
      CHK(z2_set_delay(c, std::min(tune[g].kdelay, c->kdelay_max)));
      CHK(z2_set_kin_tiles(c, tune[g].cb, tune[g].rb));
    }
    CHK(batch_events(cl.get()));
    CHK(z2_batch_pre(cl.get()));
    double te = 0.0, tg = 0.0;
    for (int mode : {Z2_EXEC_EAGER, Z2_EXEC_GRAPH}) {
      cl->exec = mode;
      CHK(z2_batch_sweep(cl.get(), 1, 0));  // graph mode: this one runs eagerly, the next ones replay
      CHK(z2_batch_sync(cl.get()));
      const double t0 = now_s();
      for (int i = 0; i < sweeps; ++i) CHK(z2_batch_sweep(cl.get(), 1, 0));
      CHK(z2_batch_sync(cl.get()));
      (mode == Z2_EXEC_EAGER ? te : tg) = now_s() - t0;
    }
    b->exec = tg < te ? Z2_EXEC_GRAPH : Z2_EXEC_EAGER;
  }
  for (size_t g = 0; g < G; ++g) CHK(apply_tune(b, (int)g, tune[g]));
  if (path && path[0]) CHK(write_tuning(b, path));
  return Z2_OK;
}

extern "C" int z2_batch_load_tuning(z2_batch *b, const char *path) {
  if (!b || !path) return set_err(Z2_EINVAL, "bad argument");
  if (b->have_pre) return set_err(Z2_ESTATE, "z2_batch_load_tuning runs before z2_batch_pre");
  FILE *f = fopen(path, "r");
  if (!f) return set_err(Z2_EINVAL, "cannot read %s", path);
  char line[512];
  std::vector<bool> done(b->grp.size(), false);
  int rc = Z2_OK;
  while (!rc && fgets(line, sizeof(line), f)) {
    int lx, ly, T, prec, nch, nw, kd, cb, rb, ex;
    if (sscanf(line, "group %d %d %d %d %d %d %d %d %d", &lx, &ly, &T, &prec, &nch, &nw, &kd, &cb, &rb) == 9) {
      for (size_t g = 0; g < b->grp.size(); ++g) {
        const z2_ctx *c = b->grp[g];
        if (done[g] || c->Lx != lx || c->Ly != ly || c->T != T || b->precision != prec || c->nch_all != nch ||
            T % nw)
          continue;
        rc = apply_tune(b, (int)g, GroupTune{nw, kd, cb, rb});
        done[g] = true;
        break;
      }
    } else if (sscanf(line, "exec %d", &ex) == 1 && (ex == Z2_EXEC_EAGER || ex == Z2_EXEC_GRAPH)) {
      b->exec = ex;
    }
  }
  fclose(f);
  return rc;
}
