// C-ABI of the batched power-flow engine (include/batchpf.h): handle, per-
// instance setup on the host, device buffers and the Newton driver.
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/batchpf.h"
#include "kernels.hip"
#include "symbolic.hpp"

using namespace bpf;

namespace {

struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct TimedEvent {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace

static unsigned div_up(size_t a, unsigned b) { return (unsigned)((a + b - 1) / b); }

// form of the sweeps (k_lu_solve's ep): 2 slot form (X in the solve LDS,
// workgroups of 1024, 512 or 256 lanes), 1 entry-parallel (X and the
// product buffer both in the solve LDS), 0 level rows.  form: 0 the fastest
// that applies, 1 (BPF_SWEEP_ENTRY) entry-parallel at most, 2
// (BPF_SWEEP_ROWS) level rows.
static int sweep_mode(const Symbolic &S, size_t scalar, int use_lds, int form,
                      int threads) {
  if (!use_lds || form == 2) return 0;
  if (form == 0 && S.sl_ok && (threads == 1024 || threads == 512 || threads == 256))
    return 2;
  if (!S.sw_ok) return 0;
  return ((size_t)ep_offset(S.nb) + 2 * (size_t)SWEEP_CHUNK) * scalar <=
                 (size_t)SOLVE_LDS_BYTES
             ? 1
             : 0;
}
// 1-D grid of the kernels that keep an instance's workgroups on one XCD
// (xcd_slot): 8 * ceil(T / 8) * parts workgroups
static dim3 xgrid(int T, size_t count, unsigned threads) {
  return dim3(8u * div_up((size_t)T, 8) * div_up(count, threads));
}

struct bpf_engine {
  bpf_options opt;
  std::string err;
  int dev = 0;
  hipStream_t stream = nullptr;
  // case (host copies)
  int n = 0, m = 0, ng = 0;
  std::vector<int> bus_type, br_from, br_to, gen_bus;
  std::vector<double> Pd, Qd, Gs, Bs, Vm, Va, R, X, B, tap, shift, Pg, Qg,
      Qmax, Qmin, VG;
  std::vector<uint8_t> br_status, gen_status;
  Symbolic S;
  // shared device arrays
  std::vector<DBuf> owned;
  DevTopo g{};
  FactorPlan fplan{};            // factorization schedule (device arrays)
  SolvePlan splan{};             // sweep schedule (device arrays)
  // per-instance device arrays
  std::vector<DBuf> inst_owned;
  int T = 0;
  DevInst d{};
  const int8_t *d_type0 = nullptr;
  const double *d_Qinj0 = nullptr;
  void *d_LU = nullptr;         // B4<float> or B4<double> [T][nnzLU]
  int lu_fp64 = -1;             // precision the LU buffer was sized for
  void *d_xglob = nullptr;      // solve vectors when LDS is too small
  void *d_TL = nullptr;         // inverse tail Schur complements [T][tn*tn]
  cd *d_Sf = nullptr, *d_St = nullptr;
  double *d_loss = nullptr;
  int *h_count = nullptr;       // pinned
  // launch shapes
  int lu_threads = 1024, solve_threads = 1024, mis_threads = 256;
  int max_lds = 65536;
  // graph of one Newton iteration
  hipGraphExec_t graph = nullptr;
  int graph_key = -1;
  bool graph_broken = false;   // a replay failed: direct launches from then on
  std::string graph_err;
  std::string graph_note;      // runtime oddities seen around replays
  int graph_used = 0;          // the last solve replayed graphs
  // timing
  int timing = 0;
  std::vector<TimedEvent> events;
  std::string launch_err;            // first failed launch (kernel name)
  const char *cur_kernel = "";
  int debug_sync = 0;                // BPF_DEBUG_SYNC: sync after launches
  int sweep_form = 0;                // sweep_mode's form (environment)
  bool capturing = false;            // recording the iteration graph
  std::map<std::string, std::pair<int, double>> ktime;
  bool solved = false;
  double *scratch = nullptr;         // this device's allocation-free scratch
};

static bpf_status scratch_setup(bpf_engine *e);

#define HIPCHK(expr)                                                        \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      e->err = std::string(#expr) + ": " + hipGetErrorString(e_);           \
      if (!e->launch_err.empty()) e->err += " [first failed launch: " +     \
                                            e->launch_err + "]";            \
      return BPF_ERR_HIP;                                                   \
    }                                                                       \
  } while (0)

static bpf_status dalloc(bpf_engine *e, std::vector<DBuf> &pool, size_t bytes,
                         void **out) {
  void *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t r = hipMalloc(&p, bytes);
  if (r != hipSuccess) {
    e->err = "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(r);
    return r == hipErrorOutOfMemory ? BPF_ERR_OOM : BPF_ERR_HIP;
  }
  pool.push_back({p, bytes});
  *out = p;
  return BPF_OK;
}

// plan arrays the kernels index with clamped indices: never empty on the
// device (a zero element of `width` ints stands in for an empty one)
static std::vector<int> nonempty(const std::vector<int> &v, int width) {
  return v.empty() ? std::vector<int>(width, 0) : v;
}

template <typename V>
static bpf_status upload(bpf_engine *e, std::vector<DBuf> &pool, const V &v,
                         const typename V::value_type **out) {
  void *p;
  size_t bytes = v.size() * sizeof(typename V::value_type);
  bpf_status s = dalloc(e, pool, bytes, &p);
  if (s) return s;
  if (bytes) {
    hipError_t r = hipMemcpy(p, v.data(), bytes, hipMemcpyHostToDevice);
    if (r != hipSuccess) {
      e->err = std::string("hipMemcpy: ") + hipGetErrorString(r);
      return BPF_ERR_HIP;
    }
  }
  *out = (const typename V::value_type *)p;
  return BPF_OK;
}

static void free_pool(std::vector<DBuf> &pool) {
  for (auto &b : pool) hipFree(b.p);
  pool.clear();
}

extern "C" void bpf_default_options(bpf_options *o) {
  memset(o, 0, sizeof(*o));
  o->alg = BPF_NR;
  o->precision = BPF_FP64;
  o->refine_steps = 1;
  o->maxit_nr = 10;
  o->maxit_fd = 30;
  o->enforce_qlim = 0;
  o->max_qlim_rounds = 10;
  o->use_graph = 1;
  o->tol = 1e-8;
  o->qlim_tol = 1e-6;
}

extern "C" const char *bpf_last_error(const bpf_engine *e) {
  if (!e) return "null engine";
  if (e->err.empty() && e->graph_broken) return e->graph_err.c_str();
  if (e->err.empty()) return e->graph_note.c_str();
  return e->err.c_str();
}

extern "C" void bpf_destroy(bpf_engine *e) {
  if (!e) return;
  if (e->graph) hipGraphExecDestroy(e->graph);
  for (auto &ev : e->events) { hipEventDestroy(ev.a); hipEventDestroy(ev.b); }
  free_pool(e->inst_owned);
  free_pool(e->owned);
  if (e->h_count) hipHostFree(e->h_count);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
}

template <typename T>
static std::vector<T> vcopy(const T *p, int n) {
  return p ? std::vector<T>(p, p + n) : std::vector<T>();
}

extern "C" bpf_status bpf_create(const bpf_case *c, const bpf_options *o,
                                 bpf_engine **out) {
  if (!out) return BPF_ERR_INVALID_ARG;
  *out = nullptr;
  bpf_engine *e = new bpf_engine();
  *out = e;
  if (!c || c->nbus <= 0 || c->nbranch < 0 || c->ngen < 0) {
    e->err = "invalid case sizes";
    return BPF_ERR_INVALID_ARG;
  }
  const void *need[] = {c->bus_type, c->Pd, c->Qd, c->Gs, c->Bs, c->Vm, c->Va};
  for (auto p : need)
    if (!p) { e->err = "null bus array"; return BPF_ERR_INVALID_ARG; }
  if (c->nbranch && (!c->br_from || !c->br_to || !c->R || !c->X || !c->B ||
                     !c->tap || !c->shift || !c->br_status)) {
    e->err = "null branch array";
    return BPF_ERR_INVALID_ARG;
  }
  if (c->ngen && (!c->gen_bus || !c->Pg || !c->Qg || !c->Qmax || !c->Qmin ||
                  !c->VG || !c->gen_status)) {
    e->err = "null generator array";
    return BPF_ERR_INVALID_ARG;
  }
  if (o) e->opt = *o; else bpf_default_options(&e->opt);
  if (e->opt.alg < 0 || e->opt.alg > 2 || e->opt.precision < 0 ||
      e->opt.precision > 1 || e->opt.refine_steps < 0 || e->opt.maxit_nr < 0 ||
      e->opt.maxit_fd < 0 || !(e->opt.tol > 0)) {
    e->err = "invalid options";
    return BPF_ERR_INVALID_ARG;
  }
  e->n = c->nbus; e->m = c->nbranch; e->ng = c->ngen;
  e->bus_type = vcopy(c->bus_type, e->n);
  e->Pd = vcopy(c->Pd, e->n); e->Qd = vcopy(c->Qd, e->n);
  e->Gs = vcopy(c->Gs, e->n); e->Bs = vcopy(c->Bs, e->n);
  e->Vm = vcopy(c->Vm, e->n); e->Va = vcopy(c->Va, e->n);
  e->br_from = vcopy(c->br_from, e->m); e->br_to = vcopy(c->br_to, e->m);
  e->R = vcopy(c->R, e->m); e->X = vcopy(c->X, e->m); e->B = vcopy(c->B, e->m);
  e->tap = vcopy(c->tap, e->m); e->shift = vcopy(c->shift, e->m);
  e->br_status = vcopy(c->br_status, e->m);
  e->gen_bus = vcopy(c->gen_bus, e->ng);
  e->Pg = vcopy(c->Pg, e->ng); e->Qg = vcopy(c->Qg, e->ng);
  e->Qmax = vcopy(c->Qmax, e->ng); e->Qmin = vcopy(c->Qmin, e->ng);
  e->VG = vcopy(c->VG, e->ng); e->gen_status = vcopy(c->gen_status, e->ng);
  for (int k = 0; k < e->ng; k++)
    if (e->gen_bus[k] < 0 || e->gen_bus[k] >= e->n) {
      e->err = "generator bus out of range";
      return BPF_ERR_INVALID_ARG;
    }
  if (!analyse(e->n, e->m, e->br_from.data(), e->br_to.data(),
               e->bus_type.data(), e->S, &e->err))
    return BPF_ERR_INVALID_ARG;
  // device
  const char *dbg = getenv("BPF_DEBUG_SYNC");
  e->debug_sync = (dbg && dbg[0] == '1') ? 1 : 0;
  HIPCHK(hipGetDevice(&e->dev));
  {
    bpf_status s0 = scratch_setup(e);
    if (s0) return s0;
  }
  HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIPCHK(hipDeviceGetAttribute(&e->max_lds,
                               hipDeviceAttributeMaxSharedMemoryPerBlock, e->dev));
  HIPCHK(hipHostMalloc((void **)&e->h_count, sizeof(int), hipHostMallocDefault));
  const Symbolic &S = e->S;
  DevTopo &g = e->g;
  g.n = e->n; g.m = e->m; g.nb = S.nb;
  g.nnzY = (int)S.ycol.size();
  g.nnzLU = (int)S.lcol.size();
  bpf_status s;
#define UP(dst, vec) \
  if ((s = upload(e, e->owned, vec, &dst))) return s;
  UP(g.yrow, S.yrow); UP(g.ycol, S.ycol); UP(g.yrowof, S.yrowof);
  {
    std::vector<int> yblk(S.ycol.size());
    for (size_t p = 0; p < yblk.size(); p++) yblk[p] = S.blk_of_bus[S.ycol[p]];
    UP(g.yblk, nonempty(yblk, 1));
  }
  UP(g.ydiag_bus, S.ydiag_bus); UP(g.yf_ptr, S.yf_ptr);
  UP(g.yf_item, S.yf_item); UP(g.yt_ptr, S.yt_ptr); UP(g.yt_item, S.yt_item);
  UP(g.blk_of_bus, S.blk_of_bus); UP(g.bus_of_blk, S.bus_of_blk);
  UP(g.y2lu, S.y2lu); UP(g.jac_lu, S.jac_lu); UP(g.jac_y, S.jac_y);
  {
    std::vector<int> j4(4 * S.jac_lu.size());
    for (size_t r = 0; r < S.jac_lu.size(); r++) {
      const int p = S.jac_y[r];
      j4[4 * r] = S.jac_lu[r];
      j4[4 * r + 1] = p;
      j4[4 * r + 2] = S.yrowof[p];
      j4[4 * r + 3] = S.ycol[p];
    }
    const int *jp;
    UP(jp, nonempty(j4, 4));
    g.jac4 = (const int4 *)jp;
  }
  UP(g.jac_l0, nonempty(S.jac_l0, 1));
  g.nJb = (int)S.jac_lu.size(); UP(g.lrow, S.lrow); UP(g.lcol, S.lcol);
  UP(g.ldiag, S.ldiag); UP(g.mirror, S.mirror);
  UP(g.br_from, e->br_from); UP(g.br_to, e->br_to);
  UP(g.bbr_ptr, S.bbr_ptr); UP(g.bbr, S.bbr);
  UP(g.R, e->R); UP(g.X, e->X); UP(g.B, e->B); UP(g.shift, e->shift);
  UP(g.Gs, e->Gs); UP(g.Qd, e->Qd); UP(g.Va0, e->Va);
  UP(g.base_type, e->bus_type);
  FactorPlan &F = e->fplan;
  F.nchunk = S.nchunk;
  UP(F.ch_task, S.ch_task); UP(F.ch_pair, S.ch_pair); UP(F.ch_max, S.ch_max);
  {
    const int *ta, *tb, *pr;
    UP(ta, nonempty(S.tk_a, 4)); UP(tb, nonempty(S.tk_b, 4));
    UP(pr, nonempty(S.pairs, 2));
    F.tk_a = (const int4 *)ta;
    F.tk_b = (const int4 *)tb;
    F.pairs = (const int2 *)pr;
  }
  F.tn = S.tail_n;
  F.tnent = (int)S.tl_pos.size();
  F.tmax = S.tl_max;
  UP(F.tdense, S.tl_dense); UP(F.tpos, S.tl_pos); UP(F.tpp, S.tl_pp);
  UP(F.tpairL, S.tl_pairL); UP(F.tpairU, S.tl_pairU);
  F.tstage_n = (int)S.tl_stage.size();
  {
    const char *tv = getenv("BPF_TAIL_VALU");
    F.tail_mfma = !(tv && tv[0] == '1');
  }
  UP(F.tstage, nonempty(S.tl_stage, 1)); UP(F.tpairS, nonempty(S.tl_pairS, 1));
  SolvePlan &SP = e->splan;
  SP.nfw = S.nfw;
  SP.nbw = S.nbw;
  UP(SP.fw_ptr, S.fw_ptr); UP(SP.fw_rows, S.fw_rows); UP(SP.fw_max, S.fw_max);
  UP(SP.bw_ptr, S.bw_ptr); UP(SP.bw_rows, S.bw_rows); UP(SP.bw_max, S.bw_max);
  SP.tn = S.tail_n;
  SP.tsmax = S.ts_max;
  UP(SP.tpiv, S.tail_piv); UP(SP.ts_ptr, S.ts_ptr); UP(SP.ts_pos, S.ts_pos);
  UP(SP.ts_col, S.ts_col); UP(SP.ts_mir, S.ts_mir);
  SP.sw_nfw = S.sw_nfw;
  SP.sw_nbw = S.sw_nbw;
  UP(SP.sw_cmax, S.sw_cmax);
  {
    const int *ch, *rw, *en;
    UP(ch, S.sw_chunk); UP(rw, nonempty(S.sw_row, 4));
    UP(en, nonempty(S.sw_ent, 2));
    SP.sw_chunk = (const int4 *)ch;
    SP.sw_row = (const int4 *)rw;
    SP.sw_ent = (const int2 *)en;
  }
  SP.sl_nfw = S.sl_nfw;
  SP.sl_nbw = S.sl_nbw;
  {
    const int *ch, *sl;
    UP(ch, nonempty(S.sl_chunk, 4)); UP(sl, nonempty(S.sl_slot, 4));
    SP.sl_chunk = (const int4 *)ch;
    SP.sl_slot = (const int4 *)sl;
  }
  const char *swr = getenv("BPF_SWEEP_ROWS"), *swe = getenv("BPF_SWEEP_ENTRY");
  e->sweep_form = (swr && swr[0] == '1') ? 2 : (swe && swe[0] == '1') ? 1 : 0;
#undef UP
  return BPF_OK;
}

extern "C" bpf_status bpf_set_launch(bpf_engine *e, int32_t lu, int32_t sv,
                                     int32_t mis) {
  if (!e) return BPF_ERR_INVALID_ARG;
  auto ok = [](int v) { return v == 0 || (v >= 64 && v <= 1024 && v % 64 == 0); };
  if (!ok(lu) || !ok(sv) || !ok(mis)) {
    e->err = "launch sizes must be multiples of 64 in [64, 1024]";
    return BPF_ERR_INVALID_ARG;
  }
  if (lu) e->lu_threads = lu;
  if (sv) e->solve_threads = sv;
  if (mis) e->mis_threads = mis;
  if (e->graph) { hipGraphExecDestroy(e->graph); e->graph = nullptr; }
  return BPF_OK;
}

extern "C" bpf_status bpf_set_instances(bpf_engine *e, int32_t T,
                                        const uint8_t *br_on,
                                        const uint8_t *gen_on, int32_t nctl,
                                        const bpf_ctl *ctl,
                                        const double *ctl_val) {
  if (!e) return BPF_ERR_INVALID_ARG;
  if (T <= 0 || T > 65535 || nctl < 0 || (nctl > 0 && (!ctl || !ctl_val))) {
    e->err = "invalid instance count or controls";
    return BPF_ERR_INVALID_ARG;
  }
  const int n = e->n, m = e->m, ng = e->ng;
  for (int c = 0; c < nctl; c++) {
    int k = ctl[c].kind, id = ctl[c].id;
    bool okk = (k == BPF_CTL_AVR || k == BPF_CTL_SHC) ? (id >= 0 && id < n)
               : (k == BPF_CTL_OLTC) ? (id >= 0 && id < m) : false;
    if (!okk) { e->err = "control " + std::to_string(c) + " out of range"; return BPF_ERR_INVALID_ARG; }
  }
  if (e->graph) { hipGraphExecDestroy(e->graph); e->graph = nullptr; }
  free_pool(e->inst_owned);
  // the factor storage lived in that pool too: forget it, so that ensure_lu
  // allocates it for the new T instead of freeing the stale address again
  // (which a new buffer of this call may have been given)
  e->d_LU = e->d_xglob = e->d_TL = nullptr;
  e->T = T;
  e->solved = false;
  e->lu_fp64 = -1;
  const size_t Tn = (size_t)T * n, Tm = (size_t)T * m;
  std::vector<uint8_t> on(Tm);
  std::vector<double> tapT(Tm), BsT(Tn), vset(Tn), P(Tn), Q(Tn), Qmx(Tn), Qmn(Tn);
  std::vector<int8_t> ty(Tn);
  std::vector<int> ngon(n);
  std::vector<double> vg(n);
  for (int t = 0; t < T; t++) {
    const size_t on_ = (size_t)t * n, om = (size_t)t * m;
    std::fill(ngon.begin(), ngon.end(), 0);
    for (int i = 0; i < n; i++) {
      ty[on_ + i] = (int8_t)e->bus_type[i];
      P[on_ + i] = -e->Pd[i];
      Q[on_ + i] = -e->Qd[i];
      Qmx[on_ + i] = 0.0;
      Qmn[on_ + i] = 0.0;
      BsT[on_ + i] = e->Bs[i];
      vg[i] = NAN;
    }
    for (int k = 0; k < ng; k++) {
      int b = e->gen_bus[k];
      bool g_on = e->gen_status[k] && (!gen_on || gen_on[(size_t)t * ng + k]);
      if (!g_on) continue;
      ngon[b]++;
      P[on_ + b] += e->Pg[k];
      Q[on_ + b] += e->Qg[k];
      Qmx[on_ + b] += e->Qmax[k];
      Qmn[on_ + b] += e->Qmin[k];
      vg[b] = e->VG[k];   // the last in-service generator row sets VG
    }
    for (int i = 0; i < n; i++) {
      if (e->bus_type[i] == BPF_PV && ngon[i] == 0) ty[on_ + i] = BPF_PQ;
      double v = e->Vm[i];
      if (e->bus_type[i] == BPF_PV && !std::isnan(vg[i])) v = vg[i];
      vset[on_ + i] = v;
    }
    for (int b = 0; b < m; b++) {
      on[om + b] = e->br_status[b] && (!br_on || br_on[om + b]);
      tapT[om + b] = e->tap[b];
    }
    for (int c = 0; c < nctl; c++) {
      double v = ctl_val[(size_t)t * nctl + c];
      int id = ctl[c].id;
      if (std::isnan(v)) continue;
      if (ctl[c].kind == BPF_CTL_AVR) vset[on_ + id] = v;
      else if (ctl[c].kind == BPF_CTL_SHC) BsT[on_ + id] = v;
      else if (e->tap[id] != 0.0) tapT[om + id] = v;
    }
  }
  DevInst &d = e->d;
  d.T = T;
  bpf_status s;
  const uint8_t *pu8;
  const double *pd;
  const int8_t *pi8;
#define UPI(dst, vec, tmp) \
  if ((s = upload(e, e->inst_owned, vec, &tmp))) return s; dst = tmp;
  UPI(d.br_on, on, pu8);
  UPI(d.tap, tapT, pd);
  UPI(d.Bs, BsT, pd);
  UPI(d.vset, vset, pd);
  UPI(d.Qmax, Qmx, pd);
  UPI(d.Qmin, Qmn, pd);
  UPI(e->d_type0, ty, pi8);   // read-only after upload
  UPI(e->d_Qinj0, Q, pd);
  if ((s = upload(e, e->inst_owned, P, &pd))) return s;
  d.Pinj = const_cast<double *>(pd);
#undef UPI
  void *p;
  const int nb = e->S.nb, nnzY = e->g.nnzY;
#define AL(dst, type, count) \
  if ((s = dalloc(e, e->inst_owned, sizeof(type) * (count), &p))) return s; dst = (type *)p;
  AL(d.type, int8_t, Tn);
  AL(d.Qinj, double, Tn);
  AL(d.Y, cd, (size_t)T * nnzY);
  AL(d.V, cd, Tn);
  AL(d.Ib, cd, Tn);
  AL(d.F, double, (size_t)T * 2 * nb);
  AL(d.dx, double, (size_t)T * 2 * nb);
  AL(d.rhs, double, (size_t)T * 2 * nb);
  AL(d.st, InstState, T);
  AL(d.active_count, int, 1);
  AL(e->d_Sf, cd, Tm);
  AL(e->d_St, cd, Tm);
  AL(e->d_loss, double, T);
  AL(d.part, double, (size_t)T * div_up(e->n, 64));
  AL(d.partv, int, (size_t)T * div_up(e->n, 64));
  AL(d.jsing, int, T);
#undef AL
  HIPCHK(hipMemset(d.jsing, 0, sizeof(int) * (size_t)T));
  return BPF_OK;
}

// (re)allocate the factor storage for the requested precision
static bpf_status ensure_lu(bpf_engine *e, int fp64) {
  if (e->lu_fp64 == fp64 && e->d_LU) return BPF_OK;
  if (e->d_LU) {
    for (auto it = e->inst_owned.begin(); it != e->inst_owned.end(); ++it)
      if (it->p == e->d_LU) { hipFree(it->p); e->inst_owned.erase(it); break; }
    for (auto it = e->inst_owned.begin(); it != e->inst_owned.end(); ++it)
      if (it->p == e->d_xglob) { hipFree(it->p); e->inst_owned.erase(it); break; }
    for (auto it = e->inst_owned.begin(); it != e->inst_owned.end(); ++it)
      if (it->p == e->d_TL) { hipFree(it->p); e->inst_owned.erase(it); break; }
    e->d_LU = e->d_xglob = e->d_TL = nullptr;
  }
  size_t el = fp64 ? sizeof(B4<double>) : sizeof(B4<float>);
  bpf_status s = dalloc(e, e->inst_owned, el * (size_t)e->T * e->g.nnzLU, &e->d_LU);
  if (s) return s;
  s = dalloc(e, e->inst_owned, el * (size_t)e->T * e->S.tail_n * e->S.tail_n,
             &e->d_TL);
  if (s) return s;
  size_t xs = (fp64 ? 8 : 4) * (size_t)2 * e->S.nb;
  if (xs > (size_t)SOLVE_LDS_BYTES) {
    s = dalloc(e, e->inst_owned, xs * e->T, &e->d_xglob);
    if (s) return s;
  }
  e->lu_fp64 = fp64;
  if (e->graph) { hipGraphExecDestroy(e->graph); e->graph = nullptr; }
  return BPF_OK;
}

// ---- launches ---------------------------------------------------------------
static void tbegin(bpf_engine *e, const char *name) {
  e->cur_kernel = name;
  if (!e->timing || e->capturing) return;
  TimedEvent t;
  t.name = name;
  hipEventCreate(&t.a);
  hipEventCreate(&t.b);
  hipEventRecord(t.a, e->stream);
  e->events.push_back(t);
}
// record the first launch error with the kernel's name; in debug mode also
// wait for the kernel so that execution errors are attributed too
static void tcheck(bpf_engine *e, const char *name) {
  hipError_t r = hipGetLastError();
  if (r == hipSuccess && e->debug_sync && !e->capturing) {
    r = hipStreamSynchronize(e->stream);
    if (r == hipSuccess) r = hipGetLastError();
  }
  if (r != hipSuccess && e->launch_err.empty())
    e->launch_err = std::string(name) + ": " + hipGetErrorString(r);
}
static void tend(bpf_engine *e) {
  tcheck(e, e->cur_kernel);
  if (!e->timing || e->capturing) return;
  hipEventRecord(e->events.back().b, e->stream);
}
static void tcollect(bpf_engine *e) {
  for (auto &ev : e->events) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, ev.a, ev.b);
    auto &slot = e->ktime[ev.name];
    slot.first++;
    slot.second += ms;
    hipEventDestroy(ev.a);
    hipEventDestroy(ev.b);
  }
  e->events.clear();
}


// mismatch rows over many workgroups, then one verdict workgroup per instance
static void mismatch_pair(bpf_engine *e, int mode, int maxit, int qlim,
                          int fd_norm) {
  const unsigned nparts = div_up(e->n, e->mis_threads);
  hipLaunchKernelGGL(k_mismatch_mb, xgrid(e->T, e->n, e->mis_threads),
                     dim3(e->mis_threads), 0, e->stream, e->g, e->d, mode,
                     fd_norm, e->opt.qlim_tol, qlim);
  // 1024 lanes: the reactive-limit switching walks every bus of an instance
  hipLaunchKernelGGL(k_verdict, dim3(e->T), dim3(1024), 0, e->stream, e->g, e->d,
                     (int)nparts, mode, maxit, e->opt.tol, qlim,
                     e->opt.max_qlim_rounds, e->opt.qlim_tol, fd_norm);
}

static void launch_mismatch(bpf_engine *e, int mode) {
  tbegin(e, "mismatch");
  const bool fd = e->opt.alg != BPF_NR;
  mismatch_pair(e, mode, fd ? e->opt.maxit_fd : e->opt.maxit_nr,
                fd ? 0 : e->opt.enforce_qlim, fd ? 1 : 0);
  tend(e);
}

// one sweep pass of k_lu_solve; the slot form (ep 2) has one kernel per
// workgroup size (SLOT_CHUNK / threads slots per lane)
template <typename T>
static void launch_solve(bpf_engine *e, int T_, const void *LU, int rhs_mode,
                         int accumulate, int use_lds, int ep) {
  const int th = e->solve_threads;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(T_), dim3(th), 0, e->stream, e->g, e->d,
                       (const B4<T> *)LU, (const B4<T> *)e->d_TL, rhs_mode,
                       accumulate, e->splan, use_lds, ep, (T *)e->d_xglob);
  };
  if (ep == 2 && th == 1024) go(k_lu_solve<T, SLOT_CHUNK / 1024>);
  else if (ep == 2 && th == 512) go(k_lu_solve<T, SLOT_CHUNK / 512>);
  else if (ep == 2 && th == 256) go(k_lu_solve<T, SLOT_CHUNK / 256>);
  else go(k_lu_solve<T, 0>);
}

template <typename T>
static void launch_newton_step(bpf_engine *e) {
  B4<T> *LU = (B4<T> *)e->d_LU;
  // the fill also finishes level 0 when the analysis allows it
  const int c0 = e->S.l0_chunks;
  tbegin(e, "jacobian");
  hipLaunchKernelGGL(k_jacobian<T>, xgrid(e->T, ((size_t)e->g.nJb + 1) / 2, 256),
                     dim3(256), 0, e->stream, e->g, e->d, LU, c0 > 0 ? 1 : 0);
  tend(e);
  tbegin(e, "lu_factor");
  hipLaunchKernelGGL(k_lu_factor<T>, dim3(e->T), dim3(e->lu_threads), 0,
                     e->stream, e->g, e->d.st, 1, e->fplan, LU,
                     (B4<T> *)e->d_TL, c0, e->d.jsing);
  tend(e);
  size_t xs = sizeof(T) * 2 * (size_t)e->S.nb;
  int use_lds = xs <= (size_t)SOLVE_LDS_BYTES;
  const int refine = (sizeof(T) == 4) ? e->opt.refine_steps : 0;
  const int ep = sweep_mode(e->S, sizeof(T), use_lds, e->sweep_form, e->solve_threads);
  // one solve on the mismatch, then per refinement step the FP64 residual
  // (many workgroups per instance) and a solve on it, accumulated into dx
  for (int s = 0; s <= refine; s++) {
    if (s > 0) {
      tbegin(e, "residual");
      hipLaunchKernelGGL(k_residual, xgrid(e->T, e->n, 256), dim3(256), 0,
                         e->stream, e->g, e->d);
      tend(e);
    }
    tbegin(e, "lu_solve");
    launch_solve<T>(e, e->T, LU, s == 0 ? 0 : 3, s > 0 ? 1 : 0, use_lds, ep);
    tend(e);
  }
}

static void launch_iteration(bpf_engine *e) {
  if (e->lu_fp64) launch_newton_step<double>(e);
  else launch_newton_step<float>(e);
  tbegin(e, "update");
  hipLaunchKernelGGL(k_update, xgrid(e->T, e->n, 256), dim3(256), 0,
                     e->stream, e->g, e->d);
  tend(e);
  launch_mismatch(e, 1);
  // the active-instance count is read outside the captured graph
  if (!e->capturing)
    hipMemcpyAsync(e->h_count, e->d.active_count, sizeof(int),
                   hipMemcpyDeviceToHost, e->stream);
}


// Newton iterations per captured graph (each replay is followed by one read
// of the active-instance count)
constexpr int NEWTON_PER_GRAPH = 4;

static bpf_status run_newton(bpf_engine *e, bool *graph_failed) {
  bpf_status s = ensure_lu(e, e->opt.precision == BPF_FP64 ? 1 : 0);
  if (s) return s;
  const int key = e->lu_fp64 * 10000000 +
                  e->lu_threads * 10000 + e->mis_threads * 10 / 64 +
                  e->solve_threads / 64;
  // per-kernel timing uses events around direct launches: elapsed times
  // between event nodes of a replayed graph are not reliable on this runtime
  bool use_graph = e->opt.use_graph != 0 && !e->graph_broken && !e->timing;
  if (use_graph && (!e->graph || e->graph_key != key)) {
    if (e->graph) { hipGraphExecDestroy(e->graph); e->graph = nullptr; }
    hipGraph_t gr;
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    e->capturing = true;
    for (int j = 0; j < NEWTON_PER_GRAPH; j++) launch_iteration(e);
    e->capturing = false;
    HIPCHK(hipStreamEndCapture(e->stream, &gr));
    hipError_t r = hipGraphInstantiate(&e->graph, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    if (r != hipSuccess) {
      e->graph = nullptr;
      e->graph_broken = true;
      e->graph_err = std::string("hipGraphInstantiate: ") + hipGetErrorString(r);
      (void)hipGetLastError();
      use_graph = false;
    } else {
      e->graph_key = key;
    }
  }
  e->graph_used = use_graph ? 1 : 0;
  const int cap = e->opt.maxit_nr * (e->opt.max_qlim_rounds + 1) + 2;
  // an error left by an earlier call must not be read as this loop's
  hipError_t stale = hipGetLastError();
  if (stale != hipSuccess && e->graph_err.empty())
    e->graph_err = std::string("error pending before the Newton loop: ") +
                   hipGetErrorName(stale);
  // The graph holds NEWTON_PER_GRAPH iterations and the active count is read
  // back once per replay: every kernel of an iteration returns at once for a
  // finished instance, so the iterations replayed past the last one change
  // nothing (the count stays 0), and the host waits once per replay.
  const int ahead = use_graph ? NEWTON_PER_GRAPH : 1;
  for (int it = 0; it < cap; it += ahead) {
    hipError_t r = hipSuccess;
    const char *where = "";
    if (use_graph) {
      where = "hipGraphLaunch";
      r = hipGraphLaunch(e->graph, e->stream);
      if (r == hipSuccess) {
        where = "hipMemcpyAsync";
        r = hipMemcpyAsync(e->h_count, e->d.active_count, sizeof(int),
                           hipMemcpyDeviceToHost, e->stream);
      }
    } else {
      launch_iteration(e);
    }
    if (r == hipSuccess) { where = "hipStreamSynchronize"; r = hipStreamSynchronize(e->stream); }
    if (r == hipSuccess) {
      where = "hipGetLastError";
      hipError_t le = hipGetLastError();
      if (use_graph) {
        // a replay's launch errors come back from hipGraphLaunch and its
        // execution errors from the synchronization; the runtime has been
        // seen to leave an unrelated error code behind after a replay that
        // ran fine, which is noted, not taken as a failure
        if (le != hipSuccess && e->graph_note.empty())
          e->graph_note = std::string("after graph replay the runtime reported ") +
                          hipGetErrorName(le) + " (ignored)";
      } else {
        r = le;
      }
    }
    if (r != hipSuccess) {
      if (use_graph) {
        // the runtime refused the replay: remember why and let bpf_solve
        // start the solve again with direct launches
        e->graph_broken = true;
        e->graph_err = std::string("graph replay (") + where + ", iteration " +
                       std::to_string(it) + "): " + hipGetErrorName(r) + " " +
                       hipGetErrorString(r);
        (void)hipGetLastError();
        *graph_failed = true;
        return BPF_OK;
      }
      e->err = std::string("Newton iteration: ") + hipGetErrorString(r) +
               (e->launch_err.empty() ? "" : " [" + e->launch_err + "]");
      return BPF_ERR_HIP;
    }
    if (*e->h_count == 0) break;
  }
  return BPF_OK;
}

// fast decoupled (mkl_fdpf, fdpf.h:450-596): B' and B'' of every instance
// assembled and factored once in FP64, then P and Q half-steps; each half-step
// is a masked solve with the same factors, an update and the |V|-normalised
// mismatch with the reference's convergence checks
static bpf_status run_fd(bpf_engine *e) {
  bpf_status s = ensure_lu(e, 1);
  if (s) return s;
  B4<double> *LU = (B4<double> *)e->d_LU;
  const unsigned T = (unsigned)e->T;
  tbegin(e, "bmat");
  hipLaunchKernelGGL(k_bmat, dim3(div_up(e->g.nnzY, 256), T), dim3(256), 0,
                     e->stream, e->g, e->d, e->opt.alg, LU);
  tend(e);
  tbegin(e, "lu_factor");
  hipLaunchKernelGGL(k_lu_factor<double>, dim3(T), dim3(e->lu_threads), 0,
                     e->stream, e->g, e->d.st, 0, e->fplan, LU,
                     (B4<double> *)e->d_TL, 0, e->d.jsing);
  tend(e);
  size_t xs = sizeof(double) * 2 * (size_t)e->S.nb;
  int use_lds = xs <= (size_t)SOLVE_LDS_BYTES;
  for (int it = 0; it < e->opt.maxit_fd + 1; it++) {
    for (int half = 1; half <= 2; half++) {
      tbegin(e, "fd_solve");
      launch_solve<double>(e, T, LU, half, 0, use_lds,
                           sweep_mode(e->S, sizeof(double), use_lds,
                                      e->sweep_form, e->solve_threads));
      tend(e);
      tbegin(e, "update");
      hipLaunchKernelGGL(k_fd_update, xgrid(T, e->n, 256), dim3(256),
                         0, e->stream, e->g, e->d, half);
      tend(e);
      tbegin(e, "mismatch");
      hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, e->stream, e->d);
      mismatch_pair(e, half == 1 ? 3 : 4, e->opt.maxit_fd, 0, 1);
      tend(e);
      HIPCHK(hipMemcpyAsync(e->h_count, e->d.active_count, sizeof(int),
                            hipMemcpyDeviceToHost, e->stream));
      HIPCHK(hipStreamSynchronize(e->stream));
      HIPCHK(hipGetLastError());
      if (*e->h_count == 0) return BPF_OK;
    }
  }
  return BPF_OK;
}

static bpf_status solve_once(bpf_engine *e, bool *graph_failed) {
  const bool fd = e->opt.alg != BPF_NR;
  const unsigned T = (unsigned)e->T;
  tbegin(e, "ybus");
  hipLaunchKernelGGL(k_ybus, xgrid(T, e->g.nnzY, 256), dim3(256), 0,
                     e->stream, e->g, e->d);
  tend(e);
  tbegin(e, "init");
  hipLaunchKernelGGL(k_init, xgrid(T, e->n, 256), dim3(256), 0,
                     e->stream, e->g, e->d, (const int8_t *)e->d_type0,
                     (const double *)e->d_Qinj0);
  tend(e);
  launch_mismatch(e, 0);
  HIPCHK(hipMemcpyAsync(e->h_count, e->d.active_count, sizeof(int),
                        hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipGetLastError());
  if (*e->h_count > 0) {
    bpf_status s = fd ? run_fd(e) : run_newton(e, graph_failed);
    if (s || *graph_failed) return s;
  }
  tbegin(e, "loss");
  hipLaunchKernelGGL(k_loss, dim3(T), dim3(256), 0, e->stream, e->g, e->d,
                     e->d_Sf, e->d_St, e->d_loss);
  tend(e);
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipGetLastError());
  return BPF_OK;
}

extern "C" bpf_status bpf_solve(bpf_engine *e) {
  if (!e) return BPF_ERR_INVALID_ARG;
  e->launch_err.clear();
  if (e->T <= 0) { e->err = "bpf_set_instances not called"; return BPF_ERR_STATE; }
  if (e->opt.alg != BPF_NR && e->opt.enforce_qlim) {
    e->err = "reactive-limit switching is supported with Newton-Raphson only";
    return BPF_ERR_UNSUPPORTED;
  }
  e->graph_used = 0;
  bool graph_failed = false;
  bpf_status s = solve_once(e, &graph_failed);
  if (!s && graph_failed) {
    for (auto &ev : e->events) { hipEventDestroy(ev.a); hipEventDestroy(ev.b); }
    e->events.clear();
    s = solve_once(e, &graph_failed);
  }
  tcollect(e);
  if (s) return s;
  e->solved = true;
  return BPF_OK;
}

extern "C" bpf_status bpf_get_results(bpf_engine *e, double *V, int32_t *its,
                                      int32_t *verdict, double *norm,
                                      double *loss, int32_t *rounds,
                                      int32_t *bus_type) {
  if (!e) return BPF_ERR_INVALID_ARG;
  if (!e->solved) { e->err = "no solution yet"; return BPF_ERR_STATE; }
  const size_t T = e->T, n = e->n;
  if (V) HIPCHK(hipMemcpy(V, e->d.V, sizeof(cd) * T * n, hipMemcpyDeviceToHost));
  std::vector<InstState> st(T);
  HIPCHK(hipMemcpy(st.data(), e->d.st, sizeof(InstState) * T, hipMemcpyDeviceToHost));
  for (size_t t = 0; t < T; t++) {
    if (its) its[t] = st[t].it_total;
    if (verdict) verdict[t] = st[t].verdict;
    if (norm) norm[t] = st[t].norm;
    if (rounds) rounds[t] = st[t].rounds;
  }
  if (loss) HIPCHK(hipMemcpy(loss, e->d_loss, sizeof(double) * T, hipMemcpyDeviceToHost));
  if (bus_type) {
    std::vector<int8_t> ty(T * n);
    HIPCHK(hipMemcpy(ty.data(), e->d.type, T * n, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < T * n; k++) bus_type[k] = ty[k];
  }
  return BPF_OK;
}

extern "C" bpf_status bpf_get_branch_flows(bpf_engine *e, double *Sf, double *St) {
  if (!e) return BPF_ERR_INVALID_ARG;
  if (!e->solved) { e->err = "no solution yet"; return BPF_ERR_STATE; }
  size_t bytes = sizeof(cd) * (size_t)e->T * e->m;
  if (Sf) HIPCHK(hipMemcpy(Sf, e->d_Sf, bytes, hipMemcpyDeviceToHost));
  if (St) HIPCHK(hipMemcpy(St, e->d_St, bytes, hipMemcpyDeviceToHost));
  return BPF_OK;
}

extern "C" bpf_status bpf_get_branch_loading(bpf_engine *e, const double *rate,
                                             double *loading, int32_t *overloads,
                                             double *max_loading) {
  if (!e || !rate) return BPF_ERR_INVALID_ARG;
  if (!e->solved) { e->err = "no solution yet"; return BPF_ERR_STATE; }
  const int T = e->T, m = e->m;
  for (int b = 0; b < m; b++)
    if (!(rate[b] >= 0.0)) {
      e->err = "branch ratings must be >= 0 (0: no rating)";
      return BPF_ERR_INVALID_ARG;
    }
  std::vector<DBuf> tmp;
  struct Free { std::vector<DBuf> &v; ~Free() { free_pool(v); } } fr{tmp};
  void *p;
  bpf_status s;
  double *d_rate, *d_load = nullptr, *d_max;
  int *d_over;
  if ((s = dalloc(e, tmp, sizeof(double) * (size_t)std::max(m, 1), &p))) return s;
  d_rate = (double *)p;
  if (loading) {
    if ((s = dalloc(e, tmp, sizeof(double) * (size_t)T * std::max(m, 1), &p))) return s;
    d_load = (double *)p;
  }
  if ((s = dalloc(e, tmp, sizeof(int) * (size_t)T, &p))) return s;
  d_over = (int *)p;
  if ((s = dalloc(e, tmp, sizeof(double) * (size_t)T, &p))) return s;
  d_max = (double *)p;
  if (m) HIPCHK(hipMemcpy(d_rate, rate, sizeof(double) * m, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_branch_report, dim3(T), dim3(256), 0, e->stream, e->g, T,
                     e->d_Sf, e->d_St, d_rate, d_load, d_over, d_max);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  if (loading)
    HIPCHK(hipMemcpy(loading, d_load, sizeof(double) * (size_t)T * m,
                     hipMemcpyDeviceToHost));
  std::vector<int> ov(T);
  HIPCHK(hipMemcpy(ov.data(), d_over, sizeof(int) * T, hipMemcpyDeviceToHost));
  if (overloads) for (int t = 0; t < T; t++) overloads[t] = ov[t];
  if (max_loading)
    HIPCHK(hipMemcpy(max_loading, d_max, sizeof(double) * T, hipMemcpyDeviceToHost));
  return BPF_OK;
}

// ---- allocation-free branch report ------------------------------------------
// The same report as bpf_get_branch_loading for callers that cannot allocate
// device memory: the ratings and the per-instance outputs are staged in one
// file-scope device array (each GPU holds its own copy for the life of the
// process), the instances are reported in batches that fit it, and each
// batch is copied to the caller before the next one reuses the space.
// Layout in doubles: rate [round_up(m, 32)], then per batch of tb instances
// loading [tb * m] (when requested), max [tb], overloads [tb] (int32 stored
// in round_up(tb / 2) doubles).
static __device__ __align__(256) double scratch[BPF_NOALLOC_SCRATCH_BYTES / sizeof(double)];

// hipGetSymbolAddress(scratch) of each device, set by bpf_create with the
// device current; the host-side name `scratch` is never a device address
static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_ptr;

static bpf_status scratch_setup(bpf_engine *e) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(e->dev);
  if (it == g_scratch_ptr.end()) {
    double *p = nullptr;
    hipError_t r = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !p) {
      e->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(r);
      return BPF_ERR_HIP;
    }
    it = g_scratch_ptr.emplace(e->dev, p).first;
  }
  e->scratch = it->second;
  return BPF_OK;
}

extern "C" bpf_status bpf_get_branch_loading_noalloc(bpf_engine *e,
                                                     const double *rate,
                                                     double *loading,
                                                     int32_t *overloads,
                                                     double *max_loading) {
  if (!e || !rate) return BPF_ERR_INVALID_ARG;
  if (!e->solved) { e->err = "no solution yet"; return BPF_ERR_STATE; }
  if (!e->scratch) { e->err = "scratch address not set up"; return BPF_ERR_STATE; }
  const int T = e->T, m = e->m;
  for (int b = 0; b < m; b++)
    if (!(rate[b] >= 0.0)) {
      e->err = "branch ratings must be >= 0 (0: no rating)";
      return BPF_ERR_INVALID_ARG;
    }
  const size_t cap = BPF_NOALLOC_SCRATCH_BYTES / sizeof(double);
  const size_t rate_d = ((size_t)m + 31) / 32 * 32;
  // doubles per instance: loading row, max, half a double of overload count
  const size_t per = (loading ? (size_t)m : 0) + 2;
  if (rate_d + per > cap) {
    // the largest branch count that fits, for the message
    size_t mmax = (cap - 2) / (loading ? 2 : 1);
    while (mmax > 0 && (mmax + 31) / 32 * 32 + (loading ? mmax : 0) + 2 > cap) mmax--;
    e->err = "bpf_get_branch_loading_noalloc: " + std::to_string(m) +
             " branches exceed its " + std::to_string(BPF_NOALLOC_SCRATCH_BYTES) +
             "-byte device scratch (at most " + std::to_string(mmax) +
             " branches" + (loading ? " with per-branch loading" : "") + ")";
    return BPF_ERR_INVALID_ARG;
  }
  const int tb = (int)std::min<size_t>((size_t)T, (cap - rate_d) / per);
  double *d_rate = e->scratch;
  double *d_load = loading ? d_rate + rate_d : nullptr;
  double *d_max = d_rate + rate_d + (loading ? (size_t)tb * m : 0);
  int *d_over = (int *)(d_max + tb);
  if (m)
    HIPCHK(hipMemcpyAsync(d_rate, rate, sizeof(double) * m,
                          hipMemcpyHostToDevice, e->stream));
  for (int t0 = 0; t0 < T; t0 += tb) {
    const int nt = std::min(tb, T - t0);
    const size_t off = (size_t)t0 * m;
    hipLaunchKernelGGL(k_branch_report, dim3(nt), dim3(256), 0, e->stream, e->g,
                       nt, e->d_Sf + off, e->d_St + off, (const double *)d_rate,
                       d_load, d_over, d_max);
    HIPCHK(hipGetLastError());
    // stream-ordered copies: the next batch's kernel starts after them
    if (loading)
      HIPCHK(hipMemcpyAsync(loading + off, d_load, sizeof(double) * nt * m,
                            hipMemcpyDeviceToHost, e->stream));
    if (overloads)
      HIPCHK(hipMemcpyAsync(overloads + t0, d_over, sizeof(int32_t) * nt,
                            hipMemcpyDeviceToHost, e->stream));
    if (max_loading)
      HIPCHK(hipMemcpyAsync(max_loading + t0, d_max, sizeof(double) * nt,
                            hipMemcpyDeviceToHost, e->stream));
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  return BPF_OK;
}

static void fill_info(const Symbolic &S, bpf_info *info) {
  info->nbus = S.n;
  info->nbranch = S.m;
  info->nblocks = S.nb;
  info->nnz_ybus = (int)S.ycol.size();
  info->nnz_jac_blocks = S.nnz_jac_blocks;
  info->nnz_lu_blocks = (int)S.lcol.size();
  info->lu_levels = S.nlev;
  info->fw_levels = S.nfw;
  info->bw_levels = S.nbw;
  info->lu_pairs = S.npairs;
  info->crout_tasks = (int)S.tk_a.size() / 4;
  info->chunks = S.nchunk;
  info->tail_level = S.tail_level;
  info->tail_n = S.tail_n;
  info->tail_entries = (int)S.tl_pos.size();
  info->tail_pairs = (int)S.tl_pairL.size();
  info->tail_row_entries = (int)S.ts_pos.size();
  info->sweep_fw_chunks = S.sw_nfw;
  info->sweep_bw_chunks = S.sw_nbw;
  info->sweep_rows = (int)S.sw_row.size() / 4;
  info->sweep_entries = (int)S.sw_ent.size() / 2;
  info->fill_level0_chunks = S.l0_chunks;
  info->tail_stage = (int)S.tl_stage.size();
  info->slot_fw_chunks = S.sl_ok ? S.sl_nfw : 0;
  info->slot_bw_chunks = S.sl_ok ? S.sl_nbw : 0;
  info->slots = S.sl_ok ? (int)S.sl_slot.size() / 4 : 0;
}

extern "C" bpf_status bpf_analyse_sweeps(const bpf_case *c, bpf_info *info,
                                         int32_t *sw_chunk, int32_t *sw_row,
                                         int32_t *sw_ent) {
  if (!c || !info || c->nbus <= 0 || c->nbranch < 0 || !c->bus_type ||
      (c->nbranch && (!c->br_from || !c->br_to)))
    return BPF_ERR_INVALID_ARG;
  Symbolic S;
  std::string err;
  if (!analyse(c->nbus, c->nbranch, c->br_from, c->br_to, c->bus_type, S, &err))
    return BPF_ERR_INVALID_ARG;
  memset(info, 0, sizeof(*info));
  fill_info(S, info);
  auto put = [](int32_t *dst, const std::vector<int> &v) {
    if (dst && !v.empty()) memcpy(dst, v.data(), sizeof(int32_t) * v.size());
  };
  put(sw_chunk, S.sw_chunk); put(sw_row, S.sw_row); put(sw_ent, S.sw_ent);
  return BPF_OK;
}

extern "C" bpf_status bpf_analyse_slots(const bpf_case *c, bpf_info *info,
                                        int32_t *sl_chunk, int32_t *sl_slot) {
  if (!c || !info || c->nbus <= 0 || c->nbranch < 0 || !c->bus_type ||
      (c->nbranch && (!c->br_from || !c->br_to)))
    return BPF_ERR_INVALID_ARG;
  Symbolic S;
  std::string err;
  if (!analyse(c->nbus, c->nbranch, c->br_from, c->br_to, c->bus_type, S, &err))
    return BPF_ERR_INVALID_ARG;
  memset(info, 0, sizeof(*info));
  fill_info(S, info);
  if (S.sl_ok) {
    if (sl_chunk && !S.sl_chunk.empty())
      memcpy(sl_chunk, S.sl_chunk.data(), sizeof(int32_t) * S.sl_chunk.size());
    if (sl_slot && !S.sl_slot.empty())
      memcpy(sl_slot, S.sl_slot.data(), sizeof(int32_t) * S.sl_slot.size());
  }
  return BPF_OK;
}

extern "C" bpf_status bpf_analyse_fill(const bpf_case *c, bpf_info *info,
                                       int32_t *jac_lu, int32_t *jac_l0) {
  if (!c || !info || c->nbus <= 0 || c->nbranch < 0 || !c->bus_type ||
      (c->nbranch && (!c->br_from || !c->br_to)))
    return BPF_ERR_INVALID_ARG;
  Symbolic S;
  std::string err;
  if (!analyse(c->nbus, c->nbranch, c->br_from, c->br_to, c->bus_type, S, &err))
    return BPF_ERR_INVALID_ARG;
  memset(info, 0, sizeof(*info));
  fill_info(S, info);
  auto put = [](int32_t *dst, const std::vector<int> &v) {
    if (dst && !v.empty()) memcpy(dst, v.data(), sizeof(int32_t) * v.size());
  };
  put(jac_lu, S.jac_lu); put(jac_l0, S.jac_l0);
  return BPF_OK;
}

extern "C" bpf_status bpf_analyse_tail(const bpf_case *c, bpf_info *info,
                                       int32_t *tl_stage, int32_t *tl_pairS) {
  if (!c || !info || c->nbus <= 0 || c->nbranch < 0 || !c->bus_type ||
      (c->nbranch && (!c->br_from || !c->br_to)))
    return BPF_ERR_INVALID_ARG;
  Symbolic S;
  std::string err;
  if (!analyse(c->nbus, c->nbranch, c->br_from, c->br_to, c->bus_type, S, &err))
    return BPF_ERR_INVALID_ARG;
  memset(info, 0, sizeof(*info));
  fill_info(S, info);
  auto put = [](int32_t *dst, const std::vector<int> &v) {
    if (dst && !v.empty()) memcpy(dst, v.data(), sizeof(int32_t) * v.size());
  };
  put(tl_stage, S.tl_stage); put(tl_pairS, S.tl_pairS);
  return BPF_OK;
}

extern "C" bpf_status bpf_analyse(const bpf_case *c, bpf_info *info,
                                  int32_t *bus_of_blk, int32_t *lu_rowptr,
                                  int32_t *lu_colind, int32_t *ch_task,
                                  int32_t *ch_pair, int32_t *tk_a, int32_t *tk_b,
                                  int32_t *pairs, int32_t *fw_ptr,
                                  int32_t *fw_rows, int32_t *bw_ptr,
                                  int32_t *bw_rows, int32_t *tl_dense,
                                  int32_t *tl_pos, int32_t *tl_ptr,
                                  int32_t *tl_pairL, int32_t *tl_pairU,
                                  int32_t *tail_piv, int32_t *ts_ptr,
                                  int32_t *ts_pos, int32_t *ts_col) {
  if (!c || !info || c->nbus <= 0 || c->nbranch < 0 || !c->bus_type ||
      (c->nbranch && (!c->br_from || !c->br_to)))
    return BPF_ERR_INVALID_ARG;
  Symbolic S;
  std::string err;
  if (!analyse(c->nbus, c->nbranch, c->br_from, c->br_to, c->bus_type, S, &err))
    return BPF_ERR_INVALID_ARG;
  memset(info, 0, sizeof(*info));
  fill_info(S, info);
  auto put = [](int32_t *dst, const std::vector<int> &v) {
    if (dst && !v.empty()) memcpy(dst, v.data(), sizeof(int32_t) * v.size());
  };
  put(bus_of_blk, S.bus_of_blk); put(lu_rowptr, S.lrow); put(lu_colind, S.lcol);
  put(ch_task, S.ch_task); put(ch_pair, S.ch_pair); put(tk_a, S.tk_a);
  put(tk_b, S.tk_b); put(pairs, S.pairs);
  put(fw_ptr, S.fw_ptr); put(fw_rows, S.fw_rows);
  put(bw_ptr, S.bw_ptr); put(bw_rows, S.bw_rows);
  put(tl_dense, S.tl_dense); put(tl_pos, S.tl_pos); put(tl_ptr, S.tl_pp);
  put(tl_pairL, S.tl_pairL); put(tl_pairU, S.tl_pairU);
  put(tail_piv, S.tail_piv); put(ts_ptr, S.ts_ptr); put(ts_pos, S.ts_pos);
  put(ts_col, S.ts_col);
  return BPF_OK;
}

extern "C" bpf_status bpf_get_info(const bpf_engine *e, bpf_info *info) {
  if (!e || !info) return BPF_ERR_INVALID_ARG;
  const Symbolic &S = e->S;
  fill_info(S, info);
  info->instances = e->T;
  info->graph_used = e->graph_used;
  int fp64 = e->opt.precision == BPF_FP64;
  info->lds_solve = ((fp64 ? 8 : 4) * 2 * (size_t)S.nb) <= (size_t)SOLVE_LDS_BYTES;
  return BPF_OK;
}

extern "C" bpf_status bpf_set_timing(bpf_engine *e, int32_t enable) {
  if (!e) return BPF_ERR_INVALID_ARG;
  e->timing = enable ? 1 : 0;
  e->ktime.clear();
  return BPF_OK;
}

extern "C" bpf_status bpf_get_kernel_time(const bpf_engine *e, const char *name,
                                          int32_t *launches, double *total_ms) {
  if (!e || !name) return BPF_ERR_INVALID_ARG;
  auto it = e->ktime.find(name);
  if (launches) *launches = it == e->ktime.end() ? 0 : it->second.first;
  if (total_ms) *total_ms = it == e->ktime.end() ? 0.0 : it->second.second;
  return BPF_OK;
}

extern "C" bpf_status bpf_get_ybus(bpf_engine *e, int32_t t, int32_t *rowptr,
                                   int32_t *colind, double *val) {
  if (!e || t < 0 || t >= e->T) return BPF_ERR_INVALID_ARG;
  if (!e->solved) { e->err = "no solution yet"; return BPF_ERR_STATE; }
  if (rowptr) memcpy(rowptr, e->S.yrow.data(), sizeof(int) * (e->n + 1));
  if (colind) memcpy(colind, e->S.ycol.data(), sizeof(int) * e->S.ycol.size());
  if (val)
    HIPCHK(hipMemcpy(val, e->d.Y + (size_t)t * e->g.nnzY, sizeof(cd) * e->g.nnzY,
                     hipMemcpyDeviceToHost));
  return BPF_OK;
}

// upload voltages and evaluate the mismatch (mode 2: no verdicts)
static bpf_status stage_eval(bpf_engine *e, const double *V) {
  if (!e->solved) { e->err = "call bpf_solve first (Ybus is assembled there)"; return BPF_ERR_STATE; }
  HIPCHK(hipMemcpy(e->d.V, V, sizeof(cd) * (size_t)e->T * e->n, hipMemcpyHostToDevice));
  mismatch_pair(e, 2, e->opt.maxit_nr, 0, 0);
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipGetLastError());
  return BPF_OK;
}

// block-ordered vector [T][2nb] -> reference layout [P(pv); P(pq); Q(pq)]
static void to_ref_layout(const bpf_engine *e, const int8_t *ty,
                          const double *blk, double *out) {
  const Symbolic &S = e->S;
  std::vector<int> pv, pq;
  for (int i = 0; i < e->n; i++) {
    if (ty[i] == BPF_PV) pv.push_back(i);
    else if (ty[i] == BPF_PQ && S.blk_of_bus[i] >= 0) pq.push_back(i);
  }
  int npv = (int)pv.size(), npq = (int)pq.size();
  for (int k = 0; k < npv; k++) out[k] = blk[2 * S.blk_of_bus[pv[k]]];
  for (int k = 0; k < npq; k++) {
    out[npv + k] = blk[2 * S.blk_of_bus[pq[k]]];
    out[npv + npq + k] = blk[2 * S.blk_of_bus[pq[k]] + 1];
  }
}

extern "C" bpf_status bpf_stage_mismatch(bpf_engine *e, const double *V,
                                         double *F, double *norm) {
  if (!e || !V) return BPF_ERR_INVALID_ARG;
  bpf_status s = stage_eval(e, V);
  if (s) return s;
  const size_t T = e->T, n = e->n, nb2 = 2 * (size_t)e->S.nb;
  std::vector<double> Fb(T * nb2);
  std::vector<int8_t> ty(T * n);
  std::vector<InstState> st(T);
  HIPCHK(hipMemcpy(Fb.data(), e->d.F, sizeof(double) * T * nb2, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ty.data(), e->d.type, T * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(st.data(), e->d.st, sizeof(InstState) * T, hipMemcpyDeviceToHost));
  for (size_t t = 0; t < T; t++) {
    if (F) to_ref_layout(e, ty.data() + t * n, Fb.data() + t * nb2, F + t * 2 * n);
    if (norm) norm[t] = st[t].norm;
  }
  return BPF_OK;
}

extern "C" bpf_status bpf_stage_jacobian(bpf_engine *e, const double *V,
                                         int32_t t, int32_t cap, int32_t *row,
                                         int32_t *col, double *val,
                                         int32_t *nnz) {
  if (!e || !V || t < 0 || t >= e->T || !nnz) return BPF_ERR_INVALID_ARG;
  bpf_status s = stage_eval(e, V);
  if (s) return s;
  if ((s = ensure_lu(e, 1))) return s;
  // mark instances active for the fill
  std::vector<InstState> st(e->T);
  HIPCHK(hipMemcpy(st.data(), e->d.st, sizeof(InstState) * e->T, hipMemcpyDeviceToHost));
  std::vector<InstState> st2 = st;
  for (auto &x : st2) x.done = 0;
  HIPCHK(hipMemcpy(e->d.st, st2.data(), sizeof(InstState) * e->T, hipMemcpyHostToDevice));
  B4<double> *LU = (B4<double> *)e->d_LU;
  HIPCHK(hipMemset(LU, 0, sizeof(B4<double>) * (size_t)e->T * e->g.nnzLU));
  hipLaunchKernelGGL(k_jacobian<double>, xgrid(e->T, ((size_t)e->g.nJb + 1) / 2, 256),
                     dim3(256), 0, e->stream, e->g, e->d, LU, 0);
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(e->d.st, st.data(), sizeof(InstState) * e->T, hipMemcpyHostToDevice));
  std::vector<B4<double>> blk(e->g.nnzLU);
  HIPCHK(hipMemcpy(blk.data(), LU + (size_t)t * e->g.nnzLU,
                   sizeof(B4<double>) * e->g.nnzLU, hipMemcpyDeviceToHost));
  std::vector<int8_t> ty(e->n);
  HIPCHK(hipMemcpy(ty.data(), e->d.type + (size_t)t * e->n, e->n, hipMemcpyDeviceToHost));
  // reference indices: P row/angle col = indicePVPQ, Q row/magnitude col =
  // indicePVPQ + npq (newtonpf.h:256-291)
  const Symbolic &S = e->S;
  std::vector<int> idx(e->n, -1);
  int npv = 0, npq = 0;
  for (int i = 0; i < e->n; i++) if (ty[i] == BPF_PV) idx[i] = npv++;
  for (int i = 0; i < e->n; i++) if (ty[i] == BPF_PQ && S.blk_of_bus[i] >= 0) idx[i] = npv + npq++;
  int cnt = 0;
  for (int p = 0; p < e->g.nnzY; p++) {
    int q = S.y2lu[p];
    if (q < 0) continue;
    int i = S.yrowof[p], j = S.ycol[p];
    const B4<double> &b = blk[q];
    bool pqi = ty[i] == BPF_PQ, pqj = ty[j] == BPF_PQ;
    auto put = [&](int r, int c, double v) {
      if (cnt < cap && row && col && val) { row[cnt] = r; col[cnt] = c; val[cnt] = v; }
      cnt++;
    };
    put(idx[i], idx[j], b.a);
    if (pqj) put(idx[i], idx[j] + npq, b.b);
    if (pqi) put(idx[i] + npq, idx[j], b.c);
    if (pqi && pqj) put(idx[i] + npq, idx[j] + npq, b.d);
  }
  *nnz = cnt;
  return cnt <= cap ? BPF_OK : BPF_ERR_INVALID_ARG;
}

extern "C" bpf_status bpf_stage_newton_step(bpf_engine *e, const double *V,
                                            double *dx) {
  if (!e || !V) return BPF_ERR_INVALID_ARG;
  bpf_status s = stage_eval(e, V);
  if (s) return s;
  int fp64 = e->opt.precision == BPF_FP64;
  if ((s = ensure_lu(e, fp64))) return s;
  std::vector<InstState> st(e->T);
  HIPCHK(hipMemcpy(st.data(), e->d.st, sizeof(InstState) * e->T, hipMemcpyDeviceToHost));
  std::vector<InstState> st2 = st;
  for (auto &x : st2) x.done = 0;
  HIPCHK(hipMemcpy(e->d.st, st2.data(), sizeof(InstState) * e->T, hipMemcpyHostToDevice));
  if (fp64) launch_newton_step<double>(e); else launch_newton_step<float>(e);
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipGetLastError());
  tcollect(e);
  const size_t T = e->T, n = e->n, nb2 = 2 * (size_t)e->S.nb;
  std::vector<InstState> st3(T);
  HIPCHK(hipMemcpy(st3.data(), e->d.st, sizeof(InstState) * T, hipMemcpyDeviceToHost));
  std::vector<double> xb(T * nb2);
  std::vector<int8_t> ty(T * n);
  HIPCHK(hipMemcpy(xb.data(), e->d.dx, sizeof(double) * T * nb2, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ty.data(), e->d.type, T * n, hipMemcpyDeviceToHost));
  for (size_t t = 0; t < T; t++) {
    if (st3[t].done && st3[t].verdict == BPF_V_SINGULAR) {
      for (size_t k = 0; k < 2 * n; k++) dx[t * 2 * n + k] = NAN;
      continue;
    }
    to_ref_layout(e, ty.data() + t * n, xb.data() + t * nb2, dx + t * 2 * n);
  }
  HIPCHK(hipMemcpy(e->d.st, st.data(), sizeof(InstState) * T, hipMemcpyHostToDevice));
  return BPF_OK;
}

extern "C" bpf_status bpf_adjoint(bpf_engine *e, const double *rate, double w,
                                  double *gP, double *gQ, double *gx,
                                  double *obj) {
  if (!e) return BPF_ERR_INVALID_ARG;
  if (!e->solved) { e->err = "call bpf_solve first"; return BPF_ERR_STATE; }
  if (e->opt.alg != BPF_NR) {
    e->err = "the adjoint uses the Newton Jacobian: solve with BPF_NR";
    return BPF_ERR_UNSUPPORTED;
  }
  bpf_status s = ensure_lu(e, 1);
  if (s) return s;
  const size_t T = e->T, n = e->n, m = e->m, nb2 = 2 * (size_t)e->S.nb;
  std::vector<DBuf> tmp;
  void *p;
  double *d_rate = nullptr, *d_G, *d_lam, *d_gP, *d_gQ, *d_gx, *d_obj;
  auto cleanup = [&]() { free_pool(tmp); };
#define TAL(dst, count)                                                    \
  if ((s = dalloc(e, tmp, sizeof(double) * (count), &p))) { cleanup(); return s; } \
  dst = (double *)p;
  if (rate) {
    TAL(d_rate, m);
    hipMemcpy(d_rate, rate, sizeof(double) * m, hipMemcpyHostToDevice);
  }
  TAL(d_G, T * nb2);
  TAL(d_lam, T * nb2);
  TAL(d_gP, T * n);
  TAL(d_gQ, T * n);
  TAL(d_gx, T * m);
  TAL(d_obj, T);
#undef TAL
  // only converged instances take part
  std::vector<InstState> st(T), st2;
  HIPCHK(hipMemcpy(st.data(), e->d.st, sizeof(InstState) * T, hipMemcpyDeviceToHost));
  st2 = st;
  for (auto &x : st2) x.done = x.verdict == BPF_V_CONVERGED ? 0 : 1;
  HIPCHK(hipMemcpy(e->d.st, st2.data(), sizeof(InstState) * T, hipMemcpyHostToDevice));
  B4<double> *LU = (B4<double> *)e->d_LU;
  const unsigned TT = (unsigned)T;
  tbegin(e, "adjoint");
  hipLaunchKernelGGL(k_jacobian<double>, xgrid(TT, ((size_t)e->g.nJb + 1) / 2, 256),
                     dim3(256), 0, e->stream, e->g, e->d, LU, 0);
  hipLaunchKernelGGL(k_lu_factor<double>, dim3(TT), dim3(e->lu_threads), 0,
                     e->stream, e->g, e->d.st, 0, e->fplan, LU,
                     (B4<double> *)e->d_TL, 0, e->d.jsing);
  hipLaunchKernelGGL(k_adj_rhs, dim3(div_up(e->S.nb, 256), TT), dim3(256), 0,
                     e->stream, e->g, e->d, (const double *)d_rate, w, d_G);
  size_t xs = sizeof(double) * nb2;
  int use_lds = xs <= (size_t)SOLVE_LDS_BYTES;
  hipLaunchKernelGGL(k_adj_solve, dim3(TT), dim3(e->solve_threads),
                     0, e->stream, e->g, e->d, (const B4<double> *)LU,
                     (const B4<double> *)e->d_TL, (const double *)d_G, d_lam,
                     e->splan, use_lds, (double *)e->d_xglob);
  hipLaunchKernelGGL(k_adj_grad, dim3(div_up(std::max(n, m), 256), TT),
                     dim3(256), 0, e->stream, e->g, e->d,
                     (const double *)d_lam, (const double *)d_rate, w, d_gP,
                     d_gQ, d_gx);
  hipLaunchKernelGGL(k_objective, dim3(TT), dim3(256), 0, e->stream, e->g,
                     e->d, (const double *)d_rate, w, d_obj);
  tend(e);
  hipError_t r1 = hipStreamSynchronize(e->stream);
  hipError_t r2 = hipGetLastError();
  tcollect(e);
  // factorization singularities are not expected at a converged point; keep
  // the verdicts of the solve either way
  hipMemcpy(e->d.st, st.data(), sizeof(InstState) * T, hipMemcpyHostToDevice);
  if (r1 != hipSuccess || r2 != hipSuccess) {
    e->err = std::string("adjoint: ") + hipGetErrorString(r1 != hipSuccess ? r1 : r2);
    cleanup();
    return BPF_ERR_HIP;
  }
  if (gP) hipMemcpy(gP, d_gP, sizeof(double) * T * n, hipMemcpyDeviceToHost);
  if (gQ) hipMemcpy(gQ, d_gQ, sizeof(double) * T * n, hipMemcpyDeviceToHost);
  if (gx) hipMemcpy(gx, d_gx, sizeof(double) * T * m, hipMemcpyDeviceToHost);
  if (obj) hipMemcpy(obj, d_obj, sizeof(double) * T, hipMemcpyDeviceToHost);
  cleanup();
  return BPF_OK;
}
