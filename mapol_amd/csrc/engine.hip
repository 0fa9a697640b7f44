// C-ABI of the batched power-flow engine (include/batchpf.h): handle, per-
// instance setup on the host, device buffers and the Newton driver.
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <utility>
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
  // the slot form: X after a zero pair in the solve LDS
  if (form == 0 && S.sl_ok && (threads == 1024 || threads == 512 || threads == 256) &&
      (2 * (size_t)S.nb + SLOT_X_PAD) * scalar <= (size_t)SOLVE_LDS_BYTES)
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
  // case (host copies, buses in the engine's numbering)
  int n = 0, m = 0, ng = 0;
  std::vector<int> bus_type, br_from, br_to, gen_bus;
  // the engine's bus numbers: bus_ext[k] is the caller's bus k is, bus_int
  // the inverse; the caller's Ybus pattern (rows by the caller's bus, columns
  // ascending) and, for each of its entries, the engine's Ybus position
  std::vector<int> bus_ext, bus_int;
  std::vector<int> yx_row, yx_col, yx_pos;
  // the same maps on the device, and a staging buffer for per-bus rows
  // crossing the interface (grown on demand, freed with the engine)
  const int *d_bus_ext = nullptr, *d_bus_int = nullptr;
  void *d_bus_stage = nullptr;
  size_t bus_stage_bytes = 0;
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
  // case arrays of the per-instance expansion and its input staging
  CaseArrays ca{};
  uint8_t *in_bron = nullptr, *in_genon = nullptr;   // [T][m], [T][ng]
  std::vector<DBuf> ctl_owned;
  void *ctl_buf = nullptr;      // int2 [nctl] + double [T][nctl]
  size_t ctl_bytes = 0;
  void *d_LU = nullptr;         // B4<float> or B4<double> [T][nnzLU]
  int lu_fp64 = -1;             // precision the LU buffer was sized for
  void *d_xglob = nullptr;      // solve vectors when LDS is too small
  void *d_TL = nullptr;         // inverse tail Schur complements [T][tn*tn]
  // FP64 factor storage of the adjoint and the Jacobian export when the
  // Newton factors are FP32 (so that neither disturbs the Newton buffers or
  // the captured graph), and the adjoint's per-instance arrays; allocated on
  // first use for the current instance count
  struct Fp64Factors { void *LU = nullptr, *TL = nullptr, *xglob = nullptr; } f64;
  struct AdjBufs {
    double *rate = nullptr, *G = nullptr, *lam = nullptr, *gP = nullptr,
           *gQ = nullptr, *gx = nullptr, *obj = nullptr;
  } adj;
  cd *d_Sf = nullptr, *d_St = nullptr;
  double *d_loss = nullptr;
  int *h_count = nullptr;       // pinned
  // launch shapes
  int lu_threads = 1024, solve_threads = 1024, mis_threads = 256;
  int max_lds = 65536;
  // graph of NEWTON_PER_GRAPH Newton iterations; destroyed whenever what it
  // captured changes (launch shapes, factor storage, instances)
  hipGraphExec_t graph = nullptr;   // part 0 (the whole batch when not split)
  // the batch split into nsplit parts of consecutive instances, each with
  // its own stream, graph and active-instance counter, replayed side by
  // side so that one part's one-workgroup-per-instance kernels (factor,
  // sweeps) overlap another part's many-workgroup kernels
  static constexpr int MAX_SPLIT = 4;
  int nsplit = 1;
  hipStream_t sub[MAX_SPLIT] = {};   // sub[0] unused (the engine's stream)
  hipGraphExec_t subgraph[MAX_SPLIT] = {};
  int split_used = 1;                // parts of the captured graphs
  bool graph_broken = false;   // a replay failed: direct launches from then on
  std::string graph_err;
  int graph_used = 0;          // the last solve replayed graphs
  int graph_fallback = 0;      // the last solve abandoned replay (graph_err says why)
  int graph_captures = 0;      // captures since bpf_create
  int fused_forward = 0;       // the last Newton step's first solve skipped the forward sweep
  // timing
  int timing = 0;
  std::vector<TimedEvent> events;
  std::string launch_err;            // first failed launch (kernel name)
  const char *cur_kernel = "";
  int debug_sync = 0;                // BPF_DEBUG_SYNC: sync after launches
  int sweep_form = 0;                // sweep_mode's form (0 slot, 1 entry, 2 rows)
  int tail_form = 1;                 // BPF_TAIL_MFMA / BPF_TAIL_VALU
  int per_graph = 4;                 // Newton iterations per captured graph
  // split factorization (several workgroups per instance while few
  // instances are active): the plan has a split variant; split_mode 0
  // adaptive, 1 off, 2 / 4 / 8 that many workgroups (bpf_set_factor_split);
  // the compute units and the parts running side by side set the budget
  bool fplan_split = false;
  int split_mode = 0;
  int num_cu = 256;
  int parts_running = 1;
  bool capturing = false;            // recording the iteration graph
  std::map<std::string, std::pair<int, double>> ktime;
  bool solved = false;
  bool ybus_ready = false;           // Ybus of the current instances assembled
  double *scratch = nullptr;         // this device's allocation-free scratch
};

static bpf_status scratch_setup(bpf_engine *e);

// the captured Newton graphs (whole batch or its parts), dropped whenever
// what they captured changes
static void drop_graphs(bpf_engine *e) {
  if (e->graph) { hipGraphExecDestroy(e->graph); e->graph = nullptr; }
  for (int h = 1; h < bpf_engine::MAX_SPLIT; h++)
    if (e->subgraph[h]) { hipGraphExecDestroy(e->subgraph[h]); e->subgraph[h] = nullptr; }
}

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

static std::vector<double> nonemptyd(const std::vector<double> &v) {
  return v.empty() ? std::vector<double>(1, 0.0) : v;
}
static std::vector<uint8_t> nonempty8(const std::vector<uint8_t> &v) {
  return v.empty() ? std::vector<uint8_t>(1, 0) : v;
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
  o->tail_limit = 0;
  o->chunk_products = 0;
  o->slot_chunk = 0;
  o->ordering = 0;
  o->bus_numbering = 1;
}

extern "C" const char *bpf_last_error(const bpf_engine *e) {
  if (!e) return "null engine";
  if (e->err.empty() && e->graph_broken) return e->graph_err.c_str();
  return e->err.c_str();
}

extern "C" void bpf_destroy(bpf_engine *e) {
  if (!e) return;
  drop_graphs(e);
  for (int h = 1; h < bpf_engine::MAX_SPLIT; h++)
    if (e->sub[h]) hipStreamDestroy(e->sub[h]);
  for (auto &ev : e->events) { hipEventDestroy(ev.a); hipEventDestroy(ev.b); }
  free_pool(e->inst_owned);
  free_pool(e->ctl_owned);
  free_pool(e->owned);
  if (e->d_bus_stage) hipFree(e->d_bus_stage);
  if (e->h_count) hipHostFree(e->h_count);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
}

template <typename T>
static std::vector<T> vcopy(const T *p, int n) {
  return p ? std::vector<T>(p, p + n) : std::vector<T>();
}

// T rows of per-bus values between the caller's bus numbering and the
// engine's, permuted on the device through a staging buffer: out of the
// engine, dst[t][x] = src[t][bus_int[x]]; into it, dst[t][k] = src[t][bus_ext[k]]
static bpf_status bus_stage(bpf_engine *e, size_t bytes) {
  if (e->bus_stage_bytes >= bytes) return BPF_OK;
  if (e->d_bus_stage) HIPCHK(hipFree(e->d_bus_stage));
  e->d_bus_stage = nullptr;
  e->bus_stage_bytes = 0;
  HIPCHK(hipMalloc(&e->d_bus_stage, bytes));
  e->bus_stage_bytes = bytes;
  return BPF_OK;
}

template <class V>
static bpf_status bus_rows_download(bpf_engine *e, const void *dsrc, V *dst, size_t T) {
  const size_t cnt = T * e->n, bytes = sizeof(V) * cnt;
  if (cnt == 0) return BPF_OK;
  if (!e->opt.bus_numbering) {
    HIPCHK(hipMemcpyAsync(dst, dsrc, bytes, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return BPF_OK;
  }
  bpf_status s = bus_stage(e, bytes);
  if (s) return s;
  hipLaunchKernelGGL(k_bus_gather<V>, dim3(div_up(cnt, 256)), dim3(256), 0, e->stream,
                     (const V *)dsrc, (V *)e->d_bus_stage, e->d_bus_int, e->n, (int)T);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dst, e->d_bus_stage, bytes, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return BPF_OK;
}

// the caller's T rows of voltages into the engine's V (queued on the
// engine's stream; the host array is read before the call returns)
static bpf_status bus_rows_upload(bpf_engine *e, const cd *src, cd *ddst, size_t T) {
  const size_t cnt = T * e->n, bytes = sizeof(cd) * cnt;
  if (cnt == 0) return BPF_OK;
  if (!e->opt.bus_numbering) {
    HIPCHK(hipMemcpyAsync(ddst, src, bytes, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return BPF_OK;
  }
  bpf_status s = bus_stage(e, bytes);
  if (s) return s;
  HIPCHK(hipMemcpyAsync(e->d_bus_stage, src, bytes, hipMemcpyHostToDevice, e->stream));
  hipLaunchKernelGGL(k_bus_gather<cd>, dim3(div_up(cnt, 256)), dim3(256), 0, e->stream,
                     (const cd *)e->d_bus_stage, ddst, e->d_bus_ext, e->n, (int)T);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  return BPF_OK;
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
      e->opt.maxit_fd < 0 || !(e->opt.tol > 0) || e->opt.bus_numbering < 0 ||
      e->opt.bus_numbering > 1) {
    e->err = "invalid options";
    return BPF_ERR_INVALID_ARG;
  }
  const int n = c->nbus;
  e->n = n; e->m = c->nbranch; e->ng = c->ngen;
  for (int b = 0; b < e->m; b++)
    if (c->br_from[b] < 0 || c->br_from[b] >= n || c->br_to[b] < 0 ||
        c->br_to[b] >= n) {
      e->err = "branch " + std::to_string(b) + " has a bus out of range";
      return BPF_ERR_INVALID_ARG;
    }
  for (int k = 0; k < e->ng; k++)
    if (c->gen_bus[k] < 0 || c->gen_bus[k] >= n) {
      e->err = "generator bus out of range";
      return BPF_ERR_INVALID_ARG;
    }
  // The engine numbers buses by reverse Cuthill-McKee: the per-bus kernels
  // gather over each bus's branches, and near numbers put those reads in
  // few cache lines.  The elimination order is taken from the caller's
  // numbering (minimum degree breaks ties by bus number), so the factors and
  // schedules are those of the case as given.
  std::vector<int> order;
  e->bus_ext.resize(n);
  if (e->opt.bus_numbering) {
    order = elimination_order(n, e->m, c->br_from, c->br_to, c->bus_type,
                              e->opt.ordering);
    e->bus_ext = rcm_numbering(n, e->m, c->br_from, c->br_to);
  } else {
    std::iota(e->bus_ext.begin(), e->bus_ext.end(), 0);
  }
  e->bus_int.assign(n, 0);
  for (int k = 0; k < n; k++) e->bus_int[e->bus_ext[k]] = k;
  for (int &b : order) b = e->bus_int[b];
  auto bus_vec = [&](auto *p) {
    std::vector<std::remove_const_t<std::remove_pointer_t<decltype(p)>>> v(n);
    for (int k = 0; k < n; k++) v[k] = p[e->bus_ext[k]];
    return v;
  };
  auto bus_ids = [&](const int *p, int cnt) {
    std::vector<int> v(cnt);
    for (int k = 0; k < cnt; k++) v[k] = e->bus_int[p[k]];
    return v;
  };
  e->bus_type = bus_vec(c->bus_type);
  e->Pd = bus_vec(c->Pd); e->Qd = bus_vec(c->Qd);
  e->Gs = bus_vec(c->Gs); e->Bs = bus_vec(c->Bs);
  e->Vm = bus_vec(c->Vm); e->Va = bus_vec(c->Va);
  e->br_from = bus_ids(c->br_from, e->m); e->br_to = bus_ids(c->br_to, e->m);
  e->R = vcopy(c->R, e->m); e->X = vcopy(c->X, e->m); e->B = vcopy(c->B, e->m);
  e->tap = vcopy(c->tap, e->m); e->shift = vcopy(c->shift, e->m);
  e->br_status = vcopy(c->br_status, e->m);
  e->gen_bus = bus_ids(c->gen_bus, e->ng);
  e->Pg = vcopy(c->Pg, e->ng); e->Qg = vcopy(c->Qg, e->ng);
  e->Qmax = vcopy(c->Qmax, e->ng); e->Qmin = vcopy(c->Qmin, e->ng);
  e->VG = vcopy(c->VG, e->ng); e->gen_status = vcopy(c->gen_status, e->ng);
  {
    AnalysisParams prm;
    prm.tail_limit = e->opt.tail_limit;
    prm.chunk_pairs = e->opt.chunk_products;
    prm.slot_chunk = e->opt.slot_chunk;
    prm.ordering = e->opt.ordering;
    prm.order_bus = order;
    if (!analyse(e->n, e->m, e->br_from.data(), e->br_to.data(),
                 e->bus_type.data(), e->S, &e->err, prm))
      return BPF_ERR_INVALID_ARG;
  }
  {
    // the caller's Ybus pattern and where each of its entries sits here
    const Symbolic &S = e->S;
    std::vector<std::pair<int, int>> rc;
    e->yx_row.assign(n + 1, 0);
    e->yx_col.clear();
    e->yx_pos.clear();
    for (int r = 0; r < n; r++) {
      const int R = e->bus_int[r];
      rc.clear();
      for (int p = S.yrow[R]; p < S.yrow[R + 1]; p++)
        rc.push_back({e->bus_ext[S.ycol[p]], p});
      std::sort(rc.begin(), rc.end());
      for (auto &x : rc) { e->yx_col.push_back(x.first); e->yx_pos.push_back(x.second); }
      e->yx_row[r + 1] = (int)e->yx_col.size();
    }
  }
  // device
  const char *dbg = getenv("BPF_DEBUG_SYNC");
  e->debug_sync = (dbg && dbg[0] == '1') ? 1 : 0;
  HIPCHK(hipGetDevice(&e->dev));
  {
    bpf_status s0 = scratch_setup(e);
    if (s0) return s0;
  }
  HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  for (int h = 1; h < bpf_engine::MAX_SPLIT; h++)
    HIPCHK(hipStreamCreateWithFlags(&e->sub[h], hipStreamNonBlocking));
  HIPCHK(hipDeviceGetAttribute(&e->max_lds,
                               hipDeviceAttributeMaxSharedMemoryPerBlock, e->dev));
  HIPCHK(hipDeviceGetAttribute(&e->num_cu, hipDeviceAttributeMultiprocessorCount,
                               e->dev));
  HIPCHK(hipHostMalloc((void **)&e->h_count, sizeof(int) * bpf_engine::MAX_SPLIT,
                       hipHostMallocDefault));
  const Symbolic &S = e->S;
  DevTopo &g = e->g;
  g.n = e->n; g.m = e->m; g.nb = S.nb;
  g.nnzY = (int)S.ycol.size();
  g.rhs0 = (int)S.lcol.size();
  g.nnzLU = g.rhs0 + S.nb;
  bpf_status s;
#define UP(dst, vec) \
  if ((s = upload(e, e->owned, vec, &dst))) return s;
  UP(g.yrow, S.yrow); UP(g.ycol, S.ycol); UP(g.yrowof, S.yrowof);
  UP(e->d_bus_ext, e->bus_ext); UP(e->d_bus_int, e->bus_int);
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
  {
    // instance-independent branch quotients, computed on the device
    void *p1, *p2;
    if ((s = dalloc(e, e->owned, sizeof(cd) * std::max(e->m, 1), &p1))) return s;
    if ((s = dalloc(e, e->owned, sizeof(cd) * std::max(e->m, 1), &p2))) return s;
    if (e->m) {
      hipLaunchKernelGGL(k_branch_pre, dim3(div_up(e->m, 256)), dim3(256), 0,
                         e->stream, g, (cd *)p1, (cd *)p2);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(e->stream));
    }
    g.ys = (const cd *)p1;
    g.ejs = (const cd *)p2;
  }
  UP(g.Gs, e->Gs); UP(g.Qd, e->Qd); UP(g.Va0, e->Va);
  UP(g.base_type, e->bus_type);
  {
    // case arrays of the per-instance expansion; generators grouped by bus
    // in ascending generator order
    CaseArrays &ca = e->ca;
    std::vector<int> gptr(e->n + 1, 0), ggen(e->ng);
    for (int k = 0; k < e->ng; k++) gptr[e->gen_bus[k] + 1]++;
    for (int i = 0; i < e->n; i++) gptr[i + 1] += gptr[i];
    {
      std::vector<int> w(gptr.begin(), gptr.end() - 1);
      for (int k = 0; k < e->ng; k++) ggen[w[e->gen_bus[k]]++] = k;
    }
    std::vector<uint8_t> bst(e->br_status), gst(e->gen_status);
    for (auto &x : bst) x = x ? 1 : 0;
    for (auto &x : gst) x = x ? 1 : 0;
    UP(ca.Pd, e->Pd); UP(ca.Qd, e->Qd); UP(ca.Bs, e->Bs); UP(ca.Vm, e->Vm);
    UP(ca.tap, nonemptyd(e->tap)); UP(ca.br_status, nonempty8(bst));
    UP(ca.gb_ptr, gptr); UP(ca.gb_gen, nonempty(ggen, 1));
    UP(ca.Pg, nonemptyd(e->Pg)); UP(ca.Qg, nonemptyd(e->Qg));
    UP(ca.Qmax, nonemptyd(e->Qmax)); UP(ca.Qmin, nonemptyd(e->Qmin));
    UP(ca.VG, nonemptyd(e->VG)); UP(ca.gen_status, nonempty8(gst));
    ca.ng = e->ng;
  }
  FactorPlan &F = e->fplan;
  F.nchunk = S.nchunk;
  {
    // every chunk as {first task, end task, first product, end product} and
    // the task count and longest reduction that set its lane groups: the
    // main chunks, then the split schedules' (their tasks and products
    // after the main ones)
    const int nt0 = (int)S.tk_a.size() / 4, np0 = (int)S.pairs.size() / 2;
    const int nsp = (int)S.sp_lg.size() / 2, nall = S.nchunk + nsp;
    std::vector<int> ch(4 * nall), lg(2 * nall), rhs(2 * nall);
    std::vector<int> tk(S.tk_a.size() + S.sp_tk_a.size());
    // one descriptor per output; an L output's pivot range packed as its
    // offset from the chunk's first product | length << TK_LEN_SHIFT
    auto pack = [&](const std::vector<int> &ta, const std::vector<int> &tb, int k0,
                    int k1, int q0, int toff, int poff) {
      for (int k = k0; k < k1; k++) {
        int *d = &tk[4 * (size_t)(toff + k)];
        d[0] = ta[4 * k];
        d[1] = ta[4 * k + 1] + poff;
        d[2] = ta[4 * k + 2] + poff;
        d[3] = 0;
        if (ta[4 * k] & TASK_LCOL) {
          const int v0 = tb[4 * k + 1] - q0, len = tb[4 * k + 2] - tb[4 * k + 1];
          if (v0 < 0 || v0 >= (1 << TK_LEN_SHIFT) || len < 0 || len >= (1 << TK_LEN_SHIFT))
            return false;
          d[3] = v0 | (len << TK_LEN_SHIFT);
        }
      }
      return true;
    };
    bool ok = true;
    for (int c = 0; c < S.nchunk; c++) {
      const int r[4] = {S.ch_task[c], S.ch_task[c + 1], S.ch_pair[c], S.ch_pair[c + 1]};
      std::copy(r, r + 4, &ch[4 * c]);
      lg[2 * c] = r[1] - r[0];
      lg[2 * c + 1] = S.ch_max[c];
      rhs[2 * c] = S.ch_rhs[2 * c];
      rhs[2 * c + 1] = S.ch_rhs[2 * c + 1];
      ok = ok && pack(S.tk_a, S.tk_b, r[0], r[1], r[2], 0, 0);
    }
    for (int c = 0; c < nsp; c++) {
      const int *r = &S.sp_ch[4 * c];
      int *d = &ch[4 * (S.nchunk + c)];
      d[0] = r[0] + nt0; d[1] = r[1] + nt0; d[2] = r[2] + np0; d[3] = r[3] + np0;
      lg[2 * (S.nchunk + c)] = S.sp_lg[2 * c];
      lg[2 * (S.nchunk + c) + 1] = S.sp_lg[2 * c + 1];
      rhs[2 * (S.nchunk + c)] = S.sp_rhs[2 * c];
      rhs[2 * (S.nchunk + c) + 1] = S.sp_rhs[2 * c + 1];
      ok = ok && pack(S.sp_tk_a, S.sp_tk_b, r[0], r[1], r[2], nt0, np0);
    }
    // the lane-group key {tasks, longest reduction} of a chunk carries the
    // count of its forward-substitution outputs and products in the upper
    // 16 bits (the LDS chunk tables have no room for another array)
    for (size_t q = 0; q < lg.size(); q++) {
      if (lg[q] < 0 || lg[q] >= (1 << 16) || rhs[q] < 0 || rhs[q] >= (1 << 15)) {
        e->err = "internal: a factorization chunk's counts exceed 16 bits";
        return BPF_ERR_UNSUPPORTED;
      }
      lg[q] |= rhs[q] << 16;
    }
    if (!ok) {
      e->err = "internal: a pivot's product range outside its factorization chunk";
      return BPF_ERR_UNSUPPORTED;
    }
    std::vector<int> pairs(S.pairs);
    pairs.insert(pairs.end(), S.sp_pairs.begin(), S.sp_pairs.end());
    const int *cp, *lp, *tp, *pr;
    UP(cp, nonempty(ch, 4));
    UP(lp, nonempty(lg, 2));
    UP(tp, nonempty(tk, 4));
    UP(pr, nonempty(pairs, 2));
    F.chunks = (const int4 *)cp;
    F.chunk_lg = (const int2 *)lp;
    F.tk = (const int4 *)tp;
    F.pairs = (const int2 *)pr;
    e->fplan_split = false;
    for (int v = 0; v < SPLIT_VARIANTS; v++) {
      const bool has = S.sp_level[v] >= 0;
      F.sp_rest[v] = has ? S.sp_rest[v] : -1;
      for (int j = 0; j <= SPLIT_MAX; j++)
        F.sp_grp[v][j] = S.nchunk + S.sp_grp[v * (SPLIT_MAX + 1) + j];
      e->fplan_split = e->fplan_split || has;
    }
  }
  F.tn = S.tail_n;
  F.tnent = (int)S.tl_pos.size();
  UP(F.tdense, S.tl_dense); UP(F.tpos, S.tl_pos); UP(F.tpp, S.tl_pp);
  UP(F.tpairL, S.tl_pairL); UP(F.tpairU, S.tl_pairU);
  F.tstage_n = (int)S.tl_stage.size();
  F.tail_mfma = 1;   // bpf_set_schedule chooses
  F.fused_limit = FUSED_LIMIT_DEFAULT;   // bpf_set_fused_forward chooses
  UP(F.tstage, nonempty(S.tl_stage, 1)); UP(F.tpairS, nonempty(S.tl_pairS, 1));
  {
    const std::vector<int> js = tail_jobs(S.tl_pp, 12), jg = tail_jobs(S.tl_pp, 4);
    F.tjobs_s = (int)js.size() / 2;
    F.tjobs_g = (int)jg.size() / 2;
    const int *p;
    UP(p, nonempty(js, 2)); F.tjob_s = (const int2 *)p;
    UP(p, nonempty(jg, 2)); F.tjob_g = (const int2 *)p;
  }
  SolvePlan &SP = e->splan;
  SP.fused_limit = FUSED_LIMIT_DEFAULT;
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
    const int *cht, *slt;
    UP(ch, nonempty(S.sl_chunk, 4)); UP(sl, nonempty(S.sl_slot, 4));
    UP(cht, nonempty(S.sl_chunk_t, 4)); UP(slt, nonempty(S.sl_slot_t, 4));
    SP.sl_chunk = (const int4 *)ch;
    SP.sl_slot = (const int4 *)sl;
    SP.sl_chunk_t = (const int4 *)cht;
    SP.sl_slot_t = (const int4 *)slt;
    SP.sl_nfw_t = S.sl_nfw_t;
    SP.sl_nbw_t = S.sl_nbw_t;
  }
  e->sweep_form = 0;
  e->tail_form = BPF_TAIL_MFMA;
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
  drop_graphs(e);
  return BPF_OK;
}

extern "C" bpf_status bpf_set_factor_split(bpf_engine *e, int32_t workgroups) {
  if (!e) return BPF_ERR_INVALID_ARG;
  if (workgroups != 0 && workgroups != 1 && workgroups != 2 && workgroups != 4 &&
      workgroups != 8) {
    e->err = "factorization split: 0 (adaptive), 1, 2, 4 or 8 workgroups";
    return BPF_ERR_INVALID_ARG;
  }
  e->split_mode = workgroups;
  drop_graphs(e);
  return BPF_OK;
}

extern "C" bpf_status bpf_set_fused_forward(bpf_engine *e, int32_t max_active) {
  if (!e) return BPF_ERR_INVALID_ARG;
  if (max_active < -1) {
    e->err = "fused forward substitution: -1 (never) or the most active instances";
    return BPF_ERR_INVALID_ARG;
  }
  e->fplan.fused_limit = max_active;
  e->splan.fused_limit = max_active;
  drop_graphs(e);
  return BPF_OK;
}

extern "C" bpf_status bpf_set_schedule(bpf_engine *e, int32_t per_graph,
                                       int32_t tail_inverse, int32_t sweep_form,
                                       int32_t parts) {
  if (!e) return BPF_ERR_INVALID_ARG;
  if (per_graph < 0 || per_graph > 16 || tail_inverse < 0 || tail_inverse > 2 ||
      sweep_form < 0 || sweep_form > 3 || parts < 0 ||
      parts > bpf_engine::MAX_SPLIT) {
    e->err = "schedule out of range (newton_per_graph 1..16, tail_inverse 1..2, "
             "sweep_form 1..3, parts 1..4)";
    return BPF_ERR_INVALID_ARG;
  }
  if (per_graph) e->per_graph = per_graph;
  if (parts) e->nsplit = parts;
  if (tail_inverse) {
    e->tail_form = tail_inverse;
    e->fplan.tail_mfma = tail_inverse == BPF_TAIL_MFMA ? 1 : 0;
  }
  if (sweep_form) e->sweep_form = sweep_form - 1;
  drop_graphs(e);
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
  e->solved = false;
  e->ybus_ready = false;
  bpf_status s;
  void *p;
  const size_t Tn = (size_t)T * n, Tm = (size_t)T * m;
  DevInst &d = e->d;
  // Per-instance buffers are kept while the instance count stays the same,
  // so a new batch of contingencies costs the upload of its masks and one
  // expansion kernel, and the captured Newton graph stays valid.
  if (T != e->T || !d.Y) {
    drop_graphs(e);
    free_pool(e->inst_owned);
    // the factor storage lived in that pool too: forget it, so that
    // ensure_lu allocates it for the new T instead of freeing the stale
    // address again (which a new buffer of this call may have been given)
    e->d_LU = e->d_xglob = e->d_TL = nullptr;
    e->f64 = bpf_engine::Fp64Factors();
    e->adj = bpf_engine::AdjBufs();
    e->lu_fp64 = -1;
    e->T = 0;
    d = DevInst();
    d.T = T;
    const int nb = e->S.nb, nnzY = e->g.nnzY;
#define AL(dst, type, count) \
  if ((s = dalloc(e, e->inst_owned, sizeof(type) * (count), &p))) return s; dst = (type *)p;
    AL(e->in_bron, uint8_t, std::max<size_t>(Tm, 1));
    AL(e->in_genon, uint8_t, (size_t)T * std::max(ng, 1));
    AL(d.br_on, uint8_t, std::max<size_t>(Tm, 1));
    AL(d.tap, double, std::max<size_t>(Tm, 1));
    AL(d.Bs, double, Tn);
    AL(d.vset, double, Tn);
    AL(d.Qmax, double, Tn);
    AL(d.Qmin, double, Tn);
    AL(e->d_type0, int8_t, Tn);
    AL(e->d_Qinj0, double, Tn);
    AL(d.Pinj, double, Tn);
    AL(d.type, int8_t, Tn);
    AL(d.Qinj, double, Tn);
    AL(d.Y, cd, (size_t)T * nnzY);
    AL(d.V, cd, Tn);
    AL(d.Ib, cd, Tn);
    AL(d.F, double, (size_t)T * 2 * nb);
    AL(d.dx, double, (size_t)T * 2 * nb);
    AL(d.rhs, double, (size_t)T * 2 * nb);
    AL(d.st, InstState, T);
    AL(d.active_count, int, bpf_engine::MAX_SPLIT);
    AL(e->d_Sf, cd, std::max<size_t>(Tm, 1));
    AL(e->d_St, cd, std::max<size_t>(Tm, 1));
    AL(e->d_loss, double, T);
    AL(d.part, double, (size_t)T * div_up(e->n, 64));
    AL(d.partv, int, (size_t)T * div_up(e->n, 64));
    AL(d.jsing, int, T);
#undef AL
    e->T = T;
  }
  // controls: descriptors and values, reallocated when they grow
  if (nctl > 0) {
    const size_t need = sizeof(int2) * nctl + sizeof(double) * (size_t)T * nctl;
    if (need > e->ctl_bytes) {
      free_pool(e->ctl_owned);
      e->ctl_bytes = 0;
      if ((s = dalloc(e, e->ctl_owned, need, &p))) return s;
      e->ctl_buf = p;
      e->ctl_bytes = need;
    }
  }
  if (br_on && m)
    HIPCHK(hipMemcpyAsync(e->in_bron, br_on, Tm, hipMemcpyHostToDevice, e->stream));
  if (gen_on && ng)
    HIPCHK(hipMemcpyAsync(e->in_genon, gen_on, (size_t)T * ng, hipMemcpyHostToDevice,
                          e->stream));
  int2 *d_ctl = nullptr;
  double *d_ctlv = nullptr;
  if (nctl > 0) {
    std::vector<int2> cv(nctl);
    for (int c = 0; c < nctl; c++) {
      const int k = ctl[c].kind;
      cv[c].x = k;
      cv[c].y = (k == BPF_CTL_AVR || k == BPF_CTL_SHC) ? e->bus_int[ctl[c].id]
                                                        : ctl[c].id;
    }
    d_ctl = (int2 *)e->ctl_buf;
    d_ctlv = (double *)(d_ctl + nctl);
    HIPCHK(hipMemcpyAsync(d_ctl, cv.data(), sizeof(int2) * nctl,
                          hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(d_ctlv, ctl_val, sizeof(double) * (size_t)T * nctl,
                          hipMemcpyHostToDevice, e->stream));
    // the host vector must outlive the copy
    HIPCHK(hipStreamSynchronize(e->stream));
  }
  hipLaunchKernelGGL(k_expand, dim3(div_up(std::max(n, m), 256), T), dim3(256), 0,
                     e->stream, e->g, d, e->ca, (br_on && m) ? e->in_bron : nullptr,
                     (gen_on && ng) ? e->in_genon : nullptr, (int8_t *)e->d_type0,
                     (double *)e->d_Qinj0, d.Pinj, (uint8_t *)d.br_on,
                     (double *)d.tap, (double *)d.Bs, (double *)d.vset,
                     (double *)d.Qmax, (double *)d.Qmin);
  HIPCHK(hipGetLastError());
  if (nctl > 0) {
    hipLaunchKernelGGL(k_expand_ctl, dim3(div_up(T, 64)), dim3(64), 0, e->stream,
                       e->g, d, e->ca.tap, nctl, (const int2 *)d_ctl,
                       (const double *)d_ctlv, (double *)d.tap, (double *)d.Bs,
                       (double *)d.vset);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemsetAsync(d.jsing, 0, sizeof(int) * (size_t)T, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
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
  drop_graphs(e);
  return BPF_OK;
}

// FP64 factor storage for the adjoint and the Jacobian export: the Newton
// buffers when they are FP64, else a second set kept for the instance count
static bpf_status fp64_factors(bpf_engine *e, B4<double> **LU, B4<double> **TL,
                               double **xglob) {
  if (e->opt.precision == BPF_FP64) {
    bpf_status s = ensure_lu(e, 1);
    if (s) return s;
    *LU = (B4<double> *)e->d_LU;
    *TL = (B4<double> *)e->d_TL;
    *xglob = (double *)e->d_xglob;
    return BPF_OK;
  }
  auto &f = e->f64;
  if (!f.LU) {
    const size_t el = sizeof(B4<double>);
    bpf_status s = dalloc(e, e->inst_owned, el * (size_t)e->T * e->g.nnzLU, &f.LU);
    if (!s) s = dalloc(e, e->inst_owned,
                       el * (size_t)e->T * e->S.tail_n * e->S.tail_n, &f.TL);
    const size_t xs = sizeof(double) * 2 * (size_t)e->S.nb;
    if (!s && xs > (size_t)SOLVE_LDS_BYTES)
      s = dalloc(e, e->inst_owned, xs * e->T, &f.xglob);
    if (s) { f = bpf_engine::Fp64Factors(); return s; }
  }
  *LU = (B4<double> *)f.LU;
  *TL = (B4<double> *)f.TL;
  *xglob = (double *)f.xglob;
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

// slots per lane of the slot form's kernel (0: another form runs)
static int slot_lanes(const bpf_engine *e, int ep) {
  int sp = 1;
  while (sp < 8 && sp * e->solve_threads < e->S.prm.slot_chunk) sp *= 2;
  return ep == 2 && sp * e->solve_threads >= e->S.prm.slot_chunk ? sp : 0;
}

// one sweep pass of k_lu_solve; the slot form (ep 2) has one kernel per
// count of slots per lane: the fewest (1, 2, 4 or 8) whose lanes cover the
// plan's largest chunk (fewer slots per lane, fewer instructions per chunk)
template <typename T>
static void launch_solve(bpf_engine *e, int T_, const void *LU, int rhs_mode,
                         int accumulate, int use_lds, int ep,
                         const void *TL = nullptr, void *xglob = nullptr) {
  const int th = e->solve_threads;
  if (!TL) { TL = e->d_TL; xglob = e->d_xglob; }
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(T_), dim3(th), 0, e->stream, e->g, e->d,
                       (const B4<T> *)LU, (const B4<T> *)TL, rhs_mode,
                       accumulate, e->splan, use_lds, ep, (T *)xglob);
  };
  const int sp = slot_lanes(e, ep);
  if (sp > 0) {
    if (sp == 1) go(k_lu_solve<T, 1>);
    else if (sp == 2) go(k_lu_solve<T, 2>);
    else if (sp == 4) go(k_lu_solve<T, 4>);
    else go(k_lu_solve<T, 8>);
  } else {
    go(k_lu_solve<T, 0>);
  }
}

template <typename T>
static void launch_newton_step(bpf_engine *e) {
  B4<T> *LU = (B4<T> *)e->d_LU;
  // the fill also finishes level 0 when the analysis allows it
  const int c0 = e->S.l0_chunks;
  size_t xs = sizeof(T) * 2 * (size_t)e->S.nb;
  int use_lds = xs <= (size_t)SOLVE_LDS_BYTES;
  const int ep = sweep_mode(e->S, sizeof(T), use_lds, e->sweep_form, e->solve_threads);
  // with the slot-form sweeps the factorization can also form the forward
  // substitution of the mismatch (the fill writes its right-hand side, and
  // the first solve starts from it; the kernels leave it out while more
  // than the plan's limit of instances are active)
  const int fused = slot_lanes(e, ep) > 0 && e->fplan.fused_limit >= 0 ? 1 : 0;
  FactorPlan fp = e->fplan;
  if (!fused) fp.fused_limit = -1;
  e->fused_forward = fused ? fp.fused_limit : -1;
  tbegin(e, "jacobian");
  hipLaunchKernelGGL(k_jacobian<T>, xgrid(e->T, ((size_t)e->g.nJb + 1) / 2, 256),
                     dim3(256), 0, e->stream, e->g, e->d, LU, c0 > 0 ? 1 : 0,
                     fp.fused_limit);
  tend(e);
  tbegin(e, "lu_factor");
  if (e->split_mode != 1 && e->fplan_split) {
    // split launch pair: the subtrees below the cut with several workgroups
    // per instance when few instances are active, then the top; the
    // workgroup count follows the batch's active-instance counter (the
    // compute units shared among the parts that run side by side)
    // the kernel's choice: -2 adaptive, else the variant of 2, 4 or 8 groups
    int force = -2;
    if (e->split_mode == 2) force = 0;
    else if (e->split_mode == 4) force = 1;
    else if (e->split_mode == 8) force = 2;
    const int budget = std::max(1, e->num_cu / std::max(1, e->parts_running));
    hipLaunchKernelGGL(k_lu_factor<T>, dim3(e->T * SPLIT_MAX), dim3(e->lu_threads), 0,
                       e->stream, e->g, e->d.st, 1, fp, LU,
                       (B4<T> *)e->d_TL, c0, e->d.jsing, 1,
                       (const int *)e->d.active_count, budget, force);
    hipLaunchKernelGGL(k_lu_factor_top<T>, dim3(e->T), dim3(e->lu_threads), 0,
                       e->stream, e->g, e->d.st, 1, fp, LU,
                       (B4<T> *)e->d_TL, c0, e->d.jsing,
                       (const int *)e->d.active_count, budget, force);
  } else {
    hipLaunchKernelGGL(k_lu_factor<T>, dim3(e->T), dim3(e->lu_threads), 0,
                       e->stream, e->g, e->d.st, 1, fp, LU,
                       (B4<T> *)e->d_TL, c0, e->d.jsing, 0,
                       (const int *)e->d.active_count, 0, -1);
  }
  tend(e);
  const int refine = (sizeof(T) == 4) ? e->opt.refine_steps : 0;
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
    launch_solve<T>(e, e->T, LU, s == 0 ? (fused ? 4 : 0) : 3, s > 0 ? 1 : 0, use_lds, ep);
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



// A part of the batch, the instances [t0, t1), as a view of the engine's
// per-instance arrays: every pointer advanced by t0 instances, its own
// active-instance counter and stream.
struct PartView {
  DevInst d;
  int T;
  void *LU, *TL, *xg;
  hipStream_t s;
  int *hc;
};

static PartView part_view(bpf_engine *e, int h, int K) {
  const int t0 = (int)((long long)e->T * h / K), t1 = (int)((long long)e->T * (h + 1) / K);
  const size_t n = e->n, m = e->m, nb2 = 2 * (size_t)e->S.nb, nY = e->g.nnzY;
  const size_t np = div_up(e->n, e->mis_threads), o = t0;
  PartView v;
  DevInst d = e->d;
  d.T = t1 - t0;
  d.br_on += o * m; d.tap += o * m;
  d.Bs += o * n; d.vset += o * n; d.type += o * n; d.Pinj += o * n;
  d.Qinj += o * n; d.Qmax += o * n; d.Qmin += o * n;
  d.Y += o * nY; d.V += o * n; d.Ib += o * n;
  d.F += o * nb2; d.dx += o * nb2; d.rhs += o * nb2;
  d.st += o; d.jsing += o; d.active_count += h;
  d.part += o * np; d.partv += o * np;
  v.d = d;
  v.T = t1 - t0;
  const size_t el = e->lu_fp64 ? sizeof(B4<double>) : sizeof(B4<float>);
  const size_t tn = e->S.tail_n;
  v.LU = (char *)e->d_LU + o * el * e->g.nnzLU;
  v.TL = (char *)e->d_TL + o * el * tn * tn;
  v.xg = e->d_xglob ? (char *)e->d_xglob + o * (e->lu_fp64 ? 8 : 4) * nb2 : nullptr;
  v.s = h == 0 ? e->stream : e->sub[h];
  v.hc = e->h_count + h;
  return v;
}

// the engine's batch fields replaced by a part view for the launches made
// in scope (they read e->d, e->T, the factor storage and e->stream)
struct PartScope {
  bpf_engine *e;
  DevInst d;
  int T;
  void *LU, *TL, *xg;
  hipStream_t s;
  int *hc;
  PartScope(bpf_engine *e_, const PartView &v)
      : e(e_), d(e_->d), T(e_->T), LU(e_->d_LU), TL(e_->d_TL), xg(e_->d_xglob),
        s(e_->stream), hc(e_->h_count) {
    e->d = v.d; e->T = v.T; e->d_LU = v.LU; e->d_TL = v.TL; e->d_xglob = v.xg;
    e->stream = v.s; e->h_count = v.hc;
  }
  ~PartScope() {
    e->d = d; e->T = T; e->d_LU = LU; e->d_TL = TL; e->d_xglob = xg;
    e->stream = s; e->h_count = hc;
  }
};

static bpf_status run_newton(bpf_engine *e, bool *graph_failed) {
  bpf_status s = ensure_lu(e, e->opt.precision == BPF_FP64 ? 1 : 0);
  if (s) return s;
  // per-kernel timing uses events around direct launches: elapsed times
  // between event nodes of a replayed graph are not reliable on this runtime
  bool use_graph = e->opt.use_graph != 0 && !e->graph_broken && !e->timing;
  int K = use_graph ? std::max(1, std::min(e->nsplit, e->T)) : 1;
  if (use_graph && e->graph && e->split_used != K) drop_graphs(e);
  e->parts_running = K;
  hipGraphExec_t *slot[bpf_engine::MAX_SPLIT] = {&e->graph};
  for (int h = 1; h < bpf_engine::MAX_SPLIT; h++) slot[h] = &e->subgraph[h];
  PartView view[bpf_engine::MAX_SPLIT];
  for (int h = 0; h < K; h++) view[h] = part_view(e, h, K);
  if (K > 1) {
    // each part's own active count for its first iteration (the first
    // mismatch counted the whole batch into the first counter)
    for (int h = 0; h < K; h++)
      hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, e->stream, view[h].d);
    HIPCHK(hipStreamSynchronize(e->stream));
  }
  if (use_graph && !e->graph) {
    for (int h = 0; h < K && use_graph; h++) {
      PartScope ps(e, view[h]);
      hipGraph_t gr;
      HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      e->capturing = true;
      for (int j = 0; j < e->per_graph; j++) launch_iteration(e);
      e->capturing = false;
      HIPCHK(hipStreamEndCapture(e->stream, &gr));
      e->graph_captures++;
      hipError_t r = hipGraphInstantiate(slot[h], gr, nullptr, nullptr, 0);
      hipGraphDestroy(gr);
      if (r != hipSuccess) {
        *slot[h] = nullptr;
        e->graph_broken = true;
        e->graph_err = std::string("hipGraphInstantiate: ") + hipGetErrorString(r);
        e->graph_fallback = 1;
        (void)hipGetLastError();
        use_graph = false;
      }
    }
    if (!use_graph) {
      // direct launches of the whole batch from here on
      drop_graphs(e);
      K = 1;
      e->parts_running = 1;
      view[0] = part_view(e, 0, 1);
    }
    e->split_used = K;
  }
  e->graph_used = use_graph ? 1 : 0;
  const int cap = e->opt.maxit_nr * (e->opt.max_qlim_rounds + 1) + 2;
  // Each graph holds per_graph iterations (4 by default) and its part's
  // active count is read back once per replay: every kernel of an iteration
  // returns at once for a finished instance, so the iterations replayed
  // past the last one change nothing (the count stays 0), and the host
  // waits once per round of replays.  A part whose count reached 0 is not
  // replayed again.
  const int ahead = use_graph ? e->per_graph : 1;
  bool active[bpf_engine::MAX_SPLIT];
  for (int h = 0; h < K; h++) active[h] = true;
  for (int it = 0; it < cap; it += ahead) {
    hipError_t r = hipSuccess;
    const char *where = "";
    if (use_graph) {
      for (int h = 0; h < K && r == hipSuccess; h++) {
        if (!active[h]) continue;
        where = "hipGraphLaunch";
        r = hipGraphLaunch(*slot[h], view[h].s);
        if (r == hipSuccess) {
          where = "hipMemcpyAsync";
          r = hipMemcpyAsync(view[h].hc, view[h].d.active_count, sizeof(int),
                             hipMemcpyDeviceToHost, view[h].s);
        }
      }
    } else {
      launch_iteration(e);
    }
    for (int h = 0; h < K && r == hipSuccess; h++) {
      if (!active[h]) continue;
      where = "hipStreamSynchronize";
      r = hipStreamSynchronize(view[h].s);
    }
    if (r == hipSuccess) {
      // bpf_solve cleared the thread's error state on entry, so anything
      // here was raised by this loop
      where = "hipGetLastError";
      r = hipGetLastError();
    }
    if (r != hipSuccess) {
      if (use_graph) {
        // the runtime refused the replay: remember why and let bpf_solve
        // start the solve again with direct launches; the fallback is
        // reported by bpf_get_info and bpf_last_error
        e->graph_broken = true;
        e->graph_fallback = 1;
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
    bool any = false;
    for (int h = 0; h < K; h++) {
      active[h] = active[h] && view[h].hc[0] > 0;
      any = any || active[h];
    }
    if (!any) break;
  }
  return BPF_OK;
}

// fast decoupled (mkl_fdpf, fdpf.h:450-596): B' and B'' of every instance
// assembled and factored once in FP64, then P and Q half-steps; each half-step
// is a masked solve with the same factors, an update and the |V|-normalised
// mismatch with the reference's convergence checks
static bpf_status run_fd(bpf_engine *e) {
  B4<double> *LU, *TLd;
  double *xg;
  bpf_status s = fp64_factors(e, &LU, &TLd, &xg);
  if (s) return s;
  const unsigned T = (unsigned)e->T;
  tbegin(e, "bmat");
  hipLaunchKernelGGL(k_bmat, dim3(div_up(e->g.nnzY, 256), T), dim3(256), 0,
                     e->stream, e->g, e->d, e->opt.alg, LU);
  tend(e);
  tbegin(e, "lu_factor");
  hipLaunchKernelGGL(k_lu_factor<double>, dim3(T), dim3(e->lu_threads), 0,
                     e->stream, e->g, e->d.st, 0, e->fplan, LU, TLd, 0, e->d.jsing, 0, nullptr, 0, -1);
  tend(e);
  size_t xs = sizeof(double) * 2 * (size_t)e->S.nb;
  int use_lds = xs <= (size_t)SOLVE_LDS_BYTES;
  for (int it = 0; it < e->opt.maxit_fd + 1; it++) {
    for (int half = 1; half <= 2; half++) {
      tbegin(e, "fd_solve");
      launch_solve<double>(e, T, LU, half, 0, use_lds,
                           sweep_mode(e->S, sizeof(double), use_lds,
                                      e->sweep_form, e->solve_threads),
                           TLd, xg);
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
  hipLaunchKernelGGL(k_loss, dim3(T), dim3(1024), 0, e->stream, e->g, e->d,
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
  e->graph_fallback = 0;
  // an error code left in this thread by the caller's own HIP calls (or a
  // library's) is not this solve's: clear it so that the checks below read
  // only what the solve itself raises
  (void)hipGetLastError();
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
  e->ybus_ready = true;
  return BPF_OK;
}

extern "C" bpf_status bpf_make_ybus(bpf_engine *e) {
  if (!e) return BPF_ERR_INVALID_ARG;
  if (e->T <= 0) { e->err = "bpf_set_instances not called"; return BPF_ERR_STATE; }
  hipLaunchKernelGGL(k_ybus, xgrid(e->T, e->g.nnzY, 256), dim3(256), 0,
                     e->stream, e->g, e->d);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  e->ybus_ready = true;
  return BPF_OK;
}

extern "C" bpf_status bpf_get_results(bpf_engine *e, double *V, int32_t *its,
                                      int32_t *verdict, double *norm,
                                      double *loss, int32_t *rounds,
                                      int32_t *bus_type) {
  if (!e) return BPF_ERR_INVALID_ARG;
  if (!e->solved) { e->err = "no solution yet"; return BPF_ERR_STATE; }
  const size_t T = e->T, n = e->n;
  bpf_status s;
  if (V && (s = bus_rows_download(e, e->d.V, (cd *)V, T))) return s;
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
    if ((s = bus_rows_download(e, e->d.type, ty.data(), T))) return s;
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

extern "C" bpf_status bpf_analyse_slots_transposed(const bpf_case *c,
                                                   int32_t *counts,
                                                   int32_t *sl_chunk,
                                                   int32_t *sl_slot) {
  if (!c || !counts || c->nbus <= 0 || c->nbranch < 0 || !c->bus_type ||
      (c->nbranch && (!c->br_from || !c->br_to)))
    return BPF_ERR_INVALID_ARG;
  Symbolic S;
  std::string err;
  if (!analyse(c->nbus, c->nbranch, c->br_from, c->br_to, c->bus_type, S, &err))
    return BPF_ERR_INVALID_ARG;
  counts[0] = S.sl_ok_t ? S.sl_nfw_t : 0;
  counts[1] = S.sl_ok_t ? S.sl_nbw_t : 0;
  counts[2] = S.sl_ok_t ? (int)S.sl_slot_t.size() / 4 : 0;
  if (S.sl_ok_t) {
    if (sl_chunk && !S.sl_chunk_t.empty())
      memcpy(sl_chunk, S.sl_chunk_t.data(), sizeof(int32_t) * S.sl_chunk_t.size());
    if (sl_slot && !S.sl_slot_t.empty())
      memcpy(sl_slot, S.sl_slot_t.data(), sizeof(int32_t) * S.sl_slot_t.size());
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
  info->graph_fallback = e->graph_fallback;
  info->graph_captures = e->graph_captures;
  info->fused_forward = e->fused_forward;
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
  if (!e->ybus_ready) {
    e->err = "no Ybus yet: call bpf_make_ybus or bpf_solve";
    return BPF_ERR_STATE;
  }
  if (rowptr) memcpy(rowptr, e->yx_row.data(), sizeof(int) * (e->n + 1));
  if (colind) memcpy(colind, e->yx_col.data(), sizeof(int) * e->yx_col.size());
  if (val) {
    std::vector<cd> y(e->g.nnzY);
    HIPCHK(hipMemcpy(y.data(), e->d.Y + (size_t)t * e->g.nnzY, sizeof(cd) * e->g.nnzY,
                     hipMemcpyDeviceToHost));
    cd *out = (cd *)val;
    for (size_t p = 0; p < y.size(); p++) out[p] = y[e->yx_pos[p]];
  }
  return BPF_OK;
}

// upload voltages and evaluate the mismatch (mode 2: no verdicts)
static bpf_status stage_eval(bpf_engine *e, const double *V) {
  if (!e->solved) { e->err = "call bpf_solve first (Ybus is assembled there)"; return BPF_ERR_STATE; }
  bpf_status s = bus_rows_upload(e, (const cd *)V, e->d.V, e->T);
  if (s) return s;
  mismatch_pair(e, 2, e->opt.maxit_nr, 0, 0);
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipGetLastError());
  return BPF_OK;
}

// block-ordered vector [T][2nb] -> reference layout [P(pv); P(pq); Q(pq)],
// pv and pq buses ascending in the caller's numbering
static void to_ref_layout(const bpf_engine *e, const int8_t *ty,
                          const double *blk, double *out) {
  const Symbolic &S = e->S;
  std::vector<int> pv, pq;
  for (int x = 0; x < e->n; x++) {
    const int i = e->bus_int[x];
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
  B4<double> *LU, *TLd;
  double *xg;
  if ((s = fp64_factors(e, &LU, &TLd, &xg))) return s;
  // mark instances active for the fill
  std::vector<InstState> st(e->T);
  HIPCHK(hipMemcpy(st.data(), e->d.st, sizeof(InstState) * e->T, hipMemcpyDeviceToHost));
  std::vector<InstState> st2 = st;
  for (auto &x : st2) x.done = 0;
  HIPCHK(hipMemcpy(e->d.st, st2.data(), sizeof(InstState) * e->T, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(LU, 0, sizeof(B4<double>) * (size_t)e->T * e->g.nnzLU));
  hipLaunchKernelGGL(k_jacobian<double>, xgrid(e->T, ((size_t)e->g.nJb + 1) / 2, 256),
                     dim3(256), 0, e->stream, e->g, e->d, LU, 0, -1);
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
  // (indices and entries in the caller's bus order)
  std::vector<int> idx(e->n, -1);
  int npv = 0, npq = 0;
  for (int x = 0; x < e->n; x++)
    if (ty[e->bus_int[x]] == BPF_PV) idx[e->bus_int[x]] = npv++;
  for (int x = 0; x < e->n; x++) {
    const int i = e->bus_int[x];
    if (ty[i] == BPF_PQ && S.blk_of_bus[i] >= 0) idx[i] = npv + npq++;
  }
  int cnt = 0;
  for (int px = 0; px < e->g.nnzY; px++) {
    const int p = e->yx_pos[px];
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
  B4<double> *LU, *TLd;
  double *xg;
  bpf_status s = fp64_factors(e, &LU, &TLd, &xg);
  if (s) return s;
  const size_t T = e->T, n = e->n, m = e->m, nb2 = 2 * (size_t)e->S.nb;
  auto &a = e->adj;
  if (!a.G) {
    void *p;
    double **dst[] = {&a.rate, &a.G, &a.lam, &a.gP, &a.gQ, &a.gx, &a.obj};
    const size_t cnt[] = {std::max<size_t>(m, 1), T * nb2, T * nb2, T * n, T * n,
                          T * std::max<size_t>(m, 1), T};
    for (int k = 0; k < 7; k++) {
      if ((s = dalloc(e, e->inst_owned, sizeof(double) * cnt[k], &p))) {
        a = bpf_engine::AdjBufs();
        return s;
      }
      *dst[k] = (double *)p;
    }
  }
  if (rate && m)
    HIPCHK(hipMemcpyAsync(a.rate, rate, sizeof(double) * m, hipMemcpyHostToDevice,
                          e->stream));
  const double *d_rate = rate ? a.rate : nullptr;
  // only converged instances take part
  std::vector<InstState> st(T), st2;
  HIPCHK(hipMemcpy(st.data(), e->d.st, sizeof(InstState) * T, hipMemcpyDeviceToHost));
  st2 = st;
  for (auto &x : st2) x.done = x.verdict == BPF_V_CONVERGED ? 0 : 1;
  HIPCHK(hipMemcpy(e->d.st, st2.data(), sizeof(InstState) * T, hipMemcpyHostToDevice));
  const unsigned TT = (unsigned)T;
  tbegin(e, "adjoint");
  hipLaunchKernelGGL(k_jacobian<double>, xgrid(TT, ((size_t)e->g.nJb + 1) / 2, 256),
                     dim3(256), 0, e->stream, e->g, e->d, LU, 0, -1);
  hipLaunchKernelGGL(k_lu_factor<double>, dim3(TT), dim3(e->lu_threads), 0,
                     e->stream, e->g, e->d.st, 0, e->fplan, LU, TLd, 0, e->d.jsing, 0, nullptr, 0, -1);
  hipLaunchKernelGGL(k_adj_rhs, dim3(div_up(e->S.nb, 256), TT), dim3(256), 0,
                     e->stream, e->g, e->d, d_rate, w, a.G);
  size_t xs = sizeof(double) * nb2;
  int use_lds = xs <= (size_t)SOLVE_LDS_BYTES;
  // the slot form of the sweeps, transposed, where the Newton solve would
  // use it (512 lanes: in FP64 with the pivots in the forward chunks, 1024
  // lanes would spill registers); else one lane group per row by levels
  const int th = e->solve_threads;
  if (e->S.sl_ok_t &&
      sweep_mode(e->S, sizeof(double), use_lds, e->sweep_form, 512) == 2) {
    hipLaunchKernelGGL(k_adj_solve_slot<SLOT_CHUNK / 512>, dim3(TT), dim3(512), 0,
                       e->stream, e->g, e->d, (const B4<double> *)LU,
                       (const B4<double> *)TLd, (const double *)a.G, a.lam,
                       e->splan);
  } else {
    hipLaunchKernelGGL(k_adj_solve, dim3(TT), dim3(th),
                       0, e->stream, e->g, e->d, (const B4<double> *)LU,
                       (const B4<double> *)TLd, (const double *)a.G, a.lam,
                       e->splan, use_lds, xg);
  }
  hipLaunchKernelGGL(k_adj_grad, dim3(div_up(std::max(n, m), 256), TT),
                     dim3(256), 0, e->stream, e->g, e->d,
                     (const double *)a.lam, d_rate, w, a.gP, a.gQ, a.gx);
  hipLaunchKernelGGL(k_objective, dim3(TT), dim3(256), 0, e->stream, e->g,
                     e->d, d_rate, w, a.obj);
  tend(e);
  hipError_t r1 = hipStreamSynchronize(e->stream);
  hipError_t r2 = hipGetLastError();
  tcollect(e);
  // factorization singularities are not expected at a converged point; keep
  // the verdicts of the solve either way
  HIPCHK(hipMemcpy(e->d.st, st.data(), sizeof(InstState) * T, hipMemcpyHostToDevice));
  if (r1 != hipSuccess || r2 != hipSuccess) {
    e->err = std::string("adjoint: ") + hipGetErrorString(r1 != hipSuccess ? r1 : r2);
    return BPF_ERR_HIP;
  }
  if (gP && (s = bus_rows_download(e, a.gP, gP, T))) return s;
  if (gQ && (s = bus_rows_download(e, a.gQ, gQ, T))) return s;
  if (gx) HIPCHK(hipMemcpy(gx, a.gx, sizeof(double) * T * m, hipMemcpyDeviceToHost));
  if (obj) HIPCHK(hipMemcpy(obj, a.obj, sizeof(double) * T, hipMemcpyDeviceToHost));
  return BPF_OK;
}

// bus order of the reference's fast-decoupled matrices (makeB.h:311-501):
// B' over pv then pq, B'' over pq, buses ascending within each group
static void fd_index(const bpf_engine *e, const int8_t *ty, int which,
                     std::vector<int> &idx) {
  idx.assign(e->n, -1);
  int k = 0;
  if (which == 1)
    for (int x = 0; x < e->n; x++)
      if (ty[e->bus_int[x]] == BPF_PV) idx[e->bus_int[x]] = k++;
  for (int x = 0; x < e->n; x++) {
    const int i = e->bus_int[x];
    if (ty[i] == BPF_PQ && e->S.blk_of_bus[i] >= 0) idx[i] = k++;
  }
}

// instances' bus types and state as a solve starts them (k_init), then
// the caller's voltages
static bpf_status stage_start(bpf_engine *e, const double *V) {
  hipLaunchKernelGGL(k_init, xgrid(e->T, e->n, 256), dim3(256), 0, e->stream,
                     e->g, e->d, (const int8_t *)e->d_type0,
                     (const double *)e->d_Qinj0);
  HIPCHK(hipGetLastError());
  if (V) {
    bpf_status s = bus_rows_upload(e, (const cd *)V, e->d.V, e->T);
    if (s) return s;
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  e->solved = false;   // the state of the last solve is overwritten
  return BPF_OK;
}

extern "C" bpf_status bpf_stage_bmat(bpf_engine *e, int32_t t, int32_t which,
                                     int32_t cap, int32_t *row, int32_t *col,
                                     double *val, int32_t *nnz) {
  if (!e || t < 0 || t >= e->T || (which != 1 && which != 2) || !nnz)
    return BPF_ERR_INVALID_ARG;
  if (e->opt.alg == BPF_NR) {
    e->err = "B' and B'' belong to the fast-decoupled variants (BPF_FDXB / BPF_FDBX)";
    return BPF_ERR_UNSUPPORTED;
  }
  B4<double> *LU, *TLd;
  double *xg;
  bpf_status s = fp64_factors(e, &LU, &TLd, &xg);
  if (s) return s;
  if ((s = stage_start(e, nullptr))) return s;
  HIPCHK(hipMemsetAsync(LU, 0, sizeof(B4<double>) * (size_t)e->T * e->g.nnzLU,
                        e->stream));
  hipLaunchKernelGGL(k_bmat, dim3(div_up(e->g.nnzY, 256), e->T), dim3(256), 0,
                     e->stream, e->g, e->d, e->opt.alg, LU);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  std::vector<B4<double>> blk(e->g.nnzLU);
  HIPCHK(hipMemcpy(blk.data(), LU + (size_t)t * e->g.nnzLU,
                   sizeof(B4<double>) * e->g.nnzLU, hipMemcpyDeviceToHost));
  std::vector<int8_t> ty(e->n);
  HIPCHK(hipMemcpy(ty.data(), e->d.type + (size_t)t * e->n, e->n, hipMemcpyDeviceToHost));
  std::vector<int> idx;
  fd_index(e, ty.data(), which, idx);
  const Symbolic &S = e->S;
  int cnt = 0;
  for (int px = 0; px < e->g.nnzY; px++) {
    const int p = e->yx_pos[px];
    const int q = S.y2lu[p];
    if (q < 0) continue;
    const int i = idx[S.yrowof[p]], j = idx[S.ycol[p]];
    if (i < 0 || j < 0) continue;
    if (cnt < cap && row && col && val) {
      row[cnt] = i;
      col[cnt] = j;
      val[cnt] = which == 1 ? blk[q].a : blk[q].d;
    }
    cnt++;
  }
  *nnz = cnt;
  return cnt <= cap ? BPF_OK : BPF_ERR_INVALID_ARG;
}

extern "C" bpf_status bpf_stage_fd_halfstep(bpf_engine *e, const double *V,
                                            int32_t which, double *Vout,
                                            double *norm) {
  if (!e || !V || !Vout || (which != 1 && which != 2)) return BPF_ERR_INVALID_ARG;
  if (e->opt.alg == BPF_NR) {
    e->err = "the half-steps belong to the fast-decoupled variants (BPF_FDXB / BPF_FDBX)";
    return BPF_ERR_UNSUPPORTED;
  }
  if (!e->ybus_ready) {
    e->err = "no Ybus yet: call bpf_make_ybus or bpf_solve";
    return BPF_ERR_STATE;
  }
  B4<double> *LU, *TLd;
  double *xg;
  bpf_status s = fp64_factors(e, &LU, &TLd, &xg);
  if (s) return s;
  if ((s = stage_start(e, V))) return s;
  const unsigned T = (unsigned)e->T;
  // B' and B'' factored, the |V|-normalised mismatch, one masked solve and
  // the angle (P) or magnitude (Q) update: the body of fdpf.h:480-560
  hipLaunchKernelGGL(k_bmat, dim3(div_up(e->g.nnzY, 256), T), dim3(256), 0,
                     e->stream, e->g, e->d, e->opt.alg, LU);
  hipLaunchKernelGGL(k_lu_factor<double>, dim3(T), dim3(e->lu_threads), 0,
                     e->stream, e->g, e->d.st, 0, e->fplan, LU, TLd, 0, e->d.jsing, 0, nullptr, 0, -1);
  mismatch_pair(e, 2, e->opt.maxit_fd, 0, 1);
  const size_t xs = sizeof(double) * 2 * (size_t)e->S.nb;
  const int use_lds = xs <= (size_t)SOLVE_LDS_BYTES;
  launch_solve<double>(e, T, LU, which, 0, use_lds,
                       sweep_mode(e->S, sizeof(double), use_lds, e->sweep_form,
                                  e->solve_threads),
                       TLd, xg);
  hipLaunchKernelGGL(k_fd_update, xgrid(T, e->n, 256), dim3(256), 0, e->stream,
                     e->g, e->d, which);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  if ((s = bus_rows_download(e, e->d.V, (cd *)Vout, T))) return s;
  if (norm) {
    std::vector<InstState> st(T);
    HIPCHK(hipMemcpy(st.data(), e->d.st, sizeof(InstState) * T, hipMemcpyDeviceToHost));
    for (unsigned t = 0; t < T; t++) norm[t] = st[t].norm;
  }
  return BPF_OK;
}

extern "C" bpf_status bpf_stage_tail_inverse(bpf_engine *e, int32_t t,
                                             int32_t *tail_bus, double *Sinv) {
  if (!e || t < 0 || t >= e->T) return BPF_ERR_INVALID_ARG;
  const int s = e->S.tail_n;
  if (tail_bus)
    for (int a = 0; a < s; a++)
      tail_bus[a] = e->bus_ext[e->S.bus_of_blk[e->S.tail_piv[a]]];
  if (!Sinv || s == 0) return BPF_OK;
  if (!e->d_TL || e->lu_fp64 < 0) {
    e->err = "no factorization yet: call bpf_stage_newton_step or bpf_solve";
    return BPF_ERR_STATE;
  }
  const size_t ss = (size_t)s * s, w = 2 * (size_t)s;
  auto put = [&](size_t k, double a, double b, double c, double d) {
    const size_t r = 2 * (k / s), col = 2 * (k % s);
    Sinv[r * w + col] = a;
    Sinv[r * w + col + 1] = b;
    Sinv[(r + 1) * w + col] = c;
    Sinv[(r + 1) * w + col + 1] = d;
  };
  if (e->lu_fp64) {
    std::vector<B4<double>> v(ss);
    HIPCHK(hipMemcpy(v.data(), (const B4<double> *)e->d_TL + (size_t)t * ss,
                     sizeof(B4<double>) * ss, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ss; k++) put(k, v[k].a, v[k].b, v[k].c, v[k].d);
  } else {
    std::vector<B4<float>> v(ss);
    HIPCHK(hipMemcpy(v.data(), (const B4<float> *)e->d_TL + (size_t)t * ss,
                     sizeof(B4<float>) * ss, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ss; k++) put(k, v[k].a, v[k].b, v[k].c, v[k].d);
  }
  return BPF_OK;
}
