// HIP kernels of the batched AC power-flow engine (gfx950, wave64).
//
// Launch shapes:
//   * element kernels (Ybus assembly, V0, mismatch rows, Jacobian fill,
//     refinement residual, voltage update): one thread per element, several
//     workgroups per instance, all of an instance's workgroups on one XCD
//     (xcd_slot);
//   * per-instance kernels (verdicts, LU factorization, triangular solves,
//     losses): one workgroup per instance; the steps of the factorization and
//     of the sweeps are separated by workgroup barriers only, so no grid-wide
//     synchronisation and no host round trip inside a stage.
#include <float.h>

#include "kernels.hpp"
#include "symbolic.hpp"

namespace bpf {

// ---------------------------------------------------------------------------
// complex FP64 helpers (division: Smith's 1962 scaled algorithm)
// ---------------------------------------------------------------------------
__device__ __forceinline__ cd mk(double x, double y) { cd z; z.x = x; z.y = y; return z; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cmul(cd a, cd b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ cd cconj(cd a) { return mk(a.x, -a.y); }
__device__ __forceinline__ double cabs_(cd a) { return hypot(a.x, a.y); }
__device__ __forceinline__ cd cdiv(cd n, cd d) {
  if (fabs(d.x) >= fabs(d.y)) {
    double r = d.y / d.x, den = d.x + d.y * r;
    return mk((n.x + n.y * r) / den, (n.y - n.x * r) / den);
  }
  double r = d.x / d.y, den = d.x * r + d.y;
  return mk((n.x * r + n.y) / den, (n.y * r - n.x) / den);
}
__device__ __forceinline__ cd cexpj(double th) {
  double s, c;
  sincos(th, &s, &c);
  return mk(c, s);
}

// the instance-independent parts of every branch's pi-model: its series
// admittance 1 / (R + jX) and its phase-shift factor e^(j shift), computed
// once per case with the operations branch_entry would use per instance
__global__ void k_branch_pre(DevTopo g, cd *ys, cd *ejs) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.m) return;
  ys[b] = cdiv(mk(1.0, 0.0), mk(g.R[b], g.X[b]));
  ejs[b] = cexpj(M_PI / 180.0 * g.shift[b]);
}

// pi-model entry of branch b for instance arrays (makeYbus.h:63-80)
__device__ cd branch_entry(const DevTopo &g, const DevInst &d, int t, int b,
                           int kind) {
  const bool on = d.br_on[(size_t)t * g.m + b] != 0;
  const double stat = on ? 1.0 : 0.0;
  // stat / (R + jX): the precomputed quotient for an in-service branch
  const cd Ys = on ? g.ys[b] : cdiv(mk(0.0, 0.0), mk(g.R[b], g.X[b]));
  double tp = d.tap[(size_t)t * g.m + b];
  cd tap = (tp != 0.0) ? mk(tp, 0.0) : mk(1.0, 0.0);
  tap = cmul(tap, g.ejs[b]);
  switch (kind) {
    case KFF: {
      cd Bc = cmul(mk(stat, 0.0), mk(0.0, g.B[b]));
      cd Ytt = cadd(Ys, cdiv(Bc, mk(2.0, 0.0)));
      return cdiv(Ytt, cmul(tap, cconj(tap)));
    }
    case KFT: return cdiv(cmul(Ys, mk(-1.0, 0.0)), cconj(tap));
    case KTF: return cdiv(cmul(Ys, mk(-1.0, 0.0)), tap);
    default: {
      cd Bc = cmul(mk(stat, 0.0), mk(0.0, g.B[b]));
      return cadd(Ys, cdiv(Bc, mk(2.0, 0.0)));
    }
  }
}

// ---------------------------------------------------------------------------
// Placement of kernels that give one instance several workgroups: every
// workgroup of an instance runs on the same XCD, so the instance's vectors
// (V, Ibus, ...) are fetched into one L2 only.  Workgroups are dispatched to
// the 8 XCDs round-robin (w mod 8), so instance t takes the workgroups
// w = 8 * (j * parts + part) + t mod 8 with j = t / 8; the 1-D grid holds
// 8 * ceil(T / 8) * parts workgroups and the surplus ones return.  Only the
// placement depends on the dispatch order, never the result.
// ---------------------------------------------------------------------------
struct XSlot {
  int t, part;
};
__device__ __forceinline__ XSlot xcd_slot(int parts) {
  const int w = blockIdx.x, s = w >> 3;
  XSlot r;
  r.t = (w & 7) + 8 * (s / parts);
  r.part = s - (s / parts) * parts;
  return r;
}
__device__ __forceinline__ int parts_of(int count) {
  return (count + (int)blockDim.x - 1) / (int)blockDim.x;
}

// ---------------------------------------------------------------------------
// Ybus assembly with outage masks: one thread per (position, instance).
// Sum order = the reference's: from-side terms, then to-side terms, then the
// bus shunt (Cf*Yf + Ct*Yt + Ysh, makeYbus.h:196-266).
// ---------------------------------------------------------------------------
__global__ void k_ybus(DevTopo g, DevInst d) {
  const XSlot sl = xcd_slot(parts_of(g.nnzY));
  const int t = sl.t, p = sl.part * blockDim.x + threadIdx.x;
  if (t >= d.T || p >= g.nnzY) return;
  cd af = mk(0.0, 0.0), at = mk(0.0, 0.0);
  for (int q = g.yf_ptr[p]; q < g.yf_ptr[p + 1]; q++) {
    int it = g.yf_item[q];
    af = cadd(af, branch_entry(g, d, t, it >> 2, it & 3));
  }
  for (int q = g.yt_ptr[p]; q < g.yt_ptr[p + 1]; q++) {
    int it = g.yt_item[q];
    at = cadd(at, branch_entry(g, d, t, it >> 2, it & 3));
  }
  cd y = cadd(af, at);
  int i = g.ydiag_bus[p];
  if (i >= 0) y = cadd(y, mk(g.Gs[i], d.Bs[(size_t)t * g.n + i]));
  d.Y[(size_t)t * g.nnzY + p] = y;
}

// ---------------------------------------------------------------------------
// Initial voltage (runpf.h:211-226) and per-solve state reset.
// ---------------------------------------------------------------------------
__global__ void k_init(DevTopo g, DevInst d, const int8_t *type0,
                       const double *Qinj0) {
  const XSlot sl = xcd_slot(parts_of(g.n));
  const int t = sl.t, i = sl.part * blockDim.x + threadIdx.x;
  if (t >= d.T) return;
  if (i == 0 && t == 0) *d.active_count = 0;
  if (i == 0) {
    InstState s;
    s.done = 0; s.verdict = 1; s.it_round = 0; s.it_total = 0; s.rounds = 0;
    s.fd_half = 0; s.norm = 0.0;
    d.st[t] = s;
  }
  if (i >= g.n) return;
  size_t o = (size_t)t * g.n + i;
  double v = d.vset[o];
  cd V = cmul(mk(v, 0.0), cexpj(g.Va0[i]));
  int bt = g.base_type[i];
  if (bt == 2 || bt == 3) V = cmul(mk(v / cabs_(V), 0.0), V);
  d.V[o] = V;
  d.type[o] = type0[o];
  d.Qinj[o] = Qinj0[o];
}

// ---------------------------------------------------------------------------
// Per-instance inputs from the outage masks (the reference uploads one
// particle matrix, runpf.h:611-616; here the masks and controls are uploaded
// and every per-instance array is built on the device).  Thread (t, i) builds
// bus i and branch i of instance t: the bus adds the in-service generators
// of its list in ascending generator order (the last one sets the voltage
// set-point, loadMatpowerFormat.h:126-128), a PV bus without one becomes PQ.
// ---------------------------------------------------------------------------
struct CaseArrays {
  const double *Pd, *Qd, *Bs, *Vm, *tap;
  const uint8_t *br_status;
  const int *gb_ptr, *gb_gen;          // generators of each bus, ascending
  const double *Pg, *Qg, *Qmax, *Qmin, *VG;
  const uint8_t *gen_status;
  int ng;
};

__global__ void k_expand(DevTopo g, DevInst d, CaseArrays c, const uint8_t *br_on,
                         const uint8_t *gen_on, int8_t *type0, double *Qinj0,
                         double *Pinj, uint8_t *on, double *tapT, double *BsT,
                         double *vset, double *Qmx, double *Qmn) {
  const int t = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.T) return;
  if (i < g.n) {
    const size_t o = (size_t)t * g.n + i;
    const int bt = g.base_type[i];
    double P = -c.Pd[i], Q = -c.Qd[i], qmx = 0.0, qmn = 0.0, vg = 0.0;
    int ngon = 0;
    for (int k = c.gb_ptr[i]; k < c.gb_ptr[i + 1]; k++) {
      const int gk = c.gb_gen[k];
      if (!c.gen_status[gk] || (gen_on && !gen_on[(size_t)t * c.ng + gk])) continue;
      ngon++;
      P += c.Pg[gk];
      Q += c.Qg[gk];
      qmx += c.Qmax[gk];
      qmn += c.Qmin[gk];
      vg = c.VG[gk];
    }
    type0[o] = (int8_t)((bt == 2 && ngon == 0) ? 1 : bt);
    Pinj[o] = P;
    Qinj0[o] = Q;
    Qmx[o] = qmx;
    Qmn[o] = qmn;
    BsT[o] = c.Bs[i];
    vset[o] = (bt == 2 && ngon > 0 && !isnan(vg)) ? vg : c.Vm[i];
  }
  if (i < g.m) {
    const size_t o = (size_t)t * g.m + i;
    on[o] = (c.br_status[i] && (!br_on || br_on[o])) ? 1 : 0;
    tapT[o] = c.tap[i];
  }
}

// controls of the particle vector (Particula.h:26-60), applied in their
// order per instance (a later control of the same element wins); NaN keeps
// the default; an OLTC control applies only to a branch with a tap
__global__ void k_expand_ctl(DevTopo g, DevInst d, const double *tap_case,
                             int nctl, const int2 *ctl, const double *ctl_val,
                             double *tapT, double *BsT, double *vset) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.T) return;
  for (int c = 0; c < nctl; c++) {
    const double v = ctl_val[(size_t)t * nctl + c];
    const int kind = ctl[c].x, id = ctl[c].y;
    if (isnan(v)) continue;
    if (kind == 1) vset[(size_t)t * g.n + id] = v;
    else if (kind == 3) BsT[(size_t)t * g.n + id] = v;
    else if (tap_case[id] != 0.0) tapT[(size_t)t * g.m + id] = v;
  }
}

// ---------------------------------------------------------------------------
// workgroup reductions (wave64 shuffles + LDS)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_or(int v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_add(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// all threads get the result
__device__ double block_max(double v, double *sh) {
  int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = (blockDim.x + 63) >> 6;
  v = wave_max(v);
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int k = 0; k < nw; k++) r = fmax(r, sh[k]);
  return r;
}
__device__ double block_sum(double v, double *sh) {
  int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int k = 0; k < nw; k++) r += sh[k];
  return r;
}
__device__ int block_or(int v, int *sh) {
  int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = (blockDim.x + 63) >> 6;
  v = wave_or(v);
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  int r = 0;
  for (int k = 0; k < nw; k++) r |= sh[k];
  return r;
}
__device__ int block_add(int v, int *sh) {
  int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = (blockDim.x + 63) >> 6;
  v = wave_add(v);
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  int r = 0;
  for (int k = 0; k < nw; k++) r += sh[k];
  return r;
}

// ---------------------------------------------------------------------------
// Mismatch S = V o conj(Ybus V) - Sbus, inf-norm, verdicts and reactive-
// limit switching (newtonpf.h:18-95,394-396,518-520), as two launches:
// k_mismatch_mb (one thread per block row, many workgroups per instance)
// writes Ibus, F and one partial norm per workgroup; k_verdict (one
// workgroup per instance) reduces the partials and decides.  The kernel
// boundary orders the two: no device-scope fences, which on gfx950 write
// back and invalidate the per-XCD L2.
// F is kept in block order: F[2k] = P, F[2k+1] = Q (0 for PV).
// mode: 0 initial check, 1 after a Newton step, 2 evaluation only,
// 3 fast decoupled after the P half-step (counts the iteration; converged or
// go on), 4 fast decoupled after the Q half-step (converged, iteration cap or
// go on) -- the loop of fdpf.h:505-590.  fd_norm divides by |V| (fdpf.h:401).
// ---------------------------------------------------------------------------
// one bus: Ibus_i = sum_j Y_ij V_j, and for a non-slack bus its P (and, PQ,
// Q) mismatch at its block position
__device__ void mis_bus(const DevTopo &g, const DevInst &d, int t, int i,
                        int fd_norm, double qtol, int qlim, double &lmax,
                        int &lnan, int &lviol) {
  const cd *Y = d.Y + (size_t)t * g.nnzY;
  const cd *V = d.V + (size_t)t * g.n;
  cd I = mk(0.0, 0.0);
  // the row in groups of RU entries: all index and admittance loads, then
  // all voltage loads, then the sum in row order (a slot past the row end
  // repeats the last entry and leaves the sum unchanged)
  constexpr int RU = 4;
  const int pa = g.yrow[i], pe = g.yrow[i + 1];
  for (int p0 = pa; p0 < pe; p0 += RU) {
    int col[RU];
    cd y[RU], v[RU];
#pragma unroll
    for (int u = 0; u < RU; u++) {
      const int p = min(p0 + u, pe - 1);
      col[u] = g.ycol[p];
      y[u] = Y[p];
    }
#pragma unroll
    for (int u = 0; u < RU; u++) v[u] = V[col[u]];
#pragma unroll
    for (int u = 0; u < RU; u++) {
      const cd s = cadd(I, cmul(y[u], v[u]));
      I.x = p0 + u < pe ? s.x : I.x;
      I.y = p0 + u < pe ? s.y : I.y;
    }
  }
  size_t o = (size_t)t * g.n + i;
  d.Ib[o] = I;
  const int k = g.blk_of_bus[i];
  if (k < 0) return;
  const cd Sb = cmul(V[i], cconj(I));
  if (qlim && d.type[o] == 2) {
    // reactive limit of a PV bus (used by the verdict once converged)
    const double qg = Sb.y + g.Qd[i];
    lviol += (qg > d.Qmax[o] + qtol || qg < d.Qmin[o] - qtol) ? 1 : 0;
  }
  cd mis = csub(Sb, mk(d.Pinj[o], d.Qinj[o]));
  if (fd_norm) mis = cdiv(mis, mk(cabs_(V[i]), 0.0));
  bool pq = d.type[o] == 1;
  // the block's (P, Q) pair in one 16-byte store
  double2 *F2 = (double2 *)(d.F + (size_t)t * 2 * g.nb);
  F2[k] = make_double2(mis.x, pq ? mis.y : 0.0);
  lmax = fmax(lmax, fabs(mis.x));
  lnan |= isnan(mis.x);
  if (pq) { lmax = fmax(lmax, fabs(mis.y)); lnan |= isnan(mis.y); }
}

__global__ void k_mismatch_mb(DevTopo g, DevInst d, int mode, int fd_norm,
                              double qtol, int qlim) {
  __shared__ double shd[16];
  __shared__ int shi[16];
  const int parts = parts_of(g.n);
  const XSlot sl = xcd_slot(parts);
  const int t = sl.t;
  if (t >= d.T) return;
  if (mode != 2 && d.st[t].done) return;
  double lmax = 0.0;
  int lnan = 0, lviol = 0;
  const int i = sl.part * blockDim.x + threadIdx.x;
  if (i < g.n) mis_bus(g, d, t, i, fd_norm, qtol, qlim, lmax, lnan, lviol);
  double bmax = block_max(lmax, shd);
  int bnan = block_or(lnan, shi);
  int bviol = qlim ? block_add(lviol, shi) : 0;
  if (threadIdx.x == 0) {
    d.part[(size_t)t * parts + sl.part] = bnan ? (double)NAN : bmax;
    d.partv[(size_t)t * parts + sl.part] = bviol;
  }
}

__global__ void k_verdict(DevTopo g, DevInst d, int nparts, int mode, int maxit,
                          double tol, int enforce_qlim, int max_rounds,
                          double qtol, int fd_norm) {
  __shared__ double shd[16];
  __shared__ int shi[16];
  const int t = blockIdx.x;
  InstState *sp = d.st + t;
  if (mode != 2 && sp->done) return;
  double pm = 0.0;
  int pn = 0, pv = 0;
  for (int q = threadIdx.x; q < nparts; q += blockDim.x) {
    double v = d.part[(size_t)t * nparts + q];
    pn |= isnan(v);
    pm = fmax(pm, v);
    pv += d.partv[(size_t)t * nparts + q];
  }
  double nrm = block_max(pm, shd);
  int anynan = block_or(pn, shi);
  // PV buses beyond their reactive limits at this V (counted per bus by the
  // mismatch kernel); after a switch the remaining ones are recounted here
  int nviol_at_v = enforce_qlim ? block_add(pv, shi) : 0;
  InstState s = *sp;
  if (mode == 1 || mode == 3) { s.it_round++; s.it_total++; }
  int8_t *ty = d.type + (size_t)t * g.n;
  double *Qinj = d.Qinj + (size_t)t * g.n;
  double *F = d.F + (size_t)t * 2 * g.nb;
  const cd *V = d.V + (size_t)t * g.n;
  const cd *Ib = d.Ib + (size_t)t * g.n;
  bool active = false;
  for (int guard = 0; guard <= max_rounds + 1; guard++) {
    s.norm = anynan ? NAN : nrm;
    if (mode == 2) break;
    if (anynan || isinf(nrm)) { s.verdict = BPF_V_NONFINITE; s.done = 1; break; }
    if (nrm < tol) {
      if (!enforce_qlim) { s.verdict = BPF_V_CONVERGED; s.done = 1; break; }
      int nv = nviol_at_v;
      if (nv < 0) {
        int lv = 0;
        for (int i = threadIdx.x; i < g.n; i += blockDim.x) {
          if (ty[i] != 2) continue;
          double qg = cmul(V[i], cconj(Ib[i])).y + g.Qd[i];
          size_t o = (size_t)t * g.n + i;
          if (qg > d.Qmax[o] + qtol || qg < d.Qmin[o] - qtol) lv++;
        }
        nv = block_add(lv, shi);
      }
      if (nv == 0) { s.verdict = BPF_V_CONVERGED; s.done = 1; break; }
      if (s.rounds >= max_rounds) { s.verdict = BPF_V_MAX_ITER; s.done = 1; break; }
      s.rounds++;
      s.it_round = 0;
      // switch every violating bus to PQ with its Q at the limit; only those
      // buses' Q rows change (they were identity rows with F = 0)
      double lm = 0.0;
      int ln = 0;
      for (int i = threadIdx.x; i < g.n; i += blockDim.x) {
        if (ty[i] != 2) continue;
        const int b = g.blk_of_bus[i];
        cd S = cmul(V[i], cconj(Ib[i]));
        double qg = S.y + g.Qd[i];
        size_t o = (size_t)t * g.n + i;
        double q;
        if (qg > d.Qmax[o] + qtol) q = d.Qmax[o] - g.Qd[i];
        else if (qg < d.Qmin[o] - qtol) q = d.Qmin[o] - g.Qd[i];
        else continue;
        ty[i] = 1;
        Qinj[i] = q;
        double fq = S.y - q;
        if (fd_norm) fq = fq / cabs_(V[i]);
        F[2 * b + 1] = fq;
        lm = fmax(lm, fabs(fq));
        ln |= isnan(fq);
      }
      nrm = fmax(nrm, block_max(lm, shd));
      anynan |= block_or(ln, shi);
      nviol_at_v = -1;  // bus types changed: count again
      continue;  // decide again with the new bus types
    }
    if (mode != 3 && s.it_round >= maxit) { s.verdict = BPF_V_MAX_ITER; s.done = 1; break; }
    active = true;
    break;
  }
  if (threadIdx.x == 0) {
    *sp = s;
    if (active) atomicAdd(d.active_count, 1);
  }
}

// ---------------------------------------------------------------------------
// 2x2 block helpers
// ---------------------------------------------------------------------------
// component-wise select: a select of whole blocks can become a select of
// their stack addresses (scratch memory)
template <typename T>
__device__ __forceinline__ B4<T> blk_sel(bool c, const B4<T> &x, const B4<T> &y) {
  B4<T> r;
  r.a = c ? x.a : y.a;
  r.b = c ? x.b : y.b;
  r.c = c ? x.c : y.c;
  r.d = c ? x.d : y.d;
  return r;
}
// Block products as explicit fused multiply-adds: left to contraction, an
// expression such as a*b + c*d may fuse either product, and each inlined
// copy chooses for itself, so the same product formed in two places (two
// schedules of one factorization) could round differently.
template <typename T>
__device__ __forceinline__ void blk_fms(B4<T> &acc, const B4<T> &l, const B4<T> &u) {
  acc.a = fma(-l.b, u.c, fma(-l.a, u.a, acc.a));
  acc.b = fma(-l.b, u.d, fma(-l.a, u.b, acc.b));
  acc.c = fma(-l.d, u.c, fma(-l.c, u.a, acc.c));
  acc.d = fma(-l.d, u.d, fma(-l.c, u.b, acc.d));
}
template <typename T>
__device__ __forceinline__ B4<T> blk_mul(const B4<T> &x, const B4<T> &y) {
  B4<T> r;
  r.a = fma(x.b, y.c, x.a * y.a);
  r.b = fma(x.b, y.d, x.a * y.b);
  r.c = fma(x.d, y.c, x.c * y.a);
  r.d = fma(x.d, y.d, x.c * y.b);
  return r;
}

// ---------------------------------------------------------------------------
// Batched Crout LU of the 2x2-block Jacobian, static pivot order from the
// symbolic analysis.  One workgroup per instance.  Level l of the
// elimination tree: phase A computes every output of the level's pivots as
// A(p) - sum L*U over its precomputed pair list (the pivot block is stored
// inverted); phase B scales the level's L blocks by the pivot inverse.
// ---------------------------------------------------------------------------
// schedules of the factorization and of the sweeps (device arrays from the
// symbolic analysis)
constexpr int TK_LEN_SHIFT = 16;       // pivot range: offset in the chunk | length << 16
struct FactorPlan {
  int nchunk;                          // main chunks (the levels below the tail)
  // every chunk, the main ones first, then those of the split schedules:
  // {first task, end task, first product, end product}, and the task count
  // and longest reduction that set its lane groups (a split chunk keeps
  // those of its main chunk, so each output is summed in the same order)
  const int4 *chunks;
  const int2 *chunk_lg;
  // per output {dst | flags, first product, end product, pivot range}; the
  // pivot range of an L output packed as its first product's offset from
  // the chunk's first product | length << TK_LEN_SHIFT
  const int4 *tk;
  const int2 *pairs;                   // {L, U} positions (U < 0: init entry)
  int tn, tnent;                       // dense tail: pivots, entries
  const int *tdense, *tpos, *tpp, *tpairL, *tpairU;
  int tstage_n;                        // staged tail blocks (0: no staging)
  const int *tstage, *tpairS;          // their positions; pair -> L | U << 16
  // the tail entries in jobs of one wave each, {first entry, end entry << 3
  // | log2 lane group size}: for the staged form (up to 12 products per
  // lane) and for the form that reads operands from memory (up to 4)
  int tjobs_s, tjobs_g;
  const int2 *tjob_s, *tjob_g;
  int tail_mfma;                       // FP32 tail inverse on matrix cores
  // the forward substitution's outputs run while at most this many
  // instances of the batch are active (-1: never); the solve then starts
  // from them (same limit in the solve plan)
  int fused_limit;
  // split factorization: variant v (2 << v workgroups per instance) leaves
  // the main chunks from sp_rest[v] to the top workgroup (-1: no such
  // variant); its group g runs chunks sp_grp[v][g] .. sp_grp[v][g + 1] - 1
  int sp_rest[SPLIT_VARIANTS];
  int sp_grp[SPLIT_VARIANTS][SPLIT_MAX + 1];
};

// the split variant for the active instances of this batch (part): the
// most workgroups per instance with active x workgroups within budget
// (the compute units the batch may fill), -1 for one workgroup; a forced
// choice (force >= -1) instead when the plan has it
__device__ __forceinline__ int split_variant(const FactorPlan &P, const int *active,
                                             int budget, int force) {
  if (force >= -1) return (force >= 0 && P.sp_rest[force] >= 0) ? force : -1;
  const int a = max(*active, 1);
  for (int v = SPLIT_VARIANTS - 1; v >= 0; v--)
    if (P.sp_rest[v] >= 0 && a * (2 << v) <= budget) return v;
  return -1;
}

// the forward substitution runs in the factorization while at most this
// many instances of a batch (part) are active (bpf_set_fused_forward)
constexpr int FUSED_LIMIT_DEFAULT = 32;

struct SolvePlan {
  int fused_limit;                     // as in the factorization plan
  int nfw, nbw;                        // levels over the non-tail rows
  const int *fw_ptr, *fw_rows, *fw_max, *bw_ptr, *bw_rows, *bw_max;
  int tn, tsmax;                       // tail pivots, longest tail-row list
  const int *tpiv, *ts_ptr, *ts_pos, *ts_col, *ts_mir;
  // entry-parallel chunks: forward ones first, then backward
  int sw_nfw, sw_nbw;
  const int4 *sw_chunk;                // {row0, row1, entry0, entry1}
  const int *sw_cmax;                  // longest row of the chunk
  const int4 *sw_row;                  // {block row, entry0, entry1, pivot position}
  const int2 *sw_ent;                  // {block position, column block}
  // slot form: one lane per entry, a power-of-two lane group per row
  int sl_nfw, sl_nbw;
  const int4 *sl_chunk;                // {slot0, slot1, log2 largest group, 0}
  const int4 *sl_slot;                 // {position, column + 1 | log2 group, row or -1, pivot}
  // the same for the transposed sweeps (adjoint)
  int sl_nfw_t, sl_nbw_t;
  const int4 *sl_chunk_t, *sl_slot_t;
};

// Lanes that share one item (a task's pair list, a row's entries): a power of
// two up to one wavefront, no more than the longest list of the level needs,
// and few enough that every item of the level gets its group in one pass
// when the workgroup is large enough.  Groups are aligned inside a wavefront,
// so their partial sums combine with cross-lane shuffles.
// s_waitcnt immediate (gfx9 encoding) that waits for every vector memory
// operation: vmcnt 0, expcnt and lgkmcnt left at their maxima
constexpr int WAIT_VMCNT0 = 0x0F70;

// whether any lane of the wave has an index below `end`, for an index that
// grows with the lane: the first lane decides, so the answer is one scalar
// for the wave and a load under it is skipped by a scalar branch
__device__ __forceinline__ bool wave_any(int index, int end) {
  return __builtin_amdgcn_readfirstlane(index) < end;
}

__device__ __forceinline__ int lane_group(int items, int maxlen) {
  int gsz = 1;
  while (gsz < 64 && gsz < maxlen && items * 2 * gsz <= (int)blockDim.x) gsz <<= 1;
  return gsz;
}

template <typename T>
__device__ __forceinline__ T group_sum(T v, int gsz) {
  for (int off = gsz >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int CTRL>
__device__ __forceinline__ float dpp_row(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_row(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// the value of lane i + 16 (lanes 0-15 and 32-47) and of lane i + 32 (lanes
// 0-31), by gfx950's lane swaps within one register (no LDS round trip)
__device__ __forceinline__ float lane_up16(float v) {
  const unsigned u = __float_as_uint(v);
  return __uint_as_float(__builtin_amdgcn_permlane16_swap(u, u, false, false)[1]);
}
__device__ __forceinline__ float lane_up32(float v) {
  const unsigned u = __float_as_uint(v);
  return __uint_as_float(__builtin_amdgcn_permlane32_swap(u, u, false, false)[1]);
}
__device__ __forceinline__ double lane_up16(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
  const unsigned l2 = __builtin_amdgcn_permlane16_swap(lo, lo, false, false)[1];
  const unsigned h2 = __builtin_amdgcn_permlane16_swap(hi, hi, false, false)[1];
  return __longlong_as_double(((long long)h2 << 32) | l2);
}
__device__ __forceinline__ double lane_up32(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
  const unsigned l2 = __builtin_amdgcn_permlane32_swap(lo, lo, false, false)[1];
  const unsigned h2 = __builtin_amdgcn_permlane32_swap(hi, hi, false, false)[1];
  return __longlong_as_double(((long long)h2 << 32) | l2);
}

// sum of each aligned group of gsz lanes (a power of two, the same for the
// whole wave) into the group's first lane: lane i adds lane i + 2^s for
// 2^s < gsz, by DPP row shifts within 16 lanes and cross-lane reads beyond.
// Every lane of a group must be active; only the first lane's sum is whole.
template <typename T>
__device__ __forceinline__ void group_sum4_first(B4<T> &v, int gsz) {
  if (gsz > 1) {
    v.a += dpp_row<0x101>(v.a); v.b += dpp_row<0x101>(v.b);
    v.c += dpp_row<0x101>(v.c); v.d += dpp_row<0x101>(v.d);
  }
  if (gsz > 2) {
    v.a += dpp_row<0x102>(v.a); v.b += dpp_row<0x102>(v.b);
    v.c += dpp_row<0x102>(v.c); v.d += dpp_row<0x102>(v.d);
  }
  if (gsz > 4) {
    v.a += dpp_row<0x104>(v.a); v.b += dpp_row<0x104>(v.b);
    v.c += dpp_row<0x104>(v.c); v.d += dpp_row<0x104>(v.d);
  }
  if (gsz > 8) {
    v.a += dpp_row<0x108>(v.a); v.b += dpp_row<0x108>(v.b);
    v.c += dpp_row<0x108>(v.c); v.d += dpp_row<0x108>(v.d);
  }
  if (gsz > 16) {
    v.a += lane_up16(v.a); v.b += lane_up16(v.b);
    v.c += lane_up16(v.c); v.d += lane_up16(v.d);
  }
  if (gsz > 32) {
    v.a += lane_up32(v.a); v.b += lane_up32(v.b);
    v.c += lane_up32(v.c); v.d += lane_up32(v.d);
  }
}

template <typename T>
__device__ __forceinline__ void blk_acc(B4<T> &acc, const B4<T> &x) {
  acc.a += x.a; acc.b += x.b; acc.c += x.c; acc.d += x.d;
}

// a lane's share of a product range in LDS (q0 + lane, step gsz): four
// independent partial sums, so that four LDS reads are in flight at a time
// (a single running sum waits for every read), combined pairwise at the end
template <typename T>
__device__ __forceinline__ B4<T> lane_range_sum(const B4<T> *buf, int q, int q1,
                                                int gsz) {
  B4<T> e0 = {0, 0, 0, 0}, e1 = e0, e2 = e0, e3 = e0;
  for (; q + 3 * gsz < q1; q += 4 * gsz) {
    blk_acc(e0, buf[q]);
    blk_acc(e1, buf[q + gsz]);
    blk_acc(e2, buf[q + 2 * gsz]);
    blk_acc(e3, buf[q + 3 * gsz]);
  }
  for (; q < q1; q += gsz) blk_acc(e0, buf[q]);
  blk_acc(e0, e1);
  blk_acc(e2, e3);
  blk_acc(e0, e2);
  return e0;
}

// the shares of two product ranges at once (an L output's own range and its
// pivot's): two partial sums per range, the reads of both ranges in flight
// together.  A lane past the end of a range reads the zero block z instead
// (adding zero leaves a sum as it is).
template <typename T>
__device__ __forceinline__ void lane_range_sum2(const B4<T> *buf, const B4<T> *z,
                                                int q, int q1, int r, int r1,
                                                int gsz, B4<T> &e, B4<T> &f) {
  B4<T> e0 = {0, 0, 0, 0}, e1 = e0, f0 = e0, f1 = e0;
  while (q < q1 || r < r1) {
    const B4<T> a0 = *(q < q1 ? buf + q : z), a1 = *(q + gsz < q1 ? buf + q + gsz : z);
    const B4<T> b0 = *(r < r1 ? buf + r : z), b1 = *(r + gsz < r1 ? buf + r + gsz : z);
    blk_acc(e0, a0); blk_acc(e1, a1);
    blk_acc(f0, b0); blk_acc(f1, b1);
    q += 2 * gsz;
    r += 2 * gsz;
  }
  blk_acc(e0, e1);
  blk_acc(f0, f1);
  e = e0;
  f = f0;
}

// group_sum4_first on two blocks at once (independent chains interleaved)
template <typename T>
__device__ __forceinline__ void group_sum4x2_first(B4<T> &v, B4<T> &w, int gsz) {
#define BPF_GS2(STEP)                                                     \
  v.a += STEP(v.a); v.b += STEP(v.b); v.c += STEP(v.c); v.d += STEP(v.d); \
  w.a += STEP(w.a); w.b += STEP(w.b); w.c += STEP(w.c); w.d += STEP(w.d);
  if (gsz > 1) { BPF_GS2(dpp_row<0x101>) }
  if (gsz > 2) { BPF_GS2(dpp_row<0x102>) }
  if (gsz > 4) { BPF_GS2(dpp_row<0x104>) }
  if (gsz > 8) { BPF_GS2(dpp_row<0x108>) }
  if (gsz > 16) { BPF_GS2(lane_up16) }
  if (gsz > 32) { BPF_GS2(lane_up32) }
#undef BPF_GS2
}

template <typename T>
__device__ __forceinline__ B4<T> blk_zero() {
  B4<T> z;
  z.a = z.b = z.c = z.d = (T)0;
  return z;
}

template <typename T>
__device__ __forceinline__ bool blk_inv(const B4<T> &a, B4<T> &inv) {
  T det = fma(a.a, a.d, -(a.b * a.c));
  bool bad = !(det != (T)0) || !isfinite((double)det);
  T r = (T)1 / det;
  inv.a = a.d * r; inv.b = -a.b * r; inv.c = -a.c * r; inv.d = a.a * r;
  return bad;
}

// inverse of a 2x2 matrix of 2x2 blocks [[A, B], [C, D]] through the Schur
// complement S = D - C A^-1 B (the two pivots of Gauss-Jordan in sequence):
// [[A^-1 + A^-1 B S^-1 C A^-1, -A^-1 B S^-1], [-S^-1 C A^-1, S^-1]].  True
// when A or S is singular.
template <typename T>
__device__ __forceinline__ bool pair_inv(const B4<T> (&m)[2][2], B4<T> (&r)[2][2]) {
  B4<T> ai, si;
  bool bad = blk_inv(m[0][0], ai);
  const B4<T> aib = blk_mul(ai, m[0][1]), cai = blk_mul(m[1][0], ai);
  B4<T> sc = m[1][1];
  blk_fms(sc, m[1][0], aib);
  bad |= blk_inv(sc, si);
  const B4<T> u = blk_mul(aib, si);      // A^-1 B S^-1
  const B4<T> w = blk_mul(si, cai);      // S^-1 C A^-1
  const B4<T> x = blk_mul(u, cai);
  r[0][0] = {ai.a + x.a, ai.b + x.b, ai.c + x.c, ai.d + x.d};
  r[0][1] = {-u.a, -u.b, -u.c, -u.d};
  r[1][0] = {-w.a, -w.b, -w.c, -w.d};
  r[1][1] = si;
  return bad;
}

// Dense-tail inverse on matrix cores (FP32, workgroups of 1024 lanes).  S,
// 2s scalars per side, is taken as one 128 x 128 matrix (identity beyond
// S, zero coupling); wave w holds the 32 x 32 tile (w / 4, w % 4) in the
// accumulators of v_mfma_f32_32x32x2f32 for the whole elimination
// (accumulator v of lane l: tile row 8 (v / 4) + 4 (l / 32) + v % 4, tile
// column l % 32; operand lane l: A row / B column l % 32, k = l / 32).
// Gauss-Jordan by pivot quads, the block pivots 2p and 2p + 1 (scalars
// 4p..4p+3), in the order of the block form.  A step updates every entry by
// one rank-4 product,
//   new = M' + A B,  A(i, k) = -col(i, k), or [i = 4p + k] on a pivot row,
//                    B(k, j) = (Q row)(k, j), or Q(k, j - 4p) in a pivot column,
//                    M' = M, or 0 on a pivot row or column,
// which leaves Q (the pivot block's inverse) on the pivot block, Q row on the
// pivot rows, -col Q in the pivot columns and M - col Q row elsewhere: two
// MFMAs per wave.  The next quad's columns and rows, and its inverse
// (pair_inv by lane 0 of the wave holding its pivot block, after reading the
// block from its lanes with v_readlane), pass through LDS double buffers:
// one barrier per step.  The columns are kept column-major and the rows
// row-major, so that a lane writes its four consecutive column entries, and
// reads or writes its column's four row entries, as one 16-byte access.  A
// singular pivot block sets sing.
constexpr int TAIL_NS = 2 * TAIL_MAX;   // scalars per side of the dense tail
struct alignas(16) TailQuads {
  float col[2][4][TAIL_NS], row[2][TAIL_NS][4], q[2][4][4];
};

__device__ __forceinline__ void tail_inverse_mfma(const B4<float> *D, int s, int nt,
                                                  TailQuads &mq, B4<float> *Ti,
                                                  int &sing) {
  typedef float f32x16 __attribute__((ext_vector_type(16)));
  const int n2 = 2 * s;
  auto mget = [&](int i, int j) -> float {
    if (i < n2 && j < n2)
      return (&D[(i >> 1) * s + (j >> 1)].a)[2 * (i & 1) + (j & 1)];
    return i == j ? 1.f : 0.f;
  };
  auto put_q = [&](int buf, const B4<float> (&r)[2][2]) {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        mq.q[buf][2 * a][2 * b] = r[a][b].a;
        mq.q[buf][2 * a][2 * b + 1] = r[a][b].b;
        mq.q[buf][2 * a + 1][2 * b] = r[a][b].c;
        mq.q[buf][2 * a + 1][2 * b + 1] = r[a][b].d;
      }
  };
  // the wave's tile in scalar registers: every choice below that depends
  // only on the tile and the step is a scalar branch
  const int lane = (int)threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int rb = (w >> 2) * 32, cb = (w & 3) * 32;
  const int half = lane >> 5;
  const int ia = rb + (lane & 31);   // row of the lane's A operands
  const int jc = cb + (lane & 31);   // column of its B operands and accumulators
  auto acc_row = [&](int v) { return rb + 8 * (v >> 2) + 4 * half + (v & 3); };
  f32x16 acc;
#pragma unroll
  for (int v = 0; v < 16; v++) acc[v] = mget(acc_row(v), jc);
  if (threadIdx.x < TAIL_NS) {
    const int q = (int)threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      mq.col[0][k][q] = mget(q, k);
      mq.row[0][q][k] = mget(k, q);
    }
  }
  if (threadIdx.x == 0) {
    B4<float> m[2][2], r[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
        m[a][b] = {mget(2 * a, 2 * b), mget(2 * a, 2 * b + 1),
                   mget(2 * a + 1, 2 * b), mget(2 * a + 1, 2 * b + 1)};
    if (pair_inv(m, r)) sing = 1;
    put_q(0, r);
  }
  __syncthreads();
  for (int p = 0; p < nt; p++) {
    const int cur = p & 1, nxt = cur ^ 1;
    const bool prow = (ia >> 2) == p, pcol = (jc >> 2) == p;
    float a[2], b[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int k = 2 * h + half;
      a[h] = prow ? ((ia & 3) == k ? 1.f : 0.f) : -mq.col[cur][k][ia];
      const float4 rw = *(const float4 *)mq.row[cur][jc];
      const float4 qk = *(const float4 *)mq.q[cur][k];
      const float qr = qk.x * rw.x + qk.y * rw.y + qk.z * rw.z + qk.w * rw.w;
      b[h] = pcol ? mq.q[cur][k][jc & 3] : qr;
    }
#pragma unroll
    for (int v = 0; v < 16; v++)
      if (pcol || (acc_row(v) >> 2) == p) acc[v] = 0.f;
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0], b[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[1], b[1], acc, 0, 0, 0);
    if (p + 1 < nt) {
      const int nb = 4 * (p + 1);
      if (nb >= rb && nb < rb + 32) {   // the next quad's rows
        const int qq = (nb - rb) >> 2, g = qq >> 1;
        float rv[4];
        if (g == 0) {
#pragma unroll
          for (int k = 0; k < 4; k++) rv[k] = acc[k];
        } else if (g == 1) {
#pragma unroll
          for (int k = 0; k < 4; k++) rv[k] = acc[4 + k];
        } else if (g == 2) {
#pragma unroll
          for (int k = 0; k < 4; k++) rv[k] = acc[8 + k];
        } else {
#pragma unroll
          for (int k = 0; k < 4; k++) rv[k] = acc[12 + k];
        }
        if (half == (qq & 1))
          *(float4 *)mq.row[nxt][jc] = make_float4(rv[0], rv[1], rv[2], rv[3]);
        if (nb >= cb && nb < cb + 32) {   // and its pivot block
          const int src = 32 * (qq & 1) + (nb - cb);   // the same in every lane
          float e[4][4];
#pragma unroll
          for (int k = 0; k < 4; k++)
#pragma unroll
            for (int m = 0; m < 4; m++)
              e[k][m] = __int_as_float(
                  __builtin_amdgcn_readlane(__float_as_int(rv[k]), src + m));
          if (lane == 0) {
            B4<float> m[2][2], r[2][2];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
              for (int y = 0; y < 2; y++)
                m[x][y] = {e[2 * x][2 * y], e[2 * x][2 * y + 1], e[2 * x + 1][2 * y],
                           e[2 * x + 1][2 * y + 1]};
            if (pair_inv(m, r)) sing = 1;
            put_q(nxt, r);
          }
        }
      }
      if (nb >= cb && nb < cb + 32) {   // the next quad's columns
        const int k = (lane & 31) - (nb - cb);
        if (k >= 0 && k < 4)
#pragma unroll
          for (int v = 0; v < 16; v += 4)
            *(float4 *)&mq.col[nxt][k][acc_row(v)] =
                make_float4(acc[v], acc[v + 1], acc[v + 2], acc[v + 3]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int i = acc_row(v);
    if (i < n2 && jc < n2) (&Ti[(i >> 1) * s + (jc >> 1)].a)[2 * (i & 1) + (jc & 1)] = acc[v];
  }
}

// ---------------------------------------------------------------------------
// Jacobian fill into the L+U storage (compuateJacobianMatrix,
// newtonpf.h:97-186): one thread per (Jacobian block, instance) writes one
// 2x2 block.  Superset pattern: a PV bus keeps an identity Q row and a zero
// magnitude column, so PV->PQ switching never changes the pattern.
// With elim0 the fill also finishes level 0 of the elimination tree (the
// factorization then starts at its level-1 chunks): a level-0 pivot is
// stored inverted, an L block of one times the pivot's inverse (the thread
// forms the pivot's block as well), in the factor's precision and the same
// operations as the factorization, so the factors are the same either way.
// A singular level-0 pivot sets jsing[t], which the factorization reports.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ B4<T> jac_block(const DevTopo &g, const DevInst &d,
                                           int t, int4 jb) {
  const int p = jb.y, ii = jb.z, jj = jb.w;
  const cd *V = d.V + (size_t)t * g.n;
  cd adm = d.Y[(size_t)t * g.nnzY + p];
  double yr = adm.x, yi = adm.y;
  double vrj = V[jj].x, vij = V[jj].y, vri = V[ii].x, vii = V[ii].y;
  double ibr = 0.0, ibi = 0.0;
  if (ii == jj) { cd I = d.Ib[(size_t)t * g.n + ii]; ibr = I.x; ibi = I.y; }
  double re = yr * vrj - yi * vij;
  double im = yr * vij + yi * vrj;
  double ab = sqrt(vrj * vrj + vij * vij);
  double ren = yr * vrj / ab - yi * vij / ab;
  double imn = yr * vij / ab + yi * vrj / ab;
  double a = -vii * (ibr - re) - vri * (-ibi + im);
  double b = vri * ren - vii * -imn + ibr * vrj / ab + ibi * vij / ab;
  double c = -vii * (-ibi + im) + vri * (ibr - re);
  double e = vri * -imn + vii * ren + ibr * vij / ab + -ibi * vrj / ab;
  int ti = d.type[(size_t)t * g.n + ii], tj = d.type[(size_t)t * g.n + jj];
  if (tj == 2) { b = 0.0; e = 0.0; }
  if (ti == 2) { c = 0.0; e = (ii == jj) ? 1.0 : 0.0; }
  B4<T> v;
  v.a = (T)a; v.b = (T)b; v.c = (T)c; v.d = (T)e;
  return v;
}

// rhs >= 0: while at most rhs instances of the batch part are active (the
// factorization's test), also write the right-hand-side blocks
// {F_P, 0; F_Q, 0} after the factor's, where its forward substitution starts
template <typename T>
__global__ void k_jacobian(DevTopo g, DevInst d, B4<T> *LU, int elim0, int rhs) {
  // one thread per Jacobian block, in block-position order: the writes into
  // the factor slab are nearly contiguous, the reads of Ybus are gathered
  // JU blocks per thread (rows r and r + blockDim.x of a part of JU *
  // blockDim.x blocks), their loads issued together: half the workgroups
  // of one block per thread, which matters once few instances remain
  constexpr int JU = 2;
  const int bd = (int)blockDim.x;
  const XSlot sl = xcd_slot((g.nJb + JU * bd - 1) / (JU * bd));
  const int t = sl.t;
  if (t >= d.T || d.st[t].done) return;
  const bool wr = rhs >= 0 && *d.active_count <= rhs;
#pragma unroll
  for (int u = 0; u < JU; u++) {
    const int r = sl.part * JU * bd + u * bd + (int)threadIdx.x;
    if (r >= g.nJb) break;
    const int4 jb = g.jac4[r];
    B4<T> v = jac_block<T>(g, d, t, jb);
    const int role = elim0 ? g.jac_l0[r] : -1;
    if (role == -2) {
      B4<T> inv;
      if (blk_inv(v, inv)) d.jsing[t] = 1;
      v = inv;
    } else if (role >= 0) {
      B4<T> inv;
      blk_inv(jac_block<T>(g, d, t, g.jac4[role]), inv);  // flagged by the pivot's own block
      v = blk_mul(v, inv);
    }
    LU[(size_t)t * g.nnzLU + jb.x] = v;
    if (wr && r < g.nb) {
      const double2 f = ((const double2 *)(d.F + (size_t)t * 2 * g.nb))[r];
      B4<T> b;
      b.a = (T)f.x; b.b = (T)0; b.c = (T)f.y; b.d = (T)0;
      LU[(size_t)t * g.nnzLU + g.rhs0 + r] = b;
    }
  }
}

// ---------------------------------------------------------------------------
// Batched LU of the 2x2-block Jacobian, static pivot order from the symbolic
// analysis.  One workgroup per instance.
// Sparse part: the outputs of the levels below the tail (each pivot's block,
// U row and L column, A(p) - sum L(r,m) U(m,c) over its descendants m) come
// in chunks of whole pivot groups.  Per chunk, every product is formed by
// its own thread into LDS (all of them independent loads of finished
// blocks), then each output reduces its product range with a lane group;
// a pivot stores its inverse, an L output reduces its pivot's range too and
// scales by the pivot inverse.  The pivot's own Jacobian block reaches LDS
// as an init entry of its range, so nothing reads a position another
// output of the same chunk writes.
// Dense tail (the top pivots, ancestors of everything else): its Schur
// complement S = A_SS - L_SN U_NS is assembled in LDS and inverted in place
// by Gauss-Jordan with the same pivot order (one barrier per step: the next
// pivot rows, columns and inverse are captured into double buffers while
// the current step updates): FP32 on 1024 lanes by pivot quads on the matrix
// cores (tail_inverse_mfma), otherwise by pivot pairs (FP32) or single
// pivots (FP64) on the vector ALUs; S^-1 goes to TL for the sweeps.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void lu_factor_body(DevTopo g, InstState *st, int count_iter,
                                               FactorPlan P, B4<T> *LU, B4<T> *TL,
                                               int c_start, int *jsing, int mode,
                                               const int *active, int budget,
                                               int force) {
  static_assert(TAIL_MAX * TAIL_MAX == CHUNK_PAIRS,
                "the chunk products and the dense tail share one LDS buffer");
  // mode 0: one workgroup per instance runs the main chunks from c_start and
  // the dense tail.  Split (two launches): mode 1 has SPLIT_MAX workgroups
  // per instance (workgroup b: instance b mod T, group b / T); with a split
  // variant for the active count, groups below its K run their subtrees'
  // chunks and flag a singular pivot in jsing; without one, group 0 does
  // what mode 0 does.  Mode 2 (one workgroup per instance) then runs the
  // chunks above the cut and the tail when mode 1 split.
  int t = blockIdx.x, cb = c_start, ce = P.nchunk;
  bool top = true;                      // runs the tail, reports singularity
  if (mode != 0) {
    const int v = split_variant(P, active, budget, force);
    if (mode == 1) {
      const int nt = (int)gridDim.x / SPLIT_MAX, gi = (int)blockIdx.x / nt;
      t = (int)blockIdx.x - gi * nt;
      if (v >= 0) {
        if (gi >= (2 << v)) return;
        cb = P.sp_grp[v][gi];
        ce = P.sp_grp[v][gi + 1];
        top = false;
      } else if (gi != 0) {
        return;
      }
    } else {
      if (v < 0) return;
      cb = P.sp_rest[v];
    }
  }
  if (st[t].done) return;
  __shared__ B4<T> buf[CHUNK_PAIRS];
  // chunk tables in LDS (scalar loads inside a phase would delay its LDS
  // accesses: both wait on the same counter); in FP32 the same space then
  // holds the staged blocks of the dense tail's products
  constexpr int STAGE = sizeof(T) == 4 ? TAIL_STAGE_MAX : 1;
  union TabStage {
    struct { int4 ch[MAX_CHUNKS]; int2 lg[MAX_CHUNKS]; } tab;
    B4<T> stage[STAGE];
    // blocked Gauss-Jordan of the dense tail (after the stage is read): per
    // step parity the next pivot pair's columns, rows and pair inverse
    struct {
      B4<T> col[2][TAIL_MAX][2], row[2][2][TAIL_MAX], q[2][2][2];
    } gj;
    // the same on matrix cores, in scalars: the next pivot quad's columns,
    // rows and inverse
    TailQuads mq;
  };
  __shared__ TabStage sh;
  int4 *const s_ch = sh.tab.ch;
  int2 *const s_lg = sh.tab.lg;
  const int nc = ce - cb;               // this workgroup's chunks
  // chunks below c_start (level 0) were finished by the Jacobian fill, which
  // flags a singular pivot in jsing[t], as do the groups of a split (taken
  // over and cleared by the workgroup that reports)
  __shared__ int sing;
  // read by the reductions for a range that has ended
  __shared__ B4<T> zblk;
  if (threadIdx.x == 0) {
    sing = 0;
    zblk = blk_zero<T>();
    if (top && (c_start > 0 || mode == 2)) {
      sing = jsing[t];
      jsing[t] = 0;
    }
  }
  // the forward substitution's outputs (the last of each chunk, their
  // products too) are left out unless the solve starts from them
  const bool fwd = active != nullptr && P.fused_limit >= 0 && *active <= P.fused_limit;
  for (int q = threadIdx.x; q < nc; q += blockDim.x) {
    int4 c4 = P.chunks[cb + q];
    int2 lg = P.chunk_lg[cb + q];
    if (!fwd) {
      c4.y -= lg.x >> 16;
      c4.w -= lg.y >> 16;
    }
    s_ch[q] = c4;
    s_lg[q] = make_int2(lg.x & 0xFFFF, lg.y & 0xFFFF);
  }
  B4<T> *A = LU + (size_t)t * g.nnzLU;
  __syncthreads();
  // product indices of the next chunk are loaded while the current chunk
  // reduces; a thread's first output descriptors while the products form.
  // Loads past the end of a range read its last element (every plan array
  // holds at least one) and are masked where used: a load under a condition
  // whose result merges with a default makes the compiler wait for it at
  // once, and global loads complete in order, so one such wait drains every
  // prefetch in flight.
  constexpr int PF = CHUNK_PAIRS / 1024;
  constexpr int NB = sizeof(T) == 4 ? 4 : 2;  // outputs per thread per batch
  int2 pidx[PF];
  int nq0 = 0, nq1 = 0;
  auto prefetch_pairs = [&](int c) {
    nq0 = s_ch[c].z;
    nq1 = s_ch[c].w;
    const int qlast = max(nq1 - 1, 0);
#pragma unroll
    for (int j = 0; j < PF; j++) {
      const int q = nq0 + (int)threadIdx.x + j * (int)blockDim.x;
      if (wave_any(q, nq1)) pidx[j] = P.pairs[min(q, qlast)];
    }
  };
  if (nc > 0) prefetch_pairs(0);
  // Loads of a slot no lane of the wave needs are skipped by a scalar branch
  // (wave_any), so the number of loads in flight is not known to the
  // compiler and any wait for one of them waits for all (global loads
  // complete in order).  The order below makes each such wait fall where
  // everything in flight is needed anyway: one explicit wait at the top of
  // the chunk (the pair indices and the previous chunk's stores), one after
  // the operands and descriptors; the next chunk's pair indices are then
  // requested and stay in flight through the reduction, which reads only
  // LDS and registers.
  for (int c = 0; c < nc; c++) {
    __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
    const int q0 = nq0, q1 = nq1;
    const int k0 = s_ch[c].x, k1 = s_ch[c].y;
    const int gsz = __builtin_amdgcn_readfirstlane(lane_group(s_lg[c].x, s_lg[c].y));
    const int lane = threadIdx.x & (gsz - 1);
    const int stride = (int)blockDim.x / gsz;
    const int kf = k0 + (int)threadIdx.x / gsz;
    const int klast = max(k1 - 1, 0);
    // operands of the first PR slots requested before any product forms
    // (all slots in FP32; FP64 takes two rounds to stay within registers)
    constexpr int PR = sizeof(T) == 4 ? PF : PF / 2;
    B4<T> lop[PR], rop[PR];
    auto load_operands = [&](int j0) {
#pragma unroll
      for (int j = 0; j < PR; j++) {
        if (wave_any(q0 + (int)threadIdx.x + (j0 + j) * (int)blockDim.x, q1)) {
          lop[j] = A[pidx[j0 + j].x];
          rop[j] = A[pidx[j0 + j].y < 0 ? pidx[j0 + j].x : pidx[j0 + j].y];
        }
      }
    };
    auto form_products = [&](int j0) {
#pragma unroll
      for (int j = 0; j < PR; j++) {
        const int q = q0 + (int)threadIdx.x + (j0 + j) * (int)blockDim.x;
        if (wave_any(q, q1)) {
          B4<T> v = blk_zero<T>();
          blk_fms(v, lop[j], rop[j]);
          v = blk_sel(pidx[j0 + j].y < 0, lop[j], v);   // an init entry carries the block itself
          if (q < q1) buf[q - q0] = v;
        }
      }
    };
    load_operands(0);
    // the thread's first batch of outputs: descriptors (requested with the
    // operands; one int4 per output)
    int4 ta[NB];
#pragma unroll
    for (int u = 0; u < NB; u++) {
      const int k = kf + u * stride;
      if (wave_any(k, k1)) ta[u] = P.tk[min(k, klast)];
    }
    __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
    form_products(0);
#pragma unroll
    for (int j0 = PR; j0 < PF; j0 += PR) {
      load_operands(j0);
      form_products(j0);
    }
    for (int q = q0 + threadIdx.x + PF * blockDim.x; q < q1; q += blockDim.x) {
      const int2 pr = P.pairs[q];
      B4<T> v;
      if (pr.y < 0) {
        v = A[pr.x];
      } else {
        v = blk_zero<T>();
        blk_fms(v, A[pr.x], A[pr.y]);
      }
      buf[q - q0] = v;
    }
    __syncthreads();
    if (c + 1 < nc) prefetch_pairs(c + 1);
    // outputs in batches of NB per thread; every output's Jacobian value is
    // an init entry of its own product range.  The first batch uses the
    // descriptors received with the operands (its own code path, so that no
    // wait for a later batch's descriptors, which would also wait for the
    // pair indices just requested, is merged into it); a later batch's
    // descriptors are requested together before any is reduced.
    auto reduce_batch = [&](int kb) {
#pragma unroll
      for (int u = 0; u < NB; u++) {
        if (kb + u * stride >= k1) break;
        const int p = ta[u].x & TASK_POS_MASK;
        const bool lcol = ta[u].x & TASK_LCOL;
        // an L output also sums its pivot's range (empty for the others);
        // in FP32 both ranges are read together (FP64 has not the registers)
        const unsigned pw = (unsigned)ta[u].w;
        const int v0 = (int)(pw & ((1u << TK_LEN_SHIFT) - 1));   // in the chunk
        const int v1 = lcol ? v0 + (int)(pw >> TK_LEN_SHIFT) : v0;
        B4<T> e, pv;
        if constexpr (sizeof(T) == 4) {
          lane_range_sum2(buf, &zblk, ta[u].y + lane - q0, ta[u].z - q0,
                          v0 + lane, v1, gsz, e, pv);
          group_sum4x2_first(e, pv, gsz);
        } else {
          e = lane_range_sum(buf, ta[u].y + lane - q0, ta[u].z - q0, gsz);
          group_sum4_first(e, gsz);
          if (lcol) {
            pv = lane_range_sum(buf, v0 + lane, v1, gsz);
            group_sum4_first(pv, gsz);
          }
        }
        if (lcol) {
          if (lane == 0) {
            B4<T> inv;
            blk_inv(pv, inv);   // a singular pivot is flagged by its own output
            A[p] = blk_mul(e, inv);
          }
        } else if (lane == 0) {
          if (ta[u].x & TASK_DIAG) {
            B4<T> inv;
            if (blk_inv(e, inv)) sing = 1;
            A[p] = inv;
          } else {
            A[p] = e;
          }
        }
      }
    };
    if (kf < k1) reduce_batch(kf);
    for (int kb = kf + NB * stride; kb < k1; kb += NB * stride) {
#pragma unroll
      for (int u = 0; u < NB; u++) {
        const int k = kb + u * stride;
        if (wave_any(k, k1)) ta[u] = P.tk[min(k, klast)];
      }
      reduce_batch(kb);
    }
    __syncthreads();
  }
  if (!top) {
    if (threadIdx.x == 0 && sing) jsing[t] = 1;
    return;
  }
  if (P.tn > 0) {
    B4<T> *D = buf;
    __shared__ B4<T> colb[2][TAIL_MAX], rowb[2][TAIL_MAX], pvb[2];
    const int s = P.tn, ss = s * s;
    for (int q = threadIdx.x; q < ss; q += blockDim.x) D[q] = blk_zero<T>();
    __syncthreads();
    if (sizeof(T) == 4 && P.tstage_n > 0) {
      // FP32: the distinct L_SN and U_NS blocks are copied into LDS once
      // (each serves several products), then every entry forms its products
      // from LDS in the order of its pair list (the sums of the plain loop)
      constexpr int SU = 6;
      const int ns = P.tstage_n, bd = (int)blockDim.x;
      // an entry's descriptor is requested a job ahead: the first job's
      // with the staging loads, the next job's while a job runs
      const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
      const int nw = (int)blockDim.x >> 6, wl = (int)threadIdx.x & 63;
      int dd_n, pos_n, qa_n, q1_n, lg_n;
      auto descriptor = [&](int j) {
        const int2 jb = P.tjob_s[min(j, P.tjobs_s - 1)];
        lg_n = jb.y & 7;
        const int k = jb.x + (wl >> lg_n);
        const int kc = min(k, P.tnent - 1);
        dd_n = P.tdense[kc];
        pos_n = P.tpos[kc];
        qa_n = P.tpp[kc];
        q1_n = k < (jb.y >> 3) ? P.tpp[kc + 1] : qa_n;   // past the job: no products
        if (k >= (jb.y >> 3)) dd_n = -1;
      };
      descriptor(wv);
      for (int q0 = threadIdx.x; q0 < ns; q0 += SU * bd) {
        int sp[SU];
#pragma unroll
        for (int u = 0; u < SU; u++) sp[u] = P.tstage[min(q0 + u * bd, ns - 1)];
        B4<T> sv[SU];
#pragma unroll
        for (int u = 0; u < SU; u++) sv[u] = A[sp[u]];
#pragma unroll
        for (int u = 0; u < SU; u++)
          if (q0 + u * bd < ns) sh.stage[q0 + u * bd] = sv[u];
      }
      __syncthreads();
      // a job's lists take one round of pair-index requests (12 per lane;
      // 16 would spill) and a group sum of log2(group size) steps
      constexpr int TU = 12;
      for (int j = wv; j < P.tjobs_s; j += nw) {
        const int dd = dd_n, pos = pos_n, qa = qa_n, q1 = q1_n;
        const int gsz = __builtin_amdgcn_readfirstlane(1 << lg_n);
        descriptor(j + nw);
        const int lane = wl & (gsz - 1);
        const bool has_a = dd >= 0 && (dd & TASK_HAS_A);
        const B4<T> a0 = A[has_a ? pos : 0];
        B4<T> acc = blk_sel(lane == 0 && has_a, a0, blk_zero<T>());
        for (int q = qa + lane; q < q1; q += TU * gsz) {
          int ps[TU];
#pragma unroll
          for (int u = 0; u < TU; u++) ps[u] = P.tpairS[min(q + u * gsz, q1 - 1)];
#pragma unroll
          for (int u = 0; u < TU; u++)
            blk_fms(acc, blk_sel(q + u * gsz < q1, sh.stage[ps[u] & 0xFFFF], blk_zero<T>()),
                    sh.stage[ps[u] >> 16]);
        }
        group_sum4_first(acc, gsz);
        if (lane == 0 && dd >= 0) D[dd & TASK_POS_MASK] = acc;
      }
    } else {
      const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
      const int nw = (int)blockDim.x >> 6, wl = (int)threadIdx.x & 63;
      constexpr int TU = 4;  // products of one lane requested together
      for (int j = wv; j < P.tjobs_g; j += nw) {
        const int2 jb = P.tjob_g[j];
        const int lg = jb.y & 7, gsz = __builtin_amdgcn_readfirstlane(1 << lg);
        const int lane = wl & (gsz - 1);
        const int k = jb.x + (wl >> lg);
        const bool on = k < (jb.y >> 3);
        const int kc = min(k, P.tnent - 1);
        const int dd = P.tdense[kc];
        const int pos = P.tpos[kc];
        const int qa = P.tpp[kc], q1 = on ? P.tpp[kc + 1] : qa;
        const B4<T> a0 = A[(dd & TASK_HAS_A) ? pos : 0];
        B4<T> acc = blk_sel(on && lane == 0 && (dd & TASK_HAS_A), a0, blk_zero<T>());
        for (int q = qa + lane; q < q1; q += TU * gsz) {
          int il[TU], iu[TU];
#pragma unroll
          for (int u = 0; u < TU; u++) {
            const int qq = min(q + u * gsz, q1 - 1);
            il[u] = P.tpairL[qq];
            iu[u] = P.tpairU[qq];
          }
          B4<T> lv[TU], uv[TU];
#pragma unroll
          for (int u = 0; u < TU; u++) {
            lv[u] = A[il[u]];
            uv[u] = A[iu[u]];
          }
          // a slot past the range repeats the last product with a zero left
          // factor: the sum and its order are those of the plain loop
#pragma unroll
          for (int u = 0; u < TU; u++)
            blk_fms(acc, blk_sel(q + u * gsz < q1, lv[u], blk_zero<T>()), uv[u]);
        }
        group_sum4_first(acc, gsz);
        if (on && lane == 0) D[dd & TASK_POS_MASK] = acc;
      }
    }
    __syncthreads();
    B4<T> *Ti = TL + (size_t)t * ss;
    // 2x2 tiles of blocks per side; an odd size gets an identity pivot
    // appended (zero coupling: every other value comes out unchanged)
    const int se = s + (s & 1), nt = se / 2;
    if (sizeof(T) == 4 && P.tail_mfma && blockDim.x == 1024) {
      if constexpr (sizeof(T) == 4) tail_inverse_mfma(D, s, nt, sh.mq, Ti, sing);
    } else if (sizeof(T) == 4 && nt * nt <= (int)blockDim.x) {
      // FP32: Gauss-Jordan by pivot pairs, each thread owning a 2x2 tile for
      // the whole elimination: per step the pair's inverse Q (formed by the
      // owner of the next diagonal tile), its two columns and two rows pass
      // through LDS; one barrier per pivot pair (in FP64 the sixteen blocks
      // a step holds would spill: it eliminates one pivot per step)
      auto dget = [&](int i, int j) {
        if (i < s && j < s) return D[i * s + j];
        B4<T> z = blk_zero<T>();
        if (i == j) { z.a = (T)1; z.d = (T)1; }
        return z;
      };
      for (int q = threadIdx.x; q < se; q += blockDim.x) {
        sh.gj.col[0][q][0] = dget(q, 0);
        sh.gj.col[0][q][1] = dget(q, 1);
        sh.gj.row[0][0][q] = dget(0, q);
        sh.gj.row[0][1][q] = dget(1, q);
      }
      if (threadIdx.x == 0) {
        const B4<T> m[2][2] = {{dget(0, 0), dget(0, 1)}, {dget(1, 0), dget(1, 1)}};
        B4<T> r[2][2];
        if (pair_inv(m, r)) sing = 1;
        sh.gj.q[0][0][0] = r[0][0]; sh.gj.q[0][0][1] = r[0][1];
        sh.gj.q[0][1][0] = r[1][0]; sh.gj.q[0][1][1] = r[1][1];
      }
      const int ti = (int)threadIdx.x / nt, tj = (int)threadIdx.x - ti * nt;
      const bool own = ti < nt;
      const int i0 = 2 * ti, j0 = 2 * tj;
      B4<T> dv[2][2];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) dv[a][b] = own ? dget(i0 + a, j0 + b) : blk_zero<T>();
      __syncthreads();
      for (int p = 0; p < nt; p++) {
        const int cur = p & 1, nxt = cur ^ 1;
        if (own) {
          B4<T> Q[2][2], cl[2][2], QR[2][2];
#pragma unroll
          for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
              Q[a][b] = sh.gj.q[cur][a][b];
              cl[a][b] = sh.gj.col[cur][i0 + a][b];
            }
#pragma unroll
          for (int m = 0; m < 2; m++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
              const B4<T> r0 = sh.gj.row[cur][0][j0 + b], r1 = sh.gj.row[cur][1][j0 + b];
              B4<T> v = blk_mul(Q[m][0], r0);
              const B4<T> v1 = blk_mul(Q[m][1], r1);
              v.a += v1.a; v.b += v1.b; v.c += v1.c; v.d += v1.d;
              QR[m][b] = v;
            }
#pragma unroll
          for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
              B4<T> v;
              if (ti == p && tj == p) {
                v = Q[a][b];
              } else if (ti == p) {
                v = QR[a][b];
              } else if (tj == p) {
                v = blk_zero<T>();
                blk_fms(v, cl[a][0], Q[0][b]);
                blk_fms(v, cl[a][1], Q[1][b]);
              } else {
                v = dv[a][b];
                blk_fms(v, cl[a][0], QR[0][b]);
                blk_fms(v, cl[a][1], QR[1][b]);
              }
              dv[a][b] = v;
            }
          // the next pair (tile p + 1): its columns, rows and inverse
          if (tj == p + 1)
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
              for (int b = 0; b < 2; b++) sh.gj.col[nxt][i0 + a][b] = dv[a][b];
          if (ti == p + 1)
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
              for (int b = 0; b < 2; b++) sh.gj.row[nxt][a][j0 + b] = dv[a][b];
          if (ti == p + 1 && tj == p + 1) {
            B4<T> r[2][2];
            if (pair_inv(dv, r)) sing = 1;
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
              for (int b = 0; b < 2; b++) sh.gj.q[nxt][a][b] = r[a][b];
          }
        }
        __syncthreads();
      }
      if (own)
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++)
            if (i0 + a < s && j0 + b < s) Ti[(i0 + a) * s + j0 + b] = dv[a][b];
    } else {
      for (int q = threadIdx.x; q < s; q += blockDim.x) {
        colb[0][q] = D[q * s];
        rowb[0][q] = D[q];
      }
      if (threadIdx.x == 0) {
        B4<T> iv;
        if (blk_inv(D[0], iv)) sing = 1;
        pvb[0] = iv;
      }
      __syncthreads();
      constexpr int GJR = 4;  // entries a thread keeps in registers
      const int nt1 = (s + 1) / 2;  // 2x2 tiles per side
      if (nt1 * nt1 <= (int)blockDim.x) {
        // each thread owns a 2x2 tile of S for the whole elimination: per
        // pivot it reads two row and two column blocks (shared by its four
        // entries) and forms P*row once per column
        const int ti = (int)threadIdx.x / nt1, tj = (int)threadIdx.x - ti * nt1;
        const bool own = ti < nt1;
        const int i0 = 2 * ti, j0 = 2 * tj;
        const bool vi1 = i0 + 1 < s, vj1 = j0 + 1 < s;
        B4<T> dv[2][2];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++) {
            const int i = min(i0 + a, s - 1), j = min(j0 + b, s - 1);
            dv[a][b] = own ? D[i * s + j] : blk_zero<T>();
          }
        for (int k = 0; k < s; k++) {
          const int cur = k & 1, nxt = cur ^ 1;
          if (own) {
            const B4<T> Pk = pvb[cur];
            const B4<T> col[2] = {colb[cur][i0], colb[cur][min(i0 + 1, s - 1)]};
            const B4<T> pr[2] = {blk_mul(Pk, rowb[cur][j0]),
                                 blk_mul(Pk, rowb[cur][min(j0 + 1, s - 1)])};
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
              for (int b = 0; b < 2; b++) {
                if ((a && !vi1) || (b && !vj1)) continue;
                const int i = i0 + a, j = j0 + b;
                B4<T> v;
                if (i == k) {
                  v = (j == k) ? Pk : pr[b];
                } else if (j == k) {
                  v = blk_zero<T>();
                  blk_fms(v, col[a], Pk);
                } else {
                  v = dv[a][b];
                  blk_fms(v, col[a], pr[b]);
                }
                dv[a][b] = v;
                if (j == k + 1) colb[nxt][i] = v;
                if (i == k + 1) rowb[nxt][j] = v;
                if (i == k + 1 && j == k + 1) {
                  B4<T> iv;
                  if (blk_inv(v, iv)) sing = 1;
                  pvb[nxt] = iv;
                }
              }
          }
          __syncthreads();
        }
        if (own)
#pragma unroll
          for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
              if ((!a || vi1) && (!b || vj1)) Ti[(i0 + a) * s + j0 + b] = dv[a][b];
      } else if (ss <= GJR * (int)blockDim.x) {
        // each thread owns up to GJR entries of S for the whole elimination:
        // only the pivot row, column and inverse pass through LDS
        B4<T> dv[GJR];
        int di[GJR], dj[GJR];
#pragma unroll
        for (int m = 0; m < GJR; m++) {
          const int q = threadIdx.x + m * blockDim.x;
          di[m] = q < ss ? q / s : -1;
          dj[m] = q < ss ? q - di[m] * s : -1;
          dv[m] = q < ss ? D[q] : blk_zero<T>();
        }
        for (int k = 0; k < s; k++) {
          const int cur = k & 1, nxt = cur ^ 1;
          const B4<T> Pk = pvb[cur];
#pragma unroll
          for (int m = 0; m < GJR; m++) {
            const int i = di[m], j = dj[m];
            if (i < 0) continue;
            B4<T> v;
            if (i == k) {
              v = (j == k) ? Pk : blk_mul(Pk, rowb[cur][j]);
            } else if (j == k) {
              v = blk_zero<T>();
              blk_fms(v, colb[cur][i], Pk);
            } else {
              v = dv[m];
              blk_fms(v, colb[cur][i], blk_mul(Pk, rowb[cur][j]));
            }
            dv[m] = v;
            if (j == k + 1) colb[nxt][i] = v;
            if (i == k + 1) rowb[nxt][j] = v;
            if (i == k + 1 && j == k + 1) {
              B4<T> iv;
              if (blk_inv(v, iv)) sing = 1;
              pvb[nxt] = iv;
            }
          }
          __syncthreads();
        }
#pragma unroll
        for (int m = 0; m < GJR; m++)
          if (di[m] >= 0) Ti[threadIdx.x + m * blockDim.x] = dv[m];
      } else {
        // more entries than registers allow: S stays in LDS
        for (int k = 0; k < s; k++) {
          const int cur = k & 1, nxt = cur ^ 1;
          const B4<T> Pk = pvb[cur];
          for (int q = threadIdx.x; q < ss; q += blockDim.x) {
            const int i = q / s, j = q - i * s;
            B4<T> v;
            if (i == k) {
              v = (j == k) ? Pk : blk_mul(Pk, rowb[cur][j]);
            } else if (j == k) {
              v = blk_zero<T>();
              blk_fms(v, colb[cur][i], Pk);
            } else {
              v = D[q];
              blk_fms(v, colb[cur][i], blk_mul(Pk, rowb[cur][j]));
            }
            D[q] = v;
            if (j == k + 1) colb[nxt][i] = v;
            if (i == k + 1) rowb[nxt][j] = v;
            if (i == k + 1 && j == k + 1) {
              B4<T> iv;
              if (blk_inv(v, iv)) sing = 1;
              pvb[nxt] = iv;
            }
          }
          __syncthreads();
        }
        for (int q = threadIdx.x; q < ss; q += blockDim.x) Ti[q] = D[q];
      }
    }
  }
  if (threadIdx.x == 0 && sing) {
    InstState s = st[t];
    s.done = 1;
    s.verdict = BPF_V_SINGULAR;
    s.it_round += count_iter;
    s.it_total += count_iter;
    st[t] = s;
  }
}

// one launch per factorization (mode 0), or the first of the split pair
// (mode 1: the subtrees, or everything when the batch does not split)
template <typename T>
__global__ void k_lu_factor(DevTopo g, InstState *st, int count_iter, FactorPlan P,
                            B4<T> *LU, B4<T> *TL, int c_start, int *jsing, int mode,
                            const int *active, int budget, int force) {
  lu_factor_body<T>(g, st, count_iter, P, LU, TL, c_start, jsing, mode, active,
                    budget, force);
}

// the second of the split pair: the chunks above the cut and the dense tail
// (returns at once when the first launch did not split)
template <typename T>
__global__ void k_lu_factor_top(DevTopo g, InstState *st, int count_iter, FactorPlan P,
                                B4<T> *LU, B4<T> *TL, int c_start, int *jsing,
                                const int *active, int budget, int force) {
  lu_factor_body<T>(g, st, count_iter, P, LU, TL, c_start, jsing, 2, active,
                    budget, force);
}

// Tail rows of a sweep: rs = b_S - L_SN y_N (trans: b_S - U_NS^T y_N), then
// x_S = S^-1 rs (trans: S^-T rs), written into X.  rs: 2 * TAIL_MAX in LDS.
template <typename T>
__device__ void tail_rows(const B4<T> *A, const B4<T> *Ti, T *X, T *rs,
                          const SolvePlan &P, bool trans) {
  const int s = P.tn;
  {
    const int gsz = lane_group(s, P.tsmax);
    const int lane = threadIdx.x & (gsz - 1);
    for (int a = (int)threadIdx.x / gsz; a < s; a += (int)blockDim.x / gsz) {
      T s0 = (T)0, s1 = (T)0;
      const int q1 = P.ts_ptr[a + 1];
      // TU entries per lane requested together; a slot past the range
      // repeats the last entry with a zero block (sum and order unchanged)
      constexpr int TU = 4;
      for (int q = P.ts_ptr[a] + lane; q < q1; q += TU * gsz) {
        int cc[TU], pp[TU];
#pragma unroll
        for (int u = 0; u < TU; u++) {
          const int qq = min(q + u * gsz, q1 - 1);
          cc[u] = P.ts_col[qq];
          pp[u] = trans ? P.ts_mir[qq] : P.ts_pos[qq];
        }
        B4<T> M[TU];
#pragma unroll
        for (int u = 0; u < TU; u++) M[u] = A[pp[u]];
#pragma unroll
        for (int u = 0; u < TU; u++) {
          const B4<T> L = blk_sel(q + u * gsz < q1, M[u], blk_zero<T>());
          const T y0 = X[2 * cc[u]], y1 = X[2 * cc[u] + 1];
          if (trans) {
            s0 += L.a * y0 + L.c * y1;
            s1 += L.b * y0 + L.d * y1;
          } else {
            s0 += L.a * y0 + L.b * y1;
            s1 += L.c * y0 + L.d * y1;
          }
        }
      }
      s0 = group_sum(s0, gsz);
      s1 = group_sum(s1, gsz);
      if (lane == 0) {
        const int i = P.tpiv[a];
        rs[2 * a] = X[2 * i] - s0;
        rs[2 * a + 1] = X[2 * i + 1] - s1;
      }
    }
  }
  __syncthreads();
  {
    const int gsz = lane_group(s, s);
    const int lane = threadIdx.x & (gsz - 1);
    for (int a = (int)threadIdx.x / gsz; a < s; a += (int)blockDim.x / gsz) {
      T s0 = (T)0, s1 = (T)0;
      constexpr int TU = 4;
      for (int b = lane; b < s; b += TU * gsz) {
        B4<T> Mv[TU];
#pragma unroll
        for (int u = 0; u < TU; u++) {
          const int bb = min(b + u * gsz, s - 1);
          Mv[u] = trans ? Ti[bb * s + a] : Ti[a * s + bb];
        }
#pragma unroll
        for (int u = 0; u < TU; u++) {
          const int bb = min(b + u * gsz, s - 1);
          const B4<T> M = blk_sel(b + u * gsz < s, Mv[u], blk_zero<T>());
          const T r0 = rs[2 * bb], r1 = rs[2 * bb + 1];
          if (trans) {
            s0 += M.a * r0 + M.c * r1;
            s1 += M.b * r0 + M.d * r1;
          } else {
            s0 += M.a * r0 + M.b * r1;
            s1 += M.c * r0 + M.d * r1;
          }
        }
      }
      s0 = group_sum(s0, gsz);
      s1 = group_sum(s1, gsz);
      if (lane == 0) {
        const int i = P.tpiv[a];
        X[2 * i] = s0;
        X[2 * i + 1] = s1;
      }
    }
  }
  __syncthreads();
}

// forward (unit-lower L) and backward (U, inverted pivots) sweeps on a block
// vector X (LDS or global, reached through a generic pointer); non-tail rows
// by levels with a lane group per row, tail rows through S^-1
template <typename T>
__device__ void lu_sweeps(const DevTopo &g, const B4<T> *A, const B4<T> *Ti,
                          T *X, T *rs, const SolvePlan &P) {
  for (int l = 0; l < P.nfw; l++) {
    const int r0 = P.fw_ptr[l], r1 = P.fw_ptr[l + 1];
    const int gsz = lane_group(r1 - r0, P.fw_max[l]);
    const int lane = threadIdx.x & (gsz - 1);
    for (int r = r0 + (int)threadIdx.x / gsz; r < r1; r += (int)blockDim.x / gsz) {
      const int i = P.fw_rows[r];
      T s0 = (T)0, s1 = (T)0;
      const int pe = g.ldiag[i];
      for (int p = g.lrow[i] + lane; p < pe; p += gsz) {
        const int c = g.lcol[p];
        const B4<T> L = A[p];
        const T y0 = X[2 * c], y1 = X[2 * c + 1];
        s0 += L.a * y0 + L.b * y1;
        s1 += L.c * y0 + L.d * y1;
      }
      s0 = group_sum(s0, gsz);
      s1 = group_sum(s1, gsz);
      if (lane == 0) {
        X[2 * i] -= s0;
        X[2 * i + 1] -= s1;
      }
    }
    __syncthreads();
  }
  if (P.tn > 0) tail_rows<T>(A, Ti, X, rs, P, false);
  for (int l = 0; l < P.nbw; l++) {
    const int r0 = P.bw_ptr[l], r1 = P.bw_ptr[l + 1];
    const int gsz = lane_group(r1 - r0, P.bw_max[l]);
    const int lane = threadIdx.x & (gsz - 1);
    for (int r = r0 + (int)threadIdx.x / gsz; r < r1; r += (int)blockDim.x / gsz) {
      const int i = P.bw_rows[r];
      const int dg = g.ldiag[i];
      const B4<T> Dg = (lane == 0) ? A[dg] : blk_zero<T>();
      T s0 = (T)0, s1 = (T)0;
      const int pe = g.lrow[i + 1];
      for (int p = dg + 1 + lane; p < pe; p += gsz) {
        const int c = g.lcol[p];
        const B4<T> U = A[p];
        const T y0 = X[2 * c], y1 = X[2 * c + 1];
        s0 += U.a * y0 + U.b * y1;
        s1 += U.c * y0 + U.d * y1;
      }
      s0 = group_sum(s0, gsz);
      s1 = group_sum(s1, gsz);
      if (lane == 0) {
        const T x0 = X[2 * i] - s0, x1 = X[2 * i + 1] - s1;
        X[2 * i] = Dg.a * x0 + Dg.b * x1;
        X[2 * i + 1] = Dg.c * x0 + Dg.d * x1;
      }
    }
    __syncthreads();
  }
}

// Entry-parallel sweeps (X and a product buffer in LDS).  Per chunk of a
// level: every entry's product (L or U block) * x_col is formed by its own
// thread into LDS, then each row reduces its products with a lane group
// and updates x.  Everything a chunk reads from HBM is static for the whole
// solve, so it is loaded ahead in two stages: the chunk's descriptor, the
// thread's first row descriptor and its entry indices two chunks ahead, the
// factor blocks (and the backward pivot inverse) one chunk ahead from those
// indices.  The sweeps write only LDS, so the loads stay in flight across
// the barriers and neither phase of a chunk waits on HBM for its first row
// and first two entries.
struct EntIdx {
  int4 meta;        // {row0, row1, entry0, entry1}
  int gsz, rfirst;  // lane group of the row phase, the thread's first row
  int4 rw;          // its descriptor (row index clamped into the chunk)
  int2 e0, e1;      // the thread's two entries {position, column} (clamped)
};

template <typename T>
struct EntVal {
  EntIdx ix;
  B4<T> a0, a1;     // factor blocks of the two entries
  B4<T> D;          // pivot inverse of the first row (backward sweep)
};

// Every load of the pipeline is unconditional, so that each chunk issues a
// fixed number of loads and the compiler waits for exactly the ones a step
// needs: an index past the end of its range is clamped to the range's last
// element (or to element 0: the plan arrays are never empty) and the result
// is masked where it is used.  A load under a condition (per lane or per
// wave) leaves the count of younger loads unknown, and as global loads
// complete in order the compiler then waits for all of them, prefetches
// included.
__device__ __forceinline__ void ep_index(const SolvePlan &P, const int4 *swc,
                                         const int *swm, int c, EntIdx &ix) {
  ix.meta = swc[c];
  ix.gsz = lane_group(ix.meta.y - ix.meta.x, swm[c]);
  ix.rfirst = ix.meta.x + (int)threadIdx.x / ix.gsz;
  ix.rw = P.sw_row[max(min(ix.rfirst, ix.meta.y - 1), 0)];
  const int e0 = ix.meta.z + threadIdx.x, elast = max(ix.meta.w - 1, 0);
  ix.e0 = P.sw_ent[min(e0, elast)];
  ix.e1 = P.sw_ent[min(e0 + (int)blockDim.x, elast)];
}

template <typename T>
__device__ __forceinline__ void ep_values(const B4<T> *A, const EntIdx &ix,
                                          bool bw, EntVal<T> &v) {
  v.ix = ix;
  v.a0 = A[ix.e0.x];
  v.a1 = A[ix.e1.x];
  v.D = A[bw ? ix.rw.w : 0];
}

template <typename T>
__device__ __forceinline__ void ep_product(T *prod, int e, const B4<T> &a,
                                           const T *X, int c) {
  const T y0 = X[2 * c], y1 = X[2 * c + 1];
  prod[2 * e] = a.a * y0 + a.b * y1;
  prod[2 * e + 1] = a.c * y0 + a.d * y1;
}

// one row of the row phase: subtract its products from X (and apply the
// pivot inverse in the backward sweep)
template <typename T>
__device__ __forceinline__ void ep_row(T *X, const T *prod, int e_base,
                                       const int4 &rw, const B4<T> &Dg, bool bw,
                                       int gsz, int lane) {
  T s0 = (T)0, s1 = (T)0;
  for (int e = rw.y + lane; e < rw.z; e += gsz) {
    s0 += prod[2 * (e - e_base)];
    s1 += prod[2 * (e - e_base) + 1];
  }
  s0 = group_sum(s0, gsz);
  s1 = group_sum(s1, gsz);
  if (lane == 0) {
    const int i = rw.x;
    const T x0 = X[2 * i] - s0, x1 = X[2 * i + 1] - s1;
    if (bw) {
      X[2 * i] = Dg.a * x0 + Dg.b * x1;
      X[2 * i + 1] = Dg.c * x0 + Dg.d * x1;
    } else {
      X[2 * i] = x0;
      X[2 * i + 1] = x1;
    }
  }
}

// one chunk; before its row phase, the next chunk's values are requested
// from its loaded indices and the indices of the one after are requested
// (always: past the last chunk the caller passes the last chunk again, so
// no branch separates the loaded registers from their uses).  The thread's
// first row comes from registers; further rows (wide chunks only) load
// their descriptors there.
template <typename T>
__device__ __forceinline__ void ep_chunk(const SolvePlan &P, const int4 *swc,
                                         const int *swm, const B4<T> *A, T *X,
                                         T *prod, const EntVal<T> &cur, bool bw,
                                         bool next_bw, int c_after,
                                         EntIdx &ix_next, EntVal<T> &nxt) {
  const int4 meta = cur.ix.meta;
  const int e0 = meta.z + threadIdx.x;
  if (e0 < meta.w) ep_product(prod, threadIdx.x, cur.a0, X, cur.ix.e0.y);
  if (e0 + (int)blockDim.x < meta.w)
    ep_product(prod, threadIdx.x + blockDim.x, cur.a1, X, cur.ix.e1.y);
  for (int e = e0 + 2 * blockDim.x; e < meta.w; e += blockDim.x) {
    const int2 en = P.sw_ent[e];
    ep_product(prod, e - meta.z, A[en.x], X, en.y);
  }
  __syncthreads();
  ep_values(A, ix_next, next_bw, nxt);
  ep_index(P, swc, swm, c_after, ix_next);
  const int gsz = cur.ix.gsz;
  const int lane = threadIdx.x & (gsz - 1);
  const int rstep = (int)blockDim.x / gsz;
  if (cur.ix.rfirst < meta.y)
    ep_row(X, prod, meta.z, cur.ix.rw, cur.D, bw, gsz, lane);
  for (int r = cur.ix.rfirst + rstep; r < meta.y; r += rstep) {
    const int4 rw = P.sw_row[r];
    ep_row(X, prod, meta.z, rw, bw ? A[rw.w] : blk_zero<T>(), bw, gsz, lane);
  }
  __syncthreads();
}

// the chunk loop runs two chunks per trip with the value buffers swapping
// roles, so no loaded value is copied between registers (a copy would wait
// for the load)
template <typename T>
__device__ __forceinline__ void lu_sweeps_ep(const B4<T> *A, const B4<T> *Ti,
                                             T *X, T *prod, T *rs,
                                             const int4 *swc, const int *swm,
                                             const SolvePlan &P) {
  const int nc = P.sw_nfw + P.sw_nbw;
  const bool tail = P.tn > 0;
  if (nc == 0) {
    if (tail) tail_rows<T>(A, Ti, X, rs, P, false);
    return;
  }
  EntIdx ixn;
  EntVal<T> va, vb;
  const int cl = nc - 1;
  ep_index(P, swc, swm, 0, ixn);
  ep_values(A, ixn, 0 >= P.sw_nfw, va);
  ep_index(P, swc, swm, min(1, cl), ixn);
  for (int c = 0; c < nc; c += 2) {
    if (c == P.sw_nfw && tail) tail_rows<T>(A, Ti, X, rs, P, false);
    ep_chunk(P, swc, swm, A, X, prod, va, c >= P.sw_nfw,
             min(c + 1, cl) >= P.sw_nfw, min(c + 2, cl), ixn, vb);
    if (c + 1 >= nc) break;
    if (c + 1 == P.sw_nfw && tail) tail_rows<T>(A, Ti, X, rs, P, false);
    ep_chunk(P, swc, swm, A, X, prod, vb, c + 1 >= P.sw_nfw,
             min(c + 2, cl) >= P.sw_nfw, min(c + 3, cl), ixn, va);
  }
  if (P.sw_nbw == 0 && tail) tail_rows<T>(A, Ti, X, rs, P, false);
}

// Slot-form sweeps (X in LDS).  A chunk's rows each own an aligned lane
// group of 2^lg lanes, one lane per entry; a lane forms its entry's product
// L_ij x_j (or U_ij x_j) in registers, the group sums its lanes' products
// into its first lane with row shifts (DPP within 16 lanes, cross-lane
// reads beyond), and that lane updates x_i (backward: times the pivot
// inverse).  One barrier per chunk.  Loads run two stages ahead as in the
// entry-parallel form (slot descriptors two chunks ahead, factor blocks one
// chunk ahead); a wave skips the slot blocks it has no lane in with a scalar
// branch, so load counts vary and one explicit wait at the top of each chunk
// receives everything the previous chunk requested.
// sum of the lanes of each aligned group of 2^lg lanes into the group's
// first lane.  An unconditional tree (step s: lane i adds lane i + 2^s) leaves
// in the first lane of every aligned window of 2^(s+1) lanes that window's
// sum, whatever the groups; a group's first lane keeps the value after step
// lg - 1, its own window.  glg = the largest lg among the wave's groups.
// Every lane of the wave must take part.
template <typename T>
__device__ __forceinline__ void slot_group_sum(T &p0, T &p1, int lg, int glg) {
  // every lane of the wave in a group of the first lane's size: after glg
  // steps each group's first lane holds its sum (the same adds as below,
  // without picking each lane's step)
  if (__builtin_amdgcn_ballot_w64(lg != glg) == 0) {
    if (glg > 0) { p0 += dpp_row<0x101>(p0); p1 += dpp_row<0x101>(p1); }
    if (glg > 1) { p0 += dpp_row<0x102>(p0); p1 += dpp_row<0x102>(p1); }
    if (glg > 2) { p0 += dpp_row<0x104>(p0); p1 += dpp_row<0x104>(p1); }
    if (glg > 3) { p0 += dpp_row<0x108>(p0); p1 += dpp_row<0x108>(p1); }
    if (glg > 4) { p0 += lane_up16(p0); p1 += lane_up16(p1); }
    if (glg > 5) { p0 += lane_up32(p0); p1 += lane_up32(p1); }
    return;
  }
  T r0 = p0, r1 = p1;
  // DPP row_shl:n (0x100 + n): lane i reads lane i + n of its 16-lane row
  if (glg > 0) {
    p0 += dpp_row<0x101>(p0); p1 += dpp_row<0x101>(p1);
    r0 = lg == 1 ? p0 : r0; r1 = lg == 1 ? p1 : r1;
  }
  if (glg > 1) {
    p0 += dpp_row<0x102>(p0); p1 += dpp_row<0x102>(p1);
    r0 = lg == 2 ? p0 : r0; r1 = lg == 2 ? p1 : r1;
  }
  if (glg > 2) {
    p0 += dpp_row<0x104>(p0); p1 += dpp_row<0x104>(p1);
    r0 = lg == 3 ? p0 : r0; r1 = lg == 3 ? p1 : r1;
  }
  if (glg > 3) {
    p0 += dpp_row<0x108>(p0); p1 += dpp_row<0x108>(p1);
    r0 = lg == 4 ? p0 : r0; r1 = lg == 4 ? p1 : r1;
  }
  if (glg > 4) {
    p0 += lane_up16(p0); p1 += lane_up16(p1);
    r0 = lg == 5 ? p0 : r0; r1 = lg == 5 ? p1 : r1;
  }
  if (glg > 5) {
    p0 += lane_up32(p0); p1 += lane_up32(p1);
    r0 = lg == 6 ? p0 : r0; r1 = lg == 6 ? p1 : r1;
  }
  p0 = r0;
  p1 = r1;
}

template <int SP>
struct SlotIdx {
  int4 s[SP];       // the thread's slots of one chunk
};

// a prefetched slot field, opaque to the compiler from here on: arithmetic
// on it (the 64-bit address of a block) is then not moved up to the load,
// where it would wait for the load at once
__device__ __forceinline__ int pinned(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <typename T, int SP>
struct SlotVal {
  B4<T> a[SP];      // entry blocks
  B4<T> d[SP];      // pivot inverses (backward sweep, groups' first lanes)
};

// The slot descriptors are loaded by every wave (clamped to the chunk's last
// slot): a load that a wave may skip leaves the register's previous value
// on the skipping path, and the compiler waits for the load before merging
// the two.  The blocks are loaded only by waves with a slot in the block
// (their registers are not read otherwise, so nothing merges).
template <int SP>
__device__ __forceinline__ SlotIdx<SP> slot_index(const int4 *slot, const int4 &m) {
  SlotIdx<SP> ix;
#pragma unroll
  for (int k = 0; k < SP; k++) {
    const int q = m.x + (int)threadIdx.x + k * (int)blockDim.x;
    ix.s[k] = slot[min(q, m.y - 1)];
  }
  return ix;
}

// TR: the transposed sweeps (U^T forward with the pivots, L^T backward),
// whose slots hold mirror positions; the pivots then belong to the forward
// chunks
template <typename T, int SP, bool TR = false>
__device__ __forceinline__ SlotVal<T, SP> slot_values(const B4<T> *A,
                                                      const int4 &m, bool bw,
                                                      const SlotIdx<SP> &ix) {
  SlotVal<T, SP> v;
  const bool piv = TR ? !bw : bw;
#pragma unroll
  for (int k = 0; k < SP; k++) {
    const int q = m.x + (int)threadIdx.x + k * (int)blockDim.x;
    if (wave_any(q, m.y)) {
      v.a[k] = A[pinned(ix.s[k].x)];
      if (piv) v.d[k] = A[pinned(ix.s[k].w)];
    }
  }
  return v;
}

// chunk c (range m; slots cur, blocks val, both received): the ranges of
// chunks c + 1 (mn) and c + 2 (ma) are read first, then every x the chunk
// needs; while those reads are in flight the blocks of chunk c + 1 are
// requested from its slots (nxt) and the slots of chunk c + 2 into cur.
// No slot of a chunk reads an x that another slot of it writes (its columns
// are rows of earlier levels, and a row has one segment per chunk), so the
// entries' x_j and the rows' own x_i are all read before any is written: one
// LDS round trip per chunk.  A chunk spans whole 64-slot windows, so a wave
// whose window starts inside it has every lane on a slot of it: a slot
// without an entry (column 0) reads the zero pair X[-2], X[-1] and forms a
// zero product without a test; a lane without a row reads x_0 (a broadcast,
// discarded).  The reads are unconditional, which keeps the wait for the
// two ranges exact.  Returns mn.
template <typename T, int SP, bool TR = false>
__device__ __forceinline__ int4 slot_chunk(const int4 *slot, const int4 *swc,
                                           const B4<T> *A, T *X, const int4 m,
                                           bool bw, SlotIdx<SP> &cur,
                                           const SlotVal<T, SP> &val,
                                           int c_next, bool next_bw,
                                           const SlotIdx<SP> &nxt,
                                           SlotVal<T, SP> &nval, int c_after) {
  __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
  int4 mn = swc[c_next], ma = swc[c_after];
  int lead[SP], lg[SP];
  T y0[SP], y1[SP], x0[SP], x1[SP];
#pragma unroll
  for (int k = 0; k < SP; k++) {
    const int y = pinned(cur.s[k].y), z = pinned(cur.s[k].z);
    const int cc = (y & ((1 << SLOT_LG_SHIFT) - 1)) - 1;
    lead[k] = z;
    lg[k] = y >> SLOT_LG_SHIFT;
    const int ii = max(z, 0) & (SLOT_NOT_LAST - 1);
    y0[k] = X[2 * cc];
    y1[k] = X[2 * cc + 1];
    x0[k] = X[2 * ii];
    x1[k] = X[2 * ii + 1];
  }
  mn.x = __builtin_amdgcn_readfirstlane(mn.x);
  mn.y = __builtin_amdgcn_readfirstlane(mn.y);
  mn.z = __builtin_amdgcn_readfirstlane(mn.z);
  ma.x = __builtin_amdgcn_readfirstlane(ma.x);
  ma.y = __builtin_amdgcn_readfirstlane(ma.y);
  nval = slot_values<T, SP, TR>(A, mn, next_bw, nxt);
  cur = slot_index<SP>(slot, ma);
  T p0[SP], p1[SP];
#pragma unroll
  for (int k = 0; k < SP; k++) {
    const int q = m.x + (int)threadIdx.x + k * (int)blockDim.x;
    if (wave_any(q, m.y)) {
      const B4<T> &a = val.a[k];
      if (TR) {   // the block (col, row) used transposed
        p0[k] = a.a * y0[k] + a.c * y1[k];
        p1[k] = a.b * y0[k] + a.d * y1[k];
      } else {
        p0[k] = a.a * y0[k] + a.b * y1[k];
        p1[k] = a.c * y0[k] + a.d * y1[k];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < SP; k++) {
    const int q = m.x + (int)threadIdx.x + k * (int)blockDim.x;
    if (wave_any(q, m.y)) {
      // groups come largest first, each aligned to its size, so the wave's
      // first lane starts its largest group: the steps that group needs
      slot_group_sum(p0[k], p1[k], lg[k], __builtin_amdgcn_readfirstlane(lg[k]));
      if (lead[k] >= 0) {
        // a row's earlier segments only subtract; its last one (or its only
        // one) also applies the pivot inverse in the backward sweep
        const int i = lead[k] & (SLOT_NOT_LAST - 1);
        const T u0 = x0[k] - p0[k], u1 = x1[k] - p1[k];
        if ((TR ? !bw : bw) && !(lead[k] & SLOT_NOT_LAST)) {
          const B4<T> &D = val.d[k];
          if (TR) {   // the pivot inverse transposed
            X[2 * i] = D.a * u0 + D.c * u1;
            X[2 * i + 1] = D.b * u0 + D.d * u1;
          } else {
            X[2 * i] = D.a * u0 + D.b * u1;
            X[2 * i + 1] = D.c * u0 + D.d * u1;
          }
        } else {
          X[2 * i] = u0;
          X[2 * i + 1] = u1;
        }
      }
    }
  }
  __syncthreads();
  return mn;
}

template <typename T, int SP, bool TR = false>
__device__ __forceinline__ void lu_sweeps_slot(const B4<T> *A, const B4<T> *Ti,
                                               T *X, T *rs, const int4 *swc,
                                               const SolvePlan &P, bool skip_fw) {
  const int nfw = TR ? P.sl_nfw_t : P.sl_nfw;
  const int nc = nfw + (TR ? P.sl_nbw_t : P.sl_nbw);
  const bool tail = P.tn > 0;
  const int4 *slots = TR ? P.sl_slot_t : P.sl_slot;
  // skip_fw: X holds the forward substitution's result already (the tail
  // rows still hold their right-hand side)
  const int c0 = skip_fw ? nfw : 0;
  if (c0 >= nc) {
    if (tail) tail_rows<T>(A, Ti, X, rs, P, TR);
    return;
  }
  const int cl = nc - 1;
  SlotIdx<SP> ia, ib;
  SlotVal<T, SP> va, vb;
  int4 m = swc[c0];
  m.x = __builtin_amdgcn_readfirstlane(m.x);
  m.y = __builtin_amdgcn_readfirstlane(m.y);
  m.z = __builtin_amdgcn_readfirstlane(m.z);
  ia = slot_index<SP>(slots, m);
  __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
  va = slot_values<T, SP, TR>(A, m, c0 >= nfw, ia);
  ib = slot_index<SP>(slots, swc[min(c0 + 1, cl)]);
  for (int c = c0; c < nc; c += 2) {
    if (c == nfw && tail) tail_rows<T>(A, Ti, X, rs, P, TR);
    m = slot_chunk<T, SP, TR>(slots, swc, A, X, m, c >= nfw, ia, va,
                              min(c + 1, cl), min(c + 1, cl) >= nfw, ib, vb,
                              min(c + 2, cl));
    if (c + 1 >= nc) break;
    if (c + 1 == nfw && tail) tail_rows<T>(A, Ti, X, rs, P, TR);
    m = slot_chunk<T, SP, TR>(slots, swc, A, X, m, c + 1 >= nfw, ib, vb,
                              min(c + 2, cl), min(c + 2, cl) >= nfw, ia, va,
                              min(c + 3, cl));
  }
  if (nc == nfw && tail) tail_rows<T>(A, Ti, X, rs, P, TR);
}

// ---------------------------------------------------------------------------
// Newton step solve: dx = J^-1 rhs with the factors in precision T (one pass;
// mixed precision runs it again on the FP64 residual of k_residual and
// accumulates).  One workgroup per instance.
// ---------------------------------------------------------------------------
// LDS reserved for the solve vector of one instance: 148 KB of gfx950's
// 160 KB (the rest holds the tail right-hand side and the chunk table),
// i.e. 18944 blocks in FP32 or 9472 in FP64; larger systems keep the vector
// in global memory (use_lds = 0)
constexpr int SOLVE_LDS_BYTES = 148 * 1024;
// the slot form keeps a zero pair before X (the x of slots without an entry)
constexpr int SLOT_X_PAD = 2;
// the entry-parallel sweeps keep their product buffer (2 * SWEEP_CHUNK
// scalars) after X in the same LDS block
__host__ __device__ constexpr int ep_offset(int nb) { return (2 * nb + 3) & ~3; }

// one pass on X: right-hand side in, sweeps, dx out.  Instantiated once with
// X in LDS and once in global memory, so that the compiler addresses X
// directly: through a generic pointer every access would be a flat
// instruction, whose wait also drains the prefetched global loads.
template <typename T, bool LDS, int SP>
__device__ __forceinline__ void solve_pass(const DevTopo &g, const B4<T> *A,
                                           const B4<T> *Ti, T *X, T *prod,
                                           T *rs, const int4 *swc,
                                           const int *swm, const SolvePlan &P,
                                           int ep, const double *src,
                                           int rhs_mode, int accumulate,
                                           double *dx) {
  // rhs_mode 0: (P, Q) mismatch; 1: (P, 0) and 2: (0, Q) for the decoupled
  // half steps; 3: the refinement residual of k_residual; 4: the mismatch,
  // forward-substituted by the factorization.  Vector loads in
  // groups of VU per thread, all requested before the first is used: with
  // 1024 lanes one round trip covers 20480 scalars (10240 blocks), the
  // right-hand side and the accumulated dx of a 9241-bus grid each in one
  // (24 would spill).
  constexpr int VU = 20;
  const int n2 = 2 * g.nb;
  const int stride = (int)blockDim.x;
  if (SP > 0 && rhs_mode == 4) {
    // rhs_mode 4 (slot form only): the factorization formed the forward
    // substitution of the mismatch in the right-hand-side blocks after the
    // factor (y in their first column; a tail row's block holds its b)
    constexpr int YU = sizeof(T) == 4 ? 10 : 5;
    const B4<T> *Y = A + g.rhs0;
    for (int k0 = threadIdx.x; k0 < g.nb; k0 += YU * stride) {
      B4<T> y[YU];
#pragma unroll
      for (int u = 0; u < YU; u++) y[u] = Y[min(k0 + u * stride, g.nb - 1)];
#pragma unroll
      for (int u = 0; u < YU; u++) {
        const int k = k0 + u * stride;
        if (k < g.nb) { X[2 * k] = y[u].a; X[2 * k + 1] = y[u].c; }
      }
    }
  } else {
    for (int k0 = threadIdx.x; k0 < n2; k0 += VU * stride) {
      double v[VU];
#pragma unroll
      for (int u = 0; u < VU; u++) v[u] = src[min(k0 + u * stride, n2 - 1)];
#pragma unroll
      for (int u = 0; u < VU; u++) {
        const int k = k0 + u * stride;
        const bool keep = rhs_mode == 0 || rhs_mode == 3 || (rhs_mode == 1) == ((k & 1) == 0);
        if (k < n2) X[k] = keep ? (T)v[u] : (T)0;
      }
    }
  }
  __syncthreads();
  if (SP > 0) lu_sweeps_slot<T, (SP > 0 ? SP : 1)>(A, Ti, X, rs, swc, P, rhs_mode == 4);
  else if (LDS && ep) lu_sweeps_ep<T>(A, Ti, X, prod, rs, swc, swm, P);
  else lu_sweeps<T>(g, A, Ti, X, rs, P);
  if (accumulate) {
    for (int k0 = threadIdx.x; k0 < n2; k0 += VU * stride) {
      double v[VU];
#pragma unroll
      for (int u = 0; u < VU; u++) v[u] = dx[min(k0 + u * stride, n2 - 1)];
#pragma unroll
      for (int u = 0; u < VU; u++) {
        const int k = k0 + u * stride;
        if (k < n2) dx[k] = v[u] + (double)X[k];
      }
    }
  } else {
    for (int k = threadIdx.x; k < n2; k += stride) dx[k] = (double)X[k];
  }
}

// SP > 0: slot form with SP slots per lane, for chunks of at most SP x the
// workgroup's lanes (the launch bound lets the larger slot counts of
// narrower workgroups stay in registers)
template <typename T, int SP>
__global__ void __launch_bounds__(SP > 0 && SLOT_CHUNK / SP < 1024 ? SLOT_CHUNK / SP : 1024)
k_lu_solve(DevTopo g, DevInst d, const B4<T> *LU, const B4<T> *TL,
           int rhs_mode, int accumulate, SolvePlan P, int use_lds, int ep,
           T *xglob) {
  __shared__ T smem[SOLVE_LDS_BYTES / sizeof(T)];
  __shared__ T rs[2 * TAIL_MAX];
  // the chunk table in LDS (see the factorization)
  __shared__ int4 s_swc[MAX_CHUNKS];
  __shared__ int s_swm[MAX_CHUNKS];
  const int t = blockIdx.x;
  if (d.st[t].done) return;
  const B4<T> *A = LU + (size_t)t * g.nnzLU;
  const B4<T> *Ti = TL + (size_t)t * P.tn * P.tn;
  const double *src = (rhs_mode == 3 ? d.rhs : d.F) + (size_t)t * 2 * g.nb;
  if (SP > 0)
    for (int q = threadIdx.x; q < P.sl_nfw + P.sl_nbw; q += blockDim.x)
      s_swc[q] = P.sl_chunk[q];
  else if (ep == 1)
    for (int q = threadIdx.x; q < P.sw_nfw + P.sw_nbw; q += blockDim.x) {
      s_swc[q] = P.sw_chunk[q];
      s_swm[q] = P.sw_cmax[q];
    }
  double *dx = d.dx + (size_t)t * 2 * g.nb;
  T *prod = smem + ep_offset(g.nb);
  // the slot form keeps X after a zero pair (read by slots without an entry)
  if (SP > 0 && threadIdx.x < SLOT_X_PAD) smem[threadIdx.x] = (T)0;
  // rhs_mode 4 falls back to the mismatch (0) when the factorization left
  // the forward substitution out: the same test on the same active count
  if (rhs_mode == 4 && !(P.fused_limit >= 0 && *d.active_count <= P.fused_limit))
    rhs_mode = 0;
  if (SP > 0)
    solve_pass<T, true, SP>(g, A, Ti, smem + SLOT_X_PAD, prod, rs, s_swc, s_swm,
                            P, 0, src, rhs_mode, accumulate, dx);
  else if (use_lds)
    solve_pass<T, true, 0>(g, A, Ti, smem, prod, rs, s_swc, s_swm, P, ep,
                           src, rhs_mode, accumulate, dx);
  else
    solve_pass<T, false, 0>(g, A, Ti, xglob + (size_t)t * 2 * g.nb, nullptr, rs,
                            s_swc, s_swm, P, 0, src, rhs_mode, accumulate, dx);
}

// ---------------------------------------------------------------------------
// FP64 refinement residual r = F - J dx, matrix-free from Ybus, V and Ibus:
//   dS_i = dV_i conj(I_i) + V_i conj(sum_j Y_ij dV_j),
//   dV_j = V_j (j dtheta_j + dVm_j / |V_j|);
// P rows: F_P - Re dS; Q rows: F_Q - Im dS for PQ buses, -dVm for PV buses
// (their identity rows).  Many workgroups per instance, bus order.
// ---------------------------------------------------------------------------
__global__ void k_residual(DevTopo g, DevInst d) {
  const XSlot sl = xcd_slot(parts_of(g.n));
  const int t = sl.t, i = sl.part * blockDim.x + threadIdx.x;
  if (t >= d.T || i >= g.n) return;
  if (d.st[t].done) return;
  const int k = g.blk_of_bus[i];
  if (k < 0) return;
  // F, dx and the residual as (P, Q) / (angle, magnitude) pairs: one
  // 16-byte access per block
  const double2 *F = (const double2 *)(d.F + (size_t)t * 2 * g.nb);
  const double2 *dx = (const double2 *)(d.dx + (size_t)t * 2 * g.nb);
  const cd *Y = d.Y + (size_t)t * g.nnzY;
  const cd *V = d.V + (size_t)t * g.n;
  cd acc = mk(0.0, 0.0);
  // the row in groups of RU entries (loads of a level all requested
  // together; terms added in row order, slack columns and slots past the
  // row end leave the sum unchanged)
  constexpr int RU = 4;
  const int pa = g.yrow[i], pe = g.yrow[i + 1];
  for (int p0 = pa; p0 < pe; p0 += RU) {
    int col[RU], bj[RU];
    cd y[RU], vj[RU];
    double2 dj[RU];
#pragma unroll
    for (int u = 0; u < RU; u++) {
      const int p = min(p0 + u, pe - 1);
      col[u] = g.ycol[p];
      bj[u] = g.yblk[p];
      y[u] = Y[p];
    }
#pragma unroll
    for (int u = 0; u < RU; u++) {
      const int b = max(bj[u], 0);
      vj[u] = V[col[u]];
      dj[u] = dx[b];
    }
#pragma unroll
    for (int u = 0; u < RU; u++) {
      const double m = dj[u].y / cabs_(vj[u]);
      const cd s = cadd(acc, cmul(y[u], cmul(vj[u], mk(m, dj[u].x))));
      const bool use = p0 + u < pe && bj[u] >= 0;
      acc.x = use ? s.x : acc.x;
      acc.y = use ? s.y : acc.y;
    }
  }
  const size_t o = (size_t)t * g.n + i;
  const cd vi = V[i];
  const double2 dk = dx[k], fk = F[k];
  const cd dvi = cmul(vi, mk(dk.y / cabs_(vi), dk.x));
  const cd dS = cadd(cmul(dvi, cconj(d.Ib[o])), cmul(vi, cconj(acc)));
  double2 *R = (double2 *)(d.rhs + (size_t)t * 2 * g.nb);
  R[k] = make_double2(fk.x - dS.x, d.type[o] == 1 ? fk.y - dS.y : -dk.y);
}

// ---------------------------------------------------------------------------
// Polar update V = (|V| - dVm) e^{j(atan(Vi/Vr) - dtheta)} (newtonpf.h:188-213;
// the reference's angle is atan(y/x), complexUtils.h:21-23).
// ---------------------------------------------------------------------------
__global__ void k_update(DevTopo g, DevInst d) {
  // the active-instance count restarts here for the verdicts that follow:
  // until now it held the last verdicts' count, which the factorization of
  // this iteration reads to choose its split
  if (blockIdx.x == 0 && threadIdx.x == 0) *d.active_count = 0;
  const XSlot sl = xcd_slot(parts_of(g.n));
  const int t = sl.t, i = sl.part * blockDim.x + threadIdx.x;
  if (t >= d.T || i >= g.n) return;
  if (d.st[t].done) return;
  const int k = g.blk_of_bus[i];
  if (k < 0) return;
  size_t o = (size_t)t * g.n + i;
  const double2 dk = ((const double2 *)(d.dx + (size_t)t * 2 * g.nb))[k];
  cd v = d.V[o];
  double mag = cabs_(v);
  if (d.type[o] == 1) mag -= dk.y;
  cd e = cexpj(atan(v.y / v.x) - dk.x);
  d.V[o] = cmul(mk(mag, 0.0), e);
}

__global__ void k_zero_count(DevInst d) { *d.active_count = 0; }

// the active (not finished) instances of a batch part, into its counter:
// the first Newton iteration of a part reads it (the solve's first
// mismatch counts the whole batch into the first part's counter)
__global__ void k_count_active(DevInst d) {
  __shared__ int n;
  if (threadIdx.x == 0) n = 0;
  __syncthreads();
  int c = 0;
  for (int t = threadIdx.x; t < d.T; t += blockDim.x) c += d.st[t].done ? 0 : 1;
  if (c) atomicAdd(&n, c);
  __syncthreads();
  if (threadIdx.x == 0) *d.active_count = n;
}

// ---------------------------------------------------------------------------
// Fast decoupled: half-step updates (fdpf.h:416-448).  part 1: angles of all
// unknown buses, part 2: magnitudes of PQ buses.
// ---------------------------------------------------------------------------
__global__ void k_fd_update(DevTopo g, DevInst d, int part) {
  const XSlot sl = xcd_slot(parts_of(g.n));
  const int t = sl.t, i = sl.part * blockDim.x + threadIdx.x;
  if (t >= d.T || i >= g.n) return;
  if (d.st[t].done) return;
  const int k = g.blk_of_bus[i];
  if (k < 0) return;
  size_t o = (size_t)t * g.n + i;
  const double *dx = d.dx + (size_t)t * 2 * g.nb;
  cd v = d.V[o];
  if (part == 1) {
    d.V[o] = cmul(mk(cabs_(v), 0.0), cexpj(atan(v.y / v.x) - dx[2 * k]));
  } else if (d.type[o] == 1) {
    d.V[o] = cmul(mk(cabs_(v) - dx[2 * k + 1], 0.0), cexpj(atan(v.y / v.x)));
  }
}

// one Ybus-style entry of the modified networks behind B' and B''
// (makeB.h:311-501): B' zeroes line charging, taps and bus shunts (and R for
// XB); B'' zeroes phase shifts (and R for BX)
__device__ cd branch_entry_fd(const DevTopo &g, const DevInst &d, int t, int b,
                              int kind, int which, int zeroR) {
  double stat = d.br_on[(size_t)t * g.m + b] ? 1.0 : 0.0;
  double R = zeroR ? 0.0 : g.R[b];
  cd Ys = cdiv(mk(stat, 0.0), mk(R, g.X[b]));
  double Bch = (which == 1) ? 0.0 : g.B[b];
  double tp = (which == 1) ? 0.0 : d.tap[(size_t)t * g.m + b];
  double sh = (which == 1) ? g.shift[b] : 0.0;
  cd tap = (tp != 0.0) ? mk(tp, 0.0) : mk(1.0, 0.0);
  tap = cmul(tap, cexpj(M_PI / 180.0 * sh));
  cd Bc = cmul(mk(stat, 0.0), mk(0.0, Bch));
  switch (kind) {
    case KFF: return cdiv(cadd(Ys, cdiv(Bc, mk(2.0, 0.0))), cmul(tap, cconj(tap)));
    case KFT: return cdiv(cmul(Ys, mk(-1.0, 0.0)), cconj(tap));
    case KTF: return cdiv(cmul(Ys, mk(-1.0, 0.0)), tap);
    default: return cadd(Ys, cdiv(Bc, mk(2.0, 0.0)));
  }
}

// B' and B'' of every instance as one block-diagonal 2x2 system in the L+U
// storage: block (i,j) = [B'_ij 0; 0 B''_ij], with an identity Q row for PV
// buses; B = -Im(Ybus of the modified network).  alg: 1 XB, 2 BX.
__global__ void k_bmat(DevTopo g, DevInst d, int alg, B4<double> *LU) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  int t = blockIdx.y;
  if (p >= g.nnzY) return;
  int q = g.y2lu[p];
  if (q < 0) return;
  double v[2];
  for (int which = 1; which <= 2; which++) {
    int zeroR = (which == 1 && alg == 1) || (which == 2 && alg == 2);
    cd af = mk(0.0, 0.0), at = mk(0.0, 0.0);
    for (int r = g.yf_ptr[p]; r < g.yf_ptr[p + 1]; r++) {
      int it = g.yf_item[r];
      af = cadd(af, branch_entry_fd(g, d, t, it >> 2, it & 3, which, zeroR));
    }
    for (int r = g.yt_ptr[p]; r < g.yt_ptr[p + 1]; r++) {
      int it = g.yt_item[r];
      at = cadd(at, branch_entry_fd(g, d, t, it >> 2, it & 3, which, zeroR));
    }
    cd y = cadd(af, at);
    int i = g.ydiag_bus[p];
    if (i >= 0) {
      double bs = (which == 1) ? 0.0 : d.Bs[(size_t)t * g.n + i];
      y = cadd(y, mk(g.Gs[i], bs));
    }
    v[which - 1] = -y.y;
  }
  int ii = g.yrowof[p], jj = g.ycol[p];
  int ti = d.type[(size_t)t * g.n + ii], tj = d.type[(size_t)t * g.n + jj];
  B4<double> blk;
  blk.a = v[0];
  blk.b = 0.0;
  blk.c = 0.0;
  blk.d = (ti == 2 || tj == 2) ? ((ii == jj) ? 1.0 : 0.0) : v[1];
  LU[(size_t)t * g.nnzLU + q] = blk;
}

// ---------------------------------------------------------------------------
// Adjoint sensitivities (no counterpart in the reference).  Objective per
// instance: f = sum_b Re(Sf_b + St_b) + w sum_ends max(0, |S| - rate_b)^2,
// Sf = Vf conj(If), St = Vt conj(It) (the loss of runpf.h:228-257 plus a
// branch-flow penalty).  With F(x, p) = 0 at the solution and
// J^T lambda = df/dx:  df/dp = df/dp|explicit - lambda^T dF/dp.
// ---------------------------------------------------------------------------
struct BranchState {
  cd Vf, Vt, Sf, St, yff, yft, ytf, ytt, af, at;  // a = d f / d S (complex)
};

__device__ BranchState branch_state(const DevTopo &g, const DevInst &d, int t,
                                    int b, const double *rate, double w) {
  BranchState s;
  const cd *V = d.V + (size_t)t * g.n;
  int f = g.br_from[b], to = g.br_to[b];
  s.Vf = V[f];
  s.Vt = V[to];
  s.yff = branch_entry(g, d, t, b, KFF);
  s.yft = branch_entry(g, d, t, b, KFT);
  s.ytf = branch_entry(g, d, t, b, KTF);
  s.ytt = branch_entry(g, d, t, b, KTT);
  cd If = cadd(cmul(s.yff, s.Vf), cmul(s.yft, s.Vt));
  cd It = cadd(cmul(s.ytf, s.Vf), cmul(s.ytt, s.Vt));
  s.Sf = cmul(s.Vf, cconj(If));
  s.St = cmul(s.Vt, cconj(It));
  // d/dz [Re S + w max(0,|S|-r)^2] = Re(a dS/dz), a = 1 + 2w(|S|-r)/|S| conj(S)
  double pf = 0.0, pt = 0.0;
  if (rate) {
    double r = rate[b], mf = cabs_(s.Sf), mt = cabs_(s.St);
    if (mf > r) pf = 2.0 * w * (mf - r) / mf;
    if (mt > r) pt = 2.0 * w * (mt - r) / mt;
  }
  s.af = cadd(mk(1.0, 0.0), cmul(mk(pf, 0.0), cconj(s.Sf)));
  s.at = cadd(mk(1.0, 0.0), cmul(mk(pt, 0.0), cconj(s.St)));
  return s;
}

// right-hand side G = df/dx in block order (angle, magnitude of PQ buses)
__global__ void k_adj_rhs(DevTopo g, DevInst d, const double *rate, double w,
                          double *G) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int t = blockIdx.y;
  if (k >= g.nb) return;
  if (d.st[t].done) return;
  int i = g.bus_of_blk[k];
  double ga = 0.0, gm = 0.0;
  for (int q = g.bbr_ptr[i]; q < g.bbr_ptr[i + 1]; q++) {
    int b = g.bbr[q];
    BranchState s = branch_state(g, d, t, b, rate, w);
    cd dSf_a, dSt_a, dSf_m, dSt_m;
    const cd J = mk(0.0, 1.0);
    if (g.br_from[b] == i) {
      double vf2 = s.Vf.x * s.Vf.x + s.Vf.y * s.Vf.y, vf = sqrt(vf2);
      dSf_a = cmul(J, csub(s.Sf, cmul(mk(vf2, 0.0), cconj(s.yff))));
      dSt_a = cmul(mk(0.0, -1.0), cmul(s.Vt, cconj(cmul(s.ytf, s.Vf))));
      dSf_m = cdiv(cadd(s.Sf, cmul(mk(vf2, 0.0), cconj(s.yff))), mk(vf, 0.0));
      dSt_m = cdiv(cmul(s.Vt, cconj(cmul(s.ytf, s.Vf))), mk(vf, 0.0));
    } else {
      double vt2 = s.Vt.x * s.Vt.x + s.Vt.y * s.Vt.y, vt = sqrt(vt2);
      dSt_a = cmul(J, csub(s.St, cmul(mk(vt2, 0.0), cconj(s.ytt))));
      dSf_a = cmul(mk(0.0, -1.0), cmul(s.Vf, cconj(cmul(s.yft, s.Vt))));
      dSt_m = cdiv(cadd(s.St, cmul(mk(vt2, 0.0), cconj(s.ytt))), mk(vt, 0.0));
      dSf_m = cdiv(cmul(s.Vf, cconj(cmul(s.yft, s.Vt))), mk(vt, 0.0));
    }
    ga += cmul(s.af, dSf_a).x + cmul(s.at, dSt_a).x;
    gm += cmul(s.af, dSf_m).x + cmul(s.at, dSt_m).x;
  }
  double *Gt = G + (size_t)t * 2 * g.nb;
  Gt[2 * k] = ga;
  Gt[2 * k + 1] = (d.type[(size_t)t * g.n + i] == 1) ? gm : 0.0;
}

// J^T lambda = G with the FP64 factors: U^T z = G (forward levels, U blocks
// reached through the mirror map), then L^T lambda = z (backward levels).
// One workgroup per instance.
__global__ void k_adj_solve(DevTopo g, DevInst d, const B4<double> *LU,
                            const B4<double> *TL, const double *G, double *lam,
                            SolvePlan P, int use_lds, double *xglob) {
  __shared__ double smem[SOLVE_LDS_BYTES / sizeof(double)];
  __shared__ double rs[2 * TAIL_MAX];
  const int t = blockIdx.x;
  if (d.st[t].done) return;
  double *X = use_lds ? smem : xglob + (size_t)t * 2 * g.nb;
  const B4<double> *A = LU + (size_t)t * g.nnzLU;
  const B4<double> *Ti = TL + (size_t)t * P.tn * P.tn;
  const double *Gt = G + (size_t)t * 2 * g.nb;
  for (int k = threadIdx.x; k < 2 * g.nb; k += blockDim.x) X[k] = Gt[k];
  __syncthreads();
  // U^T z = g over the non-tail rows (forward levels)
  for (int l = 0; l < P.nfw; l++) {
    for (int r = P.fw_ptr[l] + threadIdx.x; r < P.fw_ptr[l + 1]; r += blockDim.x) {
      int i = P.fw_rows[r];
      double x0 = X[2 * i], x1 = X[2 * i + 1];
      for (int p = g.lrow[i]; p < g.ldiag[i]; p++) {
        int c = g.lcol[p];
        B4<double> U = A[g.mirror[p]];   // block (c, i) of U
        double y0 = X[2 * c], y1 = X[2 * c + 1];
        x0 -= U.a * y0 + U.c * y1;
        x1 -= U.b * y0 + U.d * y1;
      }
      B4<double> D = A[g.ldiag[i]];     // inverse pivot, used transposed
      X[2 * i] = D.a * x0 + D.c * x1;
      X[2 * i + 1] = D.b * x0 + D.d * x1;
    }
    __syncthreads();
  }
  // tail: lambda_S = S^-T (z_S - U_NS^T z_N)
  if (P.tn > 0) tail_rows<double>(A, Ti, X, rs, P, true);
  // L^T lambda = z over the non-tail rows (backward levels)
  for (int l = 0; l < P.nbw; l++) {
    for (int r = P.bw_ptr[l] + threadIdx.x; r < P.bw_ptr[l + 1]; r += blockDim.x) {
      int i = P.bw_rows[r];
      double x0 = X[2 * i], x1 = X[2 * i + 1];
      for (int p = g.ldiag[i] + 1; p < g.lrow[i + 1]; p++) {
        int c = g.lcol[p];
        B4<double> L = A[g.mirror[p]];   // block (c, i) of L
        double y0 = X[2 * c], y1 = X[2 * c + 1];
        x0 -= L.a * y0 + L.c * y1;
        x1 -= L.b * y0 + L.d * y1;
      }
      X[2 * i] = x0;
      X[2 * i + 1] = x1;
    }
    __syncthreads();
  }
  double *lt = lam + (size_t)t * 2 * g.nb;
  for (int k = threadIdx.x; k < 2 * g.nb; k += blockDim.x) lt[k] = X[k];
}

// gradients: df/dP_i = lambda_P(i), df/dQ_i = lambda_Q(i) (PQ buses), and
// df/dx_b = explicit - lambda^T dF/dx_b through dYs/dx = -j Ys^2 (in service)
// the adjoint solve J^T lambda = g in the slot form (X in LDS): the
// transposed slots of the Newton sweeps, U^T forward with the pivots, the
// tail by its transposed inverse, L^T backward; one workgroup per instance
template <int SP>
__global__ void __launch_bounds__(SLOT_CHUNK / SP)
k_adj_solve_slot(DevTopo g, DevInst d, const B4<double> *LU, const B4<double> *TL,
                 const double *G, double *lam, SolvePlan P) {
  __shared__ double Xs[SOLVE_LDS_BYTES / sizeof(double)];
  __shared__ double rs[2 * TAIL_MAX];
  __shared__ int4 s_swc[MAX_CHUNKS];
  const int t = blockIdx.x;
  if (d.st[t].done) return;
  // X after a zero pair (read by slots without an entry)
  double *X = Xs + SLOT_X_PAD;
  if (threadIdx.x < SLOT_X_PAD) Xs[threadIdx.x] = 0.0;
  const B4<double> *A = LU + (size_t)t * g.nnzLU;
  const B4<double> *Ti = TL + (size_t)t * P.tn * P.tn;
  const double *Gt = G + (size_t)t * 2 * g.nb;
  for (int q = threadIdx.x; q < P.sl_nfw_t + P.sl_nbw_t; q += blockDim.x)
    s_swc[q] = P.sl_chunk_t[q];
  for (int k = threadIdx.x; k < 2 * g.nb; k += blockDim.x) X[k] = Gt[k];
  __syncthreads();
  lu_sweeps_slot<double, SP, true>(A, Ti, X, rs, s_swc, P, false);
  __syncthreads();
  double *lt = lam + (size_t)t * 2 * g.nb;
  for (int k = threadIdx.x; k < 2 * g.nb; k += blockDim.x) lt[k] = X[k];
}

__global__ void k_adj_grad(DevTopo g, DevInst d, const double *lam,
                           const double *rate, double w, double *gP,
                           double *gQ, double *gx) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  int t = blockIdx.y;
  const bool ok = d.st[t].verdict == BPF_V_CONVERGED;
  const double *lt = lam + (size_t)t * 2 * g.nb;
  if (e < g.n) {
    int k = g.blk_of_bus[e];
    size_t o = (size_t)t * g.n + e;
    gP[o] = ok ? (k >= 0 ? lt[2 * k] : 0.0) : NAN;
    gQ[o] = ok ? ((k >= 0 && d.type[o] == 1) ? lt[2 * k + 1] : 0.0) : NAN;
  }
  if (e >= g.m) return;
  size_t ob = (size_t)t * g.m + e;
  if (!ok) { gx[ob] = NAN; return; }
  if (!d.br_on[ob]) { gx[ob] = 0.0; return; }
  BranchState s = branch_state(g, d, t, e, rate, w);
  cd Ys = cdiv(mk(1.0, 0.0), mk(g.R[e], g.X[e]));
  cd dYs = cmul(mk(0.0, -1.0), cmul(Ys, Ys));
  double tp = d.tap[ob];
  cd tap = (tp != 0.0) ? mk(tp, 0.0) : mk(1.0, 0.0);
  tap = cmul(tap, cexpj(M_PI / 180.0 * g.shift[e]));
  cd dyff = cdiv(dYs, cmul(tap, cconj(tap)));
  cd dyft = cdiv(cmul(dYs, mk(-1.0, 0.0)), cconj(tap));
  cd dytf = cdiv(cmul(dYs, mk(-1.0, 0.0)), tap);
  cd dSf = cmul(s.Vf, cconj(cadd(cmul(dyff, s.Vf), cmul(dyft, s.Vt))));
  cd dSt = cmul(s.Vt, cconj(cadd(cmul(dytf, s.Vf), cmul(dYs, s.Vt))));
  double val = cmul(s.af, dSf).x + cmul(s.at, dSt).x;
  int bf = g.blk_of_bus[g.br_from[e]], bt = g.blk_of_bus[g.br_to[e]];
  if (bf >= 0) {
    double lq = d.type[(size_t)t * g.n + g.br_from[e]] == 1 ? lt[2 * bf + 1] : 0.0;
    val -= lt[2 * bf] * dSf.x + lq * dSf.y;
  }
  if (bt >= 0) {
    double lq = d.type[(size_t)t * g.n + g.br_to[e]] == 1 ? lt[2 * bt + 1] : 0.0;
    val -= lt[2 * bt] * dSt.x + lq * dSt.y;
  }
  gx[ob] = val;
}

// objective value per instance (one workgroup per instance)
__global__ void k_objective(DevTopo g, DevInst d, const double *rate, double w,
                            double *obj) {
  __shared__ double sh[16];
  const int t = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < g.m; b += blockDim.x) {
    BranchState s = branch_state(g, d, t, b, rate, w);
    acc += cadd(s.Sf, s.St).x;
    if (rate) {
      double ef = cabs_(s.Sf) - rate[b], et = cabs_(s.St) - rate[b];
      if (ef > 0) acc += w * ef * ef;
      if (et > 0) acc += w * et * et;
    }
  }
  double tot = block_sum(acc, sh);
  if (threadIdx.x == 0)
    obj[t] = (d.st[t].verdict == BPF_V_CONVERGED) ? tot : NAN;
}

// ---------------------------------------------------------------------------
// Losses and branch flows (computeLoss + reduceLoss, runpf.h:543-604; CPU
// runpf.h:228-257).  One workgroup per instance.
// ---------------------------------------------------------------------------
__global__ void k_loss(DevTopo g, DevInst d, cd *Sf, cd *St, double *loss) {
  __shared__ double sh[16];
  const int t = blockIdx.x;
  const cd *V = d.V + (size_t)t * g.n;
  double acc = 0.0;
  for (int b = threadIdx.x; b < g.m; b += blockDim.x) {
    int f = g.br_from[b], to = g.br_to[b];
    cd yff = branch_entry(g, d, t, b, KFF), yft = branch_entry(g, d, t, b, KFT);
    cd ytf = branch_entry(g, d, t, b, KTF), ytt = branch_entry(g, d, t, b, KTT);
    cd If, It;
    if (f < to) {
      If = cadd(cmul(yff, V[f]), cmul(yft, V[to]));
      It = cadd(cmul(ytf, V[f]), cmul(ytt, V[to]));
    } else {
      If = cadd(cmul(yft, V[to]), cmul(yff, V[f]));
      It = cadd(cmul(ytt, V[to]), cmul(ytf, V[f]));
    }
    cd sf = cmul(cconj(If), V[f]);
    cd s2 = cmul(cconj(It), V[to]);
    Sf[(size_t)t * g.m + b] = cconj(sf);
    St[(size_t)t * g.m + b] = cconj(s2);
    acc += cadd(sf, s2).x;
  }
  double tot = block_sum(acc, sh);
  if (threadIdx.x == 0)
    loss[t] = (d.st[t].verdict == BPF_V_CONVERGED) ? tot : DBL_MAX;
}

// ---------------------------------------------------------------------------
// Branch loading report from the flows of k_loss: per branch
// 100 max(|Sf|, |St|) / rating (0 for a branch without a rating or out of
// service, whose flows are zero), per instance the number of branches above
// 100 % and the largest loading.  One workgroup per instance.
// ---------------------------------------------------------------------------
__global__ void k_branch_report(DevTopo g, int T, const cd *Sf, const cd *St,
                                const double *rate, double *loading, int *over,
                                double *maxl) {
  __shared__ double shd[16];
  __shared__ int shi[16];
  const int t = blockIdx.x;
  if (t >= T) return;
  double lmax = 0.0;
  int cnt = 0;
  for (int b = threadIdx.x; b < g.m; b += blockDim.x) {
    const size_t o = (size_t)t * g.m + b;
    const double r = rate[b];
    const double s = fmax(cabs_(Sf[o]), cabs_(St[o]));
    const double l = r > 0.0 ? 100.0 * s / r : 0.0;
    if (loading) loading[o] = l;
    lmax = fmax(lmax, l);
    cnt += l > 100.0 ? 1 : 0;
  }
  const double bm = block_max(lmax, shd);
  const int bc = block_add(cnt, shi);
  if (threadIdx.x == 0) {
    over[t] = bc;
    maxl[t] = bm;
  }
}

// T rows of per-bus values from one bus numbering to another: dst[t][k] =
// src[t][map[k]] (the interface's per-bus inputs and outputs)
template <typename V>
__global__ void k_bus_gather(const V *src, V *dst, const int *map, int n, int T) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)n * T) return;
  const size_t t = q / (size_t)n, k = q - t * (size_t)n;
  dst[q] = src[t * (size_t)n + map[k]];
}

// explicit instantiations used by the engine
template __global__ void k_bus_gather<cd>(const cd *, cd *, const int *, int, int);
template __global__ void k_bus_gather<double>(const double *, double *, const int *, int, int);
template __global__ void k_bus_gather<int8_t>(const int8_t *, int8_t *, const int *, int, int);
template __global__ void k_jacobian<float>(DevTopo, DevInst, B4<float> *, int, int);
template __global__ void k_jacobian<double>(DevTopo, DevInst, B4<double> *, int, int);
template __global__ void k_lu_factor<float>(DevTopo, InstState *, int, FactorPlan, B4<float> *, B4<float> *, int, int *, int, const int *, int, int);
template __global__ void k_lu_factor<double>(DevTopo, InstState *, int, FactorPlan, B4<double> *, B4<double> *, int, int *, int, const int *, int, int);
template __global__ void k_lu_factor_top<float>(DevTopo, InstState *, int, FactorPlan, B4<float> *, B4<float> *, int, int *, const int *, int, int);
template __global__ void k_lu_factor_top<double>(DevTopo, InstState *, int, FactorPlan, B4<double> *, B4<double> *, int, int *, const int *, int, int);
template __global__ void k_lu_solve<float, 0>(DevTopo, DevInst, const B4<float> *, const B4<float> *, int, int, SolvePlan, int, int, float *);
template __global__ void k_lu_solve<float, 1>(DevTopo, DevInst, const B4<float> *, const B4<float> *, int, int, SolvePlan, int, int, float *);
template __global__ void k_lu_solve<float, 2>(DevTopo, DevInst, const B4<float> *, const B4<float> *, int, int, SolvePlan, int, int, float *);
template __global__ void k_lu_solve<float, 4>(DevTopo, DevInst, const B4<float> *, const B4<float> *, int, int, SolvePlan, int, int, float *);
template __global__ void k_lu_solve<float, 8>(DevTopo, DevInst, const B4<float> *, const B4<float> *, int, int, SolvePlan, int, int, float *);
template __global__ void k_lu_solve<double, 0>(DevTopo, DevInst, const B4<double> *, const B4<double> *, int, int, SolvePlan, int, int, double *);
template __global__ void k_lu_solve<double, 1>(DevTopo, DevInst, const B4<double> *, const B4<double> *, int, int, SolvePlan, int, int, double *);
template __global__ void k_lu_solve<double, 2>(DevTopo, DevInst, const B4<double> *, const B4<double> *, int, int, SolvePlan, int, int, double *);
template __global__ void k_lu_solve<double, 4>(DevTopo, DevInst, const B4<double> *, const B4<double> *, int, int, SolvePlan, int, int, double *);
template __global__ void k_lu_solve<double, 8>(DevTopo, DevInst, const B4<double> *, const B4<double> *, int, int, SolvePlan, int, int, double *);
template __global__ void k_adj_solve_slot<4>(DevTopo, DevInst, const B4<double> *, const B4<double> *, const double *, double *, SolvePlan);

}  // namespace bpf
