/*
 * CPU oracle for the batched AC power-flow path (TEST INFRASTRUCTURE ONLY;
 * see pf_oracle.h for the scope rule and the reference functions restated).
 * FP64 throughout (an 80-bit build of the same source exists for gradient
 * checks, see ORC_EXTENDED below); compiled with -ffp-contract=off so every
 * expression is evaluated exactly as written.
 */
#include "pf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Working precision.  The default build is FP64 like the reference; the
 * ORC_EXTENDED build (liborcpf_x.so) runs the same code in x87 80-bit long
 * double, used only where the tests need a re-solve whose rounding is far
 * below FP64 (finite-difference checks of gradients).  The public entry
 * points keep FP64 arrays in both builds. */
#ifdef ORC_EXTENDED
typedef long double real;
#define R_hypot hypotl
#define R_fabs fabsl
#define R_cos cosl
#define R_sin sinl
#define R_atan atanl
#define R_sqrt sqrtl
#define R_fmax fmaxl
#define R_PI M_PIl
#else
typedef double real;
#define R_hypot hypot
#define R_fabs fabs
#define R_cos cos
#define R_sin sin
#define R_atan atan
#define R_sqrt sqrt
#define R_fmax fmax
#define R_PI M_PI
#endif

/* ------------------------------------------------------------------------- */
/* complex helpers (the reference uses cuComplex; these are written from the
 * definitions; division is Smith's 1962 scaled algorithm)                     */
/* ------------------------------------------------------------------------- */
typedef struct { real re, im; } cpx;

static cpx c_make(real re, real im) { cpx z; z.re = re; z.im = im; return z; }
static cpx c_add(cpx a, cpx b) { return c_make(a.re + b.re, a.im + b.im); }
static cpx c_sub(cpx a, cpx b) { return c_make(a.re - b.re, a.im - b.im); }
static cpx c_mul(cpx a, cpx b) {
  return c_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static cpx c_conj(cpx a) { return c_make(a.re, -a.im); }
static real c_abs(cpx a) { return R_hypot(a.re, a.im); }
static cpx c_div(cpx n, cpx d) {
  if (R_fabs(d.re) >= R_fabs(d.im)) {
    real r = d.im / d.re, den = d.re + d.im * r;
    return c_make((n.re + n.im * r) / den, (n.im - n.re * r) / den);
  } else {
    real r = d.re / d.im, den = d.re * r + d.im;
    return c_make((n.re * r + n.im) / den, (n.im * r - n.re) / den);
  }
}
static cpx c_expj(real th) { return c_make(R_cos(th), R_sin(th)); }
/* the reference's angle is R_atan(y/x), not atan2 (complexUtils.h:21-23) */
static real c_angle(cpx a) { return R_atan(a.im / a.re); }

/* ------------------------------------------------------------------------- */
/* per-instance resolved data                                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
  int n, m, g;
  int *type;        /* effective bus type */
  real *P, *Q;    /* net injections (pu) */
  real *vset;     /* set-point magnitude used for V0 */
  unsigned char *on;/* effective branch status */
  real *tap;      /* effective tap ratio (tap field, 0 => 1) */
  real *shift;    /* degrees */
  real *R, *X, *B;
  real *Gs, *Bs;
  real *Qmax, *Qmin; /* per bus, summed over in-service generators */
  real *Qd;
} inst_t;

static void inst_free(inst_t *I) {
  free(I->type); free(I->P); free(I->Q); free(I->vset); free(I->on);
  free(I->tap); free(I->shift); free(I->R); free(I->X); free(I->B);
  free(I->Gs); free(I->Bs); free(I->Qmax); free(I->Qmin); free(I->Qd);
}

static void inst_build(const orc_case *c, const orc_inst *in, inst_t *I) {
  int n = c->nbus, m = c->nbranch, g = c->ngen;
  I->n = n; I->m = m; I->g = g;
  I->type = (int *)malloc(sizeof(int) * n);
  I->P = (real *)calloc(n, sizeof(real));
  I->Q = (real *)calloc(n, sizeof(real));
  I->vset = (real *)malloc(sizeof(real) * n);
  I->on = (unsigned char *)malloc(m);
  I->tap = (real *)malloc(sizeof(real) * m);
  I->shift = (real *)malloc(sizeof(real) * m);
  I->R = (real *)malloc(sizeof(real) * m);
  I->X = (real *)malloc(sizeof(real) * m);
  I->B = (real *)malloc(sizeof(real) * m);
  I->Gs = (real *)malloc(sizeof(real) * n);
  I->Bs = (real *)malloc(sizeof(real) * n);
  I->Qmax = (real *)calloc(n, sizeof(real));
  I->Qmin = (real *)calloc(n, sizeof(real));
  I->Qd = (real *)malloc(sizeof(real) * n);
  int *ngen_on = (int *)calloc(n, sizeof(int));
  int *ngen_any = (int *)calloc(n, sizeof(int));
  real *vg = (real *)malloc(sizeof(real) * n);
  for (int i = 0; i < n; i++) {
    I->type[i] = c->bus_type[i];
    I->P[i] = -c->Pd[i];
    I->Q[i] = -c->Qd[i];
    I->Qd[i] = c->Qd[i];
    I->Gs[i] = c->Gs[i];
    I->Bs[i] = (in && in->bsh && !isnan(in->bsh[i])) ? in->bsh[i] : c->Bs[i];
    vg[i] = NAN;
  }
  /* generation is added to the bus after the load, as the MATPOWER loader
   * does (loadMatpowerFormat.h:88,126-128) */
  for (int k = 0; k < g; k++) {
    int b = c->gen_bus[k];
    int on = c->gen_status[k] && (!in || !in->gen_on || in->gen_on[k]);
    ngen_any[b]++;
    if (!on) continue;
    ngen_on[b]++;
    I->P[b] += c->Pg[k];
    I->Q[b] += c->Qg[k];
    I->Qmax[b] += c->Qmax[k];
    I->Qmin[b] += c->Qmin[k];
    vg[b] = c->VG[k]; /* the last generator row wins (loadMatpowerFormat.h:139-140) */
  }
  for (int i = 0; i < n; i++) {
    /* a PV bus whose generators are all out becomes PQ */
    if (I->type[i] == ORC_PV && ngen_on[i] == 0) I->type[i] = ORC_PQ;
    real v = c->Vm[i];
    if (c->bus_type[i] == ORC_PV && !isnan(vg[i])) v = vg[i];
    if (in && in->vset && !isnan(in->vset[i])) v = in->vset[i];
    I->vset[i] = v;
  }
  for (int b = 0; b < m; b++) {
    I->on[b] = c->br_status[b] && (!in || !in->br_on || in->br_on[b]);
    real t = c->tap[b];
    if (t != 0.0 && in && in->tap && !isnan(in->tap[b])) t = in->tap[b];
    I->tap[b] = t;
    I->shift[b] = c->shift[b];
    I->R[b] = c->R[b]; I->X[b] = c->X[b]; I->B[b] = c->B[b];
  }
  free(ngen_on); free(ngen_any); free(vg);
}

/* ------------------------------------------------------------------------- */
/* Ybus, Yf, Yt (makeYbus.h:17-274)                                           */
/* ------------------------------------------------------------------------- */
typedef struct {
  int n, nnz;
  int *rowptr, *colind;
  cpx *val;
  /* branch pi-model, per branch: Yff, Yft, Ytf, Ytt */
  cpx *yff, *yft, *ytf, *ytt;
} ybus_t;

static void ybus_free(ybus_t *Y) {
  free(Y->rowptr); free(Y->colind); free(Y->val);
  free(Y->yff); free(Y->yft); free(Y->ytf); free(Y->ytt);
}

/* pi-model of one branch (makeYbus.h:63-80) */
static void branch_pi(const inst_t *I, int b, cpx *Yff, cpx *Yft, cpx *Ytf,
                      cpx *Ytt) {
  real stat = I->on[b] ? 1.0 : 0.0;
  cpx Ys = c_div(c_make(stat, 0.0), c_make(I->R[b], I->X[b]));
  cpx Bc = c_mul(c_make(stat, 0.0), c_make(0.0, I->B[b]));
  cpx tap = (I->tap[b] != 0.0) ? c_make(I->tap[b], 0.0) : c_make(1.0, 0.0);
  tap = c_mul(tap, c_expj(R_PI / 180.0 * I->shift[b]));
  *Ytt = c_add(Ys, c_div(Bc, c_make(2.0, 0.0)));
  *Yff = c_div(*Ytt, c_mul(tap, c_conj(tap)));
  *Yft = c_div(c_mul(Ys, c_make(-1.0, 0.0)), c_conj(tap));
  *Ytf = c_div(c_mul(Ys, c_make(-1.0, 0.0)), tap);
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

static int find_col(const int *colind, int lo, int hi, int j) {
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (colind[mid] < j) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* Ybus = Cf*Yf + Ct*Yt + Ysh with the reference's summation order: per entry,
 * the from-side terms in branch order, then the to-side terms in branch
 * order, then the shunt (MKL csrmultcsr row accumulation then csradd). */
static void ybus_build(const inst_t *I, const int *from, const int *to,
                       ybus_t *Y) {
  int n = I->n, m = I->m;
  Y->n = n;
  int *cnt = (int *)calloc(n + 1, sizeof(int));
  int *tmp = (int *)malloc(sizeof(int) * (n + 2 * m));
  int *start = (int *)malloc(sizeof(int) * (n + 1));
  /* candidate columns per row: diagonal + both ends */
  for (int i = 0; i < n; i++) cnt[i] = 1;
  for (int b = 0; b < m; b++) { cnt[from[b]]++; cnt[to[b]]++; }
  start[0] = 0;
  for (int i = 0; i < n; i++) start[i + 1] = start[i] + cnt[i];
  int *fill = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) { tmp[start[i]] = i; fill[i] = start[i] + 1; }
  for (int b = 0; b < m; b++) {
    tmp[fill[from[b]]++] = to[b];
    tmp[fill[to[b]]++] = from[b];
  }
  Y->rowptr = (int *)malloc(sizeof(int) * (n + 1));
  Y->rowptr[0] = 0;
  int nnz = 0;
  for (int i = 0; i < n; i++) {
    int *seg = tmp + start[i];
    int len = start[i + 1] - start[i];
    qsort(seg, len, sizeof(int), cmp_int);
    int u = 0;
    for (int k = 0; k < len; k++)
      if (k == 0 || seg[k] != seg[k - 1]) seg[u++] = seg[k];
    cnt[i] = u;
    nnz += u;
    Y->rowptr[i + 1] = nnz;
  }
  Y->nnz = nnz;
  Y->colind = (int *)malloc(sizeof(int) * nnz);
  for (int i = 0; i < n; i++)
    memcpy(Y->colind + Y->rowptr[i], tmp + start[i], sizeof(int) * cnt[i]);
  Y->yff = (cpx *)malloc(sizeof(cpx) * m);
  Y->yft = (cpx *)malloc(sizeof(cpx) * m);
  Y->ytf = (cpx *)malloc(sizeof(cpx) * m);
  Y->ytt = (cpx *)malloc(sizeof(cpx) * m);
  for (int b = 0; b < m; b++)
    branch_pi(I, b, &Y->yff[b], &Y->yft[b], &Y->ytf[b], &Y->ytt[b]);
  /* accumulate CfYf and CtYt separately, then add, then add Ysh */
  cpx *cf = (cpx *)calloc(nnz, sizeof(cpx));
  cpx *ct = (cpx *)calloc(nnz, sizeof(cpx));
  for (int b = 0; b < m; b++) {
    int f = from[b], t = to[b];
    int pff = find_col(Y->colind, Y->rowptr[f], Y->rowptr[f + 1], f);
    int pft = find_col(Y->colind, Y->rowptr[f], Y->rowptr[f + 1], t);
    int ptt = find_col(Y->colind, Y->rowptr[t], Y->rowptr[t + 1], t);
    int ptf = find_col(Y->colind, Y->rowptr[t], Y->rowptr[t + 1], f);
    cf[pff] = c_add(cf[pff], Y->yff[b]);
    cf[pft] = c_add(cf[pft], Y->yft[b]);
    ct[ptt] = c_add(ct[ptt], Y->ytt[b]);
    ct[ptf] = c_add(ct[ptf], Y->ytf[b]);
  }
  Y->val = (cpx *)malloc(sizeof(cpx) * nnz);
  for (int k = 0; k < nnz; k++) Y->val[k] = c_add(cf[k], ct[k]);
  for (int i = 0; i < n; i++) {
    int p = find_col(Y->colind, Y->rowptr[i], Y->rowptr[i + 1], i);
    Y->val[p] = c_add(Y->val[p], c_make(I->Gs[i], I->Bs[i]));
  }
  free(cf); free(ct); free(cnt); free(tmp); free(start); free(fill);
}

/* ------------------------------------------------------------------------- */
/* bus ordering (Topology.h:100-126)                                           */
/* ------------------------------------------------------------------------- */
typedef struct {
  int npv, npq;
  int *pv, *pq;  /* ascending bus index */
  int *idx;      /* indicePVPQ per bus, -1 for slack */
} order_t;

static void order_build(const int *type, int n, order_t *O) {
  O->pv = (int *)malloc(sizeof(int) * n);
  O->pq = (int *)malloc(sizeof(int) * n);
  O->idx = (int *)malloc(sizeof(int) * n);
  O->npv = O->npq = 0;
  for (int i = 0; i < n; i++) {
    if (type[i] == ORC_PV) O->pv[O->npv++] = i;
    else if (type[i] == ORC_PQ) O->pq[O->npq++] = i;
  }
  for (int i = 0; i < n; i++) O->idx[i] = -1;
  for (int k = 0; k < O->npv; k++) O->idx[O->pv[k]] = k;
  for (int k = 0; k < O->npq; k++) O->idx[O->pq[k]] = O->npv + k;
}
static void order_free(order_t *O) { free(O->pv); free(O->pq); free(O->idx); }

/* ------------------------------------------------------------------------- */
/* voltage (runpf.h:211-226), mismatch (newtonpf.h:18-71), Ibus (:73-95)       */
/* ------------------------------------------------------------------------- */
static void initial_voltage(const orc_case *c, const inst_t *I, cpx *V) {
  for (int i = 0; i < I->n; i++) {
    real Vbus = I->vset[i];
    V[i] = c_mul(c_make(Vbus, 0.0), c_expj(c->Va[i]));
    if (c->bus_type[i] == ORC_PV || c->bus_type[i] == ORC_SLACK)
      V[i] = c_mul(c_make(Vbus / c_abs(V[i]), 0.0), V[i]);
  }
}

static cpx row_current(const ybus_t *Y, const cpx *V, int i) {
  cpx acc = c_make(0.0, 0.0);
  for (int k = Y->rowptr[i]; k < Y->rowptr[i + 1]; k++)
    acc = c_add(acc, c_mul(Y->val[k], V[Y->colind[k]]));
  return acc;
}

/* NR mismatch; returns inf-norm; NaN-aware (a NaN entry yields NaN norm) */
static real mismatch_nr(const ybus_t *Y, const inst_t *I, const order_t *O,
                          const cpx *V, real *F) {
  real err = 0.0;
  int npv = O->npv, npq = O->npq;
  int nan = 0;
  for (int id = 0; id < npv + npq; id++) {
    int i = (id < npv) ? id : id - npv;
    int bus = (id < npv) ? O->pv[i] : O->pq[i];
    cpx cc = row_current(Y, V, bus);
    cpx miss = c_sub(c_mul(V[bus], c_conj(cc)), c_make(I->P[bus], I->Q[bus]));
    if (id < npv) {
      F[i] = miss.re;
      err = R_fmax(err, R_fabs(miss.re));
      nan |= isnan(miss.re);
    } else {
      F[npv + i] = miss.re;
      F[npv + npq + i] = miss.im;
      err = R_fmax(err, R_fabs(miss.re));
      err = R_fmax(err, R_fabs(miss.im));
      nan |= isnan(miss.re) | isnan(miss.im);
    }
  }
  return nan ? NAN : err;
}

/* ------------------------------------------------------------------------- */
/* Jacobian pattern (newtonpf.h:215-310) and values (newtonpf.h:97-186)       */
/* ------------------------------------------------------------------------- */
typedef struct {
  int nJ, nnz;
  int *rowptr, *colind, *ybpos; /* ybpos: Ybus position of each entry */
  real *val;
} jac_t;

static void jac_free(jac_t *J) {
  free(J->rowptr); free(J->colind); free(J->ybpos); free(J->val);
}

typedef struct { int r, c, y; } trip_t;
static int cmp_trip(const void *a, const void *b) {
  const trip_t *x = (const trip_t *)a, *y = (const trip_t *)b;
  if (x->r != y->r) return (x->r > y->r) - (x->r < y->r);
  return (x->c > y->c) - (x->c < y->c);
}

static void jac_pattern(const ybus_t *Y, const int *type, const order_t *O,
                        jac_t *J) {
  int n = Y->n, npv = O->npv, npq = O->npq;
  int nJ = npv + 2 * npq;
  int cnt = 0;
  for (int i = 0; i < n; i++)
    for (int k = Y->rowptr[i]; k < Y->rowptr[i + 1]; k++) {
      int j = Y->colind[k];
      int a = type[i], b = type[j];
      if ((a == ORC_PV || a == ORC_PQ) && (b == ORC_PV || b == ORC_PQ))
        cnt += (a == ORC_PQ ? 2 : 1) * (b == ORC_PQ ? 2 : 1);
    }
  trip_t *t = (trip_t *)malloc(sizeof(trip_t) * (cnt > 0 ? cnt : 1));
  int p = 0;
  for (int i = 0; i < n; i++)
    for (int k = Y->rowptr[i]; k < Y->rowptr[i + 1]; k++) {
      int j = Y->colind[k];
      int a = type[i], b = type[j];
      if (!((a == ORC_PV || a == ORC_PQ) && (b == ORC_PV || b == ORC_PQ)))
        continue;
      int ri = O->idx[i], cj = O->idx[j];
      t[p].r = ri; t[p].c = cj; t[p].y = k; p++;
      if (b == ORC_PQ) { t[p].r = ri; t[p].c = cj + npq; t[p].y = k; p++; }
      if (a == ORC_PQ) { t[p].r = ri + npq; t[p].c = cj; t[p].y = k; p++; }
      if (a == ORC_PQ && b == ORC_PQ) {
        t[p].r = ri + npq; t[p].c = cj + npq; t[p].y = k; p++;
      }
    }
  qsort(t, cnt, sizeof(trip_t), cmp_trip);
  J->nJ = nJ;
  J->nnz = cnt;
  J->rowptr = (int *)calloc(nJ + 1, sizeof(int));
  J->colind = (int *)malloc(sizeof(int) * (cnt > 0 ? cnt : 1));
  J->ybpos = (int *)malloc(sizeof(int) * (cnt > 0 ? cnt : 1));
  J->val = (real *)malloc(sizeof(real) * (cnt > 0 ? cnt : 1));
  for (int q = 0; q < cnt; q++) {
    J->rowptr[t[q].r + 1]++;
    J->colind[q] = t[q].c;
    J->ybpos[q] = t[q].y;
  }
  for (int r = 0; r < nJ; r++) J->rowptr[r + 1] += J->rowptr[r];
  free(t);
}

static void jac_values(const ybus_t *Y, const order_t *O, const cpx *V,
                       const cpx *Ibus, jac_t *J) {
  int npv = O->npv, npq = O->npq, length = npv + npq;
  for (int r = 0; r < J->nJ; r++) {
    int ii = (r < length) ? ((r < npv) ? O->pv[r] : O->pq[r - npv])
                          : O->pq[r - length];
    for (int q = J->rowptr[r]; q < J->rowptr[r + 1]; q++) {
      int cidx = J->colind[q];
      int jj = (cidx < length) ? ((cidx < npv) ? O->pv[cidx] : O->pq[cidx - npv])
                               : O->pq[cidx - length];
      cpx adm = Y->val[J->ybpos[q]];
      real yr = adm.re, yi = adm.im;
      real vrj = V[jj].re, vij = V[jj].im;
      real ibr = (ii == jj) ? Ibus[ii].re : 0.0;
      real ibi = (ii == jj) ? Ibus[ii].im : 0.0;
      real vri = V[ii].re, vii = V[ii].im;
      real v;
      if (r < length) {
        if (cidx < length) {
          real re = yr * vrj - yi * vij;
          real im = yr * vij + yi * vrj;
          v = -vii * (ibr - re) - vri * (-ibi + im);
        } else {
          real ab = R_sqrt(vrj * vrj + vij * vij);
          real re = yr * vrj / ab - yi * vij / ab;
          real im = yr * vij / ab + yi * vrj / ab;
          v = vri * re - vii * -im + ibr * vrj / ab + ibi * vij / ab;
        }
      } else {
        if (cidx < length) {
          real re = yr * vrj - yi * vij;
          real im = yr * vij + yi * vrj;
          v = -vii * (-ibi + im) + vri * (ibr - re);
        } else {
          real ab = R_sqrt(vrj * vrj + vij * vij);
          real re = yr * vrj / ab - yi * vij / ab;
          real im = yr * vij / ab + yi * vrj / ab;
          v = vri * -im + vii * re + ibr * vij / ab + -ibi * vrj / ab;
        }
      }
      J->val[q] = v;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* sparse LU: minimum degree on A+A^T, left-looking LU, partial pivoting     */
/* ------------------------------------------------------------------------- */
typedef struct { int *a; int len, cap; } ivec;

static void iv_push(ivec *v, int x) {
  if (v->len == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 8;
    v->a = (int *)realloc(v->a, sizeof(int) * v->cap);
  }
  v->a[v->len++] = x;
}

void orc_min_degree(int n, const int *rowptr, const int *colind, int *perm) {
  ivec *adj = (ivec *)calloc(n, sizeof(ivec));
  for (int i = 0; i < n; i++)
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      int j = colind[k];
      if (j == i) continue;
      iv_push(&adj[i], j);
      iv_push(&adj[j], i);
    }
  for (int i = 0; i < n; i++) {
    qsort(adj[i].a, adj[i].len, sizeof(int), cmp_int);
    int u = 0;
    for (int k = 0; k < adj[i].len; k++)
      if (k == 0 || adj[i].a[k] != adj[i].a[k - 1]) adj[i].a[u++] = adj[i].a[k];
    adj[i].len = u;
  }
  int *head = (int *)malloc(sizeof(int) * (n + 1));
  int *next = (int *)malloc(sizeof(int) * n);
  int *prev = (int *)malloc(sizeof(int) * n);
  int *deg = (int *)malloc(sizeof(int) * n);
  char *gone = (char *)calloc(n, 1);
  for (int d = 0; d <= n; d++) head[d] = -1;
#define MD_INSERT(v)                                                           \
  do {                                                                         \
    int d_ = deg[v];                                                           \
    next[v] = head[d_]; prev[v] = -1;                                          \
    if (head[d_] >= 0) prev[head[d_]] = v;                                     \
    head[d_] = v;                                                              \
  } while (0)
#define MD_REMOVE(v)                                                           \
  do {                                                                         \
    if (prev[v] >= 0) next[prev[v]] = next[v]; else head[deg[v]] = next[v];    \
    if (next[v] >= 0) prev[next[v]] = prev[v];                                 \
  } while (0)
  for (int v = n - 1; v >= 0; v--) { deg[v] = adj[v].len; MD_INSERT(v); }
  int mindeg = 0;
  int *merged = (int *)malloc(sizeof(int) * (n + 1));
  for (int k = 0; k < n; k++) {
    while (head[mindeg] < 0) mindeg++;
    int v = head[mindeg];
    MD_REMOVE(v);
    gone[v] = 1;
    perm[k] = v;
    ivec N = adj[v];
    for (int q = 0; q < N.len; q++) {
      int u = N.a[q];
      MD_REMOVE(u);
      /* adj[u] <- (adj[u] U N) \ {u, v} */
      int a = 0, b = 0, out = 0;
      ivec *A = &adj[u];
      while (a < A->len || b < N.len) {
        int x;
        if (b >= N.len || (a < A->len && A->a[a] < N.a[b])) x = A->a[a++];
        else if (a >= A->len || N.a[b] < A->a[a]) x = N.a[b++];
        else { x = A->a[a]; a++; b++; }
        if (x == u || x == v) continue;
        merged[out++] = x;
      }
      if (out > A->cap) { A->cap = out; A->a = (int *)realloc(A->a, sizeof(int) * out); }
      memcpy(A->a, merged, sizeof(int) * out);
      A->len = out;
      deg[u] = out;
      MD_INSERT(u);
      if (out < mindeg) mindeg = out;
    }
    free(adj[v].a);
    adj[v].a = NULL; adj[v].len = adj[v].cap = 0;
  }
#undef MD_INSERT
#undef MD_REMOVE
  free(merged); free(head); free(next); free(prev); free(deg); free(gone);
  free(adj);
}

typedef struct {
  int n;
  int *q, *pinv;
  int *Lp, *Li; real *Lx; int lcap;
  int *Up, *Ui; real *Ux; int ucap;
} lu_t;

static void lu_free(lu_t *F) {
  free(F->q); free(F->pinv); free(F->Lp); free(F->Li); free(F->Lx);
  free(F->Up); free(F->Ui); free(F->Ux);
}

/* Factor A (CSC: Ap, Ai, Ax) with column order q.  Returns 0 or ORC_SINGULAR /
 * ORC_NONFINITE. */
static int lu_factor(int n, const int *Ap, const int *Ai, const real *Ax,
                     const int *q, real thresh, lu_t *F) {
  F->n = n;
  F->q = (int *)malloc(sizeof(int) * n);
  memcpy(F->q, q, sizeof(int) * n);
  F->pinv = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) F->pinv[i] = -1;
  F->lcap = 4 * Ap[n] + n; F->ucap = 4 * Ap[n] + n;
  F->Lp = (int *)malloc(sizeof(int) * (n + 1));
  F->Li = (int *)malloc(sizeof(int) * F->lcap);
  F->Lx = (real *)malloc(sizeof(real) * F->lcap);
  F->Up = (int *)malloc(sizeof(int) * (n + 1));
  F->Ui = (int *)malloc(sizeof(int) * F->ucap);
  F->Ux = (real *)malloc(sizeof(real) * F->ucap);
  real *x = (real *)calloc(n, sizeof(real));
  int *xi = (int *)malloc(sizeof(int) * n);
  int *stk = (int *)malloc(sizeof(int) * n);
  int *pos = (int *)malloc(sizeof(int) * n);
  int *mark = (int *)calloc(n, sizeof(int));
  int lnz = 0, unz = 0, status = 0;
  for (int k = 0; k < n; k++) {
    int col = q[k];
    F->Lp[k] = lnz; F->Up[k] = unz;
    /* reach of A(:,col) in the graph of L: rows in topological order */
    int top = n;
    for (int p = Ap[col]; p < Ap[col + 1]; p++) {
      int r0 = Ai[p];
      if (mark[r0] == k + 1) continue;
      int head = 0;
      stk[0] = r0; mark[r0] = k + 1; pos[0] = -1;
      while (head >= 0) {
        int r = stk[head];
        int j = F->pinv[r];
        if (pos[head] < 0) pos[head] = (j >= 0) ? F->Lp[j] : 0;
        int end = (j >= 0) ? F->Lp[j + 1] : 0;
        int done = 1;
        if (j >= 0) {
          for (int t = pos[head]; t < end; t++) {
            int ch = F->Li[t];
            if (mark[ch] == k + 1) continue;
            pos[head] = t + 1;
            mark[ch] = k + 1;
            stk[++head] = ch; pos[head] = -1;
            done = 0;
            break;
          }
        }
        if (done) { xi[--top] = r; head--; }
      }
    }
    for (int p = top; p < n; p++) x[xi[p]] = 0.0;
    for (int p = Ap[col]; p < Ap[col + 1]; p++) x[Ai[p]] = Ax[p];
    for (int p = top; p < n; p++) {
      int r = xi[p];
      int j = F->pinv[r];
      if (j < 0) continue;
      real xr = x[r];
      for (int t = F->Lp[j]; t < F->Lp[j + 1]; t++) x[F->Li[t]] -= F->Lx[t] * xr;
    }
    /* pivot selection */
    int ipiv = -1; real amax = -1.0;
    int need = n - top + 1;
    if (unz + need > F->ucap) {
      F->ucap = 2 * (unz + need);
      F->Ui = (int *)realloc(F->Ui, sizeof(int) * F->ucap);
      F->Ux = (real *)realloc(F->Ux, sizeof(real) * F->ucap);
    }
    if (lnz + need > F->lcap) {
      F->lcap = 2 * (lnz + need);
      F->Li = (int *)realloc(F->Li, sizeof(int) * F->lcap);
      F->Lx = (real *)realloc(F->Lx, sizeof(real) * F->lcap);
    }
    for (int p = top; p < n; p++) {
      int r = xi[p];
      if (F->pinv[r] < 0) {
        real a = R_fabs(x[r]);
        if (isnan(a)) { status = ORC_NONFINITE; }
        if (a > amax) { amax = a; ipiv = r; }
      } else {
        F->Ui[unz] = F->pinv[r];
        F->Ux[unz++] = x[r];
      }
    }
    if (status) break;
    if (F->pinv[col] < 0 && mark[col] == k + 1 && R_fabs(x[col]) >= thresh * amax)
      ipiv = col;
    if (ipiv < 0 || amax <= 0.0 || !isfinite(amax)) {
      status = (ipiv >= 0 && !isfinite(amax)) ? ORC_NONFINITE : ORC_SINGULAR;
      break;
    }
    real piv = x[ipiv];
    F->Ui[unz] = k; F->Ux[unz++] = piv;
    F->pinv[ipiv] = k;
    for (int p = top; p < n; p++) {
      int r = xi[p];
      if (F->pinv[r] < 0) {
        F->Li[lnz] = r;
        F->Lx[lnz++] = x[r] / piv;
      }
      x[r] = 0.0;
    }
  }
  if (!status) {
    F->Lp[n] = lnz; F->Up[n] = unz;
    for (int p = 0; p < lnz; p++) F->Li[p] = F->pinv[F->Li[p]];
  }
  free(x); free(xi); free(stk); free(pos); free(mark);
  return status;
}

static void lu_solve(const lu_t *F, const real *b, real *xout) {
  int n = F->n;
  real *c = (real *)malloc(sizeof(real) * n);
  for (int r = 0; r < n; r++) c[F->pinv[r]] = b[r];
  for (int k = 0; k < n; k++) {
    real ck = c[k];
    for (int t = F->Lp[k]; t < F->Lp[k + 1]; t++) c[F->Li[t]] -= F->Lx[t] * ck;
  }
  for (int k = n - 1; k >= 0; k--) {
    /* diagonal is the last entry of U(:,k) */
    int last = F->Up[k + 1] - 1;
    c[k] /= F->Ux[last];
    real ck = c[k];
    for (int t = F->Up[k]; t < last; t++) c[F->Ui[t]] -= F->Ux[t] * ck;
  }
  for (int k = 0; k < n; k++) xout[F->q[k]] = c[k];
  free(c);
}

/* CSR -> CSC (values follow) */
static void csr_to_csc(int n, const int *rp, const int *ci, const real *v,
                       int *Ap, int *Ai, real *Ax) {
  int nnz = rp[n];
  for (int j = 0; j <= n; j++) Ap[j] = 0;
  for (int k = 0; k < nnz; k++) Ap[ci[k] + 1]++;
  for (int j = 0; j < n; j++) Ap[j + 1] += Ap[j];
  int *w = (int *)malloc(sizeof(int) * (n + 1));
  memcpy(w, Ap, sizeof(int) * (n + 1));
  for (int i = 0; i < n; i++)
    for (int k = rp[i]; k < rp[i + 1]; k++) {
      int d = w[ci[k]]++;
      Ai[d] = i;
      Ax[d] = v[k];
    }
  free(w);
}

/* factor a CSR matrix with a given column order; returns status */
static int csr_factor(int n, const int *rp, const int *ci, const real *v,
                      const int *q, real thresh, lu_t *F) {
  int nnz = rp[n];
  int *Ap = (int *)malloc(sizeof(int) * (n + 1));
  int *Ai = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
  real *Ax = (real *)malloc(sizeof(real) * (nnz > 0 ? nnz : 1));
  csr_to_csc(n, rp, ci, v, Ap, Ai, Ax);
  int st = lu_factor(n, Ap, Ai, Ax, q, thresh, F);
  free(Ap); free(Ai); free(Ax);
  return st;
}

int orc_sparse_solve(int n, const int *rowptr, const int *colind,
                     const double *val, const double *b, double *x,
                     double pivot_thresh) {
  int *q = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  orc_min_degree(n, rowptr, colind, q);
  int nnz = rowptr[n];
  real *v = (real *)malloc(sizeof(real) * (nnz > 0 ? nnz : 1));
  real *bb = (real *)malloc(sizeof(real) * (n > 0 ? n : 1));
  real *xx = (real *)malloc(sizeof(real) * (n > 0 ? n : 1));
  for (int k = 0; k < nnz; k++) v[k] = val[k];
  for (int k = 0; k < n; k++) bb[k] = b[k];
  lu_t F;
  memset(&F, 0, sizeof(F));
  int st = csr_factor(n, rowptr, colind, v, q, pivot_thresh, &F);
  if (!st) {
    lu_solve(&F, bb, xx);
    for (int k = 0; k < n; k++) x[k] = (double)xx[k];
  }
  lu_free(&F);
  free(q); free(v); free(bb); free(xx);
  return st;
}

#ifdef ORC_LINSOLVE_EIGEN
/* oracle/_ref build: Newton steps solved by the reference's Eigen SparseLU
 * option (newtonpf.h:341-363), implemented in ref_eigen.cpp */
int orc_ref_linsolve(int n, const int *rowptr, const int *colind,
                     const double *val, const double *b, double *x);
#endif

/* ------------------------------------------------------------------------- */
/* voltage update (newtonpf.h:188-213)                                         */
/* ------------------------------------------------------------------------- */
static void update_voltage(const order_t *O, cpx *V, const real *dx) {
  int npv = O->npv, npq = O->npq;
  for (int id = 0; id < npv + npq; id++) {
    if (id < npv) {
      int i = O->pv[id];
      cpx v = V[i];
      V[i] = c_mul(c_make(c_abs(v), 0.0), c_expj(c_angle(v) - dx[id]));
    } else {
      int i = O->pq[id - npv];
      cpx v = V[i];
      V[i] = c_mul(c_make(c_abs(v) - dx[npq + id], 0.0),
                   c_expj(c_angle(v) - dx[id]));
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Newton-Raphson (newtonpf.h:365-523)                                        */
/* ------------------------------------------------------------------------- */
static int newtonpf(const ybus_t *Y, const inst_t *I, const int *type, cpx *V,
                    const orc_opts *o, int *iters, real *norm_out) {
  order_t O;
  order_build(type, I->n, &O);
  int nJ = O.npv + 2 * O.npq;
  real *F = (real *)malloc(sizeof(real) * (nJ > 0 ? nJ : 1));
  real *dx = (real *)malloc(sizeof(real) * (nJ > 0 ? nJ : 1));
  cpx *Ibus = (cpx *)malloc(sizeof(cpx) * I->n);
  jac_t J;
  memset(&J, 0, sizeof(J));
  int *q = NULL;
  real err = mismatch_nr(Y, I, &O, V, F);
  int iter = 0, verdict = ORC_MAX_ITER;
  if (err < o->tol) verdict = ORC_CONVERGED;
  else if (isnan(err) || isinf(err)) verdict = ORC_NONFINITE;
  while (verdict == ORC_MAX_ITER && iter < o->maxit_nr) {
    iter++;
    for (int i = 0; i < I->n; i++) Ibus[i] = row_current(Y, V, i);
    if (J.nnz == 0 && !J.rowptr) {
      jac_pattern(Y, type, &O, &J);
      q = (int *)malloc(sizeof(int) * (nJ > 0 ? nJ : 1));
      orc_min_degree(nJ, J.rowptr, J.colind, q);
    }
    jac_values(Y, &O, V, Ibus, &J);
#if defined(ORC_LINSOLVE_EIGEN) && defined(ORC_EXTENDED)
#error "the Eigen-linked build is FP64 only"
#endif
#ifdef ORC_LINSOLVE_EIGEN
    int st = orc_ref_linsolve(nJ, J.rowptr, J.colind, J.val, F, dx);
    if (st) { verdict = st; break; }
#else
    lu_t LU;
    memset(&LU, 0, sizeof(LU));
    int st = csr_factor(nJ, J.rowptr, J.colind, J.val, q, o->pivot_thresh, &LU);
    if (st) { lu_free(&LU); verdict = st; break; }
    lu_solve(&LU, F, dx);
    lu_free(&LU);
#endif
    update_voltage(&O, V, dx);
    err = mismatch_nr(Y, I, &O, V, F);
    if (err < o->tol) verdict = ORC_CONVERGED;
    else if (isnan(err) || isinf(err)) verdict = ORC_NONFINITE;
  }
  *iters = iter;
  *norm_out = err;
  free(F); free(dx); free(Ibus); free(q);
  jac_free(&J);
  order_free(&O);
  return verdict;
}

/* ------------------------------------------------------------------------- */
/* fast decoupled (makeB.h:311-501, fdpf.h:369-596)                           */
/* ------------------------------------------------------------------------- */
static real mismatch_fd(const ybus_t *Y, const inst_t *I, const order_t *O,
                          const cpx *V, real *P, real *Q) {
  real errP = -1.0, errQ = -1.0;
  int npv = O->npv, npq = O->npq, nan = 0;
  for (int id = 0; id < npv + npq; id++) {
    int i = (id < npv) ? id : id - npv;
    int bus = (id < npv) ? O->pv[i] : O->pq[i];
    cpx cc = row_current(Y, V, bus);
    cpx miss = c_sub(c_mul(V[bus], c_conj(cc)), c_make(I->P[bus], I->Q[bus]));
    miss = c_div(miss, c_make(c_abs(V[bus]), 0.0));
    if (id < npv) {
      P[i] = miss.re;
      errP = R_fmax(errP, R_fabs(P[i]));
      nan |= isnan(miss.re);
    } else {
      P[npv + i] = miss.re;
      errP = R_fmax(errP, R_fabs(P[npv + i]));
      Q[i] = miss.im;
      errQ = R_fmax(errQ, R_fabs(Q[i]));
      nan |= isnan(miss.re) | isnan(miss.im);
    }
  }
  return nan ? NAN : R_fmax(errP, errQ);
}

/* B matrix over an index list (idx[bus] >= 0 selects) from a Ybus build */
static void make_bmat(const ybus_t *Y, const int *sel, int nsel, const int *idx,
                      int **rp_o, int **ci_o, real **v_o) {
  int *rp = (int *)calloc(nsel + 1, sizeof(int));
  int cnt = 0;
  for (int r = 0; r < nsel; r++) {
    int i = sel[r];
    for (int k = Y->rowptr[i]; k < Y->rowptr[i + 1]; k++)
      if (idx[Y->colind[k]] >= 0) cnt++;
    rp[r + 1] = cnt;
  }
  int *ci = (int *)malloc(sizeof(int) * (cnt > 0 ? cnt : 1));
  real *v = (real *)malloc(sizeof(real) * (cnt > 0 ? cnt : 1));
  trip_t *row = (trip_t *)malloc(sizeof(trip_t) * (Y->n + 1));
  for (int r = 0; r < nsel; r++) {
    int i = sel[r], len = 0;
    for (int k = Y->rowptr[i]; k < Y->rowptr[i + 1]; k++) {
      int j = Y->colind[k];
      if (idx[j] < 0) continue;
      row[len].r = 0; row[len].c = idx[j]; row[len].y = k; len++;
    }
    qsort(row, len, sizeof(trip_t), cmp_trip);
    for (int t = 0; t < len; t++) {
      ci[rp[r] + t] = row[t].c;
      v[rp[r] + t] = -Y->val[row[t].y].im;
    }
  }
  free(row);
  *rp_o = rp; *ci_o = ci; *v_o = v;
}

/* B' (over pv then pq, fdpf.h:481-484) and B'' (over pq) of mkl_makeB
 * (makeB.h:392-501) from two modified Ybus builds:
 *   B':  line charging 0, taps 1, bus shunts 0 (and controls off,
 *        makeB.h:311-324); XB also R = 0; phase shifts kept;
 *   B'': phase shifts 0 (makeB.h:383-390); BX also R = 0; taps, charging,
 *        shunts and controls kept.
 * (mkl_zeroBusesAndBranches clears bus shunts inside its branch loop, so a
 * grid with fewer branches than buses keeps the shunts of the last buses;
 * the reference's own Ybus build, makeYbus.h:54-59, leaves those buses'
 * shunts unset in that case, so only nbranch >= nbus is defined, and there
 * every shunt is cleared.) */
typedef struct {
  int npv, npq, np;
  int *rpP, *ciP, *rpQ, *ciQ;
  real *vP, *vQ;
} bmat_t;

static void bmat_free(bmat_t *B) {
  free(B->rpP); free(B->ciP); free(B->vP);
  free(B->rpQ); free(B->ciQ); free(B->vQ);
}

static void bmat_build(const inst_t *I, const order_t *O, const int *from,
                       const int *to, int alg, bmat_t *Bm) {
  int n = I->n, m = I->m;
  int npv = O->npv, npq = O->npq, np = npv + npq;
  Bm->npv = npv; Bm->npq = npq; Bm->np = np;
  inst_t T = *I;
  real *zB = (real *)calloc(m > 0 ? m : 1, sizeof(real));
  real *zT = (real *)calloc(m > 0 ? m : 1, sizeof(real));
  real *zR = (real *)calloc(m > 0 ? m : 1, sizeof(real));
  real *zBs = (real *)calloc(n, sizeof(real));
  T.B = zB; T.tap = zT; T.Bs = zBs;
  if (alg == ORC_FDXB) T.R = zR;
  ybus_t Yp;
  memset(&Yp, 0, sizeof(Yp));
  ybus_build(&T, from, to, &Yp);
  int *sel_p = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  for (int k = 0; k < npv; k++) sel_p[k] = O->pv[k];
  for (int k = 0; k < npq; k++) sel_p[npv + k] = O->pq[k];
  make_bmat(&Yp, sel_p, np, O->idx, &Bm->rpP, &Bm->ciP, &Bm->vP);
  inst_t T2 = *I;
  real *zS = (real *)calloc(m > 0 ? m : 1, sizeof(real));
  T2.shift = zS;
  if (alg == ORC_FDBX) T2.R = zR;
  ybus_t Ypp;
  memset(&Ypp, 0, sizeof(Ypp));
  ybus_build(&T2, from, to, &Ypp);
  int *pqidx = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) pqidx[i] = -1;
  for (int k = 0; k < npq; k++) pqidx[O->pq[k]] = k;
  make_bmat(&Ypp, O->pq, npq, pqidx, &Bm->rpQ, &Bm->ciQ, &Bm->vQ);
  free(sel_p); free(pqidx);
  ybus_free(&Yp); ybus_free(&Ypp);
  free(zB); free(zT); free(zR); free(zBs); free(zS);
}

static int fdpf(const ybus_t *Y, const inst_t *I, const int *type,
                const int *from, const int *to, cpx *V, const orc_opts *o,
                int *iters, real *norm_out) {
  int n = I->n;
  order_t O;
  order_build(type, n, &O);
  int npv = O.npv, npq = O.npq, np = npv + npq;
  bmat_t Bm;
  bmat_build(I, &O, from, to, o->alg, &Bm);
  int *rpP = Bm.rpP, *ciP = Bm.ciP, *rpQ = Bm.rpQ, *ciQ = Bm.ciQ;
  real *vP = Bm.vP, *vQ = Bm.vQ;

  real *P = (real *)malloc(sizeof(real) * (np > 0 ? np : 1));
  real *Q = (real *)malloc(sizeof(real) * (npq > 0 ? npq : 1));
  real *dVa = (real *)malloc(sizeof(real) * (np > 0 ? np : 1));
  real *dVm = (real *)malloc(sizeof(real) * (npq > 0 ? npq : 1));
  int verdict = ORC_MAX_ITER, it = 0;
  real err = mismatch_fd(Y, I, &O, V, P, Q);
  if (err < o->tol) verdict = ORC_CONVERGED;
  else if (isnan(err)) verdict = ORC_NONFINITE;
  /* B' and B'' are constant: factor once (the reference refactors the same
   * values every iteration, fdpf.h:511,558, which yields identical factors) */
  lu_t LP, LQ;
  memset(&LP, 0, sizeof(LP));
  memset(&LQ, 0, sizeof(LQ));
  int *qP = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  int *qQ = (int *)malloc(sizeof(int) * (npq > 0 ? npq : 1));
  if (verdict == ORC_MAX_ITER) {
    orc_min_degree(np, rpP, ciP, qP);
    orc_min_degree(npq, rpQ, ciQ, qQ);
    int st = csr_factor(np, rpP, ciP, vP, qP, o->pivot_thresh, &LP);
    if (!st) st = csr_factor(npq, rpQ, ciQ, vQ, qQ, o->pivot_thresh, &LQ);
    if (st) verdict = st;
  }
  while (verdict == ORC_MAX_ITER && it < o->maxit_fd) {
    it++;
    lu_solve(&LP, P, dVa);
    for (int id = 0; id < np; id++) {
      int i = (id < npv) ? O.pv[id] : O.pq[id - npv];
      cpx v = V[i];
      V[i] = c_mul(c_make(c_abs(v), 0.0), c_expj(c_angle(v) - dVa[id]));
    }
    err = mismatch_fd(Y, I, &O, V, P, Q);
    if (err < o->tol) { verdict = ORC_CONVERGED; break; }
    if (isnan(err)) { verdict = ORC_NONFINITE; break; }
    lu_solve(&LQ, Q, dVm);
    for (int id = 0; id < npq; id++) {
      int i = O.pq[id];
      cpx v = V[i];
      V[i] = c_mul(c_make(c_abs(v) - dVm[id], 0.0), c_expj(c_angle(v)));
    }
    err = mismatch_fd(Y, I, &O, V, P, Q);
    if (err < o->tol) verdict = ORC_CONVERGED;
    else if (isnan(err)) verdict = ORC_NONFINITE;
  }
  *iters = it;
  *norm_out = err;
  lu_free(&LP); lu_free(&LQ);
  free(qP); free(qQ); free(P); free(Q); free(dVa); free(dVm);
  bmat_free(&Bm);
  order_free(&O);
  return verdict;
}

/* ------------------------------------------------------------------------- */
/* loss and branch flows (runpf.h:228-257)                                     */
/* ------------------------------------------------------------------------- */
static real branch_loss(const ybus_t *Y, const int *from, const int *to,
                          int m, const cpx *V, cpx *Sf, cpx *St) {
  real sum = 0.0;
  for (int b = 0; b < m; b++) {
    int f = from[b], t = to[b];
    /* Yf/Yt rows hold two entries with sorted columns (makeYbus.h:84-98) */
    cpx If = c_make(0.0, 0.0), It = c_make(0.0, 0.0);
    if (f < t) {
      If = c_add(c_add(If, c_mul(Y->yff[b], V[f])), c_mul(Y->yft[b], V[t]));
      It = c_add(c_add(It, c_mul(Y->ytf[b], V[f])), c_mul(Y->ytt[b], V[t]));
    } else {
      If = c_add(c_add(If, c_mul(Y->yft[b], V[t])), c_mul(Y->yff[b], V[f]));
      It = c_add(c_add(It, c_mul(Y->ytt[b], V[t])), c_mul(Y->ytf[b], V[f]));
    }
    cpx sf = c_mul(c_conj(If), V[f]);
    cpx st = c_mul(c_conj(It), V[t]);
    if (Sf) Sf[b] = c_conj(sf);   /* V_f conj(I_f) */
    if (St) St[b] = c_conj(st);
    sum += c_add(sf, st).re;
  }
  return sum;
}

/* ------------------------------------------------------------------------- */
/* public entry points (FP64 arrays; converted at the boundary)               */
/* ------------------------------------------------------------------------- */
static cpx *cpx_in(const double *v, int n) {
  cpx *z = (cpx *)malloc(sizeof(cpx) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) z[i] = c_make(v[2 * i], v[2 * i + 1]);
  return z;
}
static void cpx_out(const cpx *z, int n, double *v) {
  for (int i = 0; i < n; i++) {
    v[2 * i] = (double)z[i].re;
    v[2 * i + 1] = (double)z[i].im;
  }
}

void orc_initial_voltage(const orc_case *c, const orc_inst *inst, double *Vout) {
  inst_t I;
  inst_build(c, inst, &I);
  cpx *V = (cpx *)malloc(sizeof(cpx) * c->nbus);
  initial_voltage(c, &I, V);
  cpx_out(V, c->nbus, Vout);
  free(V);
  inst_free(&I);
}

int orc_make_ybus(const orc_case *c, const orc_inst *inst, int *rowptr,
                  int *colind, double *val, int cap) {
  inst_t I;
  inst_build(c, inst, &I);
  ybus_t Y;
  memset(&Y, 0, sizeof(Y));
  ybus_build(&I, c->br_from, c->br_to, &Y);
  int nnz = Y.nnz;
  if (nnz <= cap) {
    memcpy(rowptr, Y.rowptr, sizeof(int) * (c->nbus + 1));
    memcpy(colind, Y.colind, sizeof(int) * nnz);
    cpx_out(Y.val, nnz, val);
  } else {
    nnz = -nnz;
  }
  ybus_free(&Y);
  inst_free(&I);
  return nnz;
}

int orc_make_bmat(const orc_case *c, const orc_inst *inst, int alg, int which,
                  int *rows_out, int *rowptr, int *colind, double *val, int cap) {
  inst_t I;
  inst_build(c, inst, &I);
  order_t O;
  order_build(I.type, I.n, &O);
  bmat_t Bm;
  bmat_build(&I, &O, c->br_from, c->br_to, alg, &Bm);
  int rows = which == 1 ? Bm.np : Bm.npq;
  const int *rp = which == 1 ? Bm.rpP : Bm.rpQ;
  const int *ci = which == 1 ? Bm.ciP : Bm.ciQ;
  const real *v = which == 1 ? Bm.vP : Bm.vQ;
  int nnz = rp[rows];
  *rows_out = rows;
  if (nnz <= cap) {
    memcpy(rowptr, rp, sizeof(int) * (rows + 1));
    memcpy(colind, ci, sizeof(int) * nnz);
    for (int k = 0; k < nnz; k++) val[k] = (double)v[k];
  } else {
    nnz = -nnz;
  }
  bmat_free(&Bm);
  order_free(&O);
  inst_free(&I);
  return nnz;
}

int orc_mismatch(const orc_case *c, const orc_inst *inst, const double *Vin,
                 double *F, double *norm) {
  inst_t I;
  inst_build(c, inst, &I);
  ybus_t Y;
  memset(&Y, 0, sizeof(Y));
  ybus_build(&I, c->br_from, c->br_to, &Y);
  order_t O;
  order_build(I.type, I.n, &O);
  int nJ = O.npv + 2 * O.npq;
  cpx *V = cpx_in(Vin, I.n);
  real *Fr = (real *)malloc(sizeof(real) * (nJ > 0 ? nJ : 1));
  *norm = (double)mismatch_nr(&Y, &I, &O, V, Fr);
  for (int k = 0; k < nJ; k++) F[k] = (double)Fr[k];
  free(Fr); free(V);
  order_free(&O);
  ybus_free(&Y);
  inst_free(&I);
  return nJ;
}

int orc_jacobian(const orc_case *c, const orc_inst *inst, const double *Vin,
                 int *rowptr, int *colind, double *val, int cap) {
  inst_t I;
  inst_build(c, inst, &I);
  ybus_t Y;
  memset(&Y, 0, sizeof(Y));
  ybus_build(&I, c->br_from, c->br_to, &Y);
  order_t O;
  order_build(I.type, I.n, &O);
  jac_t J;
  memset(&J, 0, sizeof(J));
  jac_pattern(&Y, I.type, &O, &J);
  cpx *V = cpx_in(Vin, I.n);
  cpx *Ib = (cpx *)malloc(sizeof(cpx) * I.n);
  for (int i = 0; i < I.n; i++) Ib[i] = row_current(&Y, V, i);
  jac_values(&Y, &O, V, Ib, &J);
  int nnz = J.nnz;
  if (nnz <= cap) {
    memcpy(rowptr, J.rowptr, sizeof(int) * (J.nJ + 1));
    memcpy(colind, J.colind, sizeof(int) * nnz);
    for (int k = 0; k < nnz; k++) val[k] = (double)J.val[k];
  } else {
    nnz = -nnz;
  }
  free(Ib); free(V);
  jac_free(&J);
  order_free(&O);
  ybus_free(&Y);
  inst_free(&I);
  return nnz;
}

/* losses + w * sum over both ends of max(0, |S| - rate)^2 at V, in the
 * working precision; optional Sf/St out */
static real objective_at(const orc_case *c, const ybus_t *Y, const cpx *V,
                         const double *rate, double w, double *Sf, double *St) {
  int m = c->nbranch;
  cpx *sf = (cpx *)malloc(sizeof(cpx) * (m > 0 ? m : 1));
  cpx *st = (cpx *)malloc(sizeof(cpx) * (m > 0 ? m : 1));
  real f = branch_loss(Y, c->br_from, c->br_to, m, V, sf, st);
  if (rate) {
    for (int b = 0; b < m; b++) {
      real ef = c_abs(sf[b]) - rate[b], et = c_abs(st[b]) - rate[b];
      if (ef > 0) f += w * ef * ef;
      if (et > 0) f += w * et * et;
    }
  }
  if (Sf) cpx_out(sf, m, Sf);
  if (St) cpx_out(st, m, St);
  free(sf); free(st);
  return f;
}

double orc_objective(const orc_case *c, const orc_inst *inst, const double *Vin,
                     const double *rate, double w, double *Sf, double *St) {
  inst_t I;
  inst_build(c, inst, &I);
  ybus_t Y;
  memset(&Y, 0, sizeof(Y));
  ybus_build(&I, c->br_from, c->br_to, &Y);
  cpx *V = cpx_in(Vin, c->nbus);
  real f = objective_at(c, &Y, V, rate, w, Sf, St);
  free(V);
  ybus_free(&Y);
  inst_free(&I);
  return (double)f;
}

/* solve, then (when f_hi is given) the objective at the final V before it is
 * rounded to FP64: f = f_hi + f_lo (whatever the verdict: floor-precision
 * re-solves run a fixed iteration count and check r->norm themselves) */
static int solve_impl(const orc_case *c, const orc_inst *inst, const orc_opts *o,
                      orc_result *r, const double *rate, double w,
                      double *f_hi, double *f_lo) {
  inst_t I;
  inst_build(c, inst, &I);
  ybus_t Y;
  memset(&Y, 0, sizeof(Y));
  ybus_build(&I, c->br_from, c->br_to, &Y);
  int n = c->nbus;
  cpx *V;
  if (inst && inst->V0) V = cpx_in(inst->V0, n);
  else { V = (cpx *)malloc(sizeof(cpx) * n); initial_voltage(c, &I, V); }
  int total = 0, rounds = 0, verdict = ORC_MAX_ITER;
  real norm = NAN;
  for (;;) {
    int it = 0;
    if (o->alg == ORC_NR)
      verdict = newtonpf(&Y, &I, I.type, V, o, &it, &norm);
    else
      verdict = fdpf(&Y, &I, I.type, c->br_from, c->br_to, V, o, &it, &norm);
    total += it;
    if (verdict != ORC_CONVERGED || !o->enforce_qlim) break;
    /* reactive limits of PV buses at the solved point */
    int nviol = 0;
    int *viol = (int *)calloc(n, sizeof(int));
    for (int i = 0; i < n; i++) {
      if (I.type[i] != ORC_PV) continue;
      cpx S = c_mul(V[i], c_conj(row_current(&Y, V, i)));
      real qg = S.im + I.Qd[i];
      if (qg > I.Qmax[i] + o->qlim_tol) { viol[i] = 1; nviol++; }
      else if (qg < I.Qmin[i] - o->qlim_tol) { viol[i] = -1; nviol++; }
    }
    if (nviol == 0) { free(viol); break; }
    if (rounds >= o->max_qlim_rounds) {
      free(viol);
      verdict = ORC_MAX_ITER;
      break;
    }
    rounds++;
    for (int i = 0; i < n; i++) {
      if (!viol[i]) continue;
      I.type[i] = ORC_PQ;
      I.Q[i] = (viol[i] > 0 ? I.Qmax[i] : I.Qmin[i]) - I.Qd[i];
    }
    free(viol);
  }
  r->iterations = total;
  r->qlim_rounds = rounds;
  r->verdict = verdict;
  r->converged = verdict == ORC_CONVERGED;
  r->norm = (double)norm;
  r->loss = r->converged
                ? (double)branch_loss(&Y, c->br_from, c->br_to, c->nbranch, V,
                                      NULL, NULL)
                : DBL_MAX;
  if (f_hi) {
    real f = objective_at(c, &Y, V, rate, w, NULL, NULL);
    *f_hi = (double)f;
    if (f_lo) *f_lo = (double)(f - (real)*f_hi);
  }
  if (r->V) cpx_out(V, n, r->V);
  if (r->bus_type) memcpy(r->bus_type, I.type, sizeof(int) * n);
  free(V);
  ybus_free(&Y);
  inst_free(&I);
  return 0;
}

int orc_solve(const orc_case *c, const orc_inst *inst, const orc_opts *o,
              orc_result *r) {
  return solve_impl(c, inst, o, r, NULL, 0.0, NULL, NULL);
}

int orc_solve_objective(const orc_case *c, const orc_inst *inst,
                        const orc_opts *o, const double *rate, double w,
                        orc_result *r, double *f_hi, double *f_lo) {
  return solve_impl(c, inst, o, r, rate, w, f_hi, f_lo);
}

int orc_precision_bits(void) {
#ifdef ORC_EXTENDED
  return LDBL_MANT_DIG;
#else
  return DBL_MANT_DIG;
#endif
}
