/*
 * batchpf.h -- C-ABI of the MI355X batched AC power-flow engine.
 *
 * The reference (MAPoL) exposes its power-flow path as C++ free functions over
 * module globals; every entry point below replaces one of them.  Plain C
 * types only: pointers, sizes, status codes.  All arrays are host memory owned
 * by the caller and are copied in or out during the call; device memory and
 * the symbolic analysis belong to the engine handle.
 *
 * Numerical failure is never a process exit (the reference exit()s on CUDA /
 * MKL errors, src/util/helper_cuda.h:967-977, newtonpf.h:321): it is a
 * per-instance verdict.  API misuse and HIP errors are status codes, with a
 * message from bpf_last_error().
 */
#ifndef BATCHPF_H
#define BATCHPF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bpf_engine bpf_engine;

typedef enum {
  BPF_OK = 0,
  BPF_ERR_INVALID_ARG = 1,
  BPF_ERR_HIP = 2,
  BPF_ERR_OOM = 3,
  BPF_ERR_STATE = 4,
  BPF_ERR_UNSUPPORTED = 5,
  BPF_ERR_IO = 6
} bpf_status;

/* bus types: MATPOWER codes as in src/powersystem/Bus.h:10-12 */
enum { BPF_PQ = 1, BPF_PV = 2, BPF_SLACK = 3 };

/* algorithms: src/powerflow/runpf.h:16-18 */
enum { BPF_NR = 0, BPF_FDXB = 1, BPF_FDBX = 2 };

/* precision of the Jacobian factorization and triangular solves; mismatch,
 * norm and voltage update are always FP64 */
enum { BPF_FP64 = 0, BPF_MIXED = 1 };

/* per-instance verdicts (the reference reports only converged or not,
 * runpf.h:337-353) */
enum {
  BPF_CONVERGED = 0,
  BPF_MAX_ITER = 1,
  BPF_SINGULAR = 2,
  BPF_NONFINITE = 3
};

/* control kinds of the reference's particle vector
 * (src/pso/Particula.h:26-60): AVR set-point, OLTC tap, switched shunt */
enum { BPF_CTL_AVR = 1, BPF_CTL_OLTC = 2, BPF_CTL_SHC = 3 };

/* Case data, per unit on baseMVA, angles in radians: the reference's Topology
 * of Bus / Branch / Generator (src/powersystem/Topology.h:18-127,
 * Bus.h:6-63, Branch.h:10-62, Generator.h:9-40) as structure-of-arrays.
 * Loads and generation are separate; the net injection the reference stores
 * in Bus.P/Bus.Q is sum(Pg of in-service generators) - Pd
 * (src/files/loadMatpowerFormat.h:88,126-128). */
typedef struct {
  int32_t nbus, nbranch, ngen;
  double baseMVA;
  const int32_t *bus_type;                 /* [nbus] */
  const double *Pd, *Qd, *Gs, *Bs;         /* [nbus] */
  const double *Vm, *Va;                   /* [nbus] initial voltage table */
  const int32_t *br_from, *br_to;          /* [nbranch], 0-based */
  const double *R, *X, *B, *tap, *shift;   /* tap 0 => 1; shift in degrees */
  const uint8_t *br_status;                /* [nbranch] */
  const int32_t *gen_bus;                  /* [ngen] */
  const double *Pg, *Qg, *Qmax, *Qmin, *VG;/* [ngen] */
  const uint8_t *gen_status;               /* [ngen] */
} bpf_case;

typedef struct {
  int32_t alg;             /* BPF_NR / BPF_FDXB / BPF_FDBX */
  int32_t precision;       /* BPF_FP64 / BPF_MIXED */
  int32_t refine_steps;    /* FP64 refinement steps of the Newton step (MIXED) */
  int32_t maxit_nr;        /* 10, runpf.h:45 */
  int32_t maxit_fd;        /* 30, runpf.h:46 */
  int32_t enforce_qlim;    /* PV->PQ switching on reactive limits */
  int32_t max_qlim_rounds;
  int32_t use_graph;       /* replay each Newton iteration as one hipGraph */
  double tol;              /* 1e-8, runpf.h:47 */
  double qlim_tol;         /* pu */
  /* tile sizes of the symbolic schedules (0: the largest the kernels hold;
   * mapol_amd/tune.py picks them per case and instance count) */
  int32_t tail_limit;      /* pivots of the dense tail, 1..64 */
  int32_t chunk_products;  /* block products per factorization chunk, <= 4096 */
  int32_t slot_chunk;      /* lane slots per slot-form sweep chunk, a multiple of 64 up to 2048 */
  int32_t ordering;        /* 0 minimum degree; 1 the same, reordered to the
                              least elimination-tree height (same fill);
                              2 nested dissection */
  int32_t bus_numbering;   /* the engine's own bus numbers: 1 (default) reverse
                              Cuthill-McKee, so a bus's neighbours sit near it
                              in every per-bus array; 0 the caller's.  The
                              elimination order, and every bus index, value
                              and matrix the interface takes or returns, stay
                              in the caller's numbering either way. */
} bpf_options;

/* one control of the reference's particle vector (Particula::Estrutura) */
typedef struct {
  int32_t kind;            /* BPF_CTL_* */
  int32_t id;              /* bus (AVR, SHC) or branch (OLTC) index */
} bpf_ctl;

typedef struct {
  int32_t nbus, nbranch, nblocks;  /* nblocks = non-slack buses */
  int32_t nnz_ybus;                /* Ybus pattern entries */
  int32_t nnz_jac_blocks;          /* 2x2 Jacobian blocks (superset pattern) */
  int32_t nnz_lu_blocks;           /* 2x2 blocks of L+U (with fill) */
  int32_t lu_levels;               /* elimination-tree levels */
  int32_t fw_levels, bw_levels;    /* triangular-solve levels */
  int64_t lu_pairs;                /* 2x2 block products per factorization */
  int32_t instances;               /* T set by bpf_set_instances */
  int32_t lds_solve;               /* 1: solve vector held in LDS */
  int32_t crout_tasks;             /* outputs below the tail (one per block) */
  int32_t chunks;                  /* factorization chunks below the tail */
  int32_t graph_used;              /* last solve replayed the iteration graph */
  int32_t tail_level;              /* levels >= tail_level: dense LDS block */
  int32_t tail_n;                  /* pivots in the dense tail (<= 64) */
  int32_t tail_entries, tail_pairs;
  int32_t tail_row_entries;        /* non-tail L entries of the tail rows */
  int32_t sweep_fw_chunks, sweep_bw_chunks;  /* entry-parallel sweep chunks */
  int32_t sweep_rows, sweep_entries;
  int32_t fill_level0_chunks;      /* level-0 chunks finished by the Jacobian
                                      fill in Newton steps (0: none) */
  int32_t tail_stage;              /* distinct blocks of the tail products
                                      staged in LDS (0: not staged) */
  int32_t slot_fw_chunks, slot_bw_chunks;    /* slot-form sweep chunks */
  int32_t slots;                   /* lane slots of the slot form (0: the
                                      slot form does not apply) */
  int32_t graph_fallback;          /* the last solve abandoned graph replay
                                      and ran with direct launches
                                      (bpf_last_error says why) */
  int32_t graph_captures;          /* Newton graphs captured since bpf_create
                                      (a new capture follows a change of
                                      instances, launch shapes or factor
                                      precision) */
  int32_t fused_forward;           /* the active-instance limit of the last
                                      Newton step: its factorization also
                                      formed the forward substitution of the
                                      mismatch, and its first solve started
                                      from it, while at most this many
                                      instances were active (-1: never; only
                                      with the slot-form sweeps) */
} bpf_info;

/* Fill an options struct with the reference's defaults (runpf.h:45-47). */
void bpf_default_options(bpf_options *o);

/* hybrid_init (src/powerflow/runpf.h:795-982): copies the case, runs the
 * symbolic analysis once (Ybus pattern, ordering, L+U pattern, schedules) and
 * uploads the shared structures.  Replaces also mkl_init (runpf.h:359). */
bpf_status bpf_create(const bpf_case *c, const bpf_options *o,
                      bpf_engine **out);

/* Per-instance inputs: the reference's particle matrix d_enxame [T][nCtl]
 * (runpf.h:611-616) generalised with branch and generator outage masks.
 * br_on [T*nbranch] and gen_on [T*ngen] may be NULL (all in service);
 * ctl/ctl_val may be NULL (defaults: VG, case tap, case Bs as main.cu:232-264
 * builds them).  Builds each instance's bus types (a PV bus without an
 * in-service generator becomes PQ), injections and set-points on the host
 * and uploads them; inputs stay resident in HBM. */
bpf_status bpf_set_instances(bpf_engine *e, int32_t T, const uint8_t *br_on,
                             const uint8_t *gen_on, int32_t nctl,
                             const bpf_ctl *ctl, const double *ctl_val);

/* hybrid_runpf (runpf.h:606-793) minus the host round trips: initial
 * voltage, Ybus assembly with outage masks, Newton-Raphson (or fast
 * decoupled) for all instances, reactive-limit rounds, losses.  Results stay
 * on the device until bpf_get_results.  Replaces mkl_runpf (runpf.h:259). */
bpf_status bpf_solve(bpf_engine *e);

/* Per instance: V (re,im) [T*nbus*2], total iterations, verdict, final
 * inf-norm, loss (DBL_MAX when not converged, runpf.h:351-353), reactive-
 * limit rounds, final bus types [T*nbus].  Any pointer may be NULL. */
bpf_status bpf_get_results(bpf_engine *e, double *V, int32_t *iterations,
                           int32_t *verdict, double *norm, double *loss,
                           int32_t *qlim_rounds, int32_t *bus_type);

/* Branch power flows Sf, St (re,im) [T*nbranch*2] at the last solution:
 * the from/to ends of the loss kernel's Yf*V, Yt*V products (runpf.h:543-572). */
bpf_status bpf_get_branch_flows(bpf_engine *e, double *Sf, double *St);

/* Branch loading and per-contingency violation report at the last solution
 * (no function of the reference: derived from the flows of runpf.h:543-572
 * for ranking contingencies).  rate[nbranch] = branch ratings in pu (MVA /
 * baseMVA, e.g. the rate_a of bpf_load_matpower; 0 = no rating).  Outputs,
 * each may be NULL: loading[T*nbranch] = 100 max(|Sf|, |St|) / rate (0 where
 * rate is 0 or the branch is out of service), overloads[T] = branches above
 * 100 %, max_loading[T] = the largest loading.  Non-converged instances are
 * reported at their last iterate (their verdict says so). */
bpf_status bpf_get_branch_loading(bpf_engine *e, const double *rate,
                                  double *loading, int32_t *overloads,
                                  double *max_loading);

/* Allocation-free variant of bpf_get_branch_loading: same arguments, same
 * results.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc/hipFree: the
 * ratings and the per-instance outputs are staged in a device array of
 * BPF_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs on
 * until the process ends, in batches of instances that fit it.  Limit: the
 * ratings plus one instance must fit, i.e. nbranch <= 262 142 with
 * loading != NULL (524 256 without).  Larger cases return
 * BPF_ERR_INVALID_ARG.  All engines on one GPU share that array: calls that
 * overlap in time on one GPU (several host threads) race; calls in sequence
 * do not. */
#define BPF_NOALLOC_SCRATCH_BYTES (4u << 20)
bpf_status bpf_get_branch_loading_noalloc(bpf_engine *e, const double *rate,
                                          double *loading, int32_t *overloads,
                                          double *max_loading);

/* hybrid_free (runpf.h:984-1078), mkl_clean (runpf.h:452). */
void bpf_destroy(bpf_engine *e);

/* Last error message; when there is none but hipGraph replay had to be
 * abandoned (the solve then ran with direct launches, bpf_info
 * graph_fallback = 1), its reason; empty otherwise. */
const char *bpf_last_error(const bpf_engine *e);
bpf_status bpf_get_info(const bpf_engine *e, bpf_info *info);

/* MATPOWER version-2 case file into a bpf_case, host only.  Replaces
 * loadMatpowerFormat (src/files/loadMatpowerFormat.h:14-209) with the same
 * conventions: loads, shunts and generation divided by baseMVA, Va in
 * radians, bus ids mapped to 0-based file order, every generator counted
 * whatever its status column says (BPF_MP_GEN_STATUS honours it), the last
 * generator of a bus sets its VG.  Loads and generation stay separate (the
 * reference stores Pg - Pd); Inf reactive limits become +-infinity (the
 * reference stores +-DBL_MAX, equal in every comparison).  An unknown bus id
 * is BPF_ERR_INVALID_ARG (the reference indexes out of range).
 * The arrays of *out (and *rate_a, branch rateA / baseMVA, if requested)
 * live in *owner until bpf_free_case; on failure *owner (if not NULL) holds
 * the message for bpf_case_error and must be freed as well. */
#define BPF_MP_GEN_STATUS 1
typedef struct bpf_case_data bpf_case_data;
bpf_status bpf_load_matpower(const char *path, int32_t flags, bpf_case *out,
                             const double **rate_a, bpf_case_data **owner);
const char *bpf_case_error(const bpf_case_data *owner);
void bpf_free_case(bpf_case_data *owner);

/* Host-only symbolic analysis (no device needed): the block order, the L+U
 * block pattern and the schedules bpf_create builds (the J pattern of
 * newtonpf.h:215-310 extended by the fill of the factorization).  Sizes come
 * back in *info; every array may be NULL, otherwise it must hold:
 * bus_of_blk[nblocks], lu_rowptr[nblocks+1], lu_colind[nnz_lu_blocks],
 * ch_task/ch_pair[chunks+1], tk_a/tk_b[4*crout_tasks], pairs[2*lu_pairs],
 * fw_ptr[fw_levels+1], fw_rows[nblocks-tail_n], bw_ptr[bw_levels+1],
 * bw_rows[nblocks-tail_n], tl_dense/tl_pos[tail_entries],
 * tl_ptr[tail_entries+1], tl_pairL/tl_pairU[tail_pairs], tail_piv[tail_n],
 * ts_ptr[tail_n+1], ts_pos/ts_col[tail_row_entries].
 * Factorization below the tail, chunk by chunk: chunk c owns the outputs
 * ch_task[c]..ch_task[c+1]-1 and the products ch_pair[c]..ch_pair[c+1]-1;
 * output k has tk_a[4k..] = {position | flags, first product, end product,
 * 0} with flags bit 30 = the position holds a Jacobian entry, bit 29 = pivot
 * block (stored inverted), bit 28 = L block (scaled by the pivot inverse,
 * the pivot's product range in tk_b[4k+1], tk_b[4k+2]); product q is
 * pairs[2q], pairs[2q+1] = L and U positions, U = -1 for an init entry (the
 * pivot's own Jacobian block).  An output is its init/Jacobian value minus
 * the sum of its L*U products.  The top tail_n pivots (blocks tail_piv) are
 * then finished as one dense block: its Schur complement (tl_dense = row *
 * tail_n + col, flag bit 30; its pairs are the non-tail contributions) is
 * inverted by Gauss-Jordan.  The sweeps solve the non-tail rows by levels
 * (fw/bw lists) and the tail rows with that inverse after subtracting their
 * non-tail L entries (ts lists: L position and column per tail row). */
bpf_status bpf_analyse(const bpf_case *c, bpf_info *info, int32_t *bus_of_blk,
                       int32_t *lu_rowptr, int32_t *lu_colind, int32_t *ch_task,
                       int32_t *ch_pair, int32_t *tk_a, int32_t *tk_b,
                       int32_t *pairs, int32_t *fw_ptr, int32_t *fw_rows,
                       int32_t *bw_ptr, int32_t *bw_rows, int32_t *tl_dense,
                       int32_t *tl_pos, int32_t *tl_ptr, int32_t *tl_pairL,
                       int32_t *tl_pairU, int32_t *tail_piv, int32_t *ts_ptr,
                       int32_t *ts_pos, int32_t *ts_col);

/* Host-only: the entry-parallel form of the sweeps (used when the solve
 * vector and a product buffer fit in LDS).  The non-tail rows of every
 * level, in the level order of fw/bw_rows, are packed into chunks of at most
 * 2048 entries; chunks 0 .. sweep_fw_chunks-1 are the forward sweep, the
 * rest the backward sweep, and the tail rows are solved between them.
 * sw_chunk[4*(fw+bw chunks)] = {first row, end row, first entry, end entry};
 * sw_row[4*sweep_rows] = {block row, first entry, end entry, pivot position
 * (backward) or -1}; sw_ent[2*sweep_entries] = {L or U position, column
 * block}.  Arrays may be NULL. */
bpf_status bpf_analyse_sweeps(const bpf_case *c, bpf_info *info,
                              int32_t *sw_chunk, int32_t *sw_row,
                              int32_t *sw_ent);

/* Host-only: the slot form of the sweeps (the default when the solve vector
 * fits in LDS and the solve workgroup has 256, 512 or 1024 lanes; the
 * reference solves with a library, newtonpf.h:312-363).  The same levels as
 * fw/bw_rows; each row owns a lane group of 2^g lanes (the power of two at
 * least its entry count, at most 64) that starts at a multiple of its size,
 * one lane per entry; a row of more than 64 entries comes in segments of 64
 * in successive chunks of its level.  sl_chunk[4*(slot_fw_chunks+
 * slot_bw_chunks)] = {first slot, end slot, largest g, 0}, at most 2048
 * slots each; sl_slot[4*slots] = {L or U position (padding: the segment's
 * first entry), column block + 1 (0: padding) | g << 27, block row (a
 * group's first slot; | 1 << 30 for a segment before the row's last) or -1,
 * pivot position (first slot of a row's last segment, backward) or the
 * position}.  Arrays may be NULL. */
bpf_status bpf_analyse_slots(const bpf_case *c, bpf_info *info,
                             int32_t *sl_chunk, int32_t *sl_slot);

/* Host-only: the slot form of the adjoint's transposed sweeps (J^T = U^T
 * L^T; the reference has no adjoint): U^T forward over the rows of L, every
 * row applying its pivot inverse (transposed) at its last segment, then
 * L^T backward over the rows of U that have entries.  The same slot layout
 * as bpf_analyse_slots with every position replaced by its mirror (the
 * block (k, i) for (i, k)), used transposed.  counts[3] = {forward chunks,
 * backward chunks, slots} (zeros when the slot form does not apply);
 * sl_chunk [4 * chunks], sl_slot [4 * slots]; arrays may be NULL. */
bpf_status bpf_analyse_slots_transposed(const bpf_case *c, int32_t *counts,
                                        int32_t *sl_chunk, int32_t *sl_slot);

/* Host-only: the Jacobian fill of a Newton step, which also finishes level
 * 0 of the elimination tree (part of the factorization of newtonpf.h:312-363
 * moved into the fill).  jac_lu[nnz_jac_blocks] = L+U position of every
 * Jacobian block; jac_l0[nnz_jac_blocks] = -1 (block unchanged), -2 (a
 * level-0 pivot: stored inverted) or the index of its level-0 pivot's
 * Jacobian block (an L block: stored times that pivot's inverse).  The
 * factorization then starts at chunk fill_level0_chunks.  Arrays may be
 * NULL. */
bpf_status bpf_analyse_fill(const bpf_case *c, bpf_info *info, int32_t *jac_lu,
                            int32_t *jac_l0);

/* Host-only: the staging of the dense tail's Schur-complement products (the
 * FP32 factorization copies their distinct L and U blocks into LDS once).
 * tl_stage[tail_stage] = L+U positions of those blocks, ascending;
 * tl_pairS[tail_pairs] = for each pair of tl_pairL/tl_pairU its two indices
 * into tl_stage, L | U << 16.  tail_stage = 0 when the blocks exceed the LDS
 * capacity (5760): then the products read the factor directly.  Arrays may
 * be NULL. */
bpf_status bpf_analyse_tail(const bpf_case *c, bpf_info *info, int32_t *tl_stage,
                            int32_t *tl_pairS);

/* ---- stage-level entry points (tests, benchmarks) ---------------------- */

/* hybrid_makeYbus / mkl_makeYbus (src/powerflow/makeYbus.h:372-510 and
 * :113-274): assemble the Ybus of every instance from the instances set by
 * bpf_set_instances, without solving (bpf_solve assembles it too). */
bpf_status bpf_make_ybus(bpf_engine *e);

/* Ybus of instance t after bpf_make_ybus or bpf_solve, CSR with sorted
 * columns; val interleaved (re, im). */
bpf_status bpf_get_ybus(bpf_engine *e, int32_t t, int32_t *rowptr,
                        int32_t *colind, double *val);

/* hybrid_makeB / mkl_makeB (src/powerflow/makeB.h:109, :392): B' (which 1,
 * rows and columns pv then pq) or B'' (which 2, over pq) of instance t for
 * the engine's fast-decoupled variant, as COO triplets in that order
 * (*nnz = entries; cap = capacity of the arrays).  Uses the instances' bus
 * types as a solve starts them; overwrites the state of the last solve. */
bpf_status bpf_stage_bmat(bpf_engine *e, int32_t t, int32_t which, int32_t cap,
                          int32_t *row, int32_t *col, double *val,
                          int32_t *nnz);

/* One half-step of the fast-decoupled iteration (hybrid_fdpf / mkl_fdpf,
 * src/powerflow/fdpf.h:99, :450; its P half-step is :480-520, the Q
 * half-step :522-560) for every instance from the voltages V [T*nbus*2]:
 * the mismatch divided by |V|, a solve with B' (which 1: angles of pv and pq
 * buses) or B'' (which 2: magnitudes of pq buses), the update.  Vout
 * receives the new voltages, norm [T] (may be NULL) the inf-norm of the
 * normalised mismatch at V.  Needs bpf_make_ybus or bpf_solve first;
 * overwrites the state of the last solve. */
bpf_status bpf_stage_fd_halfstep(bpf_engine *e, const double *V, int32_t which,
                                 double *Vout, double *norm);

/* checkConvergence / computeDiagIbus (newtonpf.h:18-95): mismatch of every
 * instance at the given voltages V [T*nbus*2] for the instances' current bus
 * types.  F is returned in the reference layout [P(pv); P(pq); Q(pq)]
 * per instance, packed at stride 2*nbus; norm [T]. */
bpf_status bpf_stage_mismatch(bpf_engine *e, const double *V, double *F,
                              double *norm);

/* compuateJacobianMatrix (newtonpf.h:97-186): Jacobian of instance t at V
 * (all instances' voltages [T*nbus*2] are uploaded), written as COO triplets
 * in the reference layout (row/col = newtonpf.h:256-291 indices); returns the
 * entry count in *nnz (cap = capacity of the arrays). */
bpf_status bpf_stage_jacobian(bpf_engine *e, const double *V, int32_t t,
                              int32_t cap, int32_t *row, int32_t *col,
                              double *val, int32_t *nnz);

/* Linear solve of the Newton step (newtonpf.h:312-363 / 1086-1198): factor
 * every instance's Jacobian at V on the device (precision of the options)
 * and solve J dx = F; dx in the reference layout, stride 2*nbus. */
bpf_status bpf_stage_newton_step(bpf_engine *e, const double *V, double *dx);

/* The inverse of the dense tail's Schur complement of instance t, as the
 * last factorization (bpf_stage_newton_step or bpf_solve) left it (no
 * counterpart in the reference, whose LU is a library call; for tests).
 * tail_bus [tail_n] = the bus of each tail pivot in elimination order;
 * Sinv [(2 tail_n)^2] row-major, scalar index 2a + r for tail pivot a with
 * r = 0 the angle / P component and r = 1 the magnitude / Q component, in
 * the factorization's precision (FP32 values widened).  Either pointer may
 * be NULL. */
bpf_status bpf_stage_tail_inverse(bpf_engine *e, int32_t t, int32_t *tail_bus,
                                  double *Sinv);

/* Adjoint sensitivities at the converged point of the last bpf_solve.
 * Objective per instance: total active losses (runpf.h:228-257) plus
 * w * sum over branch ends of max(0, |S| - rate)^2.  Solves J^T lambda =
 * df/dx with FP64 factors and returns df/dP, df/dQ of bus injections
 * [T*nbus] and df/dx of branch reactances [T*nbranch].  rate may be NULL
 * (no penalty).  No counterpart in the reference. */
bpf_status bpf_adjoint(bpf_engine *e, const double *rate, double w,
                       double *grad_P, double *grad_Q, double *grad_x,
                       double *objective);

/* Kernel timing on the engine's stream with hipEvents recorded around
 * direct launches (enable before bpf_solve).  While timing is on, the
 * Newton loop launches its kernels directly instead of replaying the
 * iteration graph, so the events bracket single kernels.  Names: ybus, init,
 * mismatch, jacobian, lu_factor, lu_solve, residual, update, loss, bmat,
 * fd_solve, adjoint. */
bpf_status bpf_set_timing(bpf_engine *e, int32_t enable);
bpf_status bpf_get_kernel_time(const bpf_engine *e, const char *name,
                               int32_t *launches, double *total_ms);

/* Launch shapes (threads per workgroup) of the per-instance kernels; the
 * autotuner in mapol_amd/tune.py sets them.  0 keeps the current value. */
bpf_status bpf_set_launch(bpf_engine *e, int32_t lu_threads,
                          int32_t solve_threads, int32_t mismatch_threads);

/* Run-time schedule choices (the autotuner sets them; 0 keeps the current
 * value): newton_per_graph = Newton iterations per captured graph (1..16;
 * the host reads the active-instance counts once per round of replays);
 * tail_inverse = 1 matrix cores, 2 vector ALUs (FP32 dense-tail inverse);
 * sweep_form = 1 slot form, 2 entry-parallel, 3 level rows (the first that
 * applies, in that order, from the chosen one); parts = 1..4: the Newton
 * loop runs the batch as that many parts of consecutive instances, each on
 * its own stream with its own graph, replayed side by side (results do not
 * depend on it). */
enum { BPF_TAIL_MFMA = 1, BPF_TAIL_VALU = 2 };
enum { BPF_SWEEP_SLOT = 1, BPF_SWEEP_ENTRY = 2, BPF_SWEEP_ROWS = 3 };
bpf_status bpf_set_schedule(bpf_engine *e, int32_t newton_per_graph,
                            int32_t tail_inverse, int32_t sweep_form,
                            int32_t parts);

/* Split factorization (no counterpart in the reference, whose CPU solver
 * factors one instance per thread): while few instances are active, the
 * Newton step's factorization runs the subtrees below a cut of the
 * elimination tree with 2, 4 or 8 workgroups per instance (subtrees share
 * no operands), then the levels above the cut and the dense tail with one.
 * workgroups = 0: adaptive, the default (the most workgroups per instance
 * with active instances x workgroups within the compute units of the
 * batch's part); 1: never; 2, 4, 8: always that many where the analysis
 * planned them.  Results do not depend on it: every output is summed in
 * the same order. */
bpf_status bpf_set_factor_split(bpf_engine *e, int32_t workgroups);

/* Forward substitution inside the factorization (no counterpart in the
 * reference, whose solves run both sweeps): while at most `max_active`
 * instances of a batch part are active, the Newton step's factorization
 * also forms L^-1 of the mismatch and the first solve skips its forward
 * sweep; above that both run as usual (-1: never; default 32).  Only the
 * slot-form sweeps use it.  Results agree to rounding either way. */
bpf_status bpf_set_fused_forward(bpf_engine *e, int32_t max_active);

#ifdef __cplusplus
}
#endif
#endif
