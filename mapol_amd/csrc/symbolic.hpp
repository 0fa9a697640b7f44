// Host-side symbolic analysis of the batched power-flow engine, done once per
// topology (bpf_create).  Everything here is shared by all instances; only
// numerical values are per instance.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace bpf {

// Ybus contribution item: branch index * 4 + kind (0 Yff, 1 Yft, 2 Ytf, 3 Ytt)
enum { KFF = 0, KFT = 1, KTF = 2, KTT = 3 };

// Crout task destination flags (bits above the block position)
constexpr int TASK_HAS_A = 1 << 30;  // the position holds a Jacobian entry
constexpr int TASK_DIAG = 1 << 29;   // pivot block: store its inverse
constexpr int TASK_LCOL = 1 << 28;   // L block: scale by the pivot inverse
constexpr int TASK_POS_MASK = (1 << 28) - 1;
constexpr int CHUNK_PAIRS = 4096;    // products per factorization chunk (LDS)
constexpr int SWEEP_CHUNK = 2048;    // entries per sweep chunk (LDS)
constexpr int SLOT_CHUNK = 2048;     // lane slots per sweep chunk, slot form
constexpr int SLOT_LG_SHIFT = 27;    // slot column word: column + 1 | log2 group << 27
constexpr int SLOT_NOT_LAST = 1 << 30;  // slot row word: a segment before the row's last
constexpr int MAX_CHUNKS = 512;      // chunk tables the kernels keep in LDS
constexpr int TAIL_MAX = 64;         // dense tail: at most 64 blocks square
constexpr int SPLIT_MAX = 8;         // workgroups per instance, split factorization
constexpr int SPLIT_VARIANTS = 3;    // splits into 2, 4 and 8 workgroups
// FP32 factorization: distinct blocks of the tail products staged in LDS
// (with the chunk tables, which the tail no longer needs, in the same space)
constexpr int TAIL_STAGE_MAX = 5760;

// tile sizes of the schedules (0 = the largest the kernels hold): pivots
// of the dense tail, products per factorization chunk, lane slots per
// slot-form sweep chunk
struct AnalysisParams {
  int tail_limit = 0, chunk_pairs = 0, slot_chunk = 0;
  // bus ordering: 0 minimum degree, 1 minimum degree reordered to the least
  // elimination-tree height with the same fill, 2 nested dissection
  int ordering = 0;
  // an elimination order given by the caller (every PV and PQ bus once,
  // first eliminated first) in place of the one `ordering` selects
  std::vector<int> order_bus;
};

struct Symbolic {
  AnalysisParams prm;   // the values used (0s replaced by the maxima)
  int n = 0, m = 0, nb = 0;
  // Ybus pattern over all buses (every branch, in service or not: outages
  // only zero values, as the reference keeps the structure, makeYbus.h:69-71)
  std::vector<int> yrow, ycol, yrowof;  // CSR + row index of each position
  std::vector<int> ydiag_bus;           // bus index if position is (i,i), else -1
  // contributions per position, in the reference's summation order: from-side
  // terms in branch order, then to-side terms in branch order
  std::vector<int> yf_ptr, yf_item, yt_ptr, yt_item;
  // branches incident to each bus (from or to end), ascending branch index
  std::vector<int> bbr_ptr, bbr;
  // non-slack buses become 2x2 blocks, in elimination order
  std::vector<int> blk_of_bus, bus_of_blk;
  // block CSR of L+U (structurally symmetric), rows in elimination order
  std::vector<int> lrow, lcol, ldiag, mirror;
  std::vector<int> y2lu;  // Ybus position -> block position (-1: slack)
  // the Jacobian blocks sorted by block position: position, Ybus position
  std::vector<int> jac_lu, jac_y;
  // Level 0 of the elimination tree (no descendants, no fill) finished by
  // the Jacobian fill: per Jacobian block -1 (nothing), -2 (a level-0 pivot:
  // store its inverse) or the Jacobian-block index of its level-0 pivot (an
  // L block: times the pivot's inverse).  l0_chunks = the factorization
  // chunks this replaces (0: the fill leaves level 0 to the factorization).
  std::vector<int> jac_l0;
  int l0_chunks = 0;
  int nnz_jac_blocks = 0;
  // Crout factorization schedule (one level = one elimination-tree height)
  int nlev = 0;
  // chunks of the levels below the tail, in execution order; level l owns
  // chunks lv_chunk[l] .. lv_chunk[l+1]-1
  int nchunk = 0;
  std::vector<int> lv_chunk;
  std::vector<int> ch_task, ch_pair;  // task / pair offsets [nchunk+1]
  std::vector<int> ch_max;            // longest reduction of the chunk
  // the forward substitution's outputs of a chunk come last, their products
  // too: per chunk {outputs, products} of them (the factorization leaves
  // them out when the solve runs its own forward sweep); sp_rhs the same
  // for the split schedules' chunks
  std::vector<int> ch_rhs, sp_rhs;
  // per task: tk_a = {dst | flags, first pair, end pair, 0};
  // tk_b = {0, the pivot task's first pair, end pair, 0} (L tasks)
  std::vector<int> tk_a, tk_b;
  // {L position, U position} per product; U position -1: init entry, the
  // value at the L position itself (a pivot's Jacobian block)
  std::vector<int> pairs;
  int64_t npairs = 0;
  // Split factorization, for few instances at a time: variant v splits
  // the levels below sp_level[v] into K = 2 << v groups of whole subtrees
  // of the elimination tree (no output of one group is an operand of
  // another), each run by a workgroup of its own; then one workgroup runs
  // the main chunks from sp_rest[v] and the dense tail.  A group's chunks
  // are the main chunks' outputs restricted to the group, in the main order
  // (sp_tk_a, sp_tk_b, sp_pairs as tk_a, tk_b, pairs); each keeps the task
  // count and longest reduction of its main chunk (sp_lg), which set the
  // lane groups, so every output is summed exactly as in the main schedule.
  // sp_ch = {first task, end task, first product, end product} per chunk;
  // group g of variant v owns chunks sp_grp[v * (SPLIT_MAX + 1) + g] up to
  // the next entry.  sp_level[v] = -1: the variant is not worth a launch.
  std::vector<int> sp_level, sp_rest, sp_grp;
  std::vector<int> sp_ch, sp_lg;
  std::vector<int> sp_tk_a, sp_tk_b, sp_pairs;
  double sp_model_us[SPLIT_VARIANTS + 1] = {0, 0, 0, 0};  // modelled chunk time, one instance
  // dense tail: the pivots of the top levels (>= tail_level, at most
  // TAIL_MAX of them, ascending = topological) are factored as one dense
  // block in LDS.  Its entries start from A minus the contributions of the
  // non-tail pivots (pairs below), then a right-looking dense LU finishes.
  int tail_level = 0;     // sparse levels 0 .. tail_level-1
  int tail_n = 0;         // pivots in the tail
  std::vector<int> tail_piv;                 // tail pivots, ascending
  std::vector<int> tl_dense, tl_pos, tl_pp;  // per tail entry (flags in tl_dense)
  std::vector<int> tl_pairL, tl_pairU;
  int tl_max = 0;                            // longest tail pair list
  // the distinct L and U blocks of the tail pairs (positions, ascending) and
  // each pair's two indices into that list (L | U << 16); empty when more
  // than TAIL_STAGE_MAX blocks
  std::vector<int> tl_stage, tl_pairS;
  // The factorization leaves the inverse of the tail's Schur complement
  // (dense, tail_n x tail_n blocks); the sweeps solve the non-tail rows by
  // levels and the tail rows with that inverse.  Tail row a (block
  // tail_piv[a]) first subtracts its L entries in non-tail columns:
  std::vector<int> ts_ptr, ts_pos, ts_col, ts_mir;  // position, column, mirror
  int ts_max = 0;
  // triangular solves: non-tail rows grouped by dependency level (tail rows
  // count as solved before the backward sweep), longest rows first
  int nfw = 0, nbw = 0;
  std::vector<int> fw_ptr, fw_rows, bw_ptr, bw_rows;
  std::vector<int> fw_max, bw_max;           // longest row per level
  // entry-parallel form of the same sweeps: the rows of each level packed
  // into chunks of at most SWEEP_CHUNK entries.  Chunks 0 .. sw_nfw-1 are
  // the forward sweep, the rest the backward sweep.
  int sw_nfw = 0, sw_nbw = 0, sw_ok = 0;
  std::vector<int> sw_chunk;   // {first row, end row, first entry, end entry}
  std::vector<int> sw_cmax;    // longest row of the chunk
  std::vector<int> sw_row;     // {block row, first entry, end entry, pivot position}
  std::vector<int> sw_ent;     // {L or U position, column block}
  // slot form of the same sweeps: every row of a chunk owns a lane group
  // (the power of two at least its length, at most 64 lanes), groups packed
  // largest first so each starts at a multiple of its size; one slot per
  // entry, padding slots carry none.  A row of more than 64 entries comes in
  // segments of 64 in successive chunks of its level (all but the last
  // marked SLOT_NOT_LAST in the row word).  A chunk holds at most SLOT_CHUNK
  // slots; chunks 0 .. sl_nfw-1 are the forward sweep.
  int sl_nfw = 0, sl_nbw = 0, sl_ok = 0;
  std::vector<int> sl_chunk;   // {first slot, end slot, log2 of the largest group, 0}
  // per slot {position (padding: the row's first entry), column + 1 (0:
  // padding) | log2 group << SLOT_LG_SHIFT, block row (the group's first
  // slot; | SLOT_NOT_LAST for an earlier segment) or -1, pivot position
  // (first slot of a row's last segment, backward) or the position}
  std::vector<int> sl_slot;
  // the slot form of the transposed sweeps (adjoint, U^T forward with the
  // pivots, L^T backward): mirror positions, chunks of their own
  int sl_nfw_t = 0, sl_nbw_t = 0, sl_ok_t = 0;
  std::vector<int> sl_chunk_t, sl_slot_t;
};

// bus_type: MATPOWER codes; buses of type PV or PQ are unknowns, every other
// bus keeps its voltage (slack).  Returns false with a message on bad input.
bool analyse(int n, int m, const int *from, const int *to,
             const int *bus_type, Symbolic &S, std::string *err,
             const AnalysisParams &prm = AnalysisParams());

// the elimination order `analyse` derives for `ordering` from the case as
// numbered here: the PV and PQ buses, first eliminated first
std::vector<int> elimination_order(int n, int m, const int *from, const int *to,
                                   const int *bus_type, int ordering);

// reverse Cuthill-McKee numbering of the bus graph (every branch, in service
// or not): entry k is the bus that takes number k.  Neighbouring buses get
// near numbers, so gathers over a bus's branches touch few cache lines.
std::vector<int> rcm_numbering(int n, int m, const int *from, const int *to);

// the tail entries (longest pair lists first, list ends `pp`) in jobs of
// one wave each: a job takes consecutive entries with one lane group size,
// the smallest power of two (up to 64) that covers its first, longest list
// with at most `per_lane` products per lane.  Packed {first entry,
// end entry << 3 | log2 group size}.
std::vector<int> tail_jobs(const std::vector<int> &pp, int per_lane);

// minimum-degree order of a symmetric graph given as adjacency lists
std::vector<int> min_degree_order(std::vector<std::vector<int>> adj);

// the same filled graph as `order`, with the least elimination-tree height
std::vector<int> min_height_reorder(const std::vector<std::vector<int>> &adj,
                                    const std::vector<int> &order);

// nested-dissection order (separators last; parts of at most `leaf`
// vertices by minimum degree)
std::vector<int> nested_dissection_order(const std::vector<std::vector<int>> &adj,
                                         int leaf);

}  // namespace bpf
