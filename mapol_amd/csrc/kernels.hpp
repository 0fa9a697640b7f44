// Device-side data types and kernel declarations of the batched power-flow
// engine.  Layout in HBM: every per-instance array is a slab [T][...] with
// the instance index outermost, so a workgroup that owns one instance reads
// one contiguous slab.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bpf {

// verdict codes (same values as BPF_CONVERGED ... in include/batchpf.h)
enum { BPF_V_CONVERGED = 0, BPF_V_MAX_ITER = 1, BPF_V_SINGULAR = 2,
       BPF_V_NONFINITE = 3 };

struct cd { double x, y; };  // complex FP64 (re, im), 16-byte element

// 2x2 Jacobian block [a b; c d]: rows (P, Q), columns (angle, magnitude)
template <typename T> struct alignas(4 * sizeof(T)) B4 { T a, b, c, d; };

// per-instance solver state
struct InstState {
  int done;      // 1: no more Newton iterations for this instance
  int verdict;   // BPF_CONVERGED / MAX_ITER / SINGULAR / NONFINITE
  int it_round;  // Newton iterations in the current reactive-limit round
  int it_total;  // all Newton iterations
  int rounds;    // reactive-limit rounds used
  int fd_half;   // FDPF: 0 before the P half-step, 1 before the Q half-step
  double norm;   // last inf-norm of the mismatch
};

// shared (topology) arrays on the device
struct DevTopo {
  // nnzLU: blocks per instance of the factor storage, the factor's own
  // (rhs0 of them) followed by nb right-hand-side blocks (the forward
  // substitution the factorization forms for the Newton step)
  int n, m, nb, nnzY, nnzLU, rhs0;
  const int *yrow, *ycol, *yrowof, *ydiag_bus;
  const int *yblk;         // per Ybus entry: block of its column bus (-1: slack)
  const int *yf_ptr, *yf_item, *yt_ptr, *yt_item;
  const int *blk_of_bus, *bus_of_blk, *y2lu;
  const int *lrow, *lcol, *ldiag, *mirror;
  const int *br_from, *br_to, *bbr_ptr, *bbr;
  const double *R, *X, *B, *shift, *Gs, *Qd, *Va0;
  const cd *ys, *ejs;      // per branch: 1 / (R + jX), e^(j shift) (k_branch_pre)
  const int *base_type;
  // Jacobian blocks in L+U position order: position, Ybus position
  int nJb;
  const int *jac_lu, *jac_y;
  const int4 *jac4;        // per Jacobian block {L+U position, Ybus entry, row bus, column bus}
  const int *jac_l0;       // level-0 role of each Jacobian block (elim0 fill)
};

// per-instance arrays on the device
struct DevInst {
  int T;
  const uint8_t *br_on;    // [T][m]
  const double *tap;       // [T][m]
  const double *Bs;        // [T][n]
  const double *vset;      // [T][n]
  int8_t *type;            // [T][n] current bus types
  double *Pinj, *Qinj;     // [T][n]
  const double *Qmax, *Qmin;  // [T][n]
  cd *Y;                   // [T][nnzY]
  cd *V;                   // [T][n]
  cd *Ib;                  // [T][n] Ybus*V at the last mismatch
  double *F;               // [T][2nb] block order (P, Q) per block
  double *dx;              // [T][2nb]
  double *rhs;             // [T][2nb] refinement residual (block order)
  InstState *st;           // [T]
  int *active_count;       // [1]
  // partial mismatch norms, one per mismatch workgroup (NaN: a NaN entry)
  double *part;            // [T][n / 64 rounded up]
  int *partv;              // [T][n / 64 rounded up] PV buses beyond their Q limits
  int *jsing;              // [T] singular level-0 pivot seen by the fill
};

}  // namespace bpf
