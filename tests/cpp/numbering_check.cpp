// The engine's own bus numbering keeps the caller's symbolic analysis: a
// case renumbered by rcm_numbering and analysed with the elimination order
// taken from the caller's numbering has the same block pattern, the same
// factorization and sweep schedules and the same dense tail as the case as
// given, array for array.  Reads "n m ordering", n bus types and m branch
// end pairs from stdin; prints "identical" (exit 0) or the arrays that
// differ (exit 1), and the mean |from - to| over branches before and after.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../mapol_amd/csrc/symbolic.hpp"

using namespace bpf;

int main() {
  int n, m, ordering;
  if (scanf("%d %d %d", &n, &m, &ordering) != 3) return 2;
  std::vector<int> ty(n), f(m), t(m);
  for (int i = 0; i < n; i++)
    if (scanf("%d", &ty[i]) != 1) return 2;
  for (int b = 0; b < m; b++)
    if (scanf("%d %d", &f[b], &t[b]) != 2) return 2;
  AnalysisParams p;
  p.ordering = ordering;
  Symbolic A, B;
  std::string err;
  if (!analyse(n, m, f.data(), t.data(), ty.data(), A, &err, p)) {
    printf("analysis failed: %s\n", err.c_str());
    return 2;
  }
  const std::vector<int> ord =
      elimination_order(n, m, f.data(), t.data(), ty.data(), ordering);
  const std::vector<int> ext = rcm_numbering(n, m, f.data(), t.data());
  std::vector<int> in(n, -1);
  if ((int)ext.size() != n) { printf("numbering has %zu entries\n", ext.size()); return 1; }
  for (int k = 0; k < n; k++) {
    if (ext[k] < 0 || ext[k] >= n || in[ext[k]] >= 0) {
      printf("numbering is not a permutation\n");
      return 1;
    }
    in[ext[k]] = k;
  }
  std::vector<int> ty2(n), f2(m), t2(m);
  for (int k = 0; k < n; k++) ty2[k] = ty[ext[k]];
  for (int b = 0; b < m; b++) { f2[b] = in[f[b]]; t2[b] = in[t[b]]; }
  AnalysisParams q = p;
  for (int b : ord) q.order_bus.push_back(in[b]);
  if (!analyse(n, m, f2.data(), t2.data(), ty2.data(), B, &err, q)) {
    printf("renumbered analysis failed: %s\n", err.c_str());
    return 1;
  }
  int bad = 0;
#define SAME(x) if (A.x != B.x) { printf("differs: %s\n", #x); bad = 1; }
  SAME(nb) SAME(lrow) SAME(lcol) SAME(ldiag) SAME(mirror)
  SAME(nchunk) SAME(ch_task) SAME(ch_pair) SAME(ch_max) SAME(tk_a) SAME(tk_b) SAME(pairs)
  SAME(tail_n) SAME(tail_piv) SAME(tl_dense) SAME(tl_pos) SAME(tl_pp)
  SAME(tl_pairL) SAME(tl_pairU) SAME(tl_stage) SAME(tl_pairS)
  SAME(ts_ptr) SAME(ts_pos) SAME(ts_col) SAME(ts_mir)
  SAME(fw_ptr) SAME(fw_rows) SAME(bw_ptr) SAME(bw_rows)
  SAME(sw_chunk) SAME(sw_row) SAME(sw_ent)
  SAME(sl_chunk) SAME(sl_slot) SAME(sl_chunk_t) SAME(sl_slot_t) SAME(jac_l0)
#undef SAME
  // the same bus behind every block, and the order analyse computes itself
  for (int k = 0; k < A.nb && !bad; k++)
    if (A.bus_of_blk[k] != ext[B.bus_of_blk[k]] || A.bus_of_blk[k] != ord[k]) {
      printf("differs: bus of block %d\n", k);
      bad = 1;
    }
  std::vector<int> ja = A.jac_lu, jb = B.jac_lu;
  std::sort(ja.begin(), ja.end());
  std::sort(jb.begin(), jb.end());
  if (ja != jb) { printf("differs: Jacobian block set\n"); bad = 1; }
  double s1 = 0, s2 = 0;
  for (int b = 0; b < m; b++) { s1 += abs(f[b] - t[b]); s2 += abs(f2[b] - t2[b]); }
  printf("%s; mean |from - to| %.1f -> %.1f\n", bad ? "MISMATCH" : "identical",
         m ? s1 / m : 0.0, m ? s2 / m : 0.0);
  return bad;
}
