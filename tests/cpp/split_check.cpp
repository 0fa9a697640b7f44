// The split factorization's schedules against the main one.  For every
// split variant the analysis planned (K = 2, 4, 8 workgroups per instance):
//
// * the groups' chunks hold every output of the main chunks below the cut
//   (from the first chunk the Jacobian fill leaves) exactly once, each with
//   the same product list in the same order, and an L output's pivot range
//   is its pivot's own products in the same chunk;
// * each split chunk sums with the lane groups of the main chunks its
//   outputs come from, for every workgroup size (256, 512, 1024), so each
//   output is summed in the same order as in the main schedule;
// * a group reads no output of another group nor of a chunk above the cut,
//   and reads an output of its own only from an earlier chunk;
// * the group chunk tables are contiguous and in range;
// * the dense tail's wave jobs cover every tail entry once, each list
//   within its job's lane groups.
//
// Reads "n m ordering", n bus types and m branch end pairs from stdin;
// prints one line per variant and "ok" (exit 0) or the first violation
// (exit 1).
#include <algorithm>
#include <array>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../mapol_amd/csrc/symbolic.hpp"

using namespace bpf;

static std::array<int, 3> lane_groups(int items, int maxlen) {
  std::array<int, 3> r;
  for (int b = 0; b < 3; b++) {
    int g = 1;
    while (g < 64 && g < maxlen && items * 2 * g <= (256 << b)) g <<= 1;
    r[b] = g;
  }
  return r;
}

static int fail(const char *what, int v, int a, int b) {
  printf("FAIL variant %d: %s (%d, %d)\n", v, what, a, b);
  return 1;
}

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
  Symbolic S;
  std::string err;
  if (!analyse(n, m, f.data(), t.data(), ty.data(), S, &err, p)) {
    printf("analysis failed: %s\n", err.c_str());
    return 2;
  }
  // main outputs: position -> (chunk, task)
  std::map<int, std::pair<int, int>> main_of;
  for (int c = 0; c < S.nchunk; c++)
    for (int k = S.ch_task[c]; k < S.ch_task[c + 1]; k++)
      main_of[S.tk_a[4 * k] & TASK_POS_MASK] = {c, k};
  const int nsp = (int)S.sp_lg.size() / 2;
  int planned = 0;
  for (int v = 0; v < SPLIT_VARIANTS; v++) {
    const int K = 2 << v, *gp = &S.sp_grp[v * (SPLIT_MAX + 1)];
    if (S.sp_level[v] < 0) {
      for (int j = 0; j < SPLIT_MAX; j++)
        if (gp[j] != gp[j + 1]) return fail("chunks in a variant not planned", v, j, gp[j]);
      printf("variant K=%d: not planned (model %.1f us against %.1f)\n", K,
             S.sp_model_us[v + 1], S.sp_model_us[0]);
      continue;
    }
    planned++;
    const int cut = S.sp_rest[v];
    if (cut < S.l0_chunks || cut > S.nchunk) return fail("cut out of range", v, cut, S.nchunk);
    for (int j = 0; j < SPLIT_MAX; j++)
      if (gp[j] > gp[j + 1] || gp[j + 1] > nsp) return fail("group table", v, j, gp[j]);
    for (int j = K; j < SPLIT_MAX; j++)
      if (gp[j] != gp[j + 1]) return fail("chunks beyond K groups", v, j, gp[j]);
    std::map<int, int> seen;             // main output position -> group
    std::map<int, std::pair<int, int>> made;   // position -> (group, chunk)
    for (int j = 0; j < K; j++) {
      for (int c = gp[j]; c < gp[j + 1]; c++) {
        const int *ch = &S.sp_ch[4 * c];
        if (ch[0] >= ch[1] || ch[3] - ch[2] > CHUNK_PAIRS)
          return fail("empty or oversized chunk", v, c, ch[3] - ch[2]);
        const std::array<int, 3> lg = lane_groups(S.sp_lg[2 * c], S.sp_lg[2 * c + 1]);
        int q = ch[2];
        for (int k = ch[0]; k < ch[1]; k++) {
          const int *a = &S.sp_tk_a[4 * k], *b = &S.sp_tk_b[4 * k];
          const int pos = a[0] & TASK_POS_MASK;
          auto it = main_of.find(pos);
          if (it == main_of.end()) return fail("output not in the main schedule", v, pos, 0);
          const int mc = it->second.first, mk = it->second.second;
          if (mc < S.l0_chunks || mc >= cut) return fail("output above the cut", v, pos, mc);
          if (seen.count(pos)) return fail("output twice", v, pos, seen[pos]);
          seen[pos] = j;
          if (a[0] != S.tk_a[4 * mk]) return fail("flags differ", v, pos, a[0]);
          if (a[1] != q || a[2] < a[1] || a[2] > ch[3]) return fail("product range", v, pos, a[1]);
          q = a[2];
          const int *ma = &S.tk_a[4 * mk];
          if (a[2] - a[1] != ma[2] - ma[1]) return fail("product count", v, pos, a[2] - a[1]);
          for (int r = 0; r < a[2] - a[1]; r++)
            if (S.sp_pairs[2 * (a[1] + r)] != S.pairs[2 * (ma[1] + r)] ||
                S.sp_pairs[2 * (a[1] + r) + 1] != S.pairs[2 * (ma[1] + r) + 1])
              return fail("product differs", v, pos, r);
          if (lane_groups(S.ch_task[mc + 1] - S.ch_task[mc], S.ch_max[mc]) != lg)
            return fail("lane groups differ from the main chunk's", v, pos, mc);
          if (a[0] & TASK_LCOL) {
            // the pivot's own products, in this chunk
            const int *mb = &S.tk_b[4 * mk];
            if (b[1] < ch[2] || b[2] > ch[3] || b[2] - b[1] != mb[2] - mb[1])
              return fail("pivot range", v, pos, b[1]);
            for (int r = 0; r < b[2] - b[1]; r++)
              if (S.sp_pairs[2 * (b[1] + r)] != S.pairs[2 * (mb[1] + r)] ||
                  S.sp_pairs[2 * (b[1] + r) + 1] != S.pairs[2 * (mb[1] + r) + 1])
                return fail("pivot product differs", v, pos, r);
          }
          // operands: an output of this group's earlier chunks, or no output
          // of the factorization at or above the fill's chunks
          for (int r = a[1]; r < a[2]; r++)
            for (int h = 0; h < 2; h++) {
              const int op = S.sp_pairs[2 * r + h];
              if (op < 0) continue;
              if (h == 1 || S.sp_pairs[2 * r + 1] >= 0 || op != pos) {
                auto mo = made.find(op);
                if (mo != made.end()) {
                  if (mo->second.first != j) return fail("reads another group", v, pos, op);
                  if (mo->second.second >= c) return fail("reads a later chunk", v, pos, op);
                } else {
                  auto w = main_of.find(op);
                  if (w != main_of.end() && w->second.first >= S.l0_chunks)
                    return fail("reads an output not yet made", v, pos, op);
                }
              }
            }
        }
        if (q != ch[3]) return fail("products not covered", v, c, q);
        for (int k = ch[0]; k < ch[1]; k++)
          made[S.sp_tk_a[4 * k] & TASK_POS_MASK] = {j, c};
      }
    }
    int below = 0;
    for (int c = S.l0_chunks; c < cut; c++) below += S.ch_task[c + 1] - S.ch_task[c];
    if ((int)seen.size() != below) return fail("outputs missing", v, (int)seen.size(), below);
    printf("variant K=%d: cut level %d, first top chunk %d, %d chunks, model %.1f us against %.1f\n",
           K, S.sp_level[v], cut, gp[K] - gp[0], S.sp_model_us[v + 1], S.sp_model_us[0]);
  }
  // the dense tail's wave jobs: consecutive entries covering every one once,
  // at most a wave of lanes each, every list within per_lane products per
  // lane of its group (or a 64-lane group), the group size the least that
  // holds the job's first list
  for (int per_lane : {12, 4}) {
    const std::vector<int> jobs = tail_jobs(S.tl_pp, per_lane);
    const int ne = (int)S.tl_pp.size() - 1;
    int next = 0;
    for (size_t j = 0; j + 1 < jobs.size(); j += 2) {
      const int e0 = jobs[j], e1 = jobs[j + 1] >> 3, lg = jobs[j + 1] & 7;
      if (e0 != next || e1 <= e0) return fail("tail jobs not contiguous", -1, e0, next);
      if (lg > 6 || e1 - e0 > (64 >> lg)) return fail("tail job wider than a wave", -1, e0, lg);
      for (int k = e0; k < e1; k++) {
        const int len = S.tl_pp[k + 1] - S.tl_pp[k];
        if (lg < 6 && (len + (1 << lg) - 1) >> lg > per_lane)
          return fail("tail list beyond its lane group", -1, k, len);
      }
      const int len0 = S.tl_pp[e0 + 1] - S.tl_pp[e0];
      if (lg > 0 && (len0 + (1 << (lg - 1)) - 1) >> (lg - 1) <= per_lane)
        return fail("tail job group larger than needed", -1, e0, lg);
      next = e1;
    }
    if (next != ne) return fail("tail entries not covered", -1, next, ne);
  }
  printf("ok, %d variants\n", planned);
  return 0;
}
