// Symbolic analysis: Ybus pattern, bus ordering, L+U block pattern and the
// level schedules of the batched factorization and triangular solves.
#include <algorithm>
#include <numeric>
#include <string>

#include "symbolic.hpp"

namespace bpf {

std::vector<int> min_degree_order(std::vector<std::vector<int>> adj) {
  const int n = (int)adj.size();
  for (auto &a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  std::vector<int> head(n + 1, -1), next(n, -1), prev(n, -1), deg(n);
  auto insert = [&](int v) {
    int d = deg[v];
    next[v] = head[d];
    prev[v] = -1;
    if (head[d] >= 0) prev[head[d]] = v;
    head[d] = v;
  };
  auto remove = [&](int v) {
    if (prev[v] >= 0) next[prev[v]] = next[v]; else head[deg[v]] = next[v];
    if (next[v] >= 0) prev[next[v]] = prev[v];
  };
  for (int v = n - 1; v >= 0; v--) { deg[v] = (int)adj[v].size(); insert(v); }
  std::vector<int> order;
  order.reserve(n);
  std::vector<int> merged;
  int mindeg = 0;
  for (int k = 0; k < n; k++) {
    while (head[mindeg] < 0) mindeg++;
    int v = head[mindeg];
    remove(v);
    order.push_back(v);
    const std::vector<int> N = std::move(adj[v]);
    adj[v].clear();
    for (int u : N) {
      remove(u);
      // adj[u] <- (adj[u] U N) \ {u, v}: eliminating v makes N a clique
      merged.clear();
      std::set_union(adj[u].begin(), adj[u].end(), N.begin(), N.end(),
                     std::back_inserter(merged));
      adj[u].clear();
      for (int x : merged)
        if (x != u && x != v) adj[u].push_back(x);
      deg[u] = (int)adj[u].size();
      insert(u);
      if (deg[u] < mindeg) mindeg = deg[u];
    }
  }
  return order;
}

static int find_pos(const std::vector<int> &rp, const std::vector<int> &ci,
                    int r, int c) {
  auto b = ci.begin() + rp[r], e = ci.begin() + rp[r + 1];
  auto it = std::lower_bound(b, e, c);
  if (it == e || *it != c) return -1;
  return (int)(it - ci.begin());
}

bool analyse(int n, int m, const int *from, const int *to, const int *type,
             Symbolic &S, std::string *err) {
  S = Symbolic();
  S.n = n;
  S.m = m;
  for (int b = 0; b < m; b++) {
    if (from[b] < 0 || from[b] >= n || to[b] < 0 || to[b] >= n) {
      if (err) *err = "branch " + std::to_string(b) + " has a bus out of range";
      return false;
    }
  }
  // ---- Ybus pattern ------------------------------------------------------
  std::vector<std::vector<int>> cols(n);
  for (int i = 0; i < n; i++) cols[i].push_back(i);
  for (int b = 0; b < m; b++) {
    cols[from[b]].push_back(to[b]);
    cols[to[b]].push_back(from[b]);
  }
  S.yrow.assign(n + 1, 0);
  for (int i = 0; i < n; i++) {
    auto &c = cols[i];
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    S.yrow[i + 1] = S.yrow[i] + (int)c.size();
  }
  const int nnzY = S.yrow[n];
  S.ycol.resize(nnzY);
  S.yrowof.resize(nnzY);
  S.ydiag_bus.assign(nnzY, -1);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < (int)cols[i].size(); k++) {
      int p = S.yrow[i] + k;
      S.ycol[p] = cols[i][k];
      S.yrowof[p] = i;
      if (cols[i][k] == i) S.ydiag_bus[p] = i;
    }
  // contributions: the from side of branch b lands in row from[b] (Yff at the
  // diagonal, Yft at column to[b]); the to side in row to[b] (Ytt, Ytf)
  std::vector<std::vector<int>> fl(nnzY), tl(nnzY);
  for (int b = 0; b < m; b++) {
    int f = from[b], t = to[b];
    fl[find_pos(S.yrow, S.ycol, f, f)].push_back(b * 4 + KFF);
    fl[find_pos(S.yrow, S.ycol, f, t)].push_back(b * 4 + KFT);
    tl[find_pos(S.yrow, S.ycol, t, t)].push_back(b * 4 + KTT);
    tl[find_pos(S.yrow, S.ycol, t, f)].push_back(b * 4 + KTF);
  }
  S.yf_ptr.assign(nnzY + 1, 0);
  S.yt_ptr.assign(nnzY + 1, 0);
  for (int p = 0; p < nnzY; p++) {
    S.yf_ptr[p + 1] = S.yf_ptr[p] + (int)fl[p].size();
    S.yt_ptr[p + 1] = S.yt_ptr[p] + (int)tl[p].size();
    S.yf_item.insert(S.yf_item.end(), fl[p].begin(), fl[p].end());
    S.yt_item.insert(S.yt_item.end(), tl[p].begin(), tl[p].end());
  }
  S.bbr_ptr.assign(n + 1, 0);
  for (int b = 0; b < m; b++) {
    S.bbr_ptr[from[b] + 1]++;
    if (to[b] != from[b]) S.bbr_ptr[to[b] + 1]++;
  }
  for (int i = 0; i < n; i++) S.bbr_ptr[i + 1] += S.bbr_ptr[i];
  S.bbr.resize(S.bbr_ptr[n]);
  {
    std::vector<int> w(S.bbr_ptr.begin(), S.bbr_ptr.end() - 1);
    for (int b = 0; b < m; b++) {
      S.bbr[w[from[b]]++] = b;
      if (to[b] != from[b]) S.bbr[w[to[b]]++] = b;
    }
  }
  // ---- blocks and ordering -------------------------------------------------
  std::vector<int> unk;
  for (int i = 0; i < n; i++)
    if (type[i] == 1 || type[i] == 2) unk.push_back(i);
  const int nb = (int)unk.size();
  S.nb = nb;
  if (nb == 0) {
    if (err) *err = "no PV or PQ bus";
    return false;
  }
  std::vector<int> loc(n, -1);
  for (int k = 0; k < nb; k++) loc[unk[k]] = k;
  std::vector<std::vector<int>> adj(nb);
  for (int b = 0; b < m; b++) {
    int a = loc[from[b]], c = loc[to[b]];
    if (a >= 0 && c >= 0 && a != c) { adj[a].push_back(c); adj[c].push_back(a); }
  }
  std::vector<int> order = min_degree_order(adj);
  S.bus_of_blk.resize(nb);
  S.blk_of_bus.assign(n, -1);
  std::vector<int> newof(nb);
  for (int k = 0; k < nb; k++) {
    S.bus_of_blk[k] = unk[order[k]];
    S.blk_of_bus[unk[order[k]]] = k;
    newof[order[k]] = k;
  }
  // ---- symbolic factorization (symmetric structure) -----------------------
  std::vector<std::vector<int>> upper(nb);  // original upper adjacency
  for (int a = 0; a < nb; a++)
    for (int c : adj[a]) {
      int x = newof[a], y = newof[c];
      if (x < y) upper[x].push_back(y);
    }
  std::vector<std::vector<int>> st(nb);  // struct(k): rows > k of column k
  std::vector<std::vector<int>> children(nb);
  std::vector<int> parent(nb, -1);
  std::vector<int> tmp;
  for (int k = 0; k < nb; k++) {
    auto &s = st[k];
    s = upper[k];
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    for (int ch : children[k]) {
      tmp.clear();
      std::set_union(s.begin(), s.end(), st[ch].begin(), st[ch].end(),
                     std::back_inserter(tmp));
      s.clear();
      for (int x : tmp)
        if (x != k) s.push_back(x);
    }
    if (!s.empty()) {
      parent[k] = s.front();
      children[s.front()].push_back(k);
    }
  }
  // Lrow(i) = {k < i : i in st[k]}
  std::vector<std::vector<int>> lrowv(nb);
  for (int k = 0; k < nb; k++)
    for (int i : st[k]) lrowv[i].push_back(k);
  S.lrow.assign(nb + 1, 0);
  for (int i = 0; i < nb; i++)
    S.lrow[i + 1] = S.lrow[i] + (int)lrowv[i].size() + 1 + (int)st[i].size();
  const int nnzLU = S.lrow[nb];
  S.lcol.resize(nnzLU);
  S.ldiag.resize(nb);
  for (int i = 0; i < nb; i++) {
    int p = S.lrow[i];
    for (int k : lrowv[i]) S.lcol[p++] = k;  // ascending (k loop ascending)
    S.ldiag[i] = p;
    S.lcol[p++] = i;
    for (int j : st[i]) S.lcol[p++] = j;
  }
  auto pos = [&](int r, int c) { return find_pos(S.lrow, S.lcol, r, c); };
  S.mirror.resize(nnzLU);
  for (int i = 0; i < nb; i++)
    for (int p = S.lrow[i]; p < S.lrow[i + 1]; p++) S.mirror[p] = pos(S.lcol[p], i);
  // Ybus positions -> block positions
  S.y2lu.assign(nnzY, -1);
  std::vector<char> inJ(nnzLU, 0);
  for (int p = 0; p < nnzY; p++) {
    int bi = S.blk_of_bus[S.yrowof[p]], bj = S.blk_of_bus[S.ycol[p]];
    if (bi < 0 || bj < 0) continue;
    int q = pos(bi, bj);
    if (q < 0) {
      if (err) *err = "internal: Jacobian block outside the L+U pattern";
      return false;
    }
    S.y2lu[p] = q;
    if (!inJ[q]) { inJ[q] = 1; S.nnz_jac_blocks++; }
  }
  // ---- Crout schedule ------------------------------------------------------
  std::vector<int> lev(nb, 0);
  for (int k = 0; k < nb; k++)
    if (parent[k] >= 0) lev[parent[k]] = std::max(lev[parent[k]], lev[k] + 1);
  S.nlev = nb ? 1 + *std::max_element(lev.begin(), lev.end()) : 0;
  std::vector<std::vector<int>> piv(S.nlev);
  for (int k = 0; k < nb; k++) piv[lev[k]].push_back(k);
  struct Task { int dst; std::vector<int> L, U; };
  S.lvA.assign(1, 0);
  S.lvB.assign(1, 0);
  S.tA_pp.assign(1, 0);
  for (int l = 0; l < S.nlev; l++) {
    std::vector<Task> tasks;
    for (int k : piv[l]) {
      // outputs of pivot k: diag, U row k, L column k (rows = st[k])
      const auto &sk = st[k];
      std::vector<Task> out(1 + 2 * sk.size());
      out[0].dst = S.ldiag[k] | TASK_DIAG;
      for (size_t a = 0; a < sk.size(); a++) {
        out[1 + a].dst = pos(k, sk[a]);
        out[1 + sk.size() + a].dst = pos(sk[a], k);
      }
      for (int mm : lrowv[k]) {          // descendants m with k in st[m]
        int pkm = pos(k, mm), pmk = pos(mm, k);
        // U_kk += L_km U_mk
        out[0].L.push_back(pkm);
        out[0].U.push_back(pmk);
        for (int j : st[mm]) {
          if (j <= k) continue;
          size_t a = std::lower_bound(sk.begin(), sk.end(), j) - sk.begin();
          // U_kj -= L_km U_mj ;  L_jk -= L_jm U_mk
          out[1 + a].L.push_back(pkm);
          out[1 + a].U.push_back(pos(mm, j));
          out[1 + sk.size() + a].L.push_back(pos(j, mm));
          out[1 + sk.size() + a].U.push_back(pmk);
        }
      }
      for (size_t a = 0; a < sk.size(); a++) {
        S.tB_dst.push_back(out[1 + sk.size() + a].dst);
        S.tB_diag.push_back(S.ldiag[k]);
      }
      for (auto &t : out) tasks.push_back(std::move(t));
    }
    // longest dot products first: better balance across the workgroup
    std::stable_sort(tasks.begin(), tasks.end(), [](const Task &a, const Task &b) {
      return a.L.size() > b.L.size();
    });
    S.lvA_max.push_back(tasks.empty() ? 0 : (int)tasks.front().L.size());
    for (auto &t : tasks) {
      int p = t.dst & TASK_POS_MASK;
      int flags = t.dst & ~TASK_POS_MASK;
      if (inJ[p]) flags |= TASK_HAS_A;
      S.tA_dst.push_back(p | flags);
      S.pairL.insert(S.pairL.end(), t.L.begin(), t.L.end());
      S.pairU.insert(S.pairU.end(), t.U.begin(), t.U.end());
      S.tA_pp.push_back((int)S.pairL.size());
    }
    S.lvA.push_back((int)S.tA_dst.size());
    S.lvB.push_back((int)S.tB_dst.size());
  }
  S.npairs = (int64_t)S.pairL.size();
  // ---- dense tail ------------------------------------------------------------
  {
    std::vector<int> cnt(S.nlev + 1, 0);  // pivots at levels >= l
    for (int k = 0; k < nb; k++) cnt[lev[k]]++;
    for (int l = S.nlev - 1; l >= 0; l--) cnt[l] += cnt[l + 1];
    int Ls = S.nlev;
    while (Ls > 0 && cnt[Ls - 1] <= TAIL_MAX) Ls--;
    // worth it only when it removes a few levels
    if (S.nlev - Ls < 4) Ls = S.nlev;
    S.tail_level = Ls;
    std::vector<int> loc(nb, -1);
    for (int k = 0; k < nb; k++)
      if (lev[k] >= Ls) { loc[k] = (int)S.tail_piv.size(); S.tail_piv.push_back(k); }
    const int s = S.tail_n = (int)S.tail_piv.size();
    S.tl_pp.assign(1, 0);
    for (int k : S.tail_piv) {
      const auto &sk = st[k];
      // outputs of tail pivot k: (k,k), (k,j), (j,k) for j in st[k] (all in
      // the tail: ancestors of a tail pivot are tail pivots)
      std::vector<std::pair<int, int>> outs;  // (row, col)
      outs.push_back({k, k});
      for (int j : sk) { outs.push_back({k, j}); outs.push_back({j, k}); }
      for (auto &rc : outs) {
        int r = rc.first, c = rc.second;
        int p = pos(r, c);
        int flags = inJ[p] ? TASK_HAS_A : 0;
        S.tl_dense.push_back((loc[r] * s + loc[c]) | flags);
        S.tl_pos.push_back(p);
        // pairs over non-tail pivots m common to Lrow(r) and Lrow(c)
        int lo = std::min(r, c);
        for (int mm : lrowv[lo]) {
          if (loc[mm] >= 0) continue;
          int hi = std::max(r, c);
          // m in Lrow(hi) <=> hi in st[m]
          if (!std::binary_search(st[mm].begin(), st[mm].end(), hi)) continue;
          S.tl_pairL.push_back(pos(r, mm));
          S.tl_pairU.push_back(pos(mm, c));
        }
        S.tl_pp.push_back((int)S.tl_pairL.size());
      }
    }
  }
  for (int q = 0; q + 1 < (int)S.tl_pp.size(); q++)
    S.tl_max = std::max(S.tl_max, S.tl_pp[q + 1] - S.tl_pp[q]);
  // ---- tail rows in the sweeps ----------------------------------------------
  std::vector<char> intail(nb, 0);
  for (int k : S.tail_piv) intail[k] = 1;
  S.ts_ptr.assign(1, 0);
  for (int k : S.tail_piv) {
    for (int p = S.lrow[k]; p < S.ldiag[k]; p++) {
      int c = S.lcol[p];
      if (intail[c]) continue;
      S.ts_pos.push_back(p);
      S.ts_col.push_back(c);
      S.ts_mir.push_back(S.mirror[p]);
    }
    S.ts_ptr.push_back((int)S.ts_pos.size());
    S.ts_max = std::max(S.ts_max, S.ts_ptr.back() - S.ts_ptr[S.ts_ptr.size() - 2]);
  }
  // ---- triangular-solve levels (non-tail rows) -------------------------------
  // a tail row is never a dependency of a non-tail row in the forward sweep
  // (its descendants are non-tail) and is solved before the backward sweep
  std::vector<int> lf(nb, 0), lb(nb, 0);
  for (int i = 0; i < nb; i++)
    for (int k : lrowv[i]) lf[i] = std::max(lf[i], lf[k] + 1);
  for (int i = nb - 1; i >= 0; i--) {
    if (intail[i]) { lb[i] = -1; continue; }
    for (int j : st[i]) lb[i] = std::max(lb[i], lb[j] + 1);
  }
  S.nfw = S.nbw = 0;
  for (int i = 0; i < nb; i++)
    if (!intail[i]) { S.nfw = std::max(S.nfw, lf[i] + 1); S.nbw = std::max(S.nbw, lb[i] + 1); }
  auto group = [&](const std::vector<int> &lv, int nl, std::vector<int> &ptr,
                   std::vector<int> &rows, std::vector<int> &mx,
                   const std::vector<int> &len) {
    ptr.assign(nl + 1, 0);
    for (int i = 0; i < nb; i++) if (!intail[i]) ptr[lv[i] + 1]++;
    for (int l = 0; l < nl; l++) ptr[l + 1] += ptr[l];
    rows.resize(ptr[nl]);
    std::vector<int> w(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < nb; i++) if (!intail[i]) rows[w[lv[i]]++] = i;
    mx.assign(nl, 0);
    for (int l = 0; l < nl; l++) {
      // longest rows first: they lead the lane groups of the level
      std::stable_sort(rows.begin() + ptr[l], rows.begin() + ptr[l + 1],
                       [&](int a, int b) { return len[a] > len[b]; });
      if (ptr[l + 1] > ptr[l]) mx[l] = len[rows[ptr[l]]];
    }
  };
  std::vector<int> lenL(nb), lenU(nb);
  for (int i = 0; i < nb; i++) {
    lenL[i] = S.ldiag[i] - S.lrow[i];
    lenU[i] = S.lrow[i + 1] - S.ldiag[i] - 1;
  }
  group(lf, S.nfw, S.fw_ptr, S.fw_rows, S.fw_max, lenL);
  group(lb, S.nbw, S.bw_ptr, S.bw_rows, S.bw_max, lenU);
  return true;
}

}  // namespace bpf
