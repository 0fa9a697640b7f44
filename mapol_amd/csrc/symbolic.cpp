// Symbolic analysis: Ybus pattern, bus ordering, L+U block pattern and the
// level schedules of the batched factorization and triangular solves.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <functional>
#include <map>
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

// Nested dissection (George 1973): split the graph by a vertex separator
// taken from a breadth-first level structure rooted at a pseudo-peripheral
// vertex, order both sides recursively and the separator last; parts of at
// most `leaf` vertices are ordered by minimum degree.  The separator is the
// smallest level whose sides both hold at least a third of the vertices,
// thinned by moving every separator vertex without a neighbour on one side
// to the other side.
std::vector<int> nested_dissection_order(const std::vector<std::vector<int>> &adj0,
                                         int leaf) {
  const int n = (int)adj0.size();
  std::vector<std::vector<int>> adj(adj0);
  for (auto &a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  std::vector<int> order;
  order.reserve(n);
  std::vector<int> part(n, -1), dist(n, -1), loc(n, -1);
  int next_part = 0;
  // BFS over the vertices of part p from s; returns the visit order
  auto bfs = [&](int s, int p, std::vector<int> &seen) {
    seen.clear();
    seen.push_back(s);
    dist[s] = 0;
    for (size_t h = 0; h < seen.size(); h++) {
      int v = seen[h];
      for (int u : adj[v])
        if (part[u] == p && dist[u] < 0) { dist[u] = dist[v] + 1; seen.push_back(u); }
    }
  };
  auto clear_dist = [&](const std::vector<int> &vs) { for (int v : vs) dist[v] = -1; };
  std::vector<int> seen, seen2;
  // work stack of vertex sets; the separators are emitted in reverse
  std::function<void(std::vector<int>)> rec = [&](std::vector<int> vs) {
    const int p = next_part++;
    for (int v : vs) part[v] = p;
    if ((int)vs.size() <= leaf) {
      std::vector<std::vector<int>> sub(vs.size());
      for (size_t a = 0; a < vs.size(); a++) loc[vs[a]] = (int)a;
      for (size_t a = 0; a < vs.size(); a++)
        for (int u : adj[vs[a]])
          if (part[u] == p) sub[a].push_back(loc[u]);
      for (int k : min_degree_order(sub)) order.push_back(vs[k]);
      return;
    }
    // components
    bfs(vs[0], p, seen);
    if (seen.size() < vs.size()) {
      std::vector<int> first = seen;
      clear_dist(seen);
      std::vector<int> rest;
      for (int v : first) part[v] = -2;
      for (int v : vs) if (part[v] == p) rest.push_back(v);
      for (int v : first) part[v] = p;
      rec(first);
      rec(rest);
      return;
    }
    // pseudo-peripheral root: repeat BFS from the farthest vertex of
    // smallest degree while the eccentricity grows
    int root = vs[0], ecc = dist[seen.back()];
    for (int rounds = 0; rounds < 8; rounds++) {
      int far = seen.back(), fd = dist[far];
      for (int v : seen)
        if (dist[v] == fd && adj[v].size() < adj[far].size()) far = v;
      clear_dist(seen);
      bfs(far, p, seen2);
      int e2 = dist[seen2.back()];
      if (e2 <= ecc) { clear_dist(seen2); bfs(root, p, seen); break; }
      root = far;
      ecc = e2;
      std::swap(seen, seen2);
    }
    const int nl = ecc + 1;
    std::vector<int> cnt(nl, 0);
    for (int v : seen) cnt[dist[v]]++;
    const int nv = (int)vs.size();
    int best = -1;
    long below = 0;
    for (int l = 0; l < nl; l++) {
      const long above = nv - below - cnt[l];
      if (l > 0 && l < nl - 1 && 3 * below >= nv && 3 * above >= nv &&
          (best < 0 || cnt[l] < cnt[best]))
        best = l;
      below += cnt[l];
    }
    if (best < 0) {   // no balanced level: the middle one
      below = 0;
      for (int l = 0; l < nl; l++) {
        if (2 * (below + cnt[l]) >= nv && l > 0 && l < nl - 1) { best = l; break; }
        below += cnt[l];
      }
    }
    if (best < 0) {   // too small to split
      clear_dist(seen);
      std::vector<std::vector<int>> sub(vs.size());
      for (size_t a = 0; a < vs.size(); a++) loc[vs[a]] = (int)a;
      for (size_t a = 0; a < vs.size(); a++)
        for (int u : adj[vs[a]])
          if (part[u] == p) sub[a].push_back(loc[u]);
      for (int k : min_degree_order(sub)) order.push_back(vs[k]);
      return;
    }
    // side: 0 = A (levels < best), 1 = separator, 2 = B
    std::vector<int> A, Sep, B;
    for (int v : seen) (dist[v] < best ? A : dist[v] == best ? Sep : B).push_back(v);
    clear_dist(seen);
    std::map<int, int> sd;
    for (int v : A) sd[v] = 0;
    for (int v : Sep) sd[v] = 1;
    for (int v : B) sd[v] = 2;
    std::vector<int> keep;
    for (int v : Sep) {
      bool nA = false, nB = false;
      for (int u : adj[v]) {
        if (part[u] != p) continue;
        int su = sd[u];
        nA |= su == 0;
        nB |= su == 2;
      }
      if (!nB) { sd[v] = 0; A.push_back(v); }
      else if (!nA) { sd[v] = 2; B.push_back(v); }
      else keep.push_back(v);
    }
    rec(A);
    rec(B);
    // the separator last, in minimum-degree order of its own graph
    const int ps = next_part++;
    for (int v : keep) part[v] = ps;
    std::vector<std::vector<int>> sub(keep.size());
    for (size_t a = 0; a < keep.size(); a++) loc[keep[a]] = (int)a;
    for (size_t a = 0; a < keep.size(); a++)
      for (int u : adj[keep[a]])
        if (part[u] == ps) sub[a].push_back(loc[u]);
    for (int k : min_degree_order(sub)) order.push_back(keep[k]);
  };
  std::vector<int> all(n);
  for (int v = 0; v < n; v++) all[v] = v;
  if (n) rec(all);
  return order;
}

// Column structures of the filled graph of `adj` under `order`: st[k] =
// positions > k of the nonzeros of column k of L (symmetric pattern), built
// from the original edges and the children's structures (George & Liu).
static std::vector<std::vector<int>> filled_structure(
    const std::vector<std::vector<int>> &adj, const std::vector<int> &order) {
  const int nb = (int)adj.size();
  std::vector<int> newof(nb);
  for (int k = 0; k < nb; k++) newof[order[k]] = k;
  std::vector<std::vector<int>> st(nb), children(nb);
  for (int a = 0; a < nb; a++)
    for (int c : adj[a])
      if (newof[a] < newof[c]) st[newof[a]].push_back(newof[c]);
  std::vector<int> tmp;
  for (int k = 0; k < nb; k++) {
    auto &s = st[k];
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
    if (!s.empty()) children[s.front()].push_back(k);
  }
  return st;
}

// Minimum-height reordering of a perfect elimination order (Jess & Kees
// 1982; Liu & Mirzaian 1989): the filled graph of `order` is chordal, and
// eliminating, round by round, a maximal independent set of its simplicial
// vertices (lowest position first) gives an equivalent order -- the same
// filled graph, so the same fill -- whose elimination tree height is the
// number of rounds, the least over all equivalent orders.  Returns the new
// order in the vertex numbering of `adj`.
std::vector<int> min_height_reorder(const std::vector<std::vector<int>> &adj,
                                    const std::vector<int> &order) {
  const int nb = (int)adj.size();
  const std::vector<std::vector<int>> st = filled_structure(adj, order);
  std::vector<std::vector<int>> nbr(nb);   // filled-graph neighbours, sorted
  for (int k = 0; k < nb; k++)
    for (int j : st[k]) { nbr[k].push_back(j); nbr[j].push_back(k); }
  for (auto &a : nbr) std::sort(a.begin(), a.end());
  std::vector<char> alive(nb, 1), taken(nb, 0), blocked(nb, 0);
  std::vector<int> out, live, round;
  out.reserve(nb);
  int remaining = nb;
  while (remaining > 0) {
    round.clear();
    for (int v = 0; v < nb; v++) {
      if (!alive[v] || blocked[v]) continue;
      live.clear();
      for (int u : nbr[v]) if (alive[u]) live.push_back(u);
      bool simplicial = true;
      for (size_t a = 0; simplicial && a < live.size(); a++)
        for (size_t b = a + 1; b < live.size(); b++)
          if (!std::binary_search(nbr[live[a]].begin(), nbr[live[a]].end(), live[b])) {
            simplicial = false;
            break;
          }
      if (!simplicial) continue;
      round.push_back(v);
      for (int u : live) blocked[u] = 1;   // independent set
    }
    for (int v : round) { alive[v] = 0; out.push_back(order[v]); }
    remaining -= (int)round.size();
    std::fill(blocked.begin(), blocked.end(), 0);
  }
  return out;
}

// the PV and PQ buses (ascending), each bus's place among them (-1 for the
// others) and their graph: one edge per branch between two of them
static void unknown_graph(int n, int m, const int *from, const int *to,
                          const int *type, std::vector<int> &unk,
                          std::vector<int> &loc,
                          std::vector<std::vector<int>> &adj) {
  unk.clear();
  for (int i = 0; i < n; i++)
    if (type[i] == 1 || type[i] == 2) unk.push_back(i);
  loc.assign(n, -1);
  for (int k = 0; k < (int)unk.size(); k++) loc[unk[k]] = k;
  adj.assign(unk.size(), {});
  for (int b = 0; b < m; b++) {
    int a = loc[from[b]], c = loc[to[b]];
    if (a >= 0 && c >= 0 && a != c) { adj[a].push_back(c); adj[c].push_back(a); }
  }
}

// minimum degree (0), then optionally its minimum-height equivalent (1, same
// fill), or nested dissection (2)
static std::vector<int> order_graph(const std::vector<std::vector<int>> &adj,
                                    int ordering) {
  if (ordering == 2) return nested_dissection_order(adj, 64);
  if (ordering == 1) return min_height_reorder(adj, min_degree_order(adj));
  return min_degree_order(adj);
}

std::vector<int> elimination_order(int n, int m, const int *from, const int *to,
                                   const int *type, int ordering) {
  std::vector<int> unk, loc;
  std::vector<std::vector<int>> adj;
  unknown_graph(n, m, from, to, type, unk, loc, adj);
  std::vector<int> ord = order_graph(adj, ordering);
  for (int &k : ord) k = unk[k];
  return ord;
}

std::vector<int> tail_jobs(const std::vector<int> &pp, int per_lane) {
  std::vector<int> jobs;
  const int ne = (int)pp.size() - 1;
  for (int k = 0; k < ne;) {
    const int len = pp[k + 1] - pp[k];
    int lg = 0;
    while (lg < 6 && (len + (1 << lg) - 1) >> lg > per_lane) lg++;
    const int e1 = std::min(ne, k + (64 >> lg));
    jobs.push_back(k);
    jobs.push_back(e1 << 3 | lg);
    k = e1;
  }
  return jobs;
}

std::vector<int> rcm_numbering(int n, int m, const int *from, const int *to) {
  std::vector<std::vector<int>> adj(n);
  for (int b = 0; b < m; b++)
    if (from[b] != to[b]) {
      adj[from[b]].push_back(to[b]);
      adj[to[b]].push_back(from[b]);
    }
  for (auto &a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  auto deg = [&](int i) { return (int)adj[i].size(); };
  std::vector<int> mark(n, -1), order, lev;
  order.reserve(n);
  int stamp = 0;
  // breadth-first levels from `r` (visits marked with a new stamp): the
  // vertices in visit order, the start of the last level, the level count
  auto levels = [&](int r, int &last0, int &depth) {
    const int s = ++stamp;
    lev.assign(1, r);
    mark[r] = s;
    last0 = 0;
    depth = 0;
    for (size_t h = 0; h < lev.size();) {
      const size_t end = lev.size();
      last0 = (int)h;
      depth++;
      for (; h < end; h++)
        for (int j : adj[lev[h]])
          if (mark[j] != s) { mark[j] = s; lev.push_back(j); }
    }
  };
  std::vector<char> done(n, 0);
  for (int r0 = 0; r0 < n; r0++) {
    if (done[r0]) continue;
    // pseudo-peripheral start (George and Liu): from a vertex of least
    // degree, move to the least-degree vertex of the last level while the
    // level count grows
    int last0, depth;
    levels(r0, last0, depth);
    int start = r0;
    for (int v : lev)
      if (deg(v) < deg(start) || (deg(v) == deg(start) && v < start)) start = v;
    levels(start, last0, depth);
    for (;;) {
      int cand = lev[last0];
      for (size_t h = last0; h < lev.size(); h++)
        if (deg(lev[h]) < deg(cand)) cand = lev[h];
      int l2, d2;
      const int keep = depth;
      levels(cand, l2, d2);
      if (d2 <= keep) break;
      start = cand;
      last0 = l2;
      depth = d2;
    }
    // Cuthill-McKee: breadth first, neighbours by ascending degree
    const size_t c0 = order.size();
    order.push_back(start);
    done[start] = 1;
    std::vector<int> nb;
    for (size_t h = c0; h < order.size(); h++) {
      nb.clear();
      for (int j : adj[order[h]])
        if (!done[j]) { done[j] = 1; nb.push_back(j); }
      std::stable_sort(nb.begin(), nb.end(),
                       [&](int a, int b) { return deg(a) < deg(b); });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
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
             Symbolic &S, std::string *err, const AnalysisParams &prm) {
  S = Symbolic();
  if (prm.tail_limit < 0 || prm.tail_limit > TAIL_MAX || prm.chunk_pairs < 0 ||
      prm.chunk_pairs > CHUNK_PAIRS || prm.slot_chunk < 0 ||
      prm.slot_chunk > SLOT_CHUNK ||
      prm.slot_chunk % 64 || prm.ordering < 0 ||
      prm.ordering > 2) {
    if (err) *err = "tile sizes out of range (tail <= " + std::to_string(TAIL_MAX) +
                    " pivots, chunk <= " + std::to_string(CHUNK_PAIRS) +
                    " products, slots a multiple of 64 up to " + std::to_string(SLOT_CHUNK) + ")";
    return false;
  }
  S.prm.tail_limit = prm.tail_limit ? prm.tail_limit : TAIL_MAX;
  S.prm.chunk_pairs = prm.chunk_pairs ? prm.chunk_pairs : CHUNK_PAIRS;
  S.prm.slot_chunk = prm.slot_chunk ? prm.slot_chunk : SLOT_CHUNK;
  S.prm.ordering = prm.ordering;
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
  std::vector<int> unk, loc;
  std::vector<std::vector<int>> adj;
  unknown_graph(n, m, from, to, type, unk, loc, adj);
  const int nb = (int)unk.size();
  S.nb = nb;
  if (nb == 0) {
    if (err) *err = "no PV or PQ bus";
    return false;
  }
  std::vector<int> order;
  if (!prm.order_bus.empty()) {
    std::vector<char> seen(nb, 0);
    bool ok = (int)prm.order_bus.size() == nb;
    for (size_t k = 0; ok && k < prm.order_bus.size(); k++) {
      const int b = prm.order_bus[k], a = (b >= 0 && b < n) ? loc[b] : -1;
      ok = a >= 0 && !seen[a];
      if (ok) { seen[a] = 1; order.push_back(a); }
    }
    if (!ok) {
      if (err) *err = "elimination order is not a permutation of the PV and PQ buses";
      return false;
    }
  } else {
    order = order_graph(adj, S.prm.ordering);
  }
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
  {
    std::vector<int> yof(nnzLU, -1);
    for (int p = 0; p < nnzY; p++)
      if (S.y2lu[p] >= 0) yof[S.y2lu[p]] = p;
    for (int q = 0; q < nnzLU; q++)
      if (yof[q] >= 0) { S.jac_lu.push_back(q); S.jac_y.push_back(yof[q]); }
  }
  // ---- Crout schedule ------------------------------------------------------
  std::vector<int> lev(nb, 0);
  for (int k = 0; k < nb; k++)
    if (parent[k] >= 0) lev[parent[k]] = std::max(lev[parent[k]], lev[k] + 1);
  S.nlev = nb ? 1 + *std::max_element(lev.begin(), lev.end()) : 0;
  std::vector<std::vector<int>> piv(S.nlev);
  for (int k = 0; k < nb; k++) piv[lev[k]].push_back(k);
  // dense tail: the top levels holding at most TAIL_MAX pivots together
  // (their pivots are ancestors of everything else, so they can go last)
  {
    std::vector<int> cnt(S.nlev + 1, 0);  // pivots at levels >= l
    for (int k = 0; k < nb; k++) cnt[lev[k]]++;
    for (int l = S.nlev - 1; l >= 0; l--) cnt[l] += cnt[l + 1];
    // the tail limit (at most TAIL_MAX, the LDS capacity) is a tuning knob
    const int limit = S.prm.tail_limit;
    int Ls = S.nlev;
    while (Ls > 0 && cnt[Ls - 1] <= limit) Ls--;
    // worth it only when it removes a few levels
    if (S.nlev - Ls < 4) Ls = S.nlev;
    S.tail_level = Ls;
  }
  // Each output of a pivot (its block, U row, L column) is A(p) minus a dot
  // product over the pivot's descendants m: sum L(r,m) U(m,c).  The outputs
  // of the levels below the tail are packed, whole pivot groups at a time,
  // into chunks of at most CHUNK_PAIRS products: the kernel forms a chunk's
  // products in parallel into LDS, then reduces them per output.  An L
  // output also reduces its pivot's products (same chunk) and scales by the
  // pivot inverse, so no separate scaling pass is needed.
  struct Task { int dst; std::vector<int> L, U; };
  std::vector<std::vector<Task>> pgroup(nb);   // each pivot's outputs
  auto gpairs = [](const std::vector<Task> &g) {
    size_t n = 0;
    for (auto &t : g) n += t.L.size();
    return n;
  };
  // one chunk from the groups of the pivots pks: the tasks heaviest first
  // (an L output also reduces its pivot's products; ties keep the order of
  // pks), then their products in task order.  Appends to the three lists;
  // returns {tasks, longest reduction}.
  int rhs_out[2] = {0, 0};   // the last chunk's forward-substitution outputs, products
  auto emit_chunk = [&](const std::vector<int> &pks, std::vector<int> &tk_a,
                        std::vector<int> &tk_b, std::vector<int> &prs) {
    rhs_out[0] = rhs_out[1] = 0;
    std::vector<std::pair<int, int>> ord;  // (pivot, index in its group)
    for (int k : pks)
      for (int q = 0; q < (int)pgroup[k].size(); q++) {
        const Task &t = pgroup[k][q];
        // a U block without products keeps its Jacobian value: nothing to do
        if (t.L.empty() && (t.dst & TASK_HAS_A) && !(t.dst & (TASK_DIAG | TASK_LCOL)))
          continue;
        ord.push_back({k, q});
      }
    auto work = [&](const std::pair<int, int> &x) {
      const Task &t = pgroup[x.first][x.second];
      size_t w = t.L.size();
      if (t.dst & TASK_LCOL) w = std::max(w, pgroup[x.first][0].L.size());
      return w;
    };
    // the forward substitution's outputs (positions from nnzLU) last
    auto rhs_task = [&](const std::pair<int, int> &x) {
      return (pgroup[x.first][x.second].dst & TASK_POS_MASK) >= nnzLU;
    };
    std::stable_sort(ord.begin(), ord.end(),
                     [&](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                       if (rhs_task(x) != rhs_task(y)) return rhs_task(y);
                       return work(x) > work(y);
                     });
    // pair offsets of every task (in chunk order)
    std::map<std::pair<int, int>, int> start;
    int off = (int)prs.size() / 2;
    for (auto &x : ord) { start[x] = off; off += (int)pgroup[x.first][x.second].L.size(); }
    int wmax = 0;
    for (auto &x : ord) {
      const Task &t = pgroup[x.first][x.second];
      const int p0 = start[x], p1 = p0 + (int)t.L.size();
      tk_a.insert(tk_a.end(), {t.dst, p0, p1, 0});
      if (t.dst & TASK_LCOL) {
        const int v0 = start[{x.first, 0}], v1 = v0 + (int)pgroup[x.first][0].L.size();
        tk_b.insert(tk_b.end(), {0, v0, v1, 0});
      } else {
        tk_b.insert(tk_b.end(), {0, 0, 0, 0});
      }
      for (size_t q = 0; q < t.L.size(); q++) {
        prs.push_back(t.L[q]);
        prs.push_back(t.U[q]);
      }
      wmax = std::max(wmax, (int)work(x));
      if (rhs_task(x)) { rhs_out[0]++; rhs_out[1] += (int)t.L.size(); }
    }
    return std::pair<int, int>((int)ord.size(), wmax);
  };
  std::vector<std::vector<int>> ch_piv;   // pivots of each main chunk, in order
  S.ch_task.assign(1, 0);
  S.ch_pair.assign(1, 0);
  S.lv_chunk.assign(1, 0);
  for (int l = 0; l < S.nlev; l++) {
    if (l >= S.tail_level) { S.lv_chunk.push_back((int)S.ch_max.size()); continue; }
    // pivot groups of the level
    std::vector<std::pair<int, std::vector<Task>>> groups;
    for (int k : piv[l]) {
      const auto &sk = st[k];
      std::vector<Task> out(1 + 2 * sk.size());
      // the pivot's own value enters its product range as an init entry
      // (U position -1): the L outputs of the group read the pivot's sum
      // from LDS while the pivot output overwrites its position
      const int pk = S.ldiag[k];
      out[0].dst = pk | TASK_DIAG;
      if (inJ[pk]) { out[0].L.push_back(pk); out[0].U.push_back(-1); }
      for (size_t a = 0; a < sk.size(); a++) {
        int pu = pos(k, sk[a]), pl = pos(sk[a], k);
        out[1 + a].dst = pu | (inJ[pu] ? TASK_HAS_A : 0);
        out[1 + sk.size() + a].dst = pl | (inJ[pl] ? TASK_HAS_A : 0) | TASK_LCOL;
      }
      for (int mm : lrowv[k]) {          // descendants m with k in st[m]
        int pkm = pos(k, mm), pmk = pos(mm, k);
        // U_kk -= L_km U_mk
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
      // every other output with a Jacobian value also starts from an init
      // entry, so the value arrives with the products' operands (a U block
      // without products keeps its value and is left out below)
      for (size_t q = 1; q < out.size(); q++) {
        Task &t = out[q];
        if (!(t.dst & TASK_HAS_A)) continue;
        if (t.L.empty() && !(t.dst & TASK_LCOL)) continue;
        t.L.insert(t.L.begin(), t.dst & TASK_POS_MASK);
        t.U.insert(t.U.begin(), -1);
        t.dst &= ~TASK_HAS_A;
      }
      // the forward substitution's entry of the pivot, y_k = b_k - sum over
      // its L row of L(k,m) y_m, as one more output of the group.  The
      // right-hand-side blocks follow the factor's (position nnzLU + k,
      // y in the first column, zero in the second) and start from b_k,
      // which the Jacobian fill writes; a pivot with an empty L row keeps b_k.
      if (!lrowv[k].empty()) {
        Task t;
        t.dst = nnzLU + k;
        t.L.push_back(nnzLU + k);
        t.U.push_back(-1);
        for (int mm : lrowv[k]) {
          t.L.push_back(pos(k, mm));
          t.U.push_back(nnzLU + mm);
        }
        out.push_back(std::move(t));
      }
      groups.push_back({k, std::move(out)});
    }
    std::stable_sort(groups.begin(), groups.end(),
                     [&](const std::pair<int, std::vector<Task>> &x,
                         const std::pair<int, std::vector<Task>> &y) {
                       return gpairs(x.second) > gpairs(y.second);
                     });
    // pack groups into chunks (first fit in order of decreasing size) of at
    // most chunk_pairs products; a group above that (but within the LDS
    // buffer of CHUNK_PAIRS) gets a chunk of its own
    const size_t cap = (size_t)S.prm.chunk_pairs;
    std::vector<std::vector<int>> chunks;   // pivots
    std::vector<size_t> fill;
    for (auto &gr : groups) {
      size_t np = gpairs(gr.second);
      if (np > (size_t)CHUNK_PAIRS) {
        if (err) *err = "a pivot's factorization products exceed one LDS chunk (" +
                        std::to_string(np) + " > " + std::to_string(CHUNK_PAIRS) + ")";
        return false;
      }
      int c = -1;
      for (int q = 0; q < (int)chunks.size(); q++)
        if (fill[q] + np <= cap) { c = q; break; }
      if (c < 0) { chunks.push_back({}); fill.push_back(0); c = (int)chunks.size() - 1; }
      chunks[c].push_back(gr.first);
      fill[c] += np;
      pgroup[gr.first] = std::move(gr.second);
    }
    for (auto &cg : chunks) {
      const auto r = emit_chunk(cg, S.tk_a, S.tk_b, S.pairs);
      S.ch_rhs.insert(S.ch_rhs.end(), {rhs_out[0], rhs_out[1]});
      S.ch_task.push_back(S.ch_task.back() + r.first);
      S.ch_pair.push_back((int)S.pairs.size() / 2);
      S.ch_max.push_back(r.second);
      ch_piv.push_back(cg);
    }
    S.lv_chunk.push_back((int)S.ch_max.size());
  }
  S.nchunk = (int)S.ch_max.size();
  S.npairs = (int64_t)S.pairs.size() / 2;
  // ---- level 0 in the Jacobian fill ------------------------------------------
  // A level-0 pivot has no descendants: its block becomes its inverse, its L
  // blocks J(r,k) J(k,k)^-1, its U blocks keep their Jacobian values.  Done
  // by the fill when level 0 lies below the tail and every such block is a
  // Jacobian block (no fill-in reaches level 0).
  {
    std::vector<int> jidx(nnzLU, -1);
    for (int r = 0; r < (int)S.jac_lu.size(); r++) jidx[S.jac_lu[r]] = r;
    S.jac_l0.assign(S.jac_lu.size(), -1);
    bool ok = S.tail_level > 0 && S.lv_chunk.size() > 1 && S.lv_chunk[1] > 0;
    for (size_t a = 0; ok && a < piv[0].size(); a++) {
      const int k = piv[0][a];
      const int rp = jidx[S.ldiag[k]];
      if (rp < 0) { ok = false; break; }
      S.jac_l0[rp] = -2;
      for (int j : st[k]) {
        const int rl = jidx[pos(j, k)];
        if (rl < 0) { ok = false; break; }
        S.jac_l0[rl] = rp;
      }
    }
    if (ok) {
      S.l0_chunks = S.lv_chunk[1];
    } else {
      S.jac_l0.assign(S.jac_lu.size(), -1);
      S.l0_chunks = 0;
    }
  }
  if (S.nchunk > MAX_CHUNKS) {
    if (err) *err = "factorization schedule has " + std::to_string(S.nchunk) +
                    " chunks, more than the " + std::to_string(MAX_CHUNKS) +
                    " the kernel keeps in LDS";
    return false;
  }
  // ---- split schedules -------------------------------------------------------
  // A pivot's outputs read only outputs of its descendants, so the subtrees
  // below a cut level share no operands and can be factored by workgroups
  // of their own.  The cut and the grouping come from a model of one
  // instance's chunk time measured on MI355X with cycle stamps (about
  // 2.6 us per chunk plus 1 ns per product) and about 4 us for the launch of
  // the top workgroup: per variant the cut with the least modelled time,
  // kept when it saves at least 15 % of the main schedule's (below that,
  // as at 1354 buses, the second launch costs more than the model says).
  {
    constexpr double CHUNK_US = 2.6, PRODUCT_US = 1.0e-3, LAUNCH_US = 4.0;
    auto chunk_us = [&](double prods) { return CHUNK_US + PRODUCT_US * prods; };
    const int c_first = S.l0_chunks;   // chunks before it are the fill's
    const int l_first = c_first > 0 ? 1 : 0;
    std::vector<double> pprod(nb, 0.0), cprod(S.nchunk);
    for (int k = 0; k < nb; k++) pprod[k] = (double)gpairs(pgroup[k]);
    for (int c = 0; c < S.nchunk; c++) cprod[c] = S.ch_pair[c + 1] - S.ch_pair[c];
    double main_us = 0;
    for (int c = c_first; c < S.nchunk; c++) main_us += chunk_us(cprod[c]);
    S.sp_model_us[0] = main_us;
    S.sp_level.assign(SPLIT_VARIANTS, -1);
    S.sp_rest.assign(SPLIT_VARIANTS, S.nchunk);
    S.sp_grp.assign(SPLIT_VARIANTS * (SPLIT_MAX + 1), 0);
    std::vector<int> rt(nb), grp(nb);
    // the lane-group size a main chunk's reduction uses for each workgroup
    // size the factorization may be launched with (the kernel's rule):
    // pieces of main chunks that agree on all three can share a chunk
    auto lane_groups = [&](int c) {
      std::array<int, 3> r;
      const int items = S.ch_task[c + 1] - S.ch_task[c], maxlen = S.ch_max[c];
      for (int b = 0; b < 3; b++) {
        int g = 1;
        while (g < 64 && g < maxlen && items * 2 * g <= (256 << b)) g <<= 1;
        r[b] = g;
      }
      return r;
    };
    std::vector<int> ch_level(S.nchunk, 0);
    for (int l = 0; l + 1 < (int)S.lv_chunk.size(); l++)
      for (int c = S.lv_chunk[l]; c < S.lv_chunk[l + 1]; c++) ch_level[c] = l;
    // group j's chunks below the cut: its share of each main chunk, shares
    // of one level with the same lane groups merged up to chunk_pairs
    // products; each chunk as {pivots, representative main chunk, products}
    struct GChunk { std::vector<int> pks; int c; double prods; };
    auto group_chunks = [&](int Lc, int j) {
      std::vector<GChunk> out;
      for (int c = c_first; c < S.lv_chunk[Lc]; c++) {
        std::vector<int> pks;
        double pr = 0;
        for (int k : ch_piv[c])
          if (grp[k] == j) { pks.push_back(k); pr += pprod[k]; }
        if (pks.empty()) continue;
        if (!out.empty()) {
          GChunk &o = out.back();
          if (ch_level[o.c] == ch_level[c] && lane_groups(o.c) == lane_groups(c) &&
              o.prods + pr <= (double)S.prm.chunk_pairs) {
            o.pks.insert(o.pks.end(), pks.begin(), pks.end());
            o.prods += pr;
            continue;
          }
        }
        out.push_back({pks, c, pr});
      }
      return out;
    };
    // the pivots below level Lc grouped by subtree into K groups (largest
    // subtrees first, each to the group with the fewest products); returns
    // the modelled time of the split factorization
    auto partition = [&](int Lc, int K) {
      for (int k = nb - 1; k >= 0; k--)   // parents come after their children
        rt[k] = (lev[k] < Lc && parent[k] >= 0 && lev[parent[k]] < Lc) ? rt[parent[k]] : k;
      std::vector<double> tp(nb, 0.0);
      std::vector<int> roots;
      for (int k = 0; k < nb; k++)
        if (lev[k] < Lc) {
          tp[rt[k]] += pprod[k];
          if (rt[k] == k) roots.push_back(k);
        }
      std::stable_sort(roots.begin(), roots.end(),
                       [&](int x, int y) { return tp[x] > tp[y]; });
      std::vector<double> load(K, 0.0);
      std::vector<int> gof(nb, -1);
      for (int r : roots) {
        const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[j] += tp[r];
        gof[r] = j;
      }
      for (int k = 0; k < nb; k++) grp[k] = lev[k] < Lc ? gof[rt[k]] : -1;
      double t1 = 0;
      for (int j = 0; j < K; j++) {
        double tj = 0;
        for (const GChunk &g : group_chunks(Lc, j)) tj += chunk_us(g.prods);
        t1 = std::max(t1, tj);
      }
      double t = t1 + LAUNCH_US;
      for (int c = S.lv_chunk[Lc]; c < S.nchunk; c++) t += chunk_us(cprod[c]);
      return t;
    };
    for (int v = 0; v < SPLIT_VARIANTS; v++) {
      const int K = 2 << v;
      int *gp = &S.sp_grp[v * (SPLIT_MAX + 1)];
      double best = 0;
      int bl = -1;
      for (int Lc = l_first + 1; Lc <= S.tail_level; Lc++) {
        const double t = partition(Lc, K);
        if (bl < 0 || t < best) { best = t; bl = Lc; }
      }
      S.sp_model_us[v + 1] = best;
      if (bl >= 0 && best < 0.85 * main_us) {
        partition(bl, K);
        S.sp_level[v] = bl;
        S.sp_rest[v] = S.lv_chunk[bl];
        for (int j = 0; j < K; j++) {
          gp[j] = (int)S.sp_lg.size() / 2;
          for (const GChunk &g : group_chunks(bl, j)) {
            const int k0 = (int)S.sp_tk_a.size() / 4, q0 = (int)S.sp_pairs.size() / 2;
            emit_chunk(g.pks, S.sp_tk_a, S.sp_tk_b, S.sp_pairs);
            S.sp_rhs.insert(S.sp_rhs.end(), {rhs_out[0], rhs_out[1]});
            S.sp_ch.insert(S.sp_ch.end(), {k0, (int)S.sp_tk_a.size() / 4, q0,
                                           (int)S.sp_pairs.size() / 2});
            S.sp_lg.insert(S.sp_lg.end(),
                           {S.ch_task[g.c + 1] - S.ch_task[g.c], S.ch_max[g.c]});
          }
        }
      }
      for (int j = S.sp_level[v] >= 0 ? K : 0; j <= SPLIT_MAX; j++)
        gp[j] = (int)S.sp_lg.size() / 2;
    }
  }
  // ---- dense tail ------------------------------------------------------------
  {
    const int Ls = S.tail_level;
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
    // longest product lists first (stable): the factorization takes the
    // entries in jobs of one wave (tail_jobs), each job's lane groups sized
    // by its first list, so the lists of a job are of about one length;
    // every entry keeps its own list and order
    const int ne = (int)S.tl_pos.size();
    std::vector<int> ord(ne);
    for (int q = 0; q < ne; q++) ord[q] = q;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
      return S.tl_pp[x + 1] - S.tl_pp[x] > S.tl_pp[y + 1] - S.tl_pp[y];
    });
    std::vector<int> dense(ne), tpos(ne), pp(1, 0), pl, pu;
    pl.reserve(S.tl_pairL.size());
    pu.reserve(S.tl_pairU.size());
    for (int q = 0; q < ne; q++) {
      const int e = ord[q];
      dense[q] = S.tl_dense[e];
      tpos[q] = S.tl_pos[e];
      for (int r = S.tl_pp[e]; r < S.tl_pp[e + 1]; r++) {
        pl.push_back(S.tl_pairL[r]);
        pu.push_back(S.tl_pairU[r]);
      }
      pp.push_back((int)pl.size());
    }
    S.tl_dense.swap(dense);
    S.tl_pos.swap(tpos);
    S.tl_pp.swap(pp);
    S.tl_pairL.swap(pl);
    S.tl_pairU.swap(pu);
  }
  for (int q = 0; q + 1 < (int)S.tl_pp.size(); q++)
    S.tl_max = std::max(S.tl_max, S.tl_pp[q + 1] - S.tl_pp[q]);
  {
    std::vector<int> blocks(S.tl_pairL);
    blocks.insert(blocks.end(), S.tl_pairU.begin(), S.tl_pairU.end());
    std::sort(blocks.begin(), blocks.end());
    blocks.erase(std::unique(blocks.begin(), blocks.end()), blocks.end());
    if (!blocks.empty() && (int)blocks.size() <= TAIL_STAGE_MAX) {
      S.tl_stage = blocks;
      auto idx = [&](int p) {
        return (int)(std::lower_bound(blocks.begin(), blocks.end(), p) - blocks.begin());
      };
      S.tl_pairS.resize(S.tl_pairL.size());
      for (size_t q = 0; q < S.tl_pairL.size(); q++)
        S.tl_pairS[q] = idx(S.tl_pairL[q]) | (idx(S.tl_pairU[q]) << 16);
    }
  }
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
  // entry-parallel chunks (rows already longest first within a level)
  auto chunks = [&](const std::vector<int> &ptr, const std::vector<int> &rows,
                    int nl, bool fw) {
    int nch = 0;
    for (int l = 0; l < nl; l++) {
      int r = ptr[l];
      while (r < ptr[l + 1]) {
        const int r0 = (int)S.sw_row.size() / 4, e0 = (int)S.sw_ent.size() / 2;
        int cmax = 0;
        while (r < ptr[l + 1]) {
          const int i = rows[r];
          const int p0 = fw ? S.lrow[i] : S.ldiag[i] + 1;
          const int p1 = fw ? S.ldiag[i] : S.lrow[i + 1];
          // a forward row without L entries keeps x = b: nothing to do
          if (fw && p1 == p0) { r++; continue; }
          const int ne = (int)S.sw_ent.size() / 2;
          if (ne > e0 && ne - e0 + (p1 - p0) > SWEEP_CHUNK) break;
          S.sw_row.insert(S.sw_row.end(), {i, ne, ne + (p1 - p0), fw ? -1 : S.ldiag[i]});
          for (int p = p0; p < p1; p++) {
            S.sw_ent.push_back(p);
            S.sw_ent.push_back(S.lcol[p]);
          }
          cmax = std::max(cmax, p1 - p0);
          r++;
        }
        if ((int)S.sw_row.size() / 4 == r0) continue;  // level left empty
        S.sw_chunk.insert(S.sw_chunk.end(), {r0, (int)S.sw_row.size() / 4, e0,
                                             (int)S.sw_ent.size() / 2});
        S.sw_cmax.push_back(cmax);
        nch++;
      }
    }
    return nch;
  };
  S.sw_nfw = chunks(S.fw_ptr, S.fw_rows, S.nfw, true);
  S.sw_nbw = chunks(S.bw_ptr, S.bw_rows, S.nbw, false);
  // a row longer than a chunk cannot take the entry-parallel path
  // (and the chunk table must fit the kernel's LDS copy)
  S.sw_ok = (int)S.sw_cmax.size() <= MAX_CHUNKS;
  for (size_t c = 0; c < S.sw_cmax.size(); c++)
    if (S.sw_chunk[4 * c + 3] - S.sw_chunk[4 * c + 2] > SWEEP_CHUNK) S.sw_ok = 0;
  // slot form: rows of a level (longest first, so lane groups shrink and
  // stay aligned) packed into chunks of at most SLOT_CHUNK slots.  A row of
  // more than 64 entries is cut into segments of 64: segment j of every
  // such row goes into the chunks of pass j of the level (a new chunk starts
  // with each pass, so no chunk holds two segments of one row); each
  // segment's first lane subtracts its partial sum from x_i, and only the
  // last one applies the pivot inverse (backward).
  struct Piece { int row, p0, p1, last; };
  // tr: the transposed sweeps of the adjoint (J^T = U^T L^T): U^T forward
  // over the rows of L (every row: each applies its pivot inverse at its
  // last segment), L^T backward over the rows of U (rows without entries
  // have nothing to do); every position is the mirror of the entry's, the
  // block (k, i) for (i, k)
  auto slots = [&](const std::vector<int> &ptr, const std::vector<int> &rows,
                   int nl, bool fw, bool tr, std::vector<int> &chunk,
                   std::vector<int> &slot) {
    int nch = 0;
    const bool pivot = tr ? fw : !fw;
    for (int l = 0; l < nl; l++) {
      std::vector<std::vector<Piece>> pass;
      for (int r = ptr[l]; r < ptr[l + 1]; r++) {
        const int i = rows[r];
        const int p0 = fw ? S.lrow[i] : S.ldiag[i] + 1;
        const int p1 = fw ? S.ldiag[i] : S.lrow[i + 1];
        if (!pivot && p1 == p0) continue;   // x = b: nothing to do
        const int nseg = std::max(1, (p1 - p0 + 63) / 64);
        if ((int)pass.size() < nseg) pass.resize(nseg);
        for (int j = 0; j < nseg; j++)
          pass[j].push_back({i, p0 + 64 * j, std::min(p0 + 64 * (j + 1), p1),
                             j == nseg - 1 ? 1 : 0});
      }
      for (auto &pieces : pass) {
        std::stable_sort(pieces.begin(), pieces.end(), [](const Piece &a, const Piece &b) {
          return a.p1 - a.p0 > b.p1 - b.p0;
        });
        size_t r = 0;
        while (r < pieces.size()) {
          const int s0 = (int)slot.size() / 4;
          int glg = -1;
          while (r < pieces.size()) {
            const Piece &pc = pieces[r];
            const int i = pc.row, len = pc.p1 - pc.p0;
            int lg = 0;
            while ((1 << lg) < len) lg++;
            const int ns = (int)slot.size() / 4;
            if (ns > s0 && ns - s0 + (1 << lg) > S.prm.slot_chunk) break;
            const int first = len > 0 ? pc.p0 : std::max(S.ldiag[i], 0);
            const int lead = i | (pc.last ? 0 : SLOT_NOT_LAST);
            for (int k = 0; k < (1 << lg); k++) {
              const int p = pc.p0 + k;
              const bool ent = p < pc.p1;
              const int pos = tr ? S.mirror[ent ? p : first] : (ent ? p : first);
              slot.insert(slot.end(),
                          {pos, (ent ? S.lcol[p] + 1 : 0) | (lg << SLOT_LG_SHIFT),
                           k == 0 ? lead : -1,
                           (k == 0 && pivot && pc.last) ? S.ldiag[i] : pos});
            }
            glg = std::max(glg, lg);
            r++;
          }
          int s1 = (int)slot.size() / 4;
          if (s1 == s0) continue;
          // the chunk ends on a multiple of 64 slots, so that a wave's 64
          // slots lie wholly inside a chunk or wholly outside it: padding
          // slots carry no entry (column 0) and no row (-1); their block
          // position repeats the last real slot's
          const int padpos = slot[slot.size() - 4];
          for (; (s1 - s0) % 64; s1++) slot.insert(slot.end(), {padpos, 0, -1, padpos});
          chunk.insert(chunk.end(), {s0, s1, glg, 0});
          nch++;
        }
      }
    }
    return nch;
  };
  S.sl_nfw = slots(S.fw_ptr, S.fw_rows, S.nfw, true, false, S.sl_chunk, S.sl_slot);
  S.sl_nbw = slots(S.bw_ptr, S.bw_rows, S.nbw, false, false, S.sl_chunk, S.sl_slot);
  S.sl_ok = S.sl_nfw + S.sl_nbw <= MAX_CHUNKS;
  S.sl_nfw_t = slots(S.fw_ptr, S.fw_rows, S.nfw, true, true, S.sl_chunk_t, S.sl_slot_t);
  S.sl_nbw_t = slots(S.bw_ptr, S.bw_rows, S.nbw, false, true, S.sl_chunk_t, S.sl_slot_t);
  S.sl_ok_t = S.sl_nfw_t + S.sl_nbw_t <= MAX_CHUNKS;
  return true;
}

}  // namespace bpf
