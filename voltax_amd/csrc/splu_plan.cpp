#include "splu_plan.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "status.hpp"

namespace vxm {
namespace {

struct DSU {
  std::vector<int> p;
  explicit DSU(int n) : p(n) { std::iota(p.begin(), p.end(), 0); }
  int find(int a) {
    while (p[a] != a) a = p[a] = p[p[a]];
    return a;
  }
  void unite(int a, int b) {
    a = find(a);
    b = find(b);
    if (a != b) p[b] = a;
  }
};

// Maximum bipartite matching of the internal unknowns of two-terminal,
// one-internal devices (rows) to distinct terminal nodes (columns), as
// voltax/sparse.py:283-297 computes it with scipy's
// maximum_bipartite_matching (scipy 1.15.3, Hopcroft-Karp).  Phases: a BFS
// layering from every unmatched row; then, for each unmatched row in order, a
// depth-first search that takes the first free column at the layer bound
// from the row on top of its stack, else pushes every matched row one layer
// further (in column order; the last pushed is explored first), and augments
// along the recorded path.  Which maximum matching comes out decides the
// pivot groups, so the search order is part of the result; the CPU tests
// compare it with scipy's on random graphs.  adj: sorted, unique columns of
// each row.  Returns the column of each row, -1 if unmatched.
std::vector<int> max_bipartite_matching(const std::vector<std::vector<int>>& adj, int ncols) {
  const int nr = (int)adj.size();
  constexpr int INF = 1 << 30;
  std::vector<int> pr(nr, -1), pc(ncols, -1), dist(nr), queue;
  struct Entry { int row, parent, via; };
  std::vector<Entry> stack, path;
  for (;;) {
    queue.clear();
    for (int u = 0; u < nr; ++u) {
      dist[u] = pr[u] < 0 ? 0 : INF;
      if (pr[u] < 0) queue.push_back(u);
    }
    int bound = INF;
    for (size_t h = 0; h < queue.size(); ++h) {
      const int u = queue[h];
      if (dist[u] >= bound) continue;
      for (int c : adj[u]) {
        const int w = pc[c];
        if (w < 0) {
          if (bound == INF) bound = dist[u] + 1;
        } else if (dist[w] == INF) {
          dist[w] = dist[u] + 1;
          queue.push_back(w);
        }
      }
    }
    if (bound == INF) break;
    bool progress = false;
    for (int k = 0; k < nr; ++k) {
      if (pr[k] >= 0) continue;
      stack.assign(1, {k, -1, -1});
      path.clear();
      bool done = false;
      while (!stack.empty() && !done) {
        const Entry e = stack.back();
        stack.pop_back();
        const int me = (int)path.size();
        path.push_back(e);
        const int dv = dist[e.row];
        dist[e.row] = INF;  // visited in this phase
        for (int c : adj[e.row]) {
          if (pc[c] < 0 && bound == dv + 1) {
            int at = me, col = c;
            while (at >= 0) {
              const Entry& p = path[at];
              pr[p.row] = col;
              pc[col] = p.row;
              col = p.via;
              at = p.parent;
            }
            done = progress = true;
            break;
          }
        }
        if (done) break;
        for (int c : adj[e.row]) {
          const int w = pc[c];
          if (w >= 0 && dist[w] == dv + 1) stack.push_back({w, me, c});
        }
      }
    }
    if (!progress) break;
  }
  return pr;
}

struct Graph {  // CSR adjacency, no self loops
  std::vector<int> ptr, adj;
  int n() const { return (int)ptr.size() - 1; }
};

// BFS distances over the vertices with mark[v] == tag, from `start`
// (lev[v] must be -1 on entry for those vertices); returns the eccentricity.
int bfs_levels(const Graph& g, int start, const std::vector<int>& mark, int tag,
               std::vector<int>& lev, std::vector<int>& order) {
  order.clear();
  order.push_back(start);
  lev[start] = 0;
  int depth = 0;
  for (size_t h = 0; h < order.size(); ++h) {
    int v = order[h];
    for (int e = g.ptr[v]; e < g.ptr[v + 1]; ++e) {
      int w = g.adj[e];
      if (mark[w] == tag && lev[w] < 0) {
        lev[w] = lev[v] + 1;
        depth = std::max(depth, lev[w]);
        order.push_back(w);
      }
    }
  }
  return depth;
}

struct NDNode {
  std::vector<int> verts;
  int parent;
};

// Nested dissection of the quotient graph with vertex weights, restating
// voltax/sparse.py:308-378 step for step (vertex sets are kept sorted, so
// "local index" order is vertex order):
//  * a set that is not connected is split into its components, pushed in
//    order of their smallest vertex (scipy's connected_components labels);
//  * a set of weight <= leaf is a leaf supernode;
//  * pseudo-peripheral start: from the smallest vertex, up to four hops to
//    the smallest farthest vertex (stopping only if it is the start itself
//    or no farther than it), then the level structure from the last start;
//  * eccentricity < 2: the set stays whole;
//  * separator level: among levels 1..depth-1 whose two sides each weigh at
//    least a third, the lightest (first on ties), else the level where the
//    cumulative weight reaches half; the separator is that level's vertices
//    with a neighbour one level further out; the near side (lower levels and
//    the rest of the level) is pushed before the far side.
void nested_dissection(const Graph& g, const std::vector<int>& w, int leaf,
                       const std::vector<char>& last, std::vector<NDNode>& nodes) {
  const int n = g.n();
  int top = -1;
  std::vector<int> rest, lastv;
  for (int v = 0; v < n; ++v) (last[v] ? lastv : rest).push_back(v);
  if (!lastv.empty()) {
    nodes.push_back({lastv, -1});
    top = 0;
  }
  std::vector<int> mark(n, -1), lev(n, -1), comp(n, -1), order;
  int tag = 0;
  std::vector<std::pair<std::vector<int>, int>> stack;
  stack.push_back({rest, top});
  auto reset = [&](const std::vector<int>& vs) {
    for (int v : vs) lev[v] = -1;
  };
  while (!stack.empty()) {
    auto [vs, par] = std::move(stack.back());
    stack.pop_back();
    if (vs.empty()) continue;
    ++tag;
    for (int v : vs) {
      mark[v] = tag;
      lev[v] = -1;
      comp[v] = -1;
    }
    // connected components, labelled in order of their smallest vertex
    int ncomp = 0;
    for (int v : vs) {
      if (comp[v] >= 0) continue;
      bfs_levels(g, v, mark, tag, lev, order);
      for (int u : order) comp[u] = ncomp;
      ++ncomp;
    }
    if (ncomp > 1) {
      std::vector<std::vector<int>> parts(ncomp);
      for (int v : vs) parts[comp[v]].push_back(v);
      for (int c = 0; c < ncomp; ++c) stack.push_back({std::move(parts[c]), par});
      continue;
    }
    long wt = 0;
    for (int v : vs) wt += w[v];
    if (wt <= leaf) {
      nodes.push_back({vs, par});
      continue;
    }
    // pseudo-peripheral start vertex
    int start = vs[0];
    for (int it = 0; it < 4; ++it) {
      reset(vs);
      const int ecc = bfs_levels(g, start, mark, tag, lev, order);
      int far = vs[0];
      for (int v : vs)
        if (lev[v] == ecc) {
          far = v;
          break;
        }
      if (lev[far] <= lev[start] || far == start) break;
      start = far;
    }
    reset(vs);
    const int depth = bfs_levels(g, start, mark, tag, lev, order);
    if (depth < 2) {
      nodes.push_back({vs, par});
      continue;
    }
    std::vector<double> lw(depth + 1, 0.0), cum(depth + 1);
    for (int v : vs) lw[lev[v]] += (double)w[v];
    std::partial_sum(lw.begin(), lw.end(), cum.begin());
    const double tot = cum[depth];
    int best = -1;
    for (int l = 1; l < depth; ++l) {
      if (cum[l - 1] >= tot / 3 && tot - cum[l] >= tot / 3) {
        if (best < 0 || lw[l] < lw[best]) best = l;
      }
    }
    if (best < 0) {
      const int mid = (int)(std::lower_bound(cum.begin(), cum.end(), tot / 2) - cum.begin());
      best = std::min(std::max(mid, 1), depth - 1);
    }
    std::vector<int> sep, lo, hi;
    for (int v : vs) {
      const int l = lev[v];
      if (l < best) {
        lo.push_back(v);
      } else if (l > best) {
        hi.push_back(v);
      } else {
        bool touches = false;
        for (int e = g.ptr[v]; e < g.ptr[v + 1] && !touches; ++e) {
          int u = g.adj[e];
          touches = (mark[u] == tag && lev[u] == best + 1);
        }
        (touches ? sep : lo).push_back(v);
      }
    }
    const int me = (int)nodes.size();
    nodes.push_back({sep, par});
    stack.push_back({lo, me});
    stack.push_back({hi, me});
  }
}

}  // namespace

int build_splu_plan(const Topology& t, int leaf, SPlanHost* out, bool compact) {
  if (leaf < 1) return VXM_ERR_INVALID;
  const int S = t.S;
  SPlanHost& P = *out;
  P = SPlanHost();
  P.S = S;
  P.nnz = t.nnz;

  // ---- 1. pivot groups (voltax/sparse.py:257-305): every device with
  // internal unknowns (branch currents, internal nodes) has them eliminated
  // with terminal nodes of its own.  A two-terminal device with one internal
  // unknown is matched to one distinct terminal; other such devices (and
  // unmatched ones) are glued whole.  Devices are visited in table order
  // (groups in sorted name order, as the reference iterates them).
  DSU dsu(S);
  std::vector<int> pair_x;
  std::vector<std::vector<int>> pair_terms, whole;
  for (int d = 0; d < t.D; ++d) {
    const int m = kind_m(t.kind[d]), T = t.nterm[d];
    if (m - T <= 0) continue;
    const int* u = t.unk.data() + (size_t)t.uw * d;
    std::vector<int> terms;
    for (int i = 0; i < T; ++i)
      if (u[i] >= 0) terms.push_back(u[i]);
    if (T == 2 && m - T == 1 && !terms.empty()) {
      pair_x.push_back(u[2]);
      pair_terms.push_back(terms);
    } else if (compact) {
      // compact grouping: the device's internal unknowns together, its
      // terminals free (a terminal shared by many devices -- a supply node --
      // would otherwise glue them all into one front)
      std::vector<int> inner;
      for (int i = T; i < m; ++i) inner.push_back(u[i]);
      whole.push_back(std::move(inner));
    } else {
      std::vector<int> all = terms;
      for (int i = T; i < m; ++i) all.push_back(u[i]);
      whole.push_back(std::move(all));
    }
  }
  if (compact) {
    // an unknown with no diagonal contribution (and no gmin: not a node),
    // other than a matched two-terminal device's, is eliminated together
    // with every unknown it couples to
    std::vector<char> paired(S, 0);
    for (int x : pair_x) paired[x] = 1;
    for (int v = t.nv; v < S; ++v) {
      const int p = t.diag[v];
      if (paired[v] || t.jptr[p + 1] != t.jptr[p]) continue;
      std::vector<int> grp{v};
      for (int q = t.indptr[v]; q < t.indptr[v + 1]; ++q) grp.push_back(t.indices[q]);
      whole.push_back(std::move(grp));
    }
  }
  if (!pair_x.empty()) {
    std::vector<std::vector<int>> adj(pair_x.size());
    for (size_t i = 0; i < pair_x.size(); ++i) {
      adj[i] = pair_terms[i];
      std::sort(adj[i].begin(), adj[i].end());
      adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
    }
    const std::vector<int> match = max_bipartite_matching(adj, t.nv);
    for (size_t i = 0; i < pair_x.size(); ++i) {
      if (match[i] >= 0) {
        dsu.unite(match[i], pair_x[i]);
      } else {
        std::vector<int> all = pair_terms[i];
        all.push_back(pair_x[i]);
        whole.push_back(std::move(all));
      }
    }
  }
  for (const auto& grp : whole)
    for (size_t i = 1; i < grp.size(); ++i) dsu.unite(grp[0], grp[i]);
  // group ids in order of their union-find roots (np.unique of the roots)
  std::vector<int> gid(S), root_id(S, -1);
  std::vector<char> is_root(S, 0);
  for (int u = 0; u < S; ++u) is_root[dsu.find(u)] = 1;
  int G = 0;
  for (int r = 0; r < S; ++r)
    if (is_root[r]) root_id[r] = G++;
  for (int u = 0; u < S; ++u) gid[u] = root_id[dsu.find(u)];
  std::vector<std::vector<int>> members(G);
  for (int u = 0; u < S; ++u) members[gid[u]].push_back(u);
  std::vector<int> wgt(G);
  for (int g = 0; g < G; ++g) wgt[g] = (int)members[g].size();

  // ---- 2. symmetric unknown adjacency and quotient graph
  std::vector<std::vector<int>> uadj(S);
  for (int r = 0; r < S; ++r)
    for (int p = t.indptr[r]; p < t.indptr[r + 1]; ++p) {
      int c = t.indices[p];
      if (c == r) continue;
      uadj[r].push_back(c);
      uadj[c].push_back(r);
    }
  for (auto& v : uadj) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  Graph q;
  q.ptr.assign(G + 1, 0);
  {
    std::vector<std::vector<int>> qa(G);
    for (int u = 0; u < S; ++u)
      for (int v : uadj[u])
        if (gid[u] != gid[v]) qa[gid[u]].push_back(gid[v]);
    for (int g = 0; g < G; ++g) {
      auto& v = qa[g];
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      q.ptr[g + 1] = q.ptr[g] + (int)v.size();
    }
    q.adj.reserve(q.ptr[G]);
    for (auto& v : qa) q.adj.insert(q.adj.end(), v.begin(), v.end());
  }

  // ---- 3. nested dissection (high-degree groups go to the root separator)
  std::vector<char> last(G, 0);
  if (S <= 2 * leaf) {
    std::fill(last.begin(), last.end(), 1);
  } else {
    const double cap = std::max(32.0, 8.0 * std::sqrt((double)G));
    for (int g = 0; g < G; ++g) last[g] = (q.ptr[g + 1] - q.ptr[g]) > cap;
  }
  std::vector<NDNode> nodes;
  nested_dissection(q, wgt, leaf, last, nodes);

  // postorder of the separator tree
  const int nn = (int)nodes.size();
  std::vector<std::vector<int>> kids(nn);
  std::vector<int> roots;
  for (int i = 0; i < nn; ++i) (nodes[i].parent >= 0 ? kids[nodes[i].parent] : roots).push_back(i);
  std::vector<int> post;
  {
    std::vector<std::pair<int, bool>> st;
    for (auto it = roots.rbegin(); it != roots.rend(); ++it) st.push_back({*it, false});
    while (!st.empty()) {
      auto [i, done] = st.back();
      st.pop_back();
      if (done) {
        post.push_back(i);
      } else {
        st.push_back({i, true});
        for (auto it = kids[i].rbegin(); it != kids[i].rend(); ++it) st.push_back({*it, false});
      }
    }
  }
  std::vector<int> newid(nn);
  for (int i = 0; i < nn; ++i) newid[post[i]] = i;

  // ---- 4. supernodes: unknown sets, parents, borders, levels
  const int ns = nn;
  P.sn.resize(ns);
  std::vector<std::vector<int>> K(ns), R(ns);
  std::vector<int> sn_of(S, -1);
  for (int i = 0; i < ns; ++i) {
    const NDNode& nd = nodes[post[i]];
    for (int g : nd.verts) K[i].insert(K[i].end(), members[g].begin(), members[g].end());
    std::sort(K[i].begin(), K[i].end());
    for (int u : K[i]) sn_of[u] = i;
    P.sn[i].parent = nd.parent >= 0 ? newid[nd.parent] : -1;
  }
  std::vector<std::vector<int>> ch(ns);
  for (int i = 0; i < ns; ++i)
    if (P.sn[i].parent >= 0) ch[P.sn[i].parent].push_back(i);
  std::vector<int> stamp(S, -1);
  for (int i = 0; i < ns; ++i) {
    std::vector<int> b;
    auto add = [&](int u) {
      if (sn_of[u] > i && stamp[u] != i) {
        stamp[u] = i;
        b.push_back(u);
      }
    };
    for (int u : K[i])
      for (int v : uadj[u]) add(v);
    for (int c : ch[i])
      for (int u : R[c]) add(u);
    std::sort(b.begin(), b.end());
    R[i] = std::move(b);
    int lv = 0;
    for (int c : ch[i]) lv = std::max(lv, P.sn[c].level + 1);
    P.sn[i].level = lv;
  }

  // ---- 5. layout and maps
  std::vector<int> where(S, -1);
  int64_t foff = 0;
  int poff = 0;
  for (int i = 0; i < ns; ++i) {
    SNode& s = P.sn[i];
    s.k = (int)K[i].size();
    s.r = (int)R[i].size();
    const int nf = s.k + s.r;
    s.front_off = foff;
    foff += (int64_t)nf * nf;
    s.piv_off = poff;
    poff += s.k;
    s.uvec_off = P.r_total;
    P.r_total += s.r;
    s.idx_off = (int)P.idx.size();
    P.idx.insert(P.idx.end(), K[i].begin(), K[i].end());
    P.idx.insert(P.idx.end(), R[i].begin(), R[i].end());
    s.child_off = (int)P.children.size();
    s.n_child = (int)ch[i].size();
    P.children.insert(P.children.end(), ch[i].begin(), ch[i].end());
    P.max_front = std::max(P.max_front, nf);
    P.max_k = std::max(P.max_k, s.k);
    const double k = s.k, r = s.r;
    P.lu_flops += 2.0 / 3.0 * k * k * k + 2.0 * k * k * r + 2.0 * k * r * r;
    P.factor_entries += k * k + 2.0 * k * r;
  }
  P.front_total = foff;
  P.piv_total = poff;
  // child border -> parent front positions
  for (int i = 0; i < ns; ++i) P.sn[i].cmap_off = -1;
  for (int p = 0; p < ns; ++p) {
    const SNode& s = P.sn[p];
    for (int j = 0; j < s.k + s.r; ++j) where[P.idx[s.idx_off + j]] = j;
    for (int c : ch[p]) {
      P.sn[c].cmap_off = (int)P.cmap.size();
      for (int u : R[c]) {
        if (where[u] < 0) return VXM_ERR_INVALID;  // structural invariant
        P.cmap.push_back(where[u]);
      }
    }
    for (int j = 0; j < s.k + s.r; ++j) where[P.idx[s.idx_off + j]] = -1;
  }
  // A entries grouped by owning supernode
  std::vector<std::vector<std::pair<int, int>>> ents(ns);
  for (int r = 0; r < S; ++r)
    for (int p = t.indptr[r]; p < t.indptr[r + 1]; ++p) {
      const int c = t.indices[p];
      ents[std::min(sn_of[r], sn_of[c])].push_back({p, r});
    }
  for (int i = 0; i < ns; ++i) {
    SNode& s = P.sn[i];
    const int nf = s.k + s.r;
    for (int j = 0; j < nf; ++j) where[P.idx[s.idx_off + j]] = j;
    s.aent_off = (int)P.aslot.size();
    s.n_aent = (int)ents[i].size();
    for (auto& e : ents[i]) {
      const int p = e.first, r = e.second;
      const int c = t.indices[p];
      if (where[r] < 0 || where[c] < 0) return VXM_ERR_INVALID;
      P.aslot.push_back(p);
      P.apos.push_back(where[r] * nf + where[c]);
    }
    for (int j = 0; j < nf; ++j) where[P.idx[s.idx_off + j]] = -1;
  }
  // level sets, heaviest fronts first within a level
  int nlev = 0;
  for (auto& s : P.sn) nlev = std::max(nlev, s.level + 1);
  std::vector<std::vector<int>> lv(nlev);
  for (int i = 0; i < ns; ++i) lv[P.sn[i].level].push_back(i);
  P.level_ptr.assign(1, 0);
  for (auto& v : lv) {
    std::stable_sort(v.begin(), v.end(), [&](int a, int b) {
      return P.sn[a].k + P.sn[a].r > P.sn[b].k + P.sn[b].r;
    });
    const int beg = (int)P.level_sn.size();
    P.level_sn.insert(P.level_sn.end(), v.begin(), v.end());
    P.level_ptr.push_back((int)P.level_sn.size());
    static const int kClassMax[4] = {1 << 30, 64, 32, 16};
    for (int c = 0; c < 4; ++c) {
      int i = beg;
      while (i < (int)P.level_sn.size() &&
             P.sn[P.level_sn[i]].k + P.sn[P.level_sn[i]].r > kClassMax[c])
        ++i;
      P.level_cls.push_back(c == 0 ? beg : i);
    }
    P.level_cls.push_back((int)P.level_sn.size());
  }
  return VXM_OK;
}

}  // namespace vxm
