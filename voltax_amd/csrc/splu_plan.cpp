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

// Kuhn's augmenting-path matching of voltage-source currents to distinct
// terminal nodes (what voltax/sparse.py:283-297 does with scipy's
// maximum_bipartite_matching).
bool try_kuhn(int v, const std::vector<std::vector<int>>& cand, std::vector<int>& owner,
              std::vector<int>& seen, int stamp) {
  for (int node : cand[v]) {
    if (seen[node] == stamp) continue;
    seen[node] = stamp;
    if (owner[node] < 0 || try_kuhn(owner[node], cand, owner, seen, stamp)) {
      owner[node] = v;
      return true;
    }
  }
  return false;
}

struct Graph {  // CSR adjacency, no self loops
  std::vector<int> ptr, adj;
  int n() const { return (int)ptr.size() - 1; }
};

// BFS over the vertices with mark[v] == tag, from `start`; returns levels.
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

// Nested dissection of the quotient graph with vertex weights.
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
  std::vector<int> mark(n, -1), lev(n, -1), order;
  int tag = 0;
  std::vector<std::pair<std::vector<int>, int>> stack;
  stack.push_back({rest, top});
  while (!stack.empty()) {
    auto [vs, par] = std::move(stack.back());
    stack.pop_back();
    if (vs.empty()) continue;
    ++tag;
    for (int v : vs) {
      mark[v] = tag;
      lev[v] = -1;
    }
    // connected components
    bfs_levels(g, vs[0], mark, tag, lev, order);
    if (order.size() < vs.size()) {
      std::vector<int> comp = order, other;
      for (int v : vs)
        if (lev[v] < 0) other.push_back(v);
      std::sort(comp.begin(), comp.end());
      stack.push_back({other, par});
      stack.push_back({comp, par});
      continue;
    }
    long wt = 0;
    for (int v : vs) wt += w[v];
    if (wt <= leaf) {
      nodes.push_back({vs, par});
      continue;
    }
    // pseudo-peripheral start: repeated farthest-vertex sweeps
    int start = vs[0];
    int ecc = bfs_levels(g, start, mark, tag, lev, order);
    for (int it = 0; it < 4; ++it) {
      int far = order.back();
      for (int v : vs) lev[v] = -1;
      int e2 = bfs_levels(g, far, mark, tag, lev, order);
      if (e2 <= ecc) {
        for (int v : vs) lev[v] = -1;
        bfs_levels(g, start, mark, tag, lev, order);
        break;
      }
      start = far;
      ecc = e2;
    }
    const int depth = ecc;
    if (depth < 2) {
      nodes.push_back({vs, par});
      continue;
    }
    std::vector<long> lw(depth + 1, 0);
    for (int v : vs) lw[lev[v]] += w[v];
    std::vector<long> cum(depth + 1);
    std::partial_sum(lw.begin(), lw.end(), cum.begin());
    const double tot = (double)cum[depth];
    int best = -1;
    for (int l = 1; l < depth; ++l) {
      if (cum[l - 1] >= tot / 3 && tot - cum[l] >= tot / 3) {
        if (best < 0 || lw[l] < lw[best]) best = l;
      }
    }
    if (best < 0) {
      int mid = (int)(std::lower_bound(cum.begin(), cum.end(), (long)std::ceil(tot / 2)) -
                      cum.begin());
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

int build_splu_plan(const Topology& t, int leaf, SPlanHost* out) {
  if (leaf < 1) return VXM_ERR_INVALID;
  const int S = t.S;
  SPlanHost& P = *out;
  P = SPlanHost();
  P.S = S;
  P.nnz = t.nnz;

  // ---- 1. pivot groups
  DSU dsu(S);
  std::vector<int> vint;
  std::vector<std::vector<int>> cand;
  std::vector<std::vector<int>> whole;
  for (int d = 0; d < t.D; ++d) {
    if (t.kind[d] != KIND_V) continue;
    const int a = t.unk[4 * d], b = t.unk[4 * d + 1], x = t.unk[4 * d + 2];
    std::vector<int> terms;
    if (a >= 0) terms.push_back(a);
    if (b >= 0) terms.push_back(b);
    if (terms.empty()) continue;
    vint.push_back(x);
    cand.push_back(terms);
  }
  if (!vint.empty()) {
    std::vector<int> owner(t.nv, -1), seen(t.nv, 0);
    for (int v = 0; v < (int)vint.size(); ++v) try_kuhn(v, cand, owner, seen, v + 1);
    std::vector<int> matched(vint.size(), -1);
    for (int node = 0; node < t.nv; ++node)
      if (owner[node] >= 0) matched[owner[node]] = node;
    for (size_t v = 0; v < vint.size(); ++v) {
      if (matched[v] >= 0) {
        dsu.unite(matched[v], vint[v]);
      } else {
        for (int term : cand[v]) dsu.unite(term, vint[v]);
      }
    }
  }
  std::vector<int> gid(S), root_id(S, -1);
  int G = 0;
  for (int u = 0; u < S; ++u) {
    int r = dsu.find(u);
    if (root_id[r] < 0) root_id[r] = G++;
    gid[u] = root_id[r];
  }
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
