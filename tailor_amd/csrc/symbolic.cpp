// symbolic.cpp -- geometric nested dissection and multifrontal symbolic
// factorization of the BiGlobal block graph.
#include "symbolic.hpp"

#include <algorithm>
#include <functional>
#include <numeric>
#include <stdexcept>

namespace tailor {

namespace {

struct Graph {
  std::vector<int> ptr, adj;  // symmetrised node adjacency (CSR, sorted)
};

int wrap(int x, int n) {
  int r = x % n;
  return r < 0 ? r + n : r;
}

Graph stencil_graph(int ny, int nz, int ssx, int sse, const std::vector<int> &xoff,
                    const std::vector<int> &eoff) {
  const int n = ny * nz;
  std::vector<std::vector<int>> nb(n);
  for (int j = 0; j < nz; ++j)
    for (int i = 0; i < ny; ++i) {
      const int u = j * ny + i;
      for (int t = 0; t < sse; ++t) {
        const int jc = wrap(j + eoff[j * sse + t], nz);
        for (int s = 0; s < ssx; ++s) {
          const int v = jc * ny + i + xoff[i * ssx + s];
          nb[u].push_back(v);
          nb[v].push_back(u);
        }
      }
    }
  Graph g;
  g.ptr.assign(n + 1, 0);
  for (int u = 0; u < n; ++u) {
    auto &l = nb[u];
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
    g.ptr[u + 1] = g.ptr[u] + static_cast<int>(l.size());
  }
  g.adj.reserve(g.ptr[n]);
  for (int u = 0; u < n; ++u) g.adj.insert(g.adj.end(), nb[u].begin(), nb[u].end());
  return g;
}

struct Tree {
  std::vector<std::vector<int>> nodes;     // per supernode, in postorder
  std::vector<std::vector<int>> children;
};

}  // namespace

int Symbolic::local_index(int f, int node) const {
  const int ps = pos[node];
  if (ps >= sn_start[f] && ps < sn_start[f + 1]) return ps - sn_start[f];
  const auto &st = strct[f];
  auto it = std::lower_bound(st.begin(), st.end(), node,
                             [&](int a, int b) { return pos[a] < pos[b]; });
  if (it == st.end() || *it != node) return -1;
  return (sn_start[f + 1] - sn_start[f]) + static_cast<int>(it - st.begin());
}

Symbolic analyse(int ny, int nz, int ssx, int sse, const std::vector<int> &xoff,
                 const std::vector<int> &eoff, int leaf_nodes) {
  const int n = ny * nz;
  const Graph g = stencil_graph(ny, nz, ssx, sse, xoff, eoff);
  Tree tree;
  std::vector<char> mark(n, 0);

  // Recursive bisection.  `ring` is true while the eta range still closes
  // on itself (the subdomain spans the whole periodic direction).
  std::function<int(std::vector<int> &, bool)> rec = [&](std::vector<int> &nodes,
                                                        bool ring) -> int {
    if (nodes.empty()) return -1;
    if (static_cast<int>(nodes.size()) <= leaf_nodes) {
      tree.nodes.push_back(nodes);
      tree.children.push_back({});
      return static_cast<int>(tree.nodes.size()) - 1;
    }
    struct Cand {
      std::vector<int> sep, left, right;
      bool xsplit = false;
      bool ok = false;
    } best;
    for (int dir = 0; dir < 2; ++dir) {
      std::vector<int> coord(nodes.size());
      for (size_t k = 0; k < nodes.size(); ++k)
        coord[k] = dir == 0 ? nodes[k] % ny : nodes[k] / ny;
      int mid;
      if (dir == 1 && ring) {
        mid = nz / 2;
      } else {
        if (dir == 1) {
          // unwrap an eta interval that crosses the periodic seam
          std::vector<char> seen(nz, 0);
          for (int v : coord) seen[v] = 1;
          const int cnt = static_cast<int>(std::count(seen.begin(), seen.end(), 1));
          const int lo = *std::min_element(coord.begin(), coord.end());
          const int hi = *std::max_element(coord.begin(), coord.end());
          if (hi - lo + 1 > cnt)
            for (int &v : coord)
              if (v < nz / 2) v += nz;
        }
        const int lo = *std::min_element(coord.begin(), coord.end());
        const int hi = *std::max_element(coord.begin(), coord.end());
        if (hi - lo < 2) continue;
        mid = (lo + hi + 1) / 2;
      }
      Cand cd;
      cd.xsplit = dir == 0;
      std::vector<int> L, R;
      for (size_t k = 0; k < nodes.size(); ++k) (coord[k] < mid ? L : R).push_back(nodes[k]);
      if (L.empty() || R.empty()) continue;
      for (int v : R) mark[v] = 1;
      for (int u : L) {
        bool adj = false;
        for (int e = g.ptr[u]; e < g.ptr[u + 1] && !adj; ++e) adj = mark[g.adj[e]] != 0;
        (adj ? cd.sep : cd.left).push_back(u);
      }
      for (int v : R) mark[v] = 0;
      cd.right = std::move(R);
      cd.ok = true;
      if (!best.ok || cd.sep.size() < best.sep.size()) best = std::move(cd);
    }
    if (!best.ok) {
      tree.nodes.push_back(nodes);
      tree.children.push_back({});
      return static_cast<int>(tree.nodes.size()) - 1;
    }
    const bool child_ring = ring && best.xsplit;
    const int c1 = rec(best.left, child_ring);
    const int c2 = rec(best.right, child_ring);
    std::vector<int> ch;
    if (c1 >= 0) ch.push_back(c1);
    if (c2 >= 0) ch.push_back(c2);
    tree.nodes.push_back(best.sep);
    tree.children.push_back(ch);
    return static_cast<int>(tree.nodes.size()) - 1;
  };
  std::vector<int> all(n);
  std::iota(all.begin(), all.end(), 0);
  rec(all, true);

  Symbolic S;
  S.ny = ny;
  S.nz = nz;
  S.nodes = n;
  S.nsn = static_cast<int>(tree.nodes.size());
  S.pos.assign(n, -1);
  S.order.reserve(n);
  S.sn_start.assign(S.nsn + 1, 0);
  for (int s = 0; s < S.nsn; ++s) {
    auto nodes = tree.nodes[s];
    std::sort(nodes.begin(), nodes.end());
    S.sn_start[s] = static_cast<int>(S.order.size());
    for (int v : nodes) {
      S.pos[v] = static_cast<int>(S.order.size());
      S.order.push_back(v);
    }
  }
  S.sn_start[S.nsn] = n;
  if (static_cast<int>(S.order.size()) != n) throw std::runtime_error("nested dissection lost nodes");
  S.children = tree.children;
  S.parent.assign(S.nsn, -1);
  for (int s = 0; s < S.nsn; ++s)
    for (int ch : S.children[s]) S.parent[ch] = s;
  S.level.assign(S.nsn, 0);
  for (int s = 0; s < S.nsn; ++s)
    for (int ch : S.children[s]) S.level[s] = std::max(S.level[s], S.level[ch] + 1);
  int maxlev = 0;
  for (int l : S.level) maxlev = std::max(maxlev, l);
  S.levels.assign(maxlev + 1, {});
  for (int s = 0; s < S.nsn; ++s) S.levels[S.level[s]].push_back(s);

  // symbolic factorization: struct(K) = ancestors' nodes adjacent to K or to
  // the update structure of K's children
  std::vector<int> sn_of(n);
  for (int s = 0; s < S.nsn; ++s)
    for (int q = S.sn_start[s]; q < S.sn_start[s + 1]; ++q) sn_of[S.order[q]] = s;
  S.strct.assign(S.nsn, {});
  std::vector<int> stamp(n, -1);
  for (int s = 0; s < S.nsn; ++s) {
    const int last = S.sn_start[s + 1] - 1;
    std::vector<int> st;
    auto add = [&](int v) {
      if (S.pos[v] > last && stamp[v] != s) {
        stamp[v] = s;
        st.push_back(v);
      }
    };
    for (int q = S.sn_start[s]; q <= last; ++q) {
      const int u = S.order[q];
      for (int e = g.ptr[u]; e < g.ptr[u + 1]; ++e) add(g.adj[e]);
    }
    for (int ch : S.children[s])
      for (int v : S.strct[ch]) add(v);
    std::sort(st.begin(), st.end(), [&](int a, int b) { return S.pos[a] < S.pos[b]; });
    S.strct[s] = std::move(st);
  }
  S.p.resize(S.nsn);
  S.c.resize(S.nsn);
  S.m.resize(S.nsn);
  for (int s = 0; s < S.nsn; ++s) {
    S.p[s] = 5 * (S.sn_start[s + 1] - S.sn_start[s]);
    S.c[s] = 5 * static_cast<int>(S.strct[s].size());
    S.m[s] = S.p[s] + S.c[s];
    const double p = S.p[s], c = S.c[s];
    S.flops += static_cast<int64_t>(8.0 * (p * p * p / 3.0 + p * p * c + p * c * c));
    S.factor_entries += static_cast<int64_t>(p * p + 2.0 * p * c);
  }
  // extend-add maps: position of each update node of a child in its parent
  S.rel.assign(S.nsn, {});
  for (int s = 0; s < S.nsn; ++s) {
    const int par = S.parent[s];
    if (par < 0) {
      if (!S.strct[s].empty()) throw std::runtime_error("root front has update rows");
      continue;
    }
    for (int v : S.strct[s]) {
      const int li = S.local_index(par, v);
      if (li < 0) throw std::runtime_error("update row missing from the parent front");
      S.rel[s].push_back(li);
    }
  }
  return S;
}

}  // namespace tailor
