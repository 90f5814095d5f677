// gen.h: synthetic graphs, initial states and update streams (host side).
//
// Graphs are built as sorted, symmetric, self-loop-free, duplicate-free CSRs and written as
// wide .egr files. Streams follow the conventions of the reference's batch generators
// (tools/generate_mixed_batches.py:94-141): one line "u v" per undirected toggle, 1-indexed,
// canonical u < v, and no pair repeated within or across batches; a line toggles the current
// graph (insert if absent, delete if present).
#pragma once

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "egr.h"
#include "rng.h"

namespace scout {

struct Csr {
  int32_t n = 0;
  std::vector<int64_t> off;
  std::vector<int32_t> col;
  int64_t entries() const { return off.empty() ? 0 : off.back(); }
};

// Builds a CSR from directed keys (u << 32 | v): sort, drop duplicates.
inline Csr csr_from_keys(int32_t n, std::vector<uint64_t>& keys) {
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  Csr g;
  g.n = n;
  g.off.assign((size_t)n + 1, 0);
  g.col.resize(keys.size());
  for (size_t i = 0; i < keys.size(); i++) {
    g.off[(size_t)(keys[i] >> 32) + 1]++;
    g.col[i] = (int32_t)(keys[i] & 0xffffffffu);
  }
  for (int32_t v = 0; v < n; v++) g.off[(size_t)v + 1] += g.off[(size_t)v];
  return g;
}

// R-MAT (Chakrabarti, Zhan and Faloutsos, SDM 2004), unpermuted: edge i of `edge_factor *
// 2^scale` draws its quadrant at every level from rmat_edge; self-loops are dropped, both
// directions kept, duplicates removed.
inline Csr gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed) {
  RmatParams p = rmat_params(a, b, c, scale);
  uint64_t key = stream_key(seed, RNG_RMAT);
  int64_t m = (int64_t)edge_factor << scale;
  std::vector<uint64_t> keys((size_t)(2 * m));
  std::vector<uint8_t> keep((size_t)m);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; i++) {
    uint64_t e = rmat_edge(p, key, (uint64_t)i);
    uint32_t u = (uint32_t)(e >> 32), v = (uint32_t)e;
    keep[(size_t)i] = u != v;
    keys[(size_t)(2 * i)] = ((uint64_t)u << 32) | v;
    keys[(size_t)(2 * i + 1)] = ((uint64_t)v << 32) | u;
  }
  size_t w = 0;
  for (int64_t i = 0; i < m; i++)
    if (keep[(size_t)i]) {
      keys[w++] = keys[(size_t)(2 * i)];
      keys[w++] = keys[(size_t)(2 * i + 1)];
    }
  keys.resize(w);
  return csr_from_keys((int32_t)(1LL << scale), keys);
}

// Erdos-Renyi G(n, m): uniform pairs u != v from draw i = 0, 1, ... until m distinct
// undirected pairs exist.
inline Csr gen_er(int32_t n, int64_t m, uint64_t seed) {
  uint64_t key = stream_key(seed, RNG_ER);
  std::unordered_set<uint64_t> seen;
  seen.reserve((size_t)m * 2);
  std::vector<uint64_t> keys;
  keys.reserve((size_t)(2 * m));
  for (uint64_t i = 0; (int64_t)seen.size() < m; i++) {
    uint64_t x = draw(key, i);
    uint32_t u = uniform_below(x, (uint32_t)n);
    uint32_t v = uniform_below(x << 32 | (x >> 32), (uint32_t)n);
    if (u == v) continue;
    uint64_t k = ((uint64_t)std::min(u, v) << 32) | std::max(u, v);
    if (!seen.insert(k).second) continue;
    keys.push_back(((uint64_t)u << 32) | v);
    keys.push_back(((uint64_t)v << 32) | u);
  }
  return csr_from_keys(n, keys);
}

// Barabasi-Albert: vertices 0..m-1 start as a clique-free seed; each new vertex t >= m
// attaches to m distinct earlier vertices chosen proportionally to degree (uniform over the
// endpoint list of existing edges, the seed vertices entered once each).
inline Csr gen_ba(int32_t n, int m, uint64_t seed) {
  uint64_t key = stream_key(seed, RNG_BA);
  std::vector<uint32_t> ends;
  ends.reserve((size_t)n * (size_t)m * 2);
  for (int i = 0; i < m && i < n; i++) ends.push_back((uint32_t)i);
  std::vector<uint64_t> keys;
  keys.reserve((size_t)n * (size_t)m * 2);
  uint64_t ctr = 0;
  std::vector<uint32_t> picked;
  for (int32_t t = m; t < n; t++) {
    picked.clear();
    while ((int)picked.size() < m) {
      uint32_t w = ends[uniform_below(draw(key, ctr++), (uint32_t)ends.size())];
      if (std::find(picked.begin(), picked.end(), w) == picked.end()) picked.push_back(w);
    }
    for (uint32_t w : picked) {
      keys.push_back(((uint64_t)t << 32) | w);
      keys.push_back(((uint64_t)w << 32) | (uint32_t)t);
      ends.push_back(w);
      ends.push_back((uint32_t)t);
    }
  }
  return csr_from_keys(n, keys);
}

inline std::string write_egr(const Csr& g, const std::string& path) {
  return egr_write_wide(path, g.n, g.off.data(), g.col.data());
}

// --- initial states (textbook) ---------------------------------------------------------

inline std::vector<int32_t> bfs_levels(const Csr& g, int32_t src) {
  std::vector<int32_t> lv((size_t)g.n, -1);
  if (src < 0 || src >= g.n) return lv;
  std::vector<int32_t> q{src};
  lv[(size_t)src] = 0;
  for (size_t h = 0; h < q.size(); h++) {
    int32_t x = q[h];
    for (int64_t i = g.off[(size_t)x]; i < g.off[(size_t)x + 1]; i++) {
      int32_t y = g.col[(size_t)i];
      if (lv[(size_t)y] < 0) {
        lv[(size_t)y] = lv[(size_t)x] + 1;
        q.push_back(y);
      }
    }
  }
  return lv;
}

inline std::vector<int32_t> cc_labels(const Csr& g) {
  std::vector<int32_t> lab((size_t)g.n, -1);
  std::vector<int32_t> q;
  for (int32_t r = 0; r < g.n; r++) {
    if (lab[(size_t)r] >= 0) continue;
    lab[(size_t)r] = r;
    q.assign(1, r);
    for (size_t h = 0; h < q.size(); h++)
      for (int64_t i = g.off[(size_t)q[h]]; i < g.off[(size_t)q[h] + 1]; i++) {
        int32_t y = g.col[(size_t)i];
        if (lab[(size_t)y] < 0) {
          lab[(size_t)y] = r;
          q.push_back(y);
        }
      }
  }
  return lab;
}

inline std::vector<uint8_t> greedy_mis(const Csr& g) {
  std::vector<uint8_t> in((size_t)g.n, 0), blocked((size_t)g.n, 0);
  for (int32_t v = 0; v < g.n; v++) {
    if (blocked[(size_t)v]) continue;
    in[(size_t)v] = 1;
    for (int64_t i = g.off[(size_t)v]; i < g.off[(size_t)v + 1]; i++) blocked[(size_t)g.col[(size_t)i]] = 1;
  }
  return in;
}

inline std::vector<int32_t> greedy_color(const Csr& g) {
  std::vector<int32_t> c((size_t)g.n, -1), stamp;
  for (int32_t v = 0; v < g.n; v++) {
    size_t d = (size_t)(g.off[(size_t)v + 1] - g.off[(size_t)v]);
    if (stamp.size() < d + 2) stamp.assign(d + 2, -1);
    for (int64_t i = g.off[(size_t)v]; i < g.off[(size_t)v + 1]; i++) {
      int32_t cw = c[(size_t)g.col[(size_t)i]];
      if (cw >= 0 && (size_t)cw < stamp.size()) stamp[(size_t)cw] = v;
    }
    int32_t k = 0;
    while (stamp[(size_t)k] == v) k++;
    c[(size_t)v] = k;
  }
  return c;
}

inline bool write_text(const std::string& path, const std::string& body) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  bool ok = fwrite(body.data(), 1, body.size(), f) == body.size();
  return fclose(f) == 0 && ok;
}

// Init files in the reference's formats (README.md:130-137).
inline std::string mis_init_text(const std::vector<uint8_t>& in) {
  std::string s;
  for (size_t v = 0; v < in.size(); v++)
    if (in[v]) s += std::to_string(v + 1) + "\n";
  return s;
}
inline std::string per_vertex_text(const std::vector<int32_t>& x) {
  std::string s;
  for (int32_t c : x) s += std::to_string(c) + "\n";
  return s;
}
inline std::string bfs_init_text(const std::vector<int32_t>& lv) {
  std::string s;
  for (size_t v = 0; v < lv.size(); v++) s += std::to_string(v + 1) + " " + std::to_string(lv[v]) + "\n";
  return s;
}

// --- update streams ---------------------------------------------------------------------

// Current undirected edge set with O(1) uniform sampling and removal.
struct EdgeSet {
  std::vector<uint64_t> list;
  std::unordered_map<uint64_t, size_t> pos;
  static uint64_t key(uint32_t u, uint32_t v) { return ((uint64_t)std::min(u, v) << 32) | std::max(u, v); }
  void build(const Csr& g) {
    list.clear();
    pos.clear();
    for (int32_t u = 0; u < g.n; u++)
      for (int64_t i = g.off[(size_t)u]; i < g.off[(size_t)u + 1]; i++) {
        uint32_t v = (uint32_t)g.col[(size_t)i];
        if ((uint32_t)u < v) add(key((uint32_t)u, v));
      }
  }
  bool has(uint64_t k) const { return pos.count(k) != 0; }
  void add(uint64_t k) {
    if (pos.emplace(k, list.size()).second) list.push_back(k);
  }
  void remove(uint64_t k) {
    auto it = pos.find(k);
    if (it == pos.end()) return;
    size_t i = it->second;
    pos.erase(it);
    if (i + 1 != list.size()) {
      list[i] = list.back();
      pos[list[i]] = i;
    }
    list.pop_back();
  }
};

struct StreamSpec {
  std::string kind = "mixed";  // mixed | tree | degree
  int batches = 1;
  int64_t size = 100;
  double delete_frac = 0.5;
  double tree_frac = 0.5;      // tree: share of deletions drawn from consecutive-level edges
  uint64_t seed = 1;
};

// Adjacency of the current graph for BFS levels during "tree" streams.
inline Csr csr_of(int32_t n, const EdgeSet& es) {
  std::vector<uint64_t> keys;
  keys.reserve(es.list.size() * 2);
  for (uint64_t k : es.list) {
    uint32_t u = (uint32_t)(k >> 32), v = (uint32_t)k;
    keys.push_back(((uint64_t)u << 32) | v);
    keys.push_back(((uint64_t)v << 32) | u);
  }
  return csr_from_keys(n, keys);
}

// Writes `spec.batches` files 1.mtx .. B.mtx under dir; returns "" or an error.
inline std::string gen_stream(const Csr& g0, const std::string& dir, const StreamSpec& spec) {
  const int32_t n = g0.n;
  EdgeSet es;
  es.build(g0);
  std::unordered_set<uint64_t> toggled;
  uint64_t kdel = stream_key(spec.seed, RNG_STREAM_DEL), kins = stream_key(spec.seed, RNG_STREAM_INS),
           ktree = stream_key(spec.seed, RNG_STREAM_TREE);
  uint64_t cdel = 0, cins = 0, ctree = 0;
  std::vector<int32_t> level;
  std::vector<double> cum;  // degree+1 prefix for the degree-biased stream
  for (int b = 1; b <= spec.batches; b++) {
    int64_t ndel = (int64_t)std::llround(spec.delete_frac * (double)spec.size);
    int64_t nins = spec.size - ndel;
    Csr cur;
    if (spec.kind == "tree" || spec.kind == "degree") cur = csr_of(n, es);
    if (spec.kind == "tree") level = bfs_levels(cur, 0);
    if (spec.kind == "degree") {
      cum.assign((size_t)n + 1, 0.0);
      for (int32_t v = 0; v < n; v++)
        cum[(size_t)v + 1] = cum[(size_t)v] + (double)(cur.off[(size_t)v + 1] - cur.off[(size_t)v] + 1);
    }
    auto degree_pick = [&](uint64_t x) -> uint32_t {
      double t = (double)(x >> 11) * (1.0 / 9007199254740992.0) * cum[(size_t)n];
      size_t i = (size_t)(std::upper_bound(cum.begin(), cum.end(), t) - cum.begin());
      return (uint32_t)std::min<size_t>(i == 0 ? 0 : i - 1, (size_t)n - 1);
    };
    std::vector<uint64_t> lines;
    std::unordered_set<uint64_t> in_batch;
    int64_t ntree = spec.kind == "tree" ? (int64_t)std::llround(spec.tree_frac * (double)ndel) : 0;
    int64_t guard = 0;
    // deletions drawn from consecutive-level edges
    for (int64_t k = 0; k < ntree && !es.list.empty() && guard < 1000 * spec.size; guard++) {
      uint64_t e = es.list[uniform_below(draw(ktree, ctree++), (uint32_t)es.list.size())];
      uint32_t u = (uint32_t)(e >> 32), v = (uint32_t)e;
      int32_t lu = level[u], lvv = level[v];
      bool tree = lu >= 0 && lvv >= 0 && (lu - lvv == 1 || lvv - lu == 1);
      if (!tree || toggled.count(e) || in_batch.count(e)) continue;
      in_batch.insert(e);
      lines.push_back(e);
      k++;
    }
    // other deletions
    int64_t rest = ndel - (int64_t)lines.size();
    for (int64_t k = 0; k < rest && !es.list.empty() && guard < 2000 * spec.size; guard++) {
      uint64_t e;
      if (spec.kind == "degree") {
        uint32_t u = degree_pick(draw(kdel, cdel++));
        int64_t d = cur.off[(size_t)u + 1] - cur.off[(size_t)u];
        if (d == 0) continue;
        uint32_t v = (uint32_t)cur.col[(size_t)(cur.off[(size_t)u] + uniform_below(draw(kdel, cdel++), (uint32_t)d))];
        e = EdgeSet::key(u, v);
        if (!es.has(e)) continue;
      } else {
        e = es.list[uniform_below(draw(kdel, cdel++), (uint32_t)es.list.size())];
      }
      if (toggled.count(e) || in_batch.count(e)) continue;
      in_batch.insert(e);
      lines.push_back(e);
      k++;
    }
    // insertions of current non-edges
    for (int64_t k = 0; k < nins && guard < 3000 * spec.size; guard++) {
      uint32_t u, v;
      if (spec.kind == "degree") {
        u = degree_pick(draw(kins, cins++));
        v = degree_pick(draw(kins, cins++));
      } else {
        uint64_t x = draw(kins, cins++);
        u = uniform_below(x, (uint32_t)n);
        v = uniform_below(x << 32 | (x >> 32), (uint32_t)n);
      }
      if (u == v) continue;
      uint64_t e = EdgeSet::key(u, v);
      if (es.has(e) || toggled.count(e) || in_batch.count(e)) continue;
      in_batch.insert(e);
      lines.push_back(e);
      k++;
    }
    // shuffle the batch deterministically so insertions and deletions interleave
    uint64_t ks = stream_key(spec.seed ^ (uint64_t)b, RNG_STREAM_SIZE);
    for (size_t i = lines.size(); i > 1; i--) {
      size_t j = uniform_below(draw(ks, i), (uint32_t)i);
      std::swap(lines[i - 1], lines[j]);
    }
    std::string body;
    body.reserve(lines.size() * 16);
    for (uint64_t e : lines) {
      uint32_t u = (uint32_t)(e >> 32), v = (uint32_t)e;
      body += std::to_string(u + 1) + " " + std::to_string(v + 1) + "\n";
      toggled.insert(e);
      if (es.has(e)) es.remove(e);
      else es.add(e);
    }
    if (!write_text(dir + "/" + std::to_string(b) + ".mtx", body)) return "cannot write batch " + std::to_string(b);
  }
  return "";
}

}  // namespace scout
