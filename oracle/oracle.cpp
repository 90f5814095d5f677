// oracle.cpp -- CPU restatement of the SCOUT reference's hot path. TEST INFRASTRUCTURE ONLY:
// loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the test oracle,
// never by the product path.
//
// Pinning (what fixes each function independently of this file):
//   LogicalGraph (scratchpad)  brute-force multiset model in tests/test_oracle.py and the
//                              reference's own operations restated 1:1 (FlashGraph.h:200-310)
//   orc_bfs / orc_cc           textbook closed forms; datasets/sample/sample_bfs.txt and
//                              sample_components.txt (golden fixtures, tests/golden/)
//   orc_greedy_mis             datasets/sample/sample_mis.txt (golden fixture)
//   orc_greedy_color           closed form of tools/greedy_color.cpp:37-52; proper colouring
//   orc_color_batch            brute force on tiny graphs + proper-colouring invariant +
//                              planted-mistake checks in tests/test_oracle.py
//   orc_mis_batch              hand-worked cases from src/mis/scout_mis.cu:290-362 in
//                              tests/test_oracle.py (the reference's own counterexample
//                              path u-v-x-y), independence invariant
//   orc_logical_csr            compactor byte layout (flash_compactor.h:30-95) on the sample
//   orc_refbfs / orc_refcc     end-of-stream equality with the closed forms (deferred
//                              fallback settles every cascade, src/bfs/scout_bfs.cu:593-602)
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include <omp.h>

#include "../scout_amd/csrc/host/egr.h"

namespace {

// Base CSR + DeltaOverlay + HybridGraph of core/FlashGraph.h:200-388, restated.
struct LogicalGraph {
  int32_t n = 0;
  std::vector<int64_t> off;
  std::vector<int32_t> col;
  std::unordered_map<int32_t, std::vector<int32_t>> added;        // FlashGraph.h:202
  std::unordered_map<int32_t, std::unordered_set<int32_t>> removed;  // FlashGraph.h:203

  int64_t base_degree(int32_t v) const { return off[(size_t)v + 1] - off[(size_t)v]; }

  // FlashGraph.h:209-225 then the linear base scan of :356-366.
  bool is_adjacent(int32_t u, int32_t v) const {
    auto a = added.find(u);
    if (a != added.end() && std::find(a->second.begin(), a->second.end(), v) != a->second.end()) return true;
    auto r = removed.find(u);
    if (r != removed.end() && r->second.count(v)) return false;
    for (int64_t i = off[(size_t)u]; i < off[(size_t)u + 1]; i++)
      if (col[(size_t)i] == v) return true;
    return false;
  }
  // FlashGraph.h:284-291
  bool un_tombstone(int32_t u, int32_t v) {
    auto ru = removed.find(u);
    if (ru == removed.end() || !ru->second.count(v)) return false;
    ru->second.erase(v);
    auto rv = removed.find(v);
    if (rv != removed.end()) rv->second.erase(u);
    return true;
  }
  // FlashGraph.h:295-310
  bool forget_delta_edge(int32_t u, int32_t v) {
    auto au = added.find(u);
    if (au == added.end()) return false;
    auto it = std::find(au->second.begin(), au->second.end(), v);
    if (it == au->second.end()) return false;
    au->second.erase(it);
    auto av = added.find(v);
    if (av != added.end()) {
      auto jt = std::find(av->second.begin(), av->second.end(), u);
      if (jt != av->second.end()) av->second.erase(jt);
    }
    return true;
  }
  // FlashGraph.h:227-243
  void insert_edge(int32_t u, int32_t v) {
    if (!un_tombstone(u, v)) {
      added[u].push_back(v);
      added[v].push_back(u);
    }
  }
  void remove_edge(int32_t u, int32_t v) {
    if (!forget_delta_edge(u, v)) {
      removed[u].insert(v);
      removed[v].insert(u);
    }
  }
  // FlashGraph.h:246-265: base minus tombstones, then added in insertion order.
  void neighbors(int32_t v, std::vector<int32_t>& out) const {
    out.clear();
    auto r = removed.find(v);
    const std::unordered_set<int32_t>* tomb = (r != removed.end() && !r->second.empty()) ? &r->second : nullptr;
    for (int64_t i = off[(size_t)v]; i < off[(size_t)v + 1]; i++) {
      int32_t w = col[(size_t)i];
      if (!tomb || !tomb->count(w)) out.push_back(w);
    }
    auto a = added.find(v);
    if (a != added.end()) out.insert(out.end(), a->second.begin(), a->second.end());
  }
  // FlashGraph.h:267-274
  int64_t degree(int32_t v) const {
    int64_t d = base_degree(v);
    auto r = removed.find(v);
    if (r != removed.end()) d -= (int64_t)r->second.size();
    auto a = added.find(v);
    if (a != added.end()) d += (int64_t)a->second.size();
    return d;
  }
  bool has_delta(int32_t v) const {
    auto r = removed.find(v);
    if (r != removed.end() && !r->second.empty()) return true;
    auto a = added.find(v);
    return a != added.end() && !a->second.empty();
  }
};

// Flat adjacency snapshot of the logical graph, for the traversals below.
struct Snapshot {
  std::vector<int64_t> off;
  std::vector<int32_t> col;
  explicit Snapshot(const LogicalGraph& g) {
    off.assign((size_t)g.n + 1, 0);
    std::vector<int32_t> row;
    col.reserve(g.col.size());
    for (int32_t v = 0; v < g.n; v++) {
      if (g.has_delta(v)) {
        g.neighbors(v, row);
        col.insert(col.end(), row.begin(), row.end());
      } else {
        col.insert(col.end(), g.col.begin() + g.off[(size_t)v], g.col.begin() + g.off[(size_t)v + 1]);
      }
      off[(size_t)v + 1] = (int64_t)col.size();
    }
  }
  int64_t b(int32_t v) const { return off[(size_t)v]; }
  int64_t e(int32_t v) const { return off[(size_t)v + 1]; }
};

}  // namespace

struct orc_graph {
  LogicalGraph g;
};

extern "C" {

orc_graph* orc_graph_open(const char* path) {
  scout::EgrFile f;
  std::string err = f.open(path);
  if (!err.empty()) {
    fprintf(stderr, "oracle: %s\n", err.c_str());
    return nullptr;
  }
  orc_graph* h = new orc_graph();
  h->g.n = f.nodes;
  f.offsets64(h->g.off);
  h->g.col.assign(f.nlist(), f.nlist() + f.edges);
  return h;
}

orc_graph* orc_graph_from_csr(int32_t n, const int64_t* off, const int32_t* col) {
  orc_graph* h = new orc_graph();
  h->g.n = n;
  h->g.off.assign(off, off + (size_t)n + 1);
  h->g.col.assign(col, col + off[n]);
  return h;
}

void orc_graph_free(orc_graph* h) { delete h; }
int32_t orc_n(const orc_graph* h) { return h->g.n; }
int64_t orc_base_edges(const orc_graph* h) { return (int64_t)h->g.col.size(); }

void orc_classify(orc_graph* h, const int32_t* uv, int64_t m, uint8_t* ins) {
  for (int64_t i = 0; i < m; i++) ins[i] = !h->g.is_adjacent(uv[2 * i], uv[2 * i + 1]);
}

void orc_apply(orc_graph* h, const int32_t* uv, const uint8_t* ins, int64_t m) {
  for (int64_t i = 0; i < m; i++) {
    if (ins[i]) h->g.insert_edge(uv[2 * i], uv[2 * i + 1]);
    else h->g.remove_edge(uv[2 * i], uv[2 * i + 1]);
  }
}

void orc_apply_conditional(orc_graph* h, const int32_t* uv, const uint8_t* ins, int64_t m) {
  for (int64_t i = 0; i < m; i++) {
    int32_t u = uv[2 * i], v = uv[2 * i + 1];
    if (ins[i] && !h->g.is_adjacent(u, v)) h->g.insert_edge(u, v);
    else if (!ins[i] && h->g.is_adjacent(u, v)) h->g.remove_edge(u, v);
  }
}

int orc_is_adjacent(orc_graph* h, int32_t u, int32_t v) { return h->g.is_adjacent(u, v) ? 1 : 0; }
int64_t orc_degree(orc_graph* h, int32_t v) { return h->g.degree(v); }

int64_t orc_row(orc_graph* h, int32_t v, int32_t* out, int64_t cap) {
  std::vector<int32_t> row;
  h->g.neighbors(v, row);
  for (int64_t i = 0; i < (int64_t)row.size() && i < cap; i++) out[i] = row[(size_t)i];
  return (int64_t)row.size();
}

// core/flash_compactor.h:30-95: degree pass (base - |removed| + |added|), then rows.
int64_t orc_logical_csr(orc_graph* h, int64_t* off, int32_t* col) {
  const LogicalGraph& g = h->g;
  off[0] = 0;
  for (int32_t v = 0; v < g.n; v++) {
    int64_t d = g.degree(v);
    if (d < 0) return -1;
    off[v + 1] = off[v] + d;
  }
  if (!col) return off[g.n];
  std::vector<int32_t> row;
  for (int32_t v = 0; v < g.n; v++) {
    g.neighbors(v, row);
    if ((int64_t)row.size() != off[v + 1] - off[v]) return -1;
    std::copy(row.begin(), row.end(), col + off[v]);
  }
  return off[g.n];
}

int orc_compact_to(orc_graph* h, const char* path) {
  std::vector<int64_t> off((size_t)h->g.n + 1);
  int64_t e = orc_logical_csr(h, off.data(), nullptr);
  if (e < 0) return 1;
  std::vector<int32_t> col((size_t)e);
  if (orc_logical_csr(h, off.data(), col.data()) < 0) return 1;
  std::string err = scout::egr_write_wide(path, h->g.n, off.data(), col.data());
  return err.empty() ? 0 : 1;
}

// Textbook BFS distances from src on the logical graph; -1 = unreached.
void orc_bfs(orc_graph* h, int32_t src, int32_t* level) {
  Snapshot s(h->g);
  int32_t n = h->g.n;
  std::fill(level, level + n, -1);
  if (src < 0 || src >= n) return;
  std::vector<int32_t> q;
  q.reserve((size_t)n);
  level[src] = 0;
  q.push_back(src);
  for (size_t head = 0; head < q.size(); head++) {
    int32_t x = q[head];
    for (int64_t i = s.b(x); i < s.e(x); i++) {
      int32_t y = s.col[(size_t)i];
      if (level[y] < 0) {
        level[y] = level[x] + 1;
        q.push_back(y);
      }
    }
  }
}

// Component label = smallest vertex id of the component (ExCC's convention,
// tools/excc_src/excc_v1.cu:193-210). Vertices are scanned in ascending order, so the first
// vertex that starts a search is its component's minimum.
int32_t orc_cc(orc_graph* h, int32_t* label) {
  Snapshot s(h->g);
  int32_t n = h->g.n;
  std::fill(label, label + n, -1);
  std::vector<int32_t> q;
  int32_t comps = 0;
  for (int32_t r = 0; r < n; r++) {
    if (label[r] >= 0) continue;
    comps++;
    label[r] = r;
    q.clear();
    q.push_back(r);
    for (size_t head = 0; head < q.size(); head++) {
      int32_t x = q[head];
      for (int64_t i = s.b(x); i < s.e(x); i++) {
        int32_t y = s.col[(size_t)i];
        if (label[y] < 0) {
          label[y] = r;
          q.push_back(y);
        }
      }
    }
  }
  return comps;
}

// src/coloring/scout_coloring.cu:243-277 with the round of :172-241, minus the 200k-vertex /
// 2e7-entry truncation. Marks: max(u,v) of every insertion line whose endpoints had equal
// pre-batch colours (a set). Each round: propose = smallest c >= 0 not the snapshot colour of
// any neighbour; a round vertex loses iff a lower-id round neighbour proposed the same colour;
// every proposal is applied; losers form the next round.
int64_t orc_color_batch(orc_graph* h, const int32_t* uv, const uint8_t* ins, int64_t m, int32_t* color,
                        int64_t* stats) {
  const LogicalGraph& g = h->g;
  std::vector<uint8_t> marked((size_t)g.n, 0);
  std::vector<int32_t> round;
  int64_t conflicts = 0;
  for (int64_t i = 0; i < m; i++) {
    if (!ins[i]) continue;
    int32_t u = uv[2 * i], v = uv[2 * i + 1];
    if (color[u] != color[v]) continue;
    conflicts++;
    int32_t loser = u > v ? u : v;
    if (!marked[(size_t)loser]) {
      marked[(size_t)loser] = 1;
      round.push_back(loser);
    }
  }
  std::sort(round.begin(), round.end());
  const int32_t UNMARKED = -2;
  std::vector<int32_t> marker((size_t)g.n, UNMARKED);
  std::vector<int32_t> prop;
  std::vector<int32_t> row;
  std::vector<std::vector<int32_t>> rows;
  int64_t rounds = 0;
  while (!round.empty()) {
    rounds++;
    prop.assign(round.size(), 0);
    rows.assign(round.size(), {});
    for (size_t k = 0; k < round.size(); k++) {
      g.neighbors(round[k], rows[k]);
      std::vector<uint8_t> used(rows[k].size() + 2, 0);
      for (int32_t w : rows[k]) {
        int32_t c = color[w];
        if (c >= 0 && (size_t)c < used.size()) used[(size_t)c] = 1;
      }
      int32_t c = 0;
      while (used[(size_t)c]) c++;
      prop[k] = c;
      marker[(size_t)round[k]] = c;
    }
    std::vector<int32_t> next;
    for (size_t k = 0; k < round.size(); k++) {
      for (int32_t w : rows[k]) {
        if (w < round[k] && marker[(size_t)w] == prop[k]) {
          next.push_back(round[k]);
          break;
        }
      }
    }
    for (size_t k = 0; k < round.size(); k++) color[round[k]] = prop[k];
    for (size_t k = 0; k < round.size(); k++) marker[(size_t)round[k]] = UNMARKED;
    round.swap(next);
  }
  if (stats) {
    stats[0] = conflicts;
    stats[1] = rounds;
  }
  return rounds;
}

// MAX_CSR_ELEMENTS of src/mis/scout_mis.cu:800-815.
int64_t orc_mis_max_csr(double avg_degree, double free_bytes) {
  const double MAX_BATCH = 1048576.0;
  long long est = (long long)(MAX_BATCH * 2.0 * avg_degree * avg_degree);
  long long mx = est > 0 ? std::min(800000000LL, est) : 50000000LL;
  mx = std::max(mx, 1000000LL);
  long long fit = (long long)(free_bytes / 4.0 / 8.0);
  if (fit > 1000000LL && mx > fit) mx = fit;
  return mx;
}

// src/mis/scout_mis.cu: the per-edge rules of process_clusters_csr_kernel (:290-362) applied
// sequentially in ascending batch-line order -- the race-free serialisation of the
// per-cluster loop -- on rows built exactly as launchBatch builds them (:500-626): the
// sub-batch is cut by the degree*1000 estimate against MAX_CSR_ELEMENTS (:519-542); every
// endpoint's row is its 2-hop ball in discovery order (get_h_hop_neighborhood, :173-215);
// every other vertex at distance 1 from an endpoint has its adjacency row; any other vertex
// has no row. `ins` is the parse-time label of each line. Returns 0, or 1 when a sub-batch
// would exceed the reference's 2^21-row pool (whose overflow rows the reference drops in
// hash order, which no restatement can reproduce).
int orc_mis_batch(orc_graph* h, const int32_t* uv, const uint8_t* ins, int64_t m, int64_t max_csr, uint8_t* mis,
                  int64_t* stats) {
  const LogicalGraph& g = h->g;
  const int64_t MAX_BATCH = 1048576;
  const size_t MAX_ROWS = (size_t)MAX_BATCH * 2;
  int64_t subs = 0, rows_total = 0, writes = 0;
  int64_t idx = 0;
  std::vector<uint8_t> visited((size_t)g.n, 0);
  std::vector<int32_t> scratch;
  while (idx < m) {
    // Sub-batch admission (:519-542).
    std::unordered_set<int32_t> uniq;
    int64_t cur = 0, se = 0, start = idx;
    while (idx < m && se < MAX_BATCH) {
      int32_t u = uv[2 * idx], v = uv[2 * idx + 1];
      int64_t est = 0;
      if (!uniq.count(u)) est += g.degree(u) * 1000;
      if (!uniq.count(v)) est += g.degree(v) * 1000;
      if (se > 0 && cur + est >= max_csr) break;
      cur += est;
      uniq.insert(u);
      uniq.insert(v);
      se++;
      idx++;
    }
    subs++;
    // Rows: endpoint -> 2-hop ball in discovery order; hop-1 vertices -> adjacency row.
    std::unordered_map<int32_t, std::vector<int32_t>> rows;
    std::unordered_map<int32_t, std::vector<int32_t>> hop1;
    for (int32_t e : uniq) {
      std::vector<int32_t> ball{e}, frontier{e}, next, tracker{e};
      visited[(size_t)e] = 1;
      for (int hop = 0; hop < 2; hop++) {
        next.clear();
        for (int32_t c : frontier) {
          g.neighbors(c, scratch);
          if (hop == 1 && !hop1.count(c)) hop1[c] = scratch;
          for (int32_t w : scratch) {
            if (!visited[(size_t)w]) {
              visited[(size_t)w] = 1;
              tracker.push_back(w);
              next.push_back(w);
              ball.push_back(w);
            }
          }
        }
        frontier.swap(next);
        if (frontier.empty()) break;
      }
      for (int32_t t : tracker) visited[(size_t)t] = 0;
      rows[e] = std::move(ball);
    }
    for (auto& kv : hop1)
      if (!uniq.count(kv.first)) rows.emplace(kv.first, std::move(kv.second));
    rows_total += (int64_t)rows.size();
    if (rows.size() > MAX_ROWS) return 1;
    auto row_of = [&](int32_t x) -> const std::vector<int32_t>* {
      auto it = rows.find(x);
      return it == rows.end() ? nullptr : &it->second;
    };
    for (int64_t k = start; k < idx; k++) {
      int32_t u = uv[2 * k], v = uv[2 * k + 1];
      bool um = mis[u] != 0, vm = mis[v] != 0;
      if (ins[k]) {
        if (um && vm) {
          int32_t r = u > v ? u : v;
          mis[r] = 0;
          writes++;
          const std::vector<int32_t>* rr = row_of(r);
          if (rr) {
            for (int32_t c : *rr) {
              if (mis[c]) continue;
              const std::vector<int32_t>* rc = row_of(c);
              if (!rc) continue;
              bool can = true;
              for (int32_t w : *rc)
                if (mis[w]) {
                  can = false;
                  break;
                }
              if (can) {
                mis[c] = 1;
                writes++;
                break;
              }
            }
          }
        }
      } else {
        int32_t keep = -1, cand = -1;
        if (um && !vm) keep = u, cand = v;
        else if (!um && vm) keep = v, cand = u;
        if (cand >= 0) {
          bool can = true;
          const std::vector<int32_t>* rc = row_of(cand);
          if (rc)
            for (int32_t w : *rc)
              if (w != keep && mis[w]) {
                can = false;
                break;
              }
          if (can) {
            mis[cand] = 1;
            writes++;
          }
        }
      }
    }
  }
  if (stats) {
    stats[0] = subs;
    stats[1] = rows_total;
    stats[2] = writes;
  }
  return 0;
}

// BFS levels are exact iff: level[src] = 0; every edge joins two unreached vertices or two
// reached ones whose levels differ by at most 1; every reached vertex other than src has a
// neighbour one level below; and src is the only vertex at level 0.
int64_t orc_check_bfs(orc_graph* h, int32_t src, const int32_t* level) {
  Snapshot s(h->g);
  int32_t n = h->g.n;
  int64_t bad = 0;
  if (src >= 0 && src < n && level[src] != 0) bad++;
  for (int32_t v = 0; v < n; v++) {
    int32_t lv = level[v];
    if (lv < -1) bad++;
    if (lv == 0 && v != src) bad++;
    bool parent = false;
    for (int64_t i = s.b(v); i < s.e(v); i++) {
      int32_t w = s.col[(size_t)i];
      int32_t lw = level[w];
      if ((lv < 0) != (lw < 0)) bad++;
      else if (lv >= 0 && (lw > lv + 1 || lw < lv - 1)) bad++;
      if (lv > 0 && lw == lv - 1) parent = true;
    }
    if (lv > 0 && !parent) bad++;
  }
  return bad;
}

// Exact min-id labels: equal to the closed form.
int64_t orc_check_cc(orc_graph* h, const int32_t* label) {
  std::vector<int32_t> truth((size_t)h->g.n);
  orc_cc(h, truth.data());
  int64_t bad = 0;
  for (int32_t v = 0; v < h->g.n; v++) bad += (truth[(size_t)v] != label[v]);
  return bad;
}

int64_t orc_check_coloring(orc_graph* h, const int32_t* color) {
  Snapshot s(h->g);
  int64_t bad = 0;
  for (int32_t v = 0; v < h->g.n; v++)
    for (int64_t i = s.b(v); i < s.e(v); i++)
      if (color[s.col[(size_t)i]] == color[v]) bad++;
  return bad;
}

int64_t orc_check_mis_independent(orc_graph* h, const uint8_t* mis) {
  Snapshot s(h->g);
  int64_t bad = 0;
  for (int32_t v = 0; v < h->g.n; v++)
    if (mis[v])
      for (int64_t i = s.b(v); i < s.e(v); i++)
        if (mis[s.col[(size_t)i]]) bad++;
  return bad;
}

// The completion of the MIS "complete" mode: after the batch's rules, every vertex within two
// hops of an endpoint of the batch (on the graph after the batch), in ascending id, joins the
// set unless it or a neighbour is already a member. Only vertices near the batch can have
// lost their last member neighbour (rule writes and deleted edges touch endpoints and their
// neighbours), so a maximal set before the batch is maximal again after it. Restated directly
// from that definition (a sequential scan); returns the vertices added.
int64_t orc_mis_complete(orc_graph* h, const int32_t* uv, int64_t m, uint8_t* mis) {
  const LogicalGraph& g = h->g;
  std::vector<uint8_t> in((size_t)g.n, 0);
  std::vector<int32_t> region, row;
  for (int64_t i = 0; i < 2 * m; i++)
    if (!in[(size_t)uv[i]]) {
      in[(size_t)uv[i]] = 1;
      region.push_back(uv[i]);
    }
  size_t begin = 0;
  for (int hop = 0; hop < 2; hop++) {  // expand the newest layer each hop
    size_t end = region.size();
    for (size_t k = begin; k < end; k++) {
      g.neighbors(region[k], row);
      for (int32_t w : row)
        if (!in[(size_t)w]) {
          in[(size_t)w] = 1;
          region.push_back(w);
        }
    }
    begin = end;
  }
  std::sort(region.begin(), region.end());
  int64_t added = 0;
  for (int32_t v : region) {
    if (mis[v]) continue;
    g.neighbors(v, row);
    bool dominated = false;
    for (int32_t w : row)
      if (mis[w]) {
        dominated = true;
        break;
      }
    if (!dominated) {
      mis[v] = 1;
      added++;
    }
  }
  return added;
}

int64_t orc_mis_maximality_violations(orc_graph* h, const uint8_t* mis) {
  Snapshot s(h->g);
  int64_t bad = 0;
  for (int32_t v = 0; v < h->g.n; v++) {
    if (mis[v]) continue;
    bool dominated = false;
    for (int64_t i = s.b(v); i < s.e(v) && !dominated; i++) dominated = mis[s.col[(size_t)i]] != 0;
    if (!dominated) bad++;
  }
  return bad;
}

// tools/compute_initial_mis.py:32-43: scan ids ascending; take v unless a neighbour is taken.
void orc_greedy_mis(orc_graph* h, uint8_t* mis) {
  Snapshot s(h->g);
  std::fill(mis, mis + h->g.n, 0);
  std::vector<uint8_t> blocked((size_t)h->g.n, 0);
  for (int32_t v = 0; v < h->g.n; v++) {
    if (blocked[(size_t)v]) continue;
    mis[v] = 1;
    for (int64_t i = s.b(v); i < s.e(v); i++) blocked[(size_t)s.col[(size_t)i]] = 1;
  }
}

// tools/greedy_color.cpp:37-52: ascending ids, smallest colour unused by coloured neighbours.
void orc_greedy_color(orc_graph* h, int32_t* color) {
  Snapshot s(h->g);
  int32_t n = h->g.n;
  std::fill(color, color + n, -1);
  std::vector<int32_t> stamp;
  for (int32_t v = 0; v < n; v++) {
    size_t d = (size_t)(s.e(v) - s.b(v));
    if (stamp.size() < d + 2) stamp.assign(d + 2, -1);
    for (int64_t i = s.b(v); i < s.e(v); i++) {
      int32_t c = color[s.col[(size_t)i]];
      if (c >= 0 && (size_t)c < stamp.size()) stamp[(size_t)c] = v;
    }
    int32_t c = 0;
    while (stamp[(size_t)c] == v) c++;
    color[v] = c;
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// The reference's per-edit BFS and CC algorithms on the CPU (the CPU baseline).
// ---------------------------------------------------------------------------------------
namespace {
const int32_t MAX_ROUND = 200000;  // MAX_ROUND_FRONTIER, src/bfs/scout_bfs.cu:40

// core/chunking.h:20-31
int64_t bound_prefix(const LogicalGraph& g, const std::vector<int32_t>& c, int64_t cap) {
  int64_t k = 0, sum = 0;
  for (int32_t v : c) {
    if (k >= cap) break;
    int64_t d = g.degree(v);
    if (sum + d > cap) break;
    sum += d;
    k++;
  }
  return k;
}
}  // namespace

struct orc_refbfs {
  LogicalGraph* g;
  std::vector<int32_t> level;
  int32_t source;
  int64_t bound;
  int defer;
  int64_t deferred = 0;
  std::vector<int32_t> row;
  std::vector<std::vector<int32_t>> rows;

  // The round's row fetch: OpenMP threads gather the rows, as the reference's host fetch does
  // before each GPU round (src/bfs/scout_bfs.cu:199,347); small rounds stay on one thread.
  void fetch(const std::vector<int32_t>& vs, int64_t k) {
    if ((int64_t)rows.size() < k) rows.resize((size_t)k);
#pragma omp parallel for schedule(dynamic, 64) if (k >= 256)
    for (int64_t i = 0; i < k; i++) g->neighbors(vs[(size_t)i], rows[(size_t)i]);
  }

  // gpu_relax_round (src/bfs/scout_bfs.cu:177-272) on the CPU: admitted prefix only.
  std::vector<int32_t> relax_round(const std::vector<int32_t>& frontier) {
    int64_t k = bound_prefix(*g, frontier, MAX_ROUND);
    std::vector<int32_t> next;
    std::vector<int32_t> lv((size_t)k);
    for (int64_t i = 0; i < k; i++) lv[(size_t)i] = level[(size_t)frontier[(size_t)i]];
    fetch(frontier, k);
    for (int64_t i = 0; i < k; i++) {
      const std::vector<int32_t>& row = rows[(size_t)i];
      int32_t cand = lv[(size_t)i] + 1;
      for (int32_t u : row) {
        int32_t old = level[(size_t)u];
        if (old == -1 || old > cand) {
          level[(size_t)u] = cand;
          if ((int64_t)next.size() < MAX_ROUND) next.push_back(u);
        }
      }
    }
    return next;
  }
  // gpu_deletion_round (:325-415): Jacobi recompute of the admitted worklist prefix.
  std::vector<int32_t> deletion_round(const std::vector<int32_t>& work) {
    int64_t k = bound_prefix(*g, work, MAX_ROUND);
    std::vector<int32_t> next, nl((size_t)k);
    fetch(work, k);
    for (int64_t i = 0; i < k; i++) {
      int32_t w = work[(size_t)i];
      int32_t old = level[(size_t)w];
      const std::vector<int32_t>& row = rows[(size_t)i];
      int32_t best = -1;
      for (int32_t x : row) {
        int32_t l = level[(size_t)x];
        if (l != -1 && (best == -1 || l < best)) best = l;
      }
      int32_t nv = best == -1 ? -1 : best + 1;
      nl[(size_t)i] = nv;
      if (nv != old)
        for (int32_t x : row) {
          int32_t l = level[(size_t)x];
          if (l != -1 && old != -1 && l == old + 1 && (int64_t)next.size() < MAX_ROUND) next.push_back(x);
        }
    }
    for (int64_t i = 0; i < k; i++) level[(size_t)work[(size_t)i]] = nl[(size_t)i];
    return next;
  }
  // gpu_full_bfs (:274-323): chunked level-synchronous rebuild; a vertex whose degree
  // exceeds the round cap is skipped, as in the reference.
  void full_bfs() {
    std::fill(level.begin(), level.end(), -1);
    level[(size_t)source] = 0;
    std::vector<int32_t> frontier{source};
    while (!frontier.empty()) {
      std::vector<int32_t> nextf;
      size_t i = 0;
      while (i < frontier.size()) {
        std::vector<int32_t> chunk;
        int64_t sum = 0;
        while (i < frontier.size()) {
          int64_t d = g->degree(frontier[i]);
          if (!chunk.empty() && (sum + d > MAX_ROUND || chunk.size() >= (size_t)MAX_ROUND)) break;
          chunk.push_back(frontier[i]);
          sum += d;
          i++;
        }
        std::vector<int32_t> nx = relax_round(chunk);
        nextf.insert(nextf.end(), nx.begin(), nx.end());
      }
      frontier.swap(nextf);
    }
  }
};

struct orc_refcc {
  LogicalGraph* g;
  std::vector<int32_t> label, lp;
  int64_t bound;
  int defer;
  int64_t deferred = 0;
  std::vector<int32_t> row;

  static int32_t find_root(std::vector<int32_t>& p, int32_t v) {
    int32_t r = v;
    while (p[(size_t)r] != r) r = p[(size_t)r];
    while (p[(size_t)v] != r) {
      int32_t nx = p[(size_t)v];
      p[(size_t)v] = r;
      v = nx;
    }
    return r;
  }
  int32_t find(int32_t v) { return find_root(lp, label[(size_t)v]); }
  bool unite(int32_t u, int32_t v) {
    int32_t ru = find(u), rv = find(v);
    if (ru == rv) return false;
    if (ru < rv) lp[(size_t)rv] = ru;
    else lp[(size_t)ru] = rv;
    return true;
  }
  void reset_from_roots(const std::vector<int32_t>& root) {
    label = root;
    lp.resize(label.size());
    for (size_t i = 0; i < lp.size(); i++) lp[i] = (int32_t)i;
  }
  // bidirectional_reachability_check (src/cc/scout_cc.cu:300-342).
  int search(int32_t u, int32_t v, std::vector<int32_t>& side) {
    side.clear();
    if (u == v) return 0;
    std::unordered_set<int32_t> vu{u}, vv{v};
    std::vector<int32_t> fu{u}, fv{v};
    size_t touched = 2;
    while (!fu.empty() && !fv.empty()) {
      bool eu = fu.size() <= fv.size();
      std::vector<int32_t>& fr = eu ? fu : fv;
      std::unordered_set<int32_t>& mine = eu ? vu : vv;
      std::unordered_set<int32_t>& other = eu ? vv : vu;
      std::vector<int32_t> next;
      for (int32_t c : fr) {
        g->neighbors(c, row);
        for (int32_t w : row) {
          if (other.count(w)) return 0;
          if (mine.count(w)) continue;
          mine.insert(w);
          next.push_back(w);
          touched++;
          if ((int64_t)touched > bound) return 2;
        }
      }
      fr.swap(next);
    }
    if (fu.empty()) side.assign(vu.begin(), vu.end());
    else side.assign(vv.begin(), vv.end());
    return 1;
  }
  // The full recompute on the host, as full_recompute_cc_streamed does when the graph does not
  // fit the device (src/cc/scout_cc.cu:76-116): OpenMP threads hook roots by relaxed CAS, the
  // lower root winning, so every root ends as its component's smallest id.
  static int32_t find_atomic(int32_t* p, int32_t x) {
    while (true) {
      int32_t q = __atomic_load_n(&p[x], __ATOMIC_RELAXED);
      if (q == x) return x;
      int32_t r = __atomic_load_n(&p[q], __ATOMIC_RELAXED);
      if (r != q) __atomic_compare_exchange_n(&p[x], &q, r, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
      x = r;  // path halving
    }
  }
  void recompute() {
    std::vector<int32_t> roots((size_t)g->n);
    int32_t* p = roots.data();
    for (int32_t v = 0; v < g->n; v++) p[v] = v;
#pragma omp parallel
    {
      std::vector<int32_t> nb;
#pragma omp for schedule(dynamic, 4096)
      for (int32_t v = 0; v < g->n; v++) {
        g->neighbors(v, nb);
        for (int32_t w : nb) {
          if (w <= v) continue;
          int32_t a = v, b = w;
          while (true) {
            a = find_atomic(p, a);
            b = find_atomic(p, b);
            if (a == b) break;
            if (a < b) std::swap(a, b);  // hook the larger root under the smaller
            int32_t expect = a;
            if (__atomic_compare_exchange_n(&p[a], &expect, b, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) break;
          }
        }
      }
#pragma omp for schedule(static)
      for (int32_t v = 0; v < g->n; v++) p[v] = find_atomic(p, v);
    }
    reset_from_roots(roots);
  }
};

extern "C" {

orc_refbfs* orc_refbfs_create(orc_graph* h, const int32_t* level, int32_t source, int64_t bound, int defer) {
  orc_refbfs* r = new orc_refbfs();
  r->g = &h->g;
  r->level.assign(level, level + h->g.n);
  r->source = source;
  r->bound = bound;
  r->defer = defer;
  return r;
}
void orc_refbfs_free(orc_refbfs* r) { delete r; }

// src/bfs/scout_bfs.cu:436-515. stats: insertions, ins_noop, ins_repaired, deletions,
// del_noop, del_local, del_fallback, deferred.
void orc_refbfs_batch(orc_refbfs* r, const int32_t* uv, const uint8_t* ins, int64_t m, int64_t* st) {
  std::vector<int32_t>& L = r->level;
  for (int64_t i = 0; i < m; i++) {
    int32_t u = uv[2 * i], v = uv[2 * i + 1];
    if (ins[i]) {
      r->g->insert_edge(u, v);
      st[0]++;
      bool iv = L[(size_t)u] != -1 && (L[(size_t)v] == -1 || L[(size_t)u] + 1 < L[(size_t)v]);
      bool iu = L[(size_t)v] != -1 && (L[(size_t)u] == -1 || L[(size_t)v] + 1 < L[(size_t)u]);
      if (!iu && !iv) {
        st[1]++;
        continue;
      }
      std::vector<int32_t> fr;
      if (iv) {
        L[(size_t)v] = L[(size_t)u] + 1;
        fr.push_back(v);
      }
      if (iu) {
        L[(size_t)u] = L[(size_t)v] + 1;
        fr.push_back(u);
      }
      while (!fr.empty()) fr = r->relax_round(fr);
      st[2]++;
    } else {
      r->g->remove_edge(u, v);
      st[3]++;
      bool tree = (L[(size_t)u] != -1 && L[(size_t)v] == L[(size_t)u] + 1) ||
                  (L[(size_t)v] != -1 && L[(size_t)u] == L[(size_t)v] + 1);
      if (!tree) {
        st[4]++;
        continue;
      }
      int32_t child = L[(size_t)u] > L[(size_t)v] ? u : v;
      std::vector<int32_t> work{child};
      size_t touched = 0;
      bool fb = false;
      while (!work.empty()) {
        touched += work.size();
        if ((int64_t)touched > r->bound) {
          fb = true;
          if (r->defer) {
            st[7]++;
            r->deferred++;
          } else {
            r->full_bfs();
          }
          break;
        }
        work = r->deletion_round(work);
      }
      if (fb) st[6]++;
      else st[5]++;
    }
  }
}

int orc_threads(void) { return omp_get_max_threads(); }

void orc_refbfs_finish(orc_refbfs* r, int32_t* level) {
  if (r->defer && r->deferred > 0) {
    r->full_bfs();
    r->deferred = 0;
  }
  std::copy(r->level.begin(), r->level.end(), level);
}

orc_refcc* orc_refcc_create(orc_graph* h, const int32_t* root, int64_t bound, int defer) {
  orc_refcc* r = new orc_refcc();
  r->g = &h->g;
  std::vector<int32_t> roots(root, root + h->g.n);
  r->reset_from_roots(roots);
  r->bound = bound;
  r->defer = defer;
  return r;
}
void orc_refcc_free(orc_refcc* r) { delete r; }

// src/cc/scout_cc.cu:392-441. stats: insertions, ins_noop, ins_merged, deletions, del_noop,
// still_connected, split, inconclusive.
void orc_refcc_batch(orc_refcc* r, const int32_t* uv, const uint8_t* ins, int64_t m, int64_t* st) {
  std::vector<int32_t> side;
  for (int64_t i = 0; i < m; i++) {
    int32_t u = uv[2 * i], v = uv[2 * i + 1];
    if (ins[i]) {
      r->g->insert_edge(u, v);
      st[0]++;
      if (r->unite(u, v)) st[2]++;
      else st[1]++;
    } else {
      r->g->remove_edge(u, v);
      st[3]++;
      if (r->find(u) != r->find(v)) {
        st[4]++;
        continue;
      }
      int res = r->search(u, v, side);
      if (res == 0) st[5]++;
      else if (res == 1) {
        st[6]++;
        int32_t fresh = (int32_t)r->lp.size();
        r->lp.push_back(fresh);
        for (int32_t w : side) r->label[(size_t)w] = fresh;
      } else {
        st[7]++;
        if (r->defer) r->deferred++;
        else r->recompute();
      }
    }
  }
}

// Settles deferred work, then canonicalises labels to the minimum vertex id per class.
void orc_refcc_finish(orc_refcc* r, int32_t* label) {
  if (r->defer && r->deferred > 0) {
    r->recompute();
    r->deferred = 0;
  }
  int32_t n = r->g->n;
  std::vector<int32_t> mn(r->lp.size(), INT32_MAX);
  for (int32_t v = 0; v < n; v++) {
    int32_t c = r->find(v);
    if (v < mn[(size_t)c]) mn[(size_t)c] = v;
  }
  for (int32_t v = 0; v < n; v++) label[v] = mn[(size_t)r->find(v)];
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// MLP of the learned scheduler (double precision). Pinned by central finite differences of
// its own loss and by an independent numpy forward pass (tests/test_oracle.py).
// ---------------------------------------------------------------------------------------
namespace {
constexpr int kIn = 12, kH = 64, kOut = 3;
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kOut, kParams = oB3 + kOut;

void mlp_fwd(const float* p, const float* x, double* z1, double* h1, double* z2, double* h2, double* out) {
  for (int j = 0; j < kH; j++) {
    double a = p[oB1 + j];
    for (int i = 0; i < kIn; i++) a += (double)x[i] * p[oW1 + i * kH + j];
    z1[j] = a;
    h1[j] = a > 0 ? a : 0;
  }
  for (int j = 0; j < kH; j++) {
    double a = p[oB2 + j];
    for (int i = 0; i < kH; i++) a += h1[i] * p[oW2 + i * kH + j];
    z2[j] = a;
    h2[j] = a > 0 ? a : 0;
  }
  for (int k = 0; k < kOut; k++) {
    double a = p[oB3 + k];
    for (int i = 0; i < kH; i++) a += h2[i] * p[oW3 + i * kOut + k];
    out[k] = a;
  }
}
}  // namespace

void orc_mlp_forward(const float* params, const float* feats, int64_t n, double* out) {
  double z1[kH], h1[kH], z2[kH], h2[kH];
  for (int64_t s = 0; s < n; s++) mlp_fwd(params, feats + s * kIn, z1, h1, z2, h2, out + s * kOut);
}

double orc_mlp_loss_grad(const float* params, const float* feats, const float* targets, int64_t n, double* grad) {
  if (grad) std::fill(grad, grad + kParams, 0.0);
  double z1[kH], h1[kH], z2[kH], h2[kH], out[kOut], d3[kOut], d2[kH], d1[kH];
  double total = 0;
  for (int64_t s = 0; s < n; s++) {
    const float* x = feats + s * kIn;
    const float* t = targets + s * kOut;
    mlp_fwd(params, x, z1, h1, z2, h2, out);
    double z = out[0];
    total += std::max(z, 0.0) - z * t[0] + std::log1p(std::exp(-std::fabs(z)));
    double e1 = out[1] - t[1], e2 = out[2] - t[2];
    total += 0.5 * (e1 * e1 + e2 * e2);
    if (!grad) continue;
    d3[0] = (1.0 / (1.0 + std::exp(-z)) - t[0]) / (double)n;
    d3[1] = e1 / (double)n;
    d3[2] = e2 / (double)n;
    for (int i = 0; i < kH; i++) {
      double a = 0;
      for (int k = 0; k < kOut; k++) a += (double)params[oW3 + i * kOut + k] * d3[k];
      d2[i] = z2[i] > 0 ? a : 0;
    }
    for (int i = 0; i < kH; i++) {
      double a = 0;
      for (int j = 0; j < kH; j++) a += (double)params[oW2 + i * kH + j] * d2[j];
      d1[i] = z1[i] > 0 ? a : 0;
    }
    for (int i = 0; i < kIn; i++)
      for (int j = 0; j < kH; j++) grad[oW1 + i * kH + j] += (double)x[i] * d1[j];
    for (int j = 0; j < kH; j++) grad[oB1 + j] += d1[j];
    for (int i = 0; i < kH; i++)
      for (int j = 0; j < kH; j++) grad[oW2 + i * kH + j] += h1[i] * d2[j];
    for (int j = 0; j < kH; j++) grad[oB2 + j] += d2[j];
    for (int i = 0; i < kH; i++)
      for (int k = 0; k < kOut; k++) grad[oW3 + i * kOut + k] += h2[i] * d3[k];
    for (int k = 0; k < kOut; k++) grad[oB3 + k] += d3[k];
  }
  return n > 0 ? total / (double)n : 0.0;
}

float orc_bf16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return x;  // NaN
  uint32_t lsb = (u >> 16) & 1u;
  u = (u + 0x7fffu + lsb) & 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

