// scout_multi: many small graphs maintained together on one GPU (MIS + colouring each).
//
// Instances are Barabasi-Albert graphs (m edges per new vertex) with vertex counts drawn
// log-uniformly from [min-n, max-n]. They are laid out as one disjoint union in HBM: instance i
// owns vertex ids [base_i, base_i + n_i). Every batch gives every instance its own toggles
// (a log-uniform count in [min-lines, max-lines], half deletions of uniform edges, half
// insertions of uniform non-edges, no pair toggled twice) and one step repairs all of them:
//   stage the batch (instance after instance) -> classify -> apply
//   -> scout_repair_mis_multi (per-instance sub-batch cuts, all instances' rounds in the same
//      launches, round groups replayed from a captured hipGraph)
//   -> scout_repair_color (the union's colouring repair is each instance's, since no edge
//      crosses instances and the lower-id rule keeps its order under the id offset)
// The colouring and MIS launch settings come from scout_color_autotune and scout_mis_autotune
// on the first batch.
//
// Checks: the same stream served instance by instance, each instance alone in its own steps,
// must give identical MIS bits and colours ("alone == in the mix"); the final colouring must be
// proper and the MIS independent on the final graph.
//
//   scout_multi [--instances 64] [--min-n 1e4] [--max-n 2e6] [--m 8] [--batches 10]
//               [--min-lines 1000] [--max-lines 100000] [--seed 1] [--no-alone] [--no-tune]
//               [--no-graph] [--gpu 0]
//
// Prints one JSON line. The reference maintains one graph per process; this driver is new.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../../include/scout_hip.h"
#include "../host/gen.h"

namespace {

[[noreturn]] void die(const char* what) {
  const char* e = scout_last_error();
  fprintf(stderr, "ERROR: %s (%s)\n", what, e ? e : "");
  fflush(stdout);
  std::_Exit(1);
}

#define CHECK(call)                     \
  do {                                  \
    if ((call) != SCOUT_OK) die(#call); \
  } while (0)

double unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

double log_uniform(uint64_t x, double lo, double hi) { return lo * std::exp(unit(x) * std::log(hi / lo)); }

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)std::ceil(p / 100.0 * (double)v.size());
  if (k == 0) k = 1;
  return v[std::min(k, v.size()) - 1];
}

bool has_edge(const scout::Csr& g, int32_t u, int32_t v) {
  auto b = g.col.begin() + g.off[(size_t)u], e = g.col.begin() + g.off[(size_t)u + 1];
  return std::binary_search(b, e, v);
}

// Toggles of one instance for one batch: `size` lines, half deletions of uniform initial
// edges, half insertions of uniform initial non-edges, never a pair toggled before. Vertex ids
// are local; the caller adds the instance base.
void instance_lines(const scout::Csr& g, int64_t size, uint64_t key, uint64_t& ctr,
                    std::unordered_set<uint64_t>& toggled, std::vector<int32_t>& out) {
  int64_t ndel = size / 2, nins = size - ndel;
  int64_t entries = g.entries();
  int32_t n = g.n;
  auto pair_key = [](uint32_t u, uint32_t v) { return ((uint64_t)std::min(u, v) << 32) | std::max(u, v); };
  for (int64_t k = 0, guard = 0; k < ndel && entries > 0 && guard < 100 * size; guard++) {
    uint64_t x = scout::draw(key, ctr++);
    int64_t e = (int64_t)(unit(x) * (double)entries);
    int32_t u = (int32_t)(std::upper_bound(g.off.begin(), g.off.end(), e) - g.off.begin()) - 1;
    int32_t v = g.col[(size_t)e];
    if (!toggled.insert(pair_key((uint32_t)u, (uint32_t)v)).second) continue;
    out.push_back(std::min(u, v));
    out.push_back(std::max(u, v));
    k++;
  }
  for (int64_t k = 0, guard = 0; k < nins && guard < 100 * size; guard++) {
    uint32_t u = scout::uniform_below(scout::draw(key, ctr++), (uint32_t)n);
    uint32_t v = scout::uniform_below(scout::draw(key, ctr++), (uint32_t)n);
    if (u == v || has_edge(g, (int32_t)u, (int32_t)v)) continue;
    if (!toggled.insert(pair_key(u, v)).second) continue;
    out.push_back((int32_t)std::min(u, v));
    out.push_back((int32_t)std::max(u, v));
    k++;
  }
}

}  // namespace

int main(int argc, char** argv) {
  int ninst = 64, mba = 8, batches = 10, gpu = 0;
  double min_n = 1e4, max_n = 2e6, min_lines = 1000, max_lines = 100000;
  uint64_t seed = 1;
  bool alone = true, tune = true, graphs = true;
  for (int i = 1; i < argc; i++) {
    std::string k = argv[i];
    if (k == "--no-alone") {
      alone = false;
      continue;
    }
    if (k == "--no-tune") {
      tune = false;
      continue;
    }
    if (k == "--no-graph") {
      graphs = false;
      continue;
    }
    if (i + 1 >= argc) k = "?";
    const char* v = i + 1 < argc ? argv[++i] : "";
    if (k == "--instances") ninst = atoi(v);
    else if (k == "--min-n") min_n = atof(v);
    else if (k == "--max-n") max_n = atof(v);
    else if (k == "--m") mba = atoi(v);
    else if (k == "--batches") batches = atoi(v);
    else if (k == "--min-lines") min_lines = atof(v);
    else if (k == "--max-lines") max_lines = atof(v);
    else if (k == "--seed") seed = strtoull(v, nullptr, 10);
    else if (k == "--gpu") gpu = atoi(v);
    else {
      fprintf(stderr, "usage: scout_multi [--instances N] [--min-n A] [--max-n B] [--m M] [--batches K] "
                      "[--min-lines L] [--max-lines L] [--seed S] [--no-alone] [--no-tune] [--no-graph] [--gpu G]\n");
      return 1;
    }
  }
  if (ninst < 1 || mba < 1 || min_n < mba + 2 || max_n < min_n || batches < 1 || min_lines < 2 ||
      max_lines < min_lines) {
    fprintf(stderr, "ERROR: need instances >= 1, m >= 1, m + 2 <= min-n <= max-n, batches >= 1, "
                    "2 <= min-lines <= max-lines\n");
    return 1;
  }
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
  auto t_gen = now();
  uint64_t key = scout::stream_key(seed, scout::RNG_MULTI);
  std::vector<scout::Csr> inst((size_t)ninst);
  std::vector<int32_t> base((size_t)ninst + 1, 0);
  for (int i = 0; i < ninst; i++) {
    int64_t n = std::llround(log_uniform(scout::draw(key, (uint64_t)i), min_n, max_n));
    base[(size_t)i + 1] = base[(size_t)i] + (int32_t)n;
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < ninst; i++)
    inst[(size_t)i] = scout::gen_ba(base[(size_t)i + 1] - base[(size_t)i], mba, scout::draw(key, 1000000 + i));
  int32_t n = base[(size_t)ninst];
  std::vector<int64_t> eoff((size_t)ninst + 1, 0);
  for (int i = 0; i < ninst; i++) eoff[(size_t)i + 1] = eoff[(size_t)i] + inst[(size_t)i].entries();
  int64_t entries = eoff[(size_t)ninst];
  std::vector<int64_t> off((size_t)n + 1);
  std::vector<int32_t> col((size_t)entries);
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < ninst; i++) {
    const scout::Csr& g = inst[(size_t)i];
    int32_t b = base[(size_t)i];
    for (int32_t v = 0; v < g.n; v++) off[(size_t)(b + v)] = eoff[(size_t)i] + g.off[(size_t)v];
    for (int64_t k = 0; k < g.entries(); k++) col[(size_t)(eoff[(size_t)i] + k)] = g.col[(size_t)k] + b;
  }
  off[(size_t)n] = entries;
  // the stream: per batch, per instance
  std::vector<std::vector<std::vector<int32_t>>> lines((size_t)batches, std::vector<std::vector<int32_t>>((size_t)ninst));
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < ninst; i++) {
    uint64_t ikey = scout::stream_key(scout::draw(key, 2000000 + i), scout::RNG_STREAM_DEL);
    uint64_t ctr = 0;
    std::unordered_set<uint64_t> toggled;
    for (int b = 0; b < batches; b++) {
      int64_t size = std::llround(log_uniform(scout::draw(key, 3000000 + (uint64_t)b * 65536 + i), min_lines, max_lines));
      auto& out = lines[(size_t)b][(size_t)i];
      instance_lines(inst[(size_t)i], size, ikey, ctr, toggled, out);
      for (int32_t& x : out) x += base[(size_t)i];
    }
  }
  double gen_s = secs(t_gen, now());
  int64_t total_lines = 0;
  std::vector<std::vector<int32_t>> batch_uv((size_t)batches);
  std::vector<std::vector<int64_t>> batch_off((size_t)batches);
  for (int b = 0; b < batches; b++) {
    auto& uv = batch_uv[(size_t)b];
    auto& bo = batch_off[(size_t)b];
    bo.push_back(0);
    for (int i = 0; i < ninst; i++) {
      auto& l = lines[(size_t)b][(size_t)i];
      uv.insert(uv.end(), l.begin(), l.end());
      bo.push_back((int64_t)uv.size() / 2);
    }
    total_lines += bo.back();
  }

  // device: union graph, greedy initial states, MAX_CSR_ELEMENTS as the MIS driver sets it
  scout_ctx* ctx = nullptr;
  CHECK(scout_ctx_create(gpu, nullptr, &ctx));
  CHECK(scout_graph_load(ctx, n, entries, off.data(), 1, col.data()));
  CHECK(scout_mis_greedy(ctx));
  CHECK(scout_color_greedy(ctx));
  double avg_degree = (double)entries / (double)n;
  int64_t free_b = 0, total_b = 0;
  CHECK(scout_mem_info(ctx, &free_b, &total_b));
  long long mx = std::max(1000000LL, std::min(800000000LL, (long long)(1048576.0 * 2.0 * avg_degree * avg_degree)));
  long long fit = (long long)((double)free_b / 4.0 / 8.0);
  if (fit > 1000000LL && mx > fit) mx = fit;
  int64_t max_csr = mx;
  CHECK(scout_snapshot(ctx));
  if (!graphs) {
    // the same kernels launched one by one instead of replayed from captured graphs
    int cur[4];
    CHECK(scout_color_get_config(ctx, cur));
    CHECK(scout_color_config(ctx, cur[0], cur[1], cur[2], 0));
    CHECK(scout_mis_config(ctx, 1024, 8, 0));
  }

  int cfg[4] = {0, 0, 0, 0}, mis_cfg[3] = {0, 0, 0};
  double tune_ms = 0, mis_tune_ms = 0;
  if (tune) {
    const auto& uv = batch_uv[0];
    CHECK(scout_stage_batch(ctx, 0, uv.data(), (int64_t)uv.size() / 2));
    CHECK(scout_classify(ctx, 0, SCOUT_CLASSIFY_CURRENT, nullptr));
    CHECK(scout_apply(ctx, 0, 0));
    CHECK(scout_color_autotune(ctx, 0, &tune_ms));
    CHECK(scout_mis_autotune(ctx, 0, ninst, batch_off[0].data(), max_csr, &mis_tune_ms));
    CHECK(scout_restore(ctx));
  }
  CHECK(scout_color_get_config(ctx, cfg));
  CHECK(scout_mis_get_config(ctx, mis_cfg));

  // the mix: one step per batch for all instances; the first pass warms up (graph capture,
  // buffer growth) and is repeated from the snapshot, the second is timed
  std::vector<double> step_ms;
  scout_stats mis_st{}, col_st{};
  for (int pass = 0; pass < 2; pass++) {
    CHECK(scout_restore(ctx));
    step_ms.clear();
    mis_st = scout_stats{};
    col_st = scout_stats{};
    for (int b = 0; b < batches; b++) {
      const auto& uv = batch_uv[(size_t)b];
      int64_t m = (int64_t)uv.size() / 2;
      CHECK(scout_sync(ctx));
      auto t0 = now();
      CHECK(scout_stage_batch(ctx, 0, uv.data(), m));
      CHECK(scout_classify(ctx, 0, SCOUT_CLASSIFY_CURRENT, nullptr));
      CHECK(scout_apply(ctx, 0, 0));
      CHECK(scout_repair_mis_multi(ctx, 0, ninst, batch_off[(size_t)b].data(), max_csr, &mis_st));
      CHECK(scout_repair_color(ctx, 0, &col_st));
      CHECK(scout_sync(ctx));
      step_ms.push_back(secs(t0, now()) * 1e3);
    }
  }
  double mix_s = 0;
  for (double x : step_ms) mix_s += x / 1e3;
  std::vector<uint8_t> mis_mix((size_t)n);
  std::vector<int32_t> col_mix((size_t)n);
  CHECK(scout_state_download(ctx, SCOUT_STATE_MIS, mis_mix.data(), n));
  CHECK(scout_state_download(ctx, SCOUT_STATE_COLOR, col_mix.data(), n));

  // checks on the final graph
  int64_t info[SCOUT_GRAPH_INFO_COUNT];
  CHECK(scout_graph_info(ctx, info));
  std::vector<int64_t> loff((size_t)n + 1);
  std::vector<int32_t> lcol((size_t)info[3]);
  CHECK(scout_graph_download(ctx, loff.data(), lcol.data()));
  int64_t bad_color = 0, bad_mis = 0;
#pragma omp parallel for reduction(+ : bad_color, bad_mis) schedule(static)
  for (int32_t u = 0; u < n; u++)
    for (int64_t k = loff[(size_t)u]; k < loff[(size_t)u + 1]; k++) {
      int32_t v = lcol[(size_t)k];
      if (u < v) {
        bad_color += col_mix[(size_t)u] == col_mix[(size_t)v];
        bad_mis += mis_mix[(size_t)u] && mis_mix[(size_t)v];
      }
    }

  // alone: the same stream, every instance in its own steps
  double alone_s = 0;
  bool same = true;
  if (alone) {
    CHECK(scout_restore(ctx));
    for (int b = 0; b < batches; b++)
      for (int i = 0; i < ninst; i++) {
        const auto& l = lines[(size_t)b][(size_t)i];
        int64_t m = (int64_t)l.size() / 2;
        if (m == 0) continue;
        CHECK(scout_sync(ctx));
        auto t0 = now();
        scout_stats s1{}, s2{};
        CHECK(scout_stage_batch(ctx, 0, l.data(), m));
        CHECK(scout_classify(ctx, 0, SCOUT_CLASSIFY_CURRENT, nullptr));
        CHECK(scout_apply(ctx, 0, 0));
        CHECK(scout_repair_mis(ctx, 0, max_csr, &s1));
        CHECK(scout_repair_color(ctx, 0, &s2));
        CHECK(scout_sync(ctx));
        alone_s += secs(t0, now());
      }
    std::vector<uint8_t> mis_alone((size_t)n);
    std::vector<int32_t> col_alone((size_t)n);
    CHECK(scout_state_download(ctx, SCOUT_STATE_MIS, mis_alone.data(), n));
    CHECK(scout_state_download(ctx, SCOUT_STATE_COLOR, col_alone.data(), n));
    same = mis_alone == mis_mix && col_alone == col_mix;
  }

  printf("{\"workload\": \"%d Barabasi-Albert graphs (m=%d, n log-uniform [%.0f, %.0f], seed %llu) in one GPU, "
         "MIS + colouring each; %d batches, per instance per batch log-uniform [%.0f, %.0f] toggles, half "
         "deletions of uniform edges, half insertions of uniform non-edges\", "
         "\"n\": %d, \"entries\": %lld, \"lines\": %lld, \"gen_s\": %.2f, \"max_csr\": %lld, "
         "\"updates_per_s\": %.1f, \"step_ms_p50\": %.3f, \"step_ms_max\": %.3f, "
         "\"mis_rounds\": %lld, \"mis_sub_batches\": %lld, \"mis_readbacks\": %lld, \"color_rounds\": %lld, "
         "\"color_readbacks\": %lld, \"color_config\": {\"thread_max\": %d, \"block_min\": %d, \"grid\": %d, "
         "\"graph\": %s}, \"autotune_best_ms\": %.3f, \"mis_config\": {\"grid\": %d, \"group_rounds\": %d, "
         "\"graph\": %s}, \"mis_autotune_best_ms\": %.3f, "
         "\"alone_updates_per_s\": %s, \"speedup_vs_alone\": %s, \"same_as_alone\": %s, "
         "\"proper_coloring\": %s, \"independent_mis\": %s}\n",
         ninst, mba, min_n, max_n, (unsigned long long)seed, batches, min_lines, max_lines, n, (long long)entries,
         (long long)total_lines, gen_s, (long long)max_csr, (double)total_lines / mix_s, pct(step_ms, 50),
         pct(step_ms, 100), (long long)mis_st.rounds, (long long)mis_st.sub_batches, (long long)mis_st.readbacks,
         (long long)col_st.rounds, (long long)col_st.readbacks, cfg[0], cfg[1], cfg[2], cfg[3] ? "true" : "false",
         tune_ms, mis_cfg[0], mis_cfg[1], mis_cfg[2] ? "true" : "false", mis_tune_ms,
         alone ? std::to_string((double)total_lines / alone_s).c_str() : "null",
         alone ? std::to_string(alone_s / mix_s).c_str() : "null", alone ? (same ? "true" : "false") : "null",
         bad_color == 0 ? "true" : "false", bad_mis == 0 ? "true" : "false");
  fflush(stdout);
  CHECK(scout_ctx_destroy(ctx));
  return same && bad_color == 0 && bad_mis == 0 ? 0 : 2;
}
