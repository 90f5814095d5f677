// scout_serve: a continuous-batching serving loop for BFS(0) + CC maintenance on one GPU.
//
// Update requests (batches of edge toggles) arrive as a Poisson process. The server keeps the
// graph and both solutions resident in HBM. Whenever the device is free it admits every
// queued request that fits the staging budget, in arrival order, as one step:
//  1. the requests' lines are staged back to back;
//  2. each request is classified and applied against the graph the earlier ones left, which
//     matches serving them one by one;
//  3. one BFS repair and one CC repair settle all of them.
// The step commits when the solutions are exact in HBM. Each request's latency runs from its
// arrival to the commit of its step.
//
// The request stream is generated first, by serving every request alone (one step each). That
// pass also gives the reference state; its rate (alone_pass_updates_per_s) includes the
// generation of the requests on the device. The served run must end in the same levels and labels
// ("alone == in the mix"), and both must equal a full recompute.
//
//   scout_serve [--scale 25] [--rate 2e6] [--min 1000] [--max 1000000] [--requests 200]
//               [--delete 0.6] [--budget-lines 4000000] [--seed 1] [--gpu 0]
//
// Prints one JSON line. The reference has no serving loop: this is new (the serving
// configuration of the benchmark plan), built on the C ABI of libscout_hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/scout_hip.h"
#include "../host/rng.h"

namespace {

[[noreturn]] void die(const char* what) {
  const char* e = scout_last_error();
  fprintf(stderr, "ERROR: %s (%s)\n", what, e ? e : "");
  fflush(stdout);
  std::_Exit(1);
}

#define CHECK(call)                   \
  do {                                \
    if ((call) != SCOUT_OK) die(#call); \
  } while (0)

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)std::ceil(p / 100.0 * (double)v.size());
  if (k == 0) k = 1;
  return v[std::min(k, v.size()) - 1];
}

double unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

int main(int argc, char** argv) {
  int scale = 25, nreq = 200, gpu = 0;
  double rate = 2e6, min_size = 1000, max_size = 1e6, del = 0.6;
  int64_t budget_lines = 4000000;
  uint64_t seed = 1;
  for (int i = 1; i + 1 < argc; i += 2) {
    std::string k = argv[i];
    const char* v = argv[i + 1];
    if (k == "--scale") scale = atoi(v);
    else if (k == "--rate") rate = atof(v);
    else if (k == "--min") min_size = atof(v);
    else if (k == "--max") max_size = atof(v);
    else if (k == "--requests") nreq = atoi(v);
    else if (k == "--delete") del = atof(v);
    else if (k == "--budget-lines") budget_lines = atoll(v);
    else if (k == "--seed") seed = strtoull(v, nullptr, 10);
    else if (k == "--gpu") gpu = atoi(v);
    else {
      fprintf(stderr, "usage: scout_serve [--scale S] [--rate EDGES_PER_S] [--min L] [--max L] [--requests N] "
                      "[--delete F] [--budget-lines L] [--seed S] [--gpu G]\n");
      return 1;
    }
  }
  if (nreq < 1 || min_size < 1 || max_size < min_size || rate <= 0 || budget_lines < (int64_t)max_size) {
    fprintf(stderr, "ERROR: need requests >= 1, 1 <= min <= max <= budget-lines, rate > 0\n");
    return 1;
  }
  // Staging budget in HBM: per line the pair (8 B), its label and done flag (2 B) and the
  // classify/apply scratch (about 40 B).
  scout_budget budget{budget_lines * 50 + ((int64_t)1 << 34), 0};
  scout_ctx* ctx = nullptr;
  CHECK(scout_ctx_create(gpu, &budget, &ctx));
  CHECK(scout_gen_rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed));
  int64_t info[6];
  CHECK(scout_graph_info(ctx, info));
  int32_t n = (int32_t)info[0];
  CHECK(scout_bfs_full(ctx, 0));
  CHECK(scout_cc_full(ctx));
  CHECK(scout_snapshot(ctx));

  // Request sizes (log-uniform) and Poisson arrival times.
  uint64_t key = scout::stream_key(seed, 0x5345525645);
  std::vector<int64_t> size((size_t)nreq);
  std::vector<double> arrive((size_t)nreq);
  double mean = (max_size - min_size) / std::log(max_size / min_size);
  if (max_size == min_size) mean = min_size;
  double lambda = rate / mean, t = 0;
  int64_t total_lines = 0;
  for (int r = 0; r < nreq; r++) {
    size[(size_t)r] = (int64_t)std::llround(min_size * std::exp(unit(scout::draw(key, 2 * r)) * std::log(max_size / min_size)));
    t += -std::log(1.0 - unit(scout::draw(key, 2 * r + 1))) / lambda;
    arrive[(size_t)r] = t;
    total_lines += size[(size_t)r];
  }

  // Pass 1: generate every request against the evolving graph, serving it alone.
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto t_alone0 = now();
  std::vector<int64_t> lines((size_t)nreq);
  for (int r = 0; r < nreq; r++) {
    int64_t m = 0;
    CHECK(scout_gen_batch(ctx, 0, size[(size_t)r], del, 0.0, seed, total_lines, &m));
    CHECK(scout_batch_store(ctx, 0, r));
    lines[(size_t)r] = m;
    int64_t seg[2] = {0, m};
    CHECK(scout_apply_segments(ctx, 0, 1, seg, SCOUT_CLASSIFY_CURRENT, 0, nullptr));
    scout_stats sb{}, sc{};
    CHECK(scout_repair_bfs(ctx, 0, SCOUT_STRATEGY_LOCAL, &sb));
    CHECK(scout_repair_cc(ctx, 0, SCOUT_STRATEGY_LOCAL, &sc));
  }
  CHECK(scout_sync(ctx));
  double alone_s = std::chrono::duration<double>(now() - t_alone0).count();
  int64_t served_lines = 0;
  for (int64_t m : lines) served_lines += m;
  std::vector<int32_t> lv_alone((size_t)n), lab_alone((size_t)n);
  CHECK(scout_state_download(ctx, SCOUT_STATE_LEVEL, lv_alone.data(), n));
  CHECK(scout_state_download(ctx, SCOUT_STATE_LABEL, lab_alone.data(), n));

  // Pass 2: the serving loop in real time.
  CHECK(scout_restore(ctx));
  std::vector<double> latency((size_t)nreq);
  int next = 0, steps = 0, max_in_step = 0;
  auto t0 = now();
  auto elapsed = [&] { return std::chrono::duration<double>(now() - t0).count(); };
  while (next < nreq) {
    double e = elapsed();
    if (arrive[(size_t)next] > e) {
      std::this_thread::sleep_for(std::chrono::duration<double>(arrive[(size_t)next] - e));
      continue;
    }
    // admit every arrived request that fits the budget (at least one)
    int first = next;
    int64_t m = 0;
    std::vector<int64_t> seg{0};
    while (next < nreq && arrive[(size_t)next] <= e && (next == first || m + lines[(size_t)next] <= budget_lines)) {
      int64_t total = 0;
      if (next == first) CHECK(scout_batch_load(ctx, 0, next, &total));
      else CHECK(scout_batch_append(ctx, 0, next, &total));
      m = total;
      seg.push_back(m);
      next++;
    }
    CHECK(scout_apply_segments(ctx, 0, (int32_t)(seg.size() - 1), seg.data(), SCOUT_CLASSIFY_CURRENT, 0, nullptr));
    scout_stats sb{}, sc{};
    CHECK(scout_repair_bfs(ctx, 0, SCOUT_STRATEGY_LOCAL, &sb));
    CHECK(scout_repair_cc(ctx, 0, SCOUT_STRATEGY_LOCAL, &sc));
    CHECK(scout_sync(ctx));
    double commit = elapsed();
    for (int r = first; r < next; r++) latency[(size_t)r] = (commit - arrive[(size_t)r]) * 1e3;
    steps++;
    max_in_step = std::max(max_in_step, next - first);
  }
  double wall = elapsed();
  std::vector<int32_t> lv((size_t)n), lab((size_t)n);
  CHECK(scout_state_download(ctx, SCOUT_STATE_LEVEL, lv.data(), n));
  CHECK(scout_state_download(ctx, SCOUT_STATE_LABEL, lab.data(), n));
  bool same_as_alone = lv == lv_alone && lab == lab_alone;
  CHECK(scout_bfs_full(ctx, 0));
  CHECK(scout_cc_full(ctx));
  std::vector<int32_t> lv_full((size_t)n), lab_full((size_t)n);
  CHECK(scout_state_download(ctx, SCOUT_STATE_LEVEL, lv_full.data(), n));
  CHECK(scout_state_download(ctx, SCOUT_STATE_LABEL, lab_full.data(), n));
  bool exact = lv == lv_full && lab == lab_full;

  printf("{\"workload\": \"serving loop, BFS(0)+CC, R-MAT scale %d ef 16, Poisson arrivals at %.6g edges/s, "
         "request sizes log-uniform [%.0f, %.0f], %.0f%% deletions (uniform over adjacency entries, so "
         "degree-biased endpoints), staging budget %lld lines\", "
         "\"requests\": %d, \"lines\": %lld, \"wall_s\": %.3f, \"updates_per_s\": %.1f, \"offered_per_s\": %.1f, "
         "\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"max_ms\": %.3f, \"steps\": %d, \"max_requests_per_step\": %d, "
         "\"alone_pass_updates_per_s\": %.1f, \"same_as_alone\": %s, \"verified_exact\": %s}\n",
         scale, rate, min_size, max_size, del * 100, (long long)budget_lines, nreq, (long long)served_lines, wall,
         (double)served_lines / wall, (double)served_lines / arrive.back(), pct(latency, 50), pct(latency, 99),
         pct(latency, 100), steps, max_in_step, (double)served_lines / alone_s, same_as_alone ? "true" : "false",
         exact ? "true" : "false");
  fflush(stdout);
  CHECK(scout_ctx_destroy(ctx));
  return same_as_alone && exact ? 0 : 2;
}
