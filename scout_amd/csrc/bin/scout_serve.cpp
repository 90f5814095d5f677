// scout_serve: a continuous-batching serving loop for BFS(0) + CC maintenance on one GPU.
//
// Update requests (batches of edge toggles, host arrays) arrive as a Poisson process. The
// server keeps the graph and both solutions resident in HBM and runs two host threads over
// the two staging channels of one context:
//   * admission: whenever a channel is free and requests have arrived, it takes the queued
//     requests in arrival order while the step fits the budgets (pinned staging: 8 B per line;
//     HBM staging: the channel's per-line device buffers), copies their lines into the channel's
//     pinned staging and starts the host-to-device copy on the copy stream -- while the device
//     thread is still repairing the previous step on the other channel;
//   * device: takes the staged steps in order; each request of a step is classified and applied
//     against the graph the earlier ones left (the same as serving them one by one), then one
//     BFS repair and one CC repair settle them all; the step commits when the solutions are
//     exact in HBM.
// Each request's latency runs from its arrival to the commit of its step.
//
// The request stream is generated first, by serving every request alone (one step each, on
// the device, then copied to host memory as the arriving payload). That pass also gives the
// reference state: the served run must end in the same levels and labels ("alone == in the
// mix"), and both must equal a full recompute. --dump DIR writes every request's lines
// (req_<r>.bin, int32 pairs), every step's request range (steps.txt) and the committed levels
// and labels after every step (lv_<k>.bin, lab_<k>.bin), for a check against the oracle.
//
//   scout_serve [--scale 25] [--rate 2e6] [--min 1000] [--max 1000000] [--requests 200] [--bfs-mode async|level]
//               [--delete 0.6] [--pinned-mb 64] [--hbm-mb 512] [--seed 1] [--gpu 0] [--dump DIR]
//
// Prints one JSON line. The reference has no serving loop: this is new (the serving
// configuration of the benchmark plan), built on the C ABI of libscout_hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <filesystem>
#include <mutex>
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

void write_file(const std::string& path, const void* data, size_t bytes) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f || fwrite(data, 1, bytes, f) != bytes) {
    fprintf(stderr, "ERROR: cannot write %s\n", path.c_str());
    std::_Exit(1);
  }
  fclose(f);
}

// Device staging bytes per admitted line: the pair (8), its label and done flag (2), the base
// range classify hands to apply (16) and apply's new-entry keys, values and sort copies (~40).
constexpr int64_t kHbmPerLine = 66;

struct Step {
  int channel = 0;
  int first = 0, last = 0;         // requests [first, last)
  std::vector<int64_t> seg;        // segment offsets, one segment per request
  double pack_ms = 0, h2d_ms = 0;  // staging: host copy into pinned memory, copy-stream transfer
  int64_t bytes = 0;
  bool staged_while_busy = false;  // the previous step was still on the device when this one was staged
};

}  // namespace

int main(int argc, char** argv) {
  int scale = 25, nreq = 200, gpu = 0;
  double rate = 2e6, min_size = 1000, max_size = 1e6, del = 0.6;
  int64_t pinned_mb = 64, hbm_mb = 512;
  uint64_t seed = 1;
  std::string dump, bfs_mode = "async";
  for (int i = 1; i + 1 < argc; i += 2) {
    std::string k = argv[i];
    const char* v = argv[i + 1];
    if (k == "--scale") scale = atoi(v);
    else if (k == "--rate") rate = atof(v);
    else if (k == "--min") min_size = atof(v);
    else if (k == "--max") max_size = atof(v);
    else if (k == "--requests") nreq = atoi(v);
    else if (k == "--delete") del = atof(v);
    else if (k == "--pinned-mb") pinned_mb = atoll(v);
    else if (k == "--hbm-mb") hbm_mb = atoll(v);
    else if (k == "--seed") seed = strtoull(v, nullptr, 10);
    else if (k == "--gpu") gpu = atoi(v);
    else if (k == "--dump") dump = v;
    else if (k == "--bfs-mode") bfs_mode = v;
    else {
      fprintf(stderr, "usage: scout_serve [--scale S] [--rate EDGES_PER_S] [--min L] [--max L] [--requests N] "
                      "[--delete F] [--pinned-mb MB] [--hbm-mb MB] [--seed S] [--gpu G] [--dump DIR] "
                      "[--bfs-mode async|level|inval]\n");
      return 1;
    }
  }
  // each of the two channels may hold half of each budget
  const int64_t pinned_bytes = pinned_mb << 20, hbm_bytes = hbm_mb << 20;
  const int64_t step_lines = std::min(pinned_bytes / 2 / 8, hbm_bytes / 2 / kHbmPerLine);
  if (nreq < 1 || min_size < 1 || max_size < min_size || rate <= 0 || step_lines < (int64_t)max_size) {
    fprintf(stderr, "ERROR: need requests >= 1, 1 <= min <= max, rate > 0, and budgets that hold a request of max "
                    "lines in half of each (a step may take %lld lines)\n", (long long)step_lines);
    return 1;
  }
  if (!dump.empty()) std::filesystem::create_directories(dump);
  scout_budget budget{0, pinned_bytes / 2};
  scout_ctx* ctx = nullptr;
  CHECK(scout_ctx_create(gpu, &budget, &ctx));
  CHECK(scout_bfs_mode(ctx, bfs_mode == "level"   ? SCOUT_BFS_LEVEL_SYNC
                            : bfs_mode == "inval" ? SCOUT_BFS_ASYNC_INVALIDATION
                                                  : SCOUT_BFS_ASYNC));
  CHECK(scout_gen_rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed));
  int64_t info[SCOUT_GRAPH_INFO_COUNT];
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

  // Pass 1: generate every request against the evolving graph, serving it alone; keep its lines
  // in host memory as the payload that will arrive.
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto t_alone0 = now();
  std::vector<std::vector<int32_t>> req((size_t)nreq);
  for (int r = 0; r < nreq; r++) {
    int64_t m = 0;
    CHECK(scout_gen_batch(ctx, 0, size[(size_t)r], del, 0.0, seed, total_lines, &m));
    req[(size_t)r].resize((size_t)(2 * m));
    CHECK(scout_batch_download(ctx, 0, req[(size_t)r].data(), m));
    int64_t seg[2] = {0, m};
    CHECK(scout_apply_segments(ctx, 0, 1, seg, SCOUT_CLASSIFY_CURRENT, 0, nullptr));
    scout_stats sb{}, sc{};
    CHECK(scout_repair_bfs(ctx, 0, SCOUT_STRATEGY_LOCAL, &sb));
    CHECK(scout_repair_cc(ctx, 0, SCOUT_STRATEGY_LOCAL, &sc));
    if (!dump.empty())
      write_file(dump + "/req_" + std::to_string(r) + ".bin", req[(size_t)r].data(), req[(size_t)r].size() * 4);
  }
  CHECK(scout_sync(ctx));
  double alone_s = std::chrono::duration<double>(now() - t_alone0).count();
  int64_t served_lines = 0;
  for (auto& q : req) served_lines += (int64_t)q.size() / 2;
  std::vector<int32_t> lv_alone((size_t)n), lab_alone((size_t)n);
  CHECK(scout_state_download(ctx, SCOUT_STATE_LEVEL, lv_alone.data(), n));
  CHECK(scout_state_download(ctx, SCOUT_STATE_LABEL, lab_alone.data(), n));

  // Pass 2: the serving loop in real time.
  CHECK(scout_restore(ctx));
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Step> ready;           // staged steps, in order
  bool channel_free[2] = {true, true};
  bool device_busy = false;
  std::vector<double> latency((size_t)nreq);
  std::vector<Step> done;
  auto t0 = now();
  auto elapsed = [&] { return std::chrono::duration<double>(now() - t0).count(); };

  std::thread admission([&] {
    int next = 0, k = 0;
    std::vector<int32_t> lines;
    while (next < nreq) {
      int c = k & 1;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return channel_free[c]; });
        channel_free[c] = false;
      }
      double e = elapsed();
      if (arrive[(size_t)next] > e) std::this_thread::sleep_for(std::chrono::duration<double>(arrive[(size_t)next] - e));
      e = elapsed();
      // every arrived request that fits the step's budget (at least one), in arrival order
      Step s;
      s.channel = c;
      s.first = next;
      s.seg.push_back(0);
      lines.clear();
      while (next < nreq && arrive[(size_t)next] <= e &&
             (next == s.first || (int64_t)(lines.size() + req[(size_t)next].size()) / 2 <= step_lines)) {
        lines.insert(lines.end(), req[(size_t)next].begin(), req[(size_t)next].end());
        s.seg.push_back((int64_t)lines.size() / 2);
        next++;
      }
      s.last = next;
      {
        std::lock_guard<std::mutex> lk(mu);
        s.staged_while_busy = device_busy;
      }
      // stage on this channel while the device thread may be repairing on the other one
      CHECK(scout_stage_batch(ctx, c, lines.data(), (int64_t)lines.size() / 2));
      CHECK(scout_stage_info(ctx, c, &s.pack_ms, &s.h2d_ms, &s.bytes));
      {
        std::lock_guard<std::mutex> lk(mu);
        ready.push_back(std::move(s));
      }
      cv.notify_all();
      k++;
    }
  });

  int served = 0, steps = 0, max_in_step = 0;
  while (served < nreq) {
    Step s;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return !ready.empty(); });
      s = std::move(ready.front());
      ready.pop_front();
      device_busy = true;
    }
    CHECK(scout_apply_segments(ctx, s.channel, (int32_t)(s.seg.size() - 1), s.seg.data(), SCOUT_CLASSIFY_CURRENT, 0,
                               nullptr));
    scout_stats sb{}, sc{};
    CHECK(scout_repair_bfs(ctx, s.channel, SCOUT_STRATEGY_LOCAL, &sb));
    CHECK(scout_repair_cc(ctx, s.channel, SCOUT_STRATEGY_LOCAL, &sc));
    CHECK(scout_sync(ctx));
    double commit = elapsed();
    for (int r = s.first; r < s.last; r++) latency[(size_t)r] = (commit - arrive[(size_t)r]) * 1e3;
    if (!dump.empty()) {
      std::vector<int32_t> lv((size_t)n), lab((size_t)n);
      CHECK(scout_state_download(ctx, SCOUT_STATE_LEVEL, lv.data(), n));
      CHECK(scout_state_download(ctx, SCOUT_STATE_LABEL, lab.data(), n));
      write_file(dump + "/lv_" + std::to_string(steps) + ".bin", lv.data(), lv.size() * 4);
      write_file(dump + "/lab_" + std::to_string(steps) + ".bin", lab.data(), lab.size() * 4);
    }
    served = s.last;
    steps++;
    max_in_step = std::max(max_in_step, s.last - s.first);
    {
      std::lock_guard<std::mutex> lk(mu);
      channel_free[s.channel] = true;
      device_busy = false;
    }
    cv.notify_all();
    done.push_back(std::move(s));
  }
  admission.join();
  double wall = elapsed();
  if (!dump.empty()) {
    std::string body;
    for (auto& s : done) body += std::to_string(s.first) + " " + std::to_string(s.last) + "\n";
    write_file(dump + "/steps.txt", body.data(), body.size());
  }
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
  double h2d_ms = 0, pack_ms = 0;
  int64_t staged_bytes = 0;
  int overlapped = 0;
  for (auto& s : done) {
    h2d_ms += s.h2d_ms;
    pack_ms += s.pack_ms;
    staged_bytes += s.bytes;
    overlapped += s.staged_while_busy;
  }

  printf("{\"workload\": \"serving loop, BFS(0)+CC, R-MAT scale %d ef 16, Poisson arrivals at %.6g edges/s, "
         "request sizes log-uniform [%.0f, %.0f], %.0f%% deletions (uniform over adjacency entries, so "
         "degree-biased endpoints), budgets %lld MB pinned + %lld MB HBM staging (%lld lines per step)\", "
         "\"requests\": %d, \"lines\": %lld, \"wall_s\": %.3f, \"updates_per_s\": %.1f, \"offered_per_s\": %.1f, "
         "\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"max_ms\": %.3f, \"steps\": %d, \"max_requests_per_step\": %d, "
         "\"steps_staged_while_device_busy\": %d, \"pcie_h2d_GBps\": %.2f, \"staged_MB\": %.1f, \"pack_ms_total\": %.2f, "
         "\"alone_pass_updates_per_s\": %.1f, \"same_as_alone\": %s, \"verified_exact\": %s}\n",
         scale, rate, min_size, max_size, del * 100, (long long)pinned_mb, (long long)hbm_mb, (long long)step_lines,
         nreq, (long long)served_lines, wall, (double)served_lines / wall, (double)served_lines / arrive.back(),
         pct(latency, 50), pct(latency, 99), pct(latency, 100), steps, max_in_step, overlapped,
         h2d_ms > 0 ? (double)staged_bytes / (h2d_ms * 1e-3) / 1e9 : 0.0, (double)staged_bytes / 1e6, pack_ms,
         (double)served_lines / alone_s, same_as_alone ? "true" : "false", exact ? "true" : "false");
  fflush(stdout);
  CHECK(scout_ctx_destroy(ctx));
  return same_as_alone && exact ? 0 : 2;
}
