// driver.h: the shared main() of the scout_{mis,coloring,bfs,cc} binaries.
//
// Same 5 positional arguments, input formats and stdout lines as the reference binaries
// (core/cli_setup.h, core/reporting.h, src/*/scout_*.cu main()). The per-batch work runs in
// libscout_hip: stage the parsed lines (pinned staging + copy stream), classify, apply the
// scratchpad rules, repair, commit. Parsing of batch k+1 runs on a host thread while batch k
// is on the device (SCOUT_NO_OVERLAP disables it), the role of the reference's 2-channel
// overlap (src/mis/scout_mis.cu:853-912).
//
// Environment, as in the reference:
//   SCOUT_NO_OVERLAP, SCOUT_NO_MADV_RANDOM, SCOUT_SKIP_COMPACTION=1 (mis);
//   SCOUT_DUMP_TOUCHED=<file>: the ids (int32) of every vertex whose row a repair read, written
//     at exit (core/FlashGraph.h:318-329);
//   SCOUT_BFS_BOUND=<b> (bfs, and the levels cc keeps; default 200, src/bfs/scout_bfs.cu:476): a
//     batch whose invalidated vertices exceed b per deleted parent edge is finished by a full
//     BFS (fallback=1), still before the batch commits;
//   SCOUT_DEFER_FALLBACK (core/local_repair.h:22-26): read and reported; every fallback runs
//     once per batch before the batch commits, so "deferred" and "immediate" coincide here;
//   SCOUT_NO_MEMO, SCOUT_FORCE_COPY: accepted with no effect (rows are resident in HBM, so there
//     is no fetch memo, and every transfer to the discrete MI355X is an explicit copy); a note
//     goes to stderr.
// Added: SCOUT_HBM_BUDGET=<bytes> caps the context's HBM (a graph whose base entries do not fit
// stays in pinned host memory); SCOUT_MIS_COMPLETE=1 (mis) completes the set after the reference's rules, keeping it
// maximal (scout_mis_mode); SCOUT_DUMP_STATE=<dir> writes the solution after every batch as raw little-endian
// arrays (<dir>/<problem>_<batch>.bin: uint8 membership / int32 otherwise);
// SCOUT_STRATEGY=local|recompute|learned (bfs, cc).
#pragma once


#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <future>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "../../../include/scout_hip.h"
#include "../host/cli.h"
#include "../host/egr.h"
#include "../host/report.h"

namespace scout {

enum class Problem { MIS, COLORING, BFS, CC };

inline double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

[[noreturn]] inline void die(const std::string& what) {
  std::cerr << "ERROR: " << what;
  const char* e = scout_last_error();
  if (e && *e) std::cerr << " (" << e << ")";
  std::cerr << std::endl;
  std::cout.flush();
  std::_Exit(1);
}

#define DRV_CHECK(call)                        \
  do {                                         \
    if ((call) != SCOUT_OK) die(#call);        \
  } while (0)

inline void dump_state(const char* dir, const char* name, int batch, const void* data, size_t bytes) {
  std::filesystem::create_directories(dir);
  std::string path = std::string(dir) + "/" + name + "_" + std::to_string(batch) + ".bin";
  FILE* f = fopen(path.c_str(), "wb");
  if (!f || fwrite(data, 1, bytes, f) != bytes) die("cannot write " + path);
  fclose(f);
}

inline int run_driver(int argc, char** argv, Problem prob) {
  const char* algo = prob == Problem::MIS ? "mis" : prob == Problem::COLORING ? "coloring" : prob == Problem::BFS ? "bfs" : "cc";
  CliArgs args;
  if (!parse_cli(argc, argv, algo, args)) return 1;
  const bool overlap = std::getenv("SCOUT_NO_OVERLAP") == nullptr;
  const char* dump_dir = std::getenv("SCOUT_DUMP_STATE");
  int strategy = SCOUT_STRATEGY_LOCAL;
  if (const char* s = std::getenv("SCOUT_STRATEGY")) {
    std::string v = s;
    if (v == "recompute") strategy = SCOUT_STRATEGY_RECOMPUTE;
    else if (v == "learned") strategy = SCOUT_STRATEGY_LEARNED;
  }

  const char* touched_path = std::getenv("SCOUT_DUMP_TOUCHED");
  long long bfs_bound = 200;
  if (const char* b = std::getenv("SCOUT_BFS_BOUND")) bfs_bound = std::max(0LL, atoll(b));
  bool defer = true;
  if (const char* d = std::getenv("SCOUT_DEFER_FALLBACK")) defer = std::string(d) != "0";
  for (const char* v : {"SCOUT_NO_MEMO", "SCOUT_FORCE_COPY"})
    if (std::getenv(v))
      std::cerr << v << ": no effect on this engine (rows are resident in HBM and every transfer is an explicit copy)"
                << std::endl;

  // SCOUT_HBM_BUDGET=<bytes>: a graph whose base entries do not fit stays storage-resident
  // (pinned host memory read over PCIe; full recomputes stream it in chunks)
  scout_budget budget{0, 0};
  if (const char* b = std::getenv("SCOUT_HBM_BUDGET")) budget.hbm_bytes = std::max(0LL, atoll(b));
  scout_ctx* ctx = nullptr;
  DRV_CHECK(scout_ctx_create(args.gpu_id, &budget, &ctx));
  char gpu_name[256] = {0};
  DRV_CHECK(scout_device_name(ctx, gpu_name, sizeof(gpu_name)));
  // device memory in use before this program's own allocations (src/mis/scout_mis.cu:51-72,789)
  int64_t mem_free0 = 0, mem_total = 0;
  DRV_CHECK(scout_mem_info(ctx, &mem_free0, &mem_total));
  double peak_delta_mb = 0.0;
  auto sample_memory = [&] {
    int64_t f = 0, t = 0;
    if (scout_mem_info(ctx, &f, &t) != SCOUT_OK) return;
    peak_delta_mb = std::max(peak_delta_mb, (double)(mem_free0 - f) / (1024.0 * 1024.0));
  };
  if (prob == Problem::BFS || prob == Problem::CC) {
    DRV_CHECK(scout_bfs_bound(ctx, bfs_bound));
    std::cout << (defer ? "[Deferred fallback: ON (default)] Batches whose cascades pass the safety bound are "
                          "settled by one full recompute each, before the batch commits."
                        : "[Deferred fallback: OFF (SCOUT_DEFER_FALLBACK=0)] Batches whose cascades pass the "
                          "safety bound are settled by one full recompute each, before the batch commits.")
              << std::endl;
  }

  auto setup_t0 = std::chrono::steady_clock::now();
  EgrFile egr;
  std::string err = egr.open(args.egr_path.c_str());
  if (!err.empty()) die("FlashGraph " + err);
  const int32_t n = egr.nodes;
  double avg_degree = n > 0 ? (double)egr.edges / n : 0.0;
  if (prob == Problem::MIS) {
    std::cout << "======================================= SCOUT MIS =========================================" << std::endl;
  }
  print_run_header(algo, args.egr_path, n, (long long)egr.edges, avg_degree, gpu_name);
  DRV_CHECK(scout_graph_load(ctx, n, (int64_t)egr.edges, egr.nindex_raw(), egr.wide ? 1 : 0, egr.nlist()));
  if (touched_path) DRV_CHECK(scout_touched_enable(ctx, 1));

  // Initial state.
  std::vector<long long> vals = read_ints(args.context_path);
  long long initial_value = 0;
  std::vector<uint8_t> mis;
  std::vector<int32_t> st32;
  if (prob == Problem::MIS) {
    decode_mis_init(vals, n, mis, initial_value);
    DRV_CHECK(scout_state_upload(ctx, SCOUT_STATE_MIS, mis.data(), n));
  } else if (prob == Problem::COLORING) {
    decode_color_init(vals, n, st32);
    int32_t mx = -1;
    for (int32_t c : st32) mx = std::max(mx, c);
    initial_value = mx + 1;
    DRV_CHECK(scout_state_upload(ctx, SCOUT_STATE_COLOR, st32.data(), n));
  } else if (prob == Problem::BFS) {
    decode_bfs_init(vals, n, st32, initial_value);
    DRV_CHECK(scout_state_upload(ctx, SCOUT_STATE_LEVEL, st32.data(), n));
  } else {
    std::vector<int32_t> root;
    if (!decode_cc_init(vals, n, root, err)) die(err);
    // canonical labels: smallest vertex id of each root class
    std::vector<int32_t> mn((size_t)n, INT32_MAX);
    for (int32_t v = 0; v < n; v++) mn[(size_t)root[(size_t)v]] = std::min(mn[(size_t)root[(size_t)v]], v);
    st32.resize((size_t)n);
    for (int32_t v = 0; v < n; v++) st32[(size_t)v] = mn[(size_t)root[(size_t)v]];
    for (int32_t v = 0; v < n; v++) initial_value += st32[(size_t)v] == v;
    DRV_CHECK(scout_state_upload(ctx, SCOUT_STATE_LABEL, st32.data(), n));
  }
  int64_t max_csr = 0;
  if (prob == Problem::MIS) {
    // MAX_CSR_ELEMENTS (src/mis/scout_mis.cu:800-815)
    int64_t free_b = 0, total_b = 0;
    DRV_CHECK(scout_mem_info(ctx, &free_b, &total_b));
    long long est = (long long)(1048576.0 * 2.0 * avg_degree * avg_degree);
    long long mx = est > 0 ? std::min(800000000LL, est) : 50000000LL;
    mx = std::max(mx, 1000000LL);
    long long fit = (long long)((double)free_b / 4.0 / 8.0);
    if (fit > 1000000LL && mx > fit) mx = fit;
    max_csr = mx;
    std::cout << "MAX_CSR_ELEMENTS set to: " << max_csr << std::endl;
    if (const char* cm = std::getenv("SCOUT_MIS_COMPLETE"))
      if (std::string(cm) == "1") DRV_CHECK(scout_mis_mode(ctx, SCOUT_MIS_COMPLETE));
  }
  double setup_ms = ms_since(setup_t0);

  std::vector<std::filesystem::path> paths;
  if (!list_batch_files(args.batch_dir, paths, err)) die(err);
  int total = std::min<int>(args.num_batches, (int)paths.size());
  if (total < 0) total = 0;

  auto parse = [&](int k) {
    std::vector<int32_t> uv;
    std::string e;
    if (!parse_batch_pairs(paths[(size_t)k].string(), n, uv, e)) die(e);
    return uv;
  };
  std::future<std::vector<int32_t>> pending;
  if (total > 0) pending = std::async(overlap ? std::launch::async : std::launch::deferred, parse, 0);

  scout_stats sum{};
  std::vector<double> batch_ms;
  long long lines_total = 0;
  double processing_ms = 0.0;
  for (int k = 0; k < total; k++) {
    auto fetch_t0 = std::chrono::steady_clock::now();
    std::vector<int32_t> uv = pending.get();
    double fetch_ms = ms_since(fetch_t0);
    if (k + 1 < total) pending = std::async(overlap ? std::launch::async : std::launch::deferred, parse, k + 1);
    int64_t m = (int64_t)uv.size() / 2;
    int ch = k & 1;
    auto t0 = std::chrono::steady_clock::now();
    scout_stats st{};
    DRV_CHECK(scout_stage_batch(ctx, ch, uv.data(), m));
    double pack_ms = 0.0, h2d_ms = 0.0;
    DRV_CHECK(scout_stage_info(ctx, ch, &pack_ms, &h2d_ms, nullptr));
    auto tc = std::chrono::steady_clock::now();
    DRV_CHECK(scout_classify(ctx, ch, prob == Problem::MIS ? SCOUT_CLASSIFY_BASE : SCOUT_CLASSIFY_CURRENT, nullptr));
    DRV_CHECK(scout_apply(ctx, ch, prob == Problem::MIS ? 1 : 0));
    DRV_CHECK(scout_sync(ctx));
    double apply_ms = ms_since(tc);
    switch (prob) {
      case Problem::MIS: DRV_CHECK(scout_repair_mis(ctx, ch, max_csr, &st)); break;
      case Problem::COLORING: DRV_CHECK(scout_repair_color(ctx, ch, &st)); break;
      case Problem::BFS: DRV_CHECK(scout_repair_bfs(ctx, ch, strategy, &st)); break;
      case Problem::CC: DRV_CHECK(scout_repair_cc(ctx, ch, strategy, &st)); break;
    }
    DRV_CHECK(scout_sync(ctx));
    double this_ms = ms_since(t0);
    processing_ms += this_ms;
    batch_ms.push_back(this_ms);
    lines_total += m;
    std::vector<KV> kv;
    if (prob == Problem::MIS) {
      // repair regions: the lines whose rule wrote the membership
      kv = {{"clusters", std::to_string(st.conflicts)}};
    } else if (prob == Problem::COLORING) {
      kv = {{"insertions", std::to_string(st.insertions)}, {"conflicts", std::to_string(st.conflicts)},
            {"rounds", std::to_string(st.rounds)}};
    } else {
      kv = {{"insertions", std::to_string(st.insertions)}, {"deletions", std::to_string(st.deletions)},
            {"fallback", std::to_string(st.fallback)}};
    }
    print_batch_result(k + 1, total, m, this_ms, kv);
    // fetch: waiting for the batch's lines; seed pack: copying them into pinned staging; h2d:
    // the copy-stream transfer; repair: classify + apply + repair; commit: the rest (read-backs
    // and the final wait)
    double repair_ms = st.repair_ms + apply_ms;
    print_stage_times(k + 1, fetch_ms, pack_ms, h2d_ms, repair_ms, std::max(0.0, this_ms - pack_ms - h2d_ms - repair_ms));
    sample_memory();
    sum.insertions += st.insertions;
    sum.deletions += st.deletions;
    sum.rounds += st.rounds;
    sum.conflicts += st.conflicts;
    sum.splits += st.splits;
    sum.fallback += st.fallback;
    sum.sub_batches += st.sub_batches;
    if (dump_dir) {
      if (prob == Problem::MIS) {
        DRV_CHECK(scout_state_download(ctx, SCOUT_STATE_MIS, mis.data(), n));
        dump_state(dump_dir, algo, k + 1, mis.data(), (size_t)n);
      } else {
        int kind = prob == Problem::COLORING ? SCOUT_STATE_COLOR : prob == Problem::BFS ? SCOUT_STATE_LEVEL : SCOUT_STATE_LABEL;
        DRV_CHECK(scout_state_download(ctx, kind, st32.data(), n));
        dump_state(dump_dir, algo, k + 1, st32.data(), (size_t)n * 4);
      }
    }
  }

  long long final_value = 0;
  double compact_ms = 0.0;
  if (prob == Problem::MIS) {
    DRV_CHECK(scout_state_download(ctx, SCOUT_STATE_MIS, mis.data(), n));
    for (int32_t v = 0; v < n; v++) final_value += mis[(size_t)v];
    std::string out_dir = "datasets/Final_MISs";
    std::filesystem::create_directories(out_dir);
    std::string out = out_dir + "/" + std::filesystem::path(args.egr_path).stem().string() + "_flash_final_mis.txt";
    std::ofstream f(out);
    if (f.is_open()) {
      for (int32_t v = 0; v < n; v++)
        if (mis[(size_t)v]) f << (v + 1) << "\n";
      f.close();
      std::cout << "Saved final MIS to " << out << std::endl;
    } else {
      std::cerr << "Warning: Could not open file to save final MIS: " << out << std::endl;
    }
    const char* skip = std::getenv("SCOUT_SKIP_COMPACTION");
    if (skip && std::string(skip) == "1") {
      std::cout << "\n[SCOUT_SKIP_COMPACTION=1] Skipping end-of-cycle compaction — on-disk .egr left untouched." << std::endl;
    } else {
      auto c0 = std::chrono::steady_clock::now();
      std::vector<int64_t> off((size_t)n + 1);
      DRV_CHECK(scout_graph_download(ctx, off.data(), nullptr));
      std::vector<int32_t> col((size_t)off[(size_t)n]);
      DRV_CHECK(scout_graph_download(ctx, off.data(), col.data()));
      err = egr_write_wide(args.egr_path, n, off.data(), col.data());
      if (!err.empty()) die("flash_compactor " + err);
      err = egr.remap(args.egr_path.c_str());
      if (!err.empty()) die(err);
      compact_ms = ms_since(c0);
    }
    std::cout << "\n=== FINAL RESULTS ===" << std::endl;
    std::cout << "  Initial MIS cardinality: " << initial_value << std::endl;
    std::cout << "  Final MIS cardinality:   " << final_value << std::endl;
    std::cout << "  Number of batches:       " << total << std::endl;
    std::cout << "\nEnd-of-cycle compaction time (NOT counted in per-batch numbers above): " << std::fixed
              << std::setprecision(3) << compact_ms << " ms" << std::endl;
    sample_memory();
    double total_mb = (double)mem_total / (1024.0 * 1024.0);
    double peak_pct = total_mb > 0.0 ? peak_delta_mb / total_mb * 100.0 : 0.0;
    std::cout << "Peak program GPU memory: " << std::setprecision(2) << peak_delta_mb << " MB (" << std::setprecision(1)
              << peak_pct << "% of " << std::setprecision(2) << total_mb << " MB total)" << std::endl;
  } else if (prob == Problem::COLORING) {
    DRV_CHECK(scout_state_download(ctx, SCOUT_STATE_COLOR, st32.data(), n));
    int32_t mx = -1;
    for (int32_t c : st32) mx = std::max(mx, c);
    final_value = mx + 1;
    std::cout << "Insertions: " << sum.insertions << " (conflict=" << sum.conflicts << ")" << std::endl;
    std::cout << "Deletions:  " << sum.deletions << " (always a structural noop for coloring)" << std::endl;
    std::cout << "Speculative rounds: total=" << sum.rounds << std::endl;
  } else if (prob == Problem::BFS) {
    DRV_CHECK(scout_state_download(ctx, SCOUT_STATE_LEVEL, st32.data(), n));
    for (int32_t v = 0; v < n; v++) final_value += st32[(size_t)v] != -1;
    std::cout << "Insertions: " << sum.insertions << std::endl;
    std::cout << "Deletions:  " << sum.deletions << " (fallback=" << sum.fallback << ")" << std::endl;
  } else {
    DRV_CHECK(scout_state_download(ctx, SCOUT_STATE_LABEL, st32.data(), n));
    for (int32_t v = 0; v < n; v++) final_value += st32[(size_t)v] == v;
    std::cout << "Insertions: " << sum.insertions << std::endl;
    std::cout << "Deletions:  " << sum.deletions << " (definitive_split=" << sum.splits << ", fallback=" << sum.fallback
              << ")" << std::endl;
  }
  const char* label = prob == Problem::MIS ? "MIS size" : prob == Problem::COLORING ? "colors used"
                      : prob == Problem::BFS ? "reached vertices" : "components";
  print_final_summary(setup_ms, processing_ms, setup_ms + processing_ms + compact_ms, label, initial_value, label,
                      final_value);
  print_throughput(lines_total, processing_ms, batch_ms);
  if (touched_path) {
    int64_t cnt = 0;
    DRV_CHECK(scout_touched_download(ctx, nullptr, 0, &cnt));
    std::vector<int32_t> ids((size_t)cnt);
    DRV_CHECK(scout_touched_download(ctx, ids.data(), cnt, &cnt));
    FILE* f = fopen(touched_path, "wb");
    if (!f || fwrite(ids.data(), sizeof(int32_t), ids.size(), f) != ids.size()) die(std::string("cannot write ") + touched_path);
    fclose(f);
  }
  std::cout.flush();
  fflush(stdout);
  std::_Exit(0);
}

}  // namespace scout
