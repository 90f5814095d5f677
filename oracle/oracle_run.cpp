// oracle_run: the CPU oracle driven over a graph, init file and batch folder (test oracle and CPU
// baseline; test infrastructure, never part of the product path).
//
//   oracle_run <mis|coloring|bfs|cc> <graph.egr> <init.txt> <batch_dir> <num_batches>
//              [--per-edit] [--dump DIR] [--max-lines K]
//
// Default: exact per-batch answers (closed forms for bfs/cc, the restated rules for
// coloring/mis). --per-edit (bfs, cc) runs the reference's own per-edit algorithms with its
// deferred fallback (src/bfs/scout_bfs.cu:436-515, src/cc/scout_cc.cu:392-441), settled after
// the stream. --max-lines K stops after K lines (a bounded sample for timing). Prints one
// "Batch i/N | edges= | time=...ms" line per batch and "[THROUGHPUT] updates= updates_per_s=".
#include <chrono>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "../scout_amd/csrc/host/cli.h"
#include "../scout_amd/csrc/host/egr.h"
#include "../scout_amd/csrc/host/report.h"
#include "oracle.h"

using namespace scout;

int main(int argc, char** argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: oracle_run <mis|coloring|bfs|cc> <graph.egr> <init> <batch_dir> <num_batches> "
                    "[--per-edit] [--dump DIR] [--max-lines K]\n");
    return 1;
  }
  std::string prob = argv[1];
  bool per_edit = false;
  std::string dump;
  long long max_lines = -1;
  for (int i = 6; i < argc; i++) {
    if (!strcmp(argv[i], "--per-edit")) per_edit = true;
    else if (!strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
    else if (!strcmp(argv[i], "--max-lines") && i + 1 < argc) max_lines = atoll(argv[++i]);
  }
  orc_graph* g = orc_graph_open(argv[2]);
  orc_graph* base = orc_graph_open(argv[2]);
  if (!g || !base) return 1;
  int32_t n = orc_n(g);
  std::vector<long long> vals = read_ints(argv[3]);
  std::vector<int32_t> st32;
  std::vector<uint8_t> mis;
  long long dummy = 0;
  std::string err;
  if (prob == "mis") decode_mis_init(vals, n, mis, dummy);
  else if (prob == "coloring") decode_color_init(vals, n, st32);
  else if (prob == "bfs") decode_bfs_init(vals, n, st32, dummy);
  else if (prob == "cc") {
    std::vector<int32_t> root;
    if (!decode_cc_init(vals, n, root, err)) {
      fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    st32 = root;
  } else {
    fprintf(stderr, "unknown problem %s\n", prob.c_str());
    return 1;
  }
  orc_refbfs* rb = nullptr;
  orc_refcc* rc = nullptr;
  if (per_edit && prob == "bfs") rb = orc_refbfs_create(g, st32.data(), 0, 200, 1);
  if (per_edit && prob == "cc") rc = orc_refcc_create(g, st32.data(), (double)orc_base_edges(g) / n < 10.0 ? n : 5000, 1);
  double avg = n > 0 ? (double)orc_base_edges(g) / n : 0.0;
  int64_t max_csr = orc_mis_max_csr(avg, 288e9);

  std::vector<std::filesystem::path> paths;
  if (!list_batch_files(argv[4], paths, err)) {
    fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  int total = std::min<int>(atoi(argv[5]), (int)paths.size());
  long long lines = 0;
  double proc_ms = 0;
  std::vector<double> bms;
  for (int k = 0; k < total; k++) {
    std::vector<int32_t> uv;
    if (!parse_batch_pairs(paths[(size_t)k].string(), n, uv, err)) {
      fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    int64_t m = (int64_t)uv.size() / 2;
    if (max_lines >= 0 && lines + m > max_lines) m = std::max<long long>(0, max_lines - lines);
    if (m == 0) break;
    std::vector<uint8_t> ins((size_t)m);
    auto t0 = std::chrono::steady_clock::now();
    orc_classify(prob == "mis" ? base : g, uv.data(), m, ins.data());
    int64_t stats[8] = {0};
    if (prob == "mis") {
      orc_apply_conditional(g, uv.data(), ins.data(), m);
      orc_mis_batch(g, uv.data(), ins.data(), m, max_csr, mis.data(), stats);
    } else if (prob == "coloring") {
      orc_apply(g, uv.data(), ins.data(), m);
      orc_color_batch(g, uv.data(), ins.data(), m, st32.data(), stats);
    } else if (rb) {
      orc_refbfs_batch(rb, uv.data(), ins.data(), m, stats);
    } else if (rc) {
      orc_refcc_batch(rc, uv.data(), ins.data(), m, stats);
    } else {
      orc_apply(g, uv.data(), ins.data(), m);
      if (prob == "bfs") orc_bfs(g, 0, st32.data());
      else orc_cc(g, st32.data());
    }
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    proc_ms += ms;
    bms.push_back(ms);
    lines += m;
    print_batch_result(k + 1, total, m, ms);
    if (!dump.empty()) {
      std::filesystem::create_directories(dump);
      std::string path = dump + "/" + prob + "_" + std::to_string(k + 1) + ".bin";
      FILE* f = fopen(path.c_str(), "wb");
      if (prob == "mis") fwrite(mis.data(), 1, mis.size(), f);
      else fwrite(st32.data(), 4, st32.size(), f);
      fclose(f);
    }
    if (max_lines >= 0 && lines >= max_lines) break;
  }
  if (rb) orc_refbfs_finish(rb, st32.data());
  if (rc) orc_refcc_finish(rc, st32.data());
  print_throughput(lines, proc_ms, bms);
  return 0;
}
