// ref_harness.cpp -- C entry points around the reference's own pure-host headers, compiled
// from the reference sources where they lie (oracle/ref_build.sh) into oracle/_ref/. Test
// infrastructure only: the tests use it to check the oracle's scratchpad restatement, the
// compactor and the report lines against the reference code itself.
//
// Wrapped: HybridGraph / DeltaOverlay (core/FlashGraph.h:200-390), compact_flash_graph
// (core/flash_compactor.h:23-119), print_* (core/reporting.h:22-62).
#include <cstdint>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "FlashGraph.h"
#include "flash_compactor.h"
#include "reporting.h"

extern "C" {

void* ref_open(const char* egr) {
  auto* g = new HybridGraph();
  g->open(egr);
  return g;
}
void ref_close(void* h) { delete static_cast<HybridGraph*>(h); }
int32_t ref_n(void* h) { return static_cast<HybridGraph*>(h)->num_nodes; }
int ref_is_adjacent(void* h, int32_t u, int32_t v) { return static_cast<HybridGraph*>(h)->isAdjacent(u, v) ? 1 : 0; }
void ref_insert(void* h, int32_t u, int32_t v) { static_cast<HybridGraph*>(h)->insertEdge(u, v); }
void ref_remove(void* h, int32_t u, int32_t v) { static_cast<HybridGraph*>(h)->removeEdge(u, v); }
int64_t ref_degree(void* h, int32_t v) { return static_cast<HybridGraph*>(h)->degree(v); }
int64_t ref_row(void* h, int32_t v, int32_t* out, int64_t cap) {
  std::vector<int> row;
  static_cast<HybridGraph*>(h)->neighbors_into(v, row);
  for (int64_t i = 0; i < (int64_t)row.size() && i < cap; i++) out[i] = row[i];
  return (int64_t)row.size();
}
// Classification of a batch as parse_batch_file does it (core/cli_setup.h:87-103): label by
// isAdjacent at parse time, then the caller applies.
void ref_classify_apply(void* h, const int32_t* uv, int64_t m, uint8_t* ins) {
  auto* g = static_cast<HybridGraph*>(h);
  for (int64_t i = 0; i < m; i++) ins[i] = g->isAdjacent(uv[2 * i], uv[2 * i + 1]) ? 0 : 1;
  for (int64_t i = 0; i < m; i++) {
    if (ins[i]) g->insertEdge(uv[2 * i], uv[2 * i + 1]);
    else g->removeEdge(uv[2 * i], uv[2 * i + 1]);
  }
}
void ref_compact(void* h, const char* out_path) {
  auto* g = static_cast<HybridGraph*>(h);
  compact_flash_graph(g->base, g->delta, out_path);
}

// A fixed sequence of report calls; the text goes into buf (returns its length).
int64_t ref_report_sample(char* buf, int64_t cap) {
  std::ostringstream os;
  std::streambuf* old = std::cout.rdbuf(os.rdbuf());
  std::ios::fmtflags flags = std::cout.flags();
  std::streamsize prec = std::cout.precision();
  print_run_header("BFS", "datasets/sample/sample.egr", 2000, 16000, 8.0, "AMD Instinct MI355X", false);
  print_batch_result(1, 3, 1000, 12.3456, {{"insertions", "10"}, {"deletions", "990"}});
  print_stage_times(1, 0.5, 0.25, 1.125, 7.0625, 0.001);
  print_batch_result(2, 3, 0, 0.0);
  print_final_summary(1234.5678, 99.9, 1334.4678, "reached", 1999, "reached", 1987);
  std::cout.flags(flags);
  std::cout.precision(prec);
  std::cout.rdbuf(old);
  std::string s = os.str();
  int64_t n = (int64_t)s.size();
  for (int64_t i = 0; i < n && i < cap; i++) buf[i] = s[i];
  return n;
}

}  // extern "C"
