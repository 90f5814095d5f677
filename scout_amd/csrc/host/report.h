// report.h: the stdout contract of the scout_* binaries, byte-compatible with
// core/reporting.h:22-62 (run header, one "Batch i/N | edges= | time=...ms | k=v" line per
// batch, the [STAGE] line, the FINAL SUMMARY block). Lines this build adds
// ([LATENCY], [THROUGHPUT]) come after the reference's own lines and use new prefixes.
#pragma once

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

namespace scout {

struct KV {
  std::string key, value;
};

inline void print_run_header(const std::string& algo, const std::string& graph, long long n, long long e,
                             double avg_degree, const std::string& gpu_name) {
  std::cout << "=== SCOUT " << algo << " ===" << std::endl;
  std::cout << "Graph: " << graph << std::endl;
  std::cout << "  nodes=" << n << " edges=" << e << " avg_degree=" << std::fixed << std::setprecision(2)
            << avg_degree << std::endl;
  std::cout << "GPU: " << gpu_name << " | discrete (explicit-transfer path)" << std::endl;
}

inline void print_batch_result(int idx, int total, long long edges, double ms, const std::vector<KV>& extra = {}) {
  std::cout << "Batch " << idx << "/" << total << " | edges=" << edges << " | time=" << std::fixed
            << std::setprecision(3) << ms << "ms";
  for (const auto& kv : extra) std::cout << " | " << kv.key << "=" << kv.value;
  std::cout << std::endl;
}

inline void print_stage_times(int idx, double fetch_ms, double seed_pack_ms, double h2d_ms, double repair_ms,
                              double commit_ms) {
  std::cout << "[STAGE] batch=" << idx << std::fixed << std::setprecision(3) << " fetch_ms=" << fetch_ms
            << " seed_pack_ms=" << seed_pack_ms << " h2d_ms=" << h2d_ms << " repair_ms=" << repair_ms
            << " commit_ms=" << commit_ms << std::endl;
}

inline void print_final_summary(double setup_ms, double processing_ms, double e2e_ms, const std::string& init_label,
                                long long init_value, const std::string& final_label, long long final_value) {
  std::cout << "\n=== FINAL SUMMARY ===" << std::endl;
  std::cout << "Setup time:      " << std::fixed << std::setprecision(3) << setup_ms << " ms" << std::endl;
  std::cout << "Processing time: " << std::fixed << std::setprecision(3) << processing_ms << " ms" << std::endl;
  std::cout << "End-to-end time: " << std::fixed << std::setprecision(3) << e2e_ms << " ms" << std::endl;
  std::cout << "Initial " << init_label << ": " << init_value << std::endl;
  std::cout << "Final " << final_label << ":   " << final_value << std::endl;
}

// Nearest-rank percentile of per-batch latencies.
inline double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)std::ceil(p / 100.0 * (double)v.size());
  if (k == 0) k = 1;
  return v[std::min(k, v.size()) - 1];
}

inline void print_throughput(long long lines, double processing_ms, const std::vector<double>& batch_ms) {
  double ups = processing_ms > 0 ? (double)lines / (processing_ms / 1000.0) : 0.0;
  std::cout << "[THROUGHPUT] updates=" << lines << " updates_per_s=" << std::fixed << std::setprecision(1) << ups
            << std::endl;
  std::cout << "[LATENCY] p50_ms=" << std::setprecision(3) << percentile(batch_ms, 50.0)
            << " p99_ms=" << percentile(batch_ms, 99.0) << std::endl;
}

}  // namespace scout
