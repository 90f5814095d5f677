// Prints a fixed sequence of report lines through scout_amd/csrc/host/report.h; the test
// compares the text with the same sequence printed by the reference's core/reporting.h.
#include "../scout_amd/csrc/host/report.h"

int main() {
  using scout::KV;
  scout::print_run_header("BFS", "datasets/sample/sample.egr", 2000, 16000, 8.0, "AMD Instinct MI355X");
  scout::print_batch_result(1, 3, 1000, 12.3456, {KV{"insertions", "10"}, KV{"deletions", "990"}});
  scout::print_stage_times(1, 0.5, 0.25, 1.125, 7.0625, 0.001);
  scout::print_batch_result(2, 3, 0, 0.0);
  scout::print_final_summary(1234.5678, 99.9, 1334.4678, "reached", 1999, "reached", 1987);
  return 0;
}
