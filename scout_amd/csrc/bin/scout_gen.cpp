// scout_gen: synthetic graphs (.egr), initial-state files and update-stream folders.
//
//   scout_gen er    <n> <m> <seed> <out.egr>
//   scout_gen rmat  <scale> <edge_factor> <seed> <out.egr> [a b c]
//   scout_gen ba    <n> <m> <seed> <out.egr>
//   scout_gen init  <mis|coloring|bfs|cc> <graph.egr> <out.txt>
//   scout_gen stream <mixed|tree|degree> <graph.egr> <out_dir> <batches> <size> <seed> [delete_frac] [tree_frac]
//   scout_gen convert <in.mtx> <out.egr>     (the reference's .mtx -> wide .egr converter)
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <string>

#include "../host/cli.h"
#include "../host/gen.h"
#include "../host/mtx.h"

using namespace scout;

static int usage() {
  fprintf(stderr,
          "usage:\n  scout_gen er <n> <m> <seed> <out.egr>\n  scout_gen rmat <scale> <edge_factor> <seed> <out.egr> [a b c]\n"
          "  scout_gen ba <n> <m> <seed> <out.egr>\n  scout_gen init <mis|coloring|bfs|cc> <graph.egr> <out.txt>\n"
          "  scout_gen stream <mixed|tree|degree> <graph.egr> <out_dir> <batches> <size> <seed> [delete_frac] [tree_frac]\n"
          "  scout_gen convert <in.mtx> <out.egr>\n");
  return 1;
}

static bool load(const char* path, Csr& g) {
  EgrFile f;
  std::string err = f.open(path);
  if (!err.empty()) {
    fprintf(stderr, "ERROR: %s\n", err.c_str());
    return false;
  }
  g.n = f.nodes;
  f.offsets64(g.off);
  g.col.assign(f.nlist(), f.nlist() + f.edges);
  return true;
}

int main(int argc, char** argv) {
  if (argc < 2) return usage();
  std::string cmd = argv[1];
  std::string err;
  if (cmd == "er" && argc >= 6) {
    Csr g = gen_er((int32_t)atol(argv[2]), atoll(argv[3]), strtoull(argv[4], nullptr, 10));
    err = write_egr(g, argv[5]);
  } else if (cmd == "rmat" && argc >= 6) {
    double a = argc >= 9 ? atof(argv[6]) : 0.57, b = argc >= 9 ? atof(argv[7]) : 0.19,
           c = argc >= 9 ? atof(argv[8]) : 0.19;
    Csr g = gen_rmat(atoi(argv[2]), atoi(argv[3]), a, b, c, strtoull(argv[4], nullptr, 10));
    err = write_egr(g, argv[5]);
  } else if (cmd == "ba" && argc >= 6) {
    Csr g = gen_ba((int32_t)atol(argv[2]), atoi(argv[3]), strtoull(argv[4], nullptr, 10));
    err = write_egr(g, argv[5]);
  } else if (cmd == "init" && argc >= 5) {
    Csr g;
    if (!load(argv[3], g)) return 1;
    std::string kind = argv[2], body;
    if (kind == "mis") body = mis_init_text(greedy_mis(g));
    else if (kind == "coloring") body = per_vertex_text(greedy_color(g));
    else if (kind == "bfs") body = bfs_init_text(bfs_levels(g, 0));
    else if (kind == "cc") body = per_vertex_text(cc_labels(g));
    else return usage();
    if (!write_text(argv[4], body)) err = std::string("cannot write ") + argv[4];
  } else if (cmd == "stream" && argc >= 8) {
    Csr g;
    if (!load(argv[3], g)) return 1;
    StreamSpec s;
    s.kind = argv[2];
    s.batches = atoi(argv[5]);
    s.size = atoll(argv[6]);
    s.seed = strtoull(argv[7], nullptr, 10);
    if (argc >= 9) s.delete_frac = atof(argv[8]);
    if (argc >= 10) s.tree_frac = atof(argv[9]);
    std::error_code ec;
    std::filesystem::create_directories(argv[4], ec);
    err = gen_stream(g, argv[4], s);
  } else if (cmd == "convert" && argc >= 4) {
    err = mtx_to_egr(argv[2], argv[3]);
  } else if (cmd == "parse" && argc >= 4) {
    // Prints the 0-indexed pairs the binaries read from a batch file (cli.h rules).
    std::vector<int32_t> uv;
    if (!parse_batch_pairs(argv[2], (int32_t)atol(argv[3]), uv, err)) {
      fprintf(stderr, "ERROR: %s\n", err.c_str());
      return 1;
    }
    for (size_t i = 0; i + 1 < uv.size(); i += 2) printf("%d %d\n", uv[i], uv[i + 1]);
  } else {
    return usage();
  }
  if (!err.empty()) {
    fprintf(stderr, "ERROR: %s\n", err.c_str());
    return 1;
  }
  return 0;
}
