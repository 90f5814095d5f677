// cli.h: the command-line and file-format boundary shared by the four scout_* binaries
// and the CPU oracle binaries.
//
//   argv:   <graph.egr> <initial_<algo>_context.txt> <batch_folder> <num_batches> <gpu_id>
//           (core/cli_setup.h:32-45: fewer than 5 arguments prints the usage line, exit 1)
//   batches: every regular file of the folder, ordered by the integer value of its stem
//           (core/cli_setup.h:64-74)
//   lines:  "u v" 1-indexed; converted to 0-indexed; out-of-range and self-loop lines are
//           dropped; anything after the second integer is ignored; a line whose first two
//           tokens are not integers is skipped (core/cli_setup.h:87-103 with the
//           std::istringstream >> int rules)
//   init:   every integer of the file in order (core/cli_setup.h:110-131)
#pragma once

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <string>
#include <vector>

namespace scout {

struct CliArgs {
  std::string egr_path, context_path, batch_dir;
  int num_batches = 0;
  int gpu_id = 0;
};

// std::stoi semantics without exceptions: leading whitespace, optional sign, digits; the
// rest of the string is ignored. Returns false when no integer can be read or it overflows.
inline bool stoi_like(const std::string& s, int& out) {
  const char* p = s.c_str();
  while (*p && std::isspace((unsigned char)*p)) p++;
  errno = 0;
  char* end = nullptr;
  long v = std::strtol(p, &end, 10);
  if (end == p || errno == ERANGE || v < INT_MIN || v > INT_MAX) return false;
  out = (int)v;
  return true;
}

inline bool parse_cli(int argc, char** argv, const char* algo, CliArgs& a) {
  if (argc < 6) {
    std::cerr << "Usage: " << argv[0] << " <graph.egr> <initial_" << algo
              << "_context.txt> <batch_folder> <num_batches> <gpu_device_id>" << std::endl;
    return false;
  }
  a.egr_path = argv[1];
  a.context_path = argv[2];
  a.batch_dir = argv[3];
  if (!stoi_like(argv[4], a.num_batches) || !stoi_like(argv[5], a.gpu_id)) {
    std::cerr << "ERROR: <num_batches> and <gpu_device_id> must be integers" << std::endl;
    return false;
  }
  return true;
}

// Regular files of `dir`, sorted by the integer value of their stem. A stem that is not an
// integer is an error (the reference's std::stoi would throw there).
inline bool list_batch_files(const std::string& dir, std::vector<std::filesystem::path>& out,
                             std::string& err) {
  out.clear();
  std::error_code ec;
  std::vector<std::pair<int, std::filesystem::path>> keyed;
  for (const auto& e : std::filesystem::directory_iterator(dir, ec)) {
    if (!e.is_regular_file()) continue;
    int k;
    if (!stoi_like(e.path().stem().string(), k)) {
      err = "batch file stem is not an integer: " + e.path().string();
      return false;
    }
    keyed.emplace_back(k, e.path());
  }
  if (ec) {
    err = "cannot list batch folder " + dir + ": " + ec.message();
    return false;
  }
  std::stable_sort(keyed.begin(), keyed.end(),
                   [](const auto& x, const auto& y) { return x.first < y.first; });
  for (auto& kv : keyed) out.push_back(kv.second);
  return true;
}

// Reads one int the way `std::istream >> int` does on a line: skip whitespace, optional
// sign, then at least one digit; the value must fit an int. Advances *pp on success.
inline bool read_int_token(const char*& p, const char* end, int& out) {
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\v' || *p == '\f' || *p == '\r')) p++;
  const char* q = p;
  bool neg = false;
  if (q < end && (*q == '+' || *q == '-')) { neg = (*q == '-'); q++; }
  if (q >= end || *q < '0' || *q > '9') return false;
  long long v = 0;
  while (q < end && *q >= '0' && *q <= '9') {
    v = v * 10 + (*q - '0');
    if (v > (long long)INT_MAX + 1) return false;
    q++;
  }
  if (neg) v = -v;
  if (v < INT_MIN || v > INT_MAX) return false;
  out = (int)v;
  p = q;
  return true;
}

// Parses a whole batch file into 0-indexed (u, v) pairs, keeping only lines whose pair is in
// range and not a self-loop. Classification (insert vs delete) is left to the caller, which
// decides it against the graph state its binary uses.
inline bool parse_batch_pairs(const std::string& path, int32_t num_nodes, std::vector<int32_t>& uv,
                              std::string& err) {
  uv.clear();
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) {
    err = "cannot open batch file " + path;
    return false;
  }
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  std::string buf(sz > 0 ? (size_t)sz : 0, '\0');
  size_t got = sz > 0 ? fread(&buf[0], 1, (size_t)sz, f) : 0;
  fclose(f);
  buf.resize(got);
  const char* p = buf.data();
  const char* end = p + buf.size();
  uv.reserve(got / 8 * 2);
  while (p < end) {
    const char* eol = (const char*)memchr(p, '\n', (size_t)(end - p));
    if (!eol) eol = end;
    const char* q = p;
    int s, d;
    if (read_int_token(q, eol, s) && read_int_token(q, eol, d)) {
      long long u = (long long)s - 1, v = (long long)d - 1;
      if (u >= 0 && u < num_nodes && v >= 0 && v < num_nodes && u != v) {
        uv.push_back((int32_t)u);
        uv.push_back((int32_t)v);
      }
    }
    p = eol + 1;
  }
  return true;
}

// Every integer of the file in order, skipping any other character (strtoll scan).
inline std::vector<long long> read_ints(const std::string& path) {
  std::vector<long long> out;
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return out;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  std::string buf(sz > 0 ? (size_t)sz : 0, '\0');
  size_t got = sz > 0 ? fread(&buf[0], 1, (size_t)sz, f) : 0;
  fclose(f);
  buf.resize(got);
  out.reserve(got / 4);
  const char* p = buf.c_str();
  char* e;
  while (*p) {
    long long v = strtoll(p, &e, 10);
    if (e == p) {
      ++p;
      continue;
    }
    out.push_back(v);
    p = e;
  }
  return out;
}

// Initial-state decoders (the reference's main() of each binary):
//   MIS   1-indexed ids of members                       (src/mis/scout_mis.cu:826-833)
//   color one colour per vertex in order, missing -> 0   (src/coloring/scout_coloring.cu:297-303)
//   BFS   "id level" pairs, 1-indexed, missing -> -1     (src/bfs/scout_bfs.cu:546-555)
//   CC    parent array in vertex order, resolved to roots (src/cc/scout_cc.cu:474-485)
inline void decode_mis_init(const std::vector<long long>& vals, int32_t n, std::vector<uint8_t>& mis,
                            long long& card) {
  mis.assign((size_t)n, 0);
  card = 0;
  for (long long id : vals) {
    long long v = id - 1;
    if (v >= 0 && v < n) {
      mis[(size_t)v] = 1;
      card++;  // counts every listed id, as the reference does
    }
  }
}
inline void decode_color_init(const std::vector<long long>& vals, int32_t n, std::vector<int32_t>& color) {
  color.assign((size_t)n, 0);
  for (int32_t i = 0; i < n && (size_t)i < vals.size(); i++) color[(size_t)i] = (int32_t)vals[(size_t)i];
}
inline void decode_bfs_init(const std::vector<long long>& vals, int32_t n, std::vector<int32_t>& level,
                            long long& reached) {
  level.assign((size_t)n, -1);
  reached = 0;
  for (size_t k = 0; k + 1 < vals.size(); k += 2) {
    long long idx = (int)vals[k] - 1;
    int lv = (int)vals[k + 1];
    if (idx >= 0 && idx < n) {
      level[(size_t)idx] = lv;
      if (lv >= 0) reached++;
    }
  }
}
// Resolves a parent array to roots with the reference's find_root (path compression).
// Out-of-range parents are an error.
inline bool decode_cc_init(const std::vector<long long>& vals, int32_t n, std::vector<int32_t>& root,
                           std::string& err) {
  std::vector<int32_t> parent((size_t)n);
  for (int32_t v = 0; v < n; v++) parent[(size_t)v] = v;
  for (int32_t i = 0; i < n && (size_t)i < vals.size(); i++) {
    long long p = vals[(size_t)i];
    if (p < 0 || p >= n) {
      err = "component file entry out of range at line " + std::to_string(i + 1);
      return false;
    }
    parent[(size_t)i] = (int32_t)p;
  }
  root.resize((size_t)n);
  for (int32_t v = 0; v < n; v++) {
    int32_t r = v;
    size_t guard = 0;
    while (parent[(size_t)r] != r) {
      r = parent[(size_t)r];
      if (++guard > (size_t)n) {
        err = "component file has a parent cycle";
        return false;
      }
    }
    int32_t x = v;
    while (parent[(size_t)x] != r) {
      int32_t nx = parent[(size_t)x];
      parent[(size_t)x] = r;
      x = nx;
    }
    root[(size_t)v] = r;
  }
  return true;
}

}  // namespace scout
