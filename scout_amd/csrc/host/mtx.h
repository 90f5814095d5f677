// mtx.h: Matrix Market edge list (.mtx) -> wide .egr, the role of the reference's converter
// (tools/efficient_converter.py:5-115).
//
// The output format and row order are the converter's:
//   * '%' lines before the size line are skipped; the first other line is the size line, and
//     the vertex count is the larger of its first two numbers;
//   * every later line with at least two fields is an entry `u v` (1-indexed); it goes to row
//     u-1 as the neighbour v-1, in the order the lines appear (rows are not sorted, and
//     nothing is symmetrised: an undirected edge must be listed in both directions);
//     lines with u beyond the vertex count are skipped;
//   * the file is always wide: int32 nodes, int32 -1, int64 entries, int64 nindex[nodes+1],
//     int32 nlist[entries].
// Two passes over the input (degrees, then placement), so the entries never need to be held in
// memory: nlist is written through a shared mapping of the output file.
// Where the converter would stop with an exception (a field that is not an integer, u < 1 in
// a data line, a missing size line), this returns an error.
#pragma once

#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <unistd.h>

namespace scout {

namespace mtx_detail {

// Splits the first two whitespace-separated fields of a line into integers; false if the line
// has fewer than two fields. *bad is set when a field is not an integer.
inline bool two_fields(const char* s, long long* a, long long* b, bool* bad) {
  const char* p = s;
  long long v[2];
  for (int k = 0; k < 2; k++) {
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') p++;
    if (!*p) return false;
    char* end = nullptr;
    errno = 0;
    v[k] = strtoll(p, &end, 10);
    bool sep = *end == 0 || *end == ' ' || *end == '\t' || *end == '\r' || *end == '\n';
    if (end == p || errno != 0 || !sep) {
      *bad = true;
      // skip the field so the count of fields stays right
      while (*p && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\n') p++;
    } else {
      p = end;
    }
  }
  *a = v[0];
  *b = v[1];
  return true;
}

// Walks the data lines after the size line: f(u, v) for every entry line (0-indexed).
template <class F>
std::string for_each_entry(const char* path, int64_t* nodes, F&& f) {
  FILE* in = fopen(path, "r");
  if (!in) return std::string("cannot open ") + path + ": " + strerror(errno);
  char* line = nullptr;
  size_t cap = 0;
  ssize_t len;
  bool header = true;
  std::string err;
  int64_t lineno = 0;
  while ((len = getline(&line, &cap, in)) >= 0) {
    lineno++;
    if (header) {
      if (line[0] == '%') continue;
      header = false;
      long long a = 0, b = 0;
      bool bad = false;
      if (!two_fields(line, &a, &b, &bad) || bad) {
        err = "line " + std::to_string(lineno) + ": the size line needs two integers";
        break;
      }
      *nodes = a > b ? a : b;
      if (*nodes < 0 || *nodes > INT32_MAX) {
        err = "line " + std::to_string(lineno) + ": vertex count out of range";
        break;
      }
      continue;
    }
    long long u = 0, v = 0;
    bool bad = false;
    if (!two_fields(line, &u, &v, &bad)) continue;
    if (bad || u < 1 || v < 1 || v - 1 > INT32_MAX) {
      err = "line " + std::to_string(lineno) + ": an entry needs two positive 32-bit vertex ids";
      break;
    }
    if (u - 1 < *nodes) f((int32_t)(u - 1), (int32_t)(v - 1));
  }
  free(line);
  fclose(in);
  if (err.empty() && header) err = std::string(path) + ": no size line";
  return err;
}

}  // namespace mtx_detail

// Converts `in` (.mtx) to the wide .egr `out`. Returns an empty string on success.
inline std::string mtx_to_egr(const char* in, const char* out) {
  int64_t nodes = 0;
  std::vector<int64_t> deg;
  // pass 1: degrees (the size line fixes the vertex count before any entry)
  std::string err = mtx_detail::for_each_entry(in, &nodes, [&](int32_t u, int32_t) {
    if (deg.empty()) deg.assign((size_t)nodes + 1, 0);
    deg[(size_t)u + 1]++;
  });
  if (!err.empty()) return err;
  if (deg.empty()) deg.assign((size_t)nodes + 1, 0);
  for (size_t i = 1; i < deg.size(); i++) deg[i] += deg[i - 1];  // deg is now nindex
  int64_t edges = deg[(size_t)nodes];
  size_t head = 16 + 8 * ((size_t)nodes + 1);
  size_t bytes = head + 4 * (size_t)edges;
  int fd = ::open(out, O_RDWR | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) return std::string("cannot create ") + out + ": " + strerror(errno);
  if (ftruncate(fd, (off_t)bytes) != 0) {
    ::close(fd);
    return std::string("cannot size ") + out + ": " + strerror(errno);
  }
  void* map = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
  ::close(fd);
  if (map == MAP_FAILED) return std::string("cannot map ") + out + ": " + strerror(errno);
  char* base = static_cast<char*>(map);
  int32_t n32 = (int32_t)nodes, sentinel = -1;
  memcpy(base, &n32, 4);
  memcpy(base + 4, &sentinel, 4);
  memcpy(base + 8, &edges, 8);
  memcpy(base + 16, deg.data(), 8 * ((size_t)nodes + 1));
  int32_t* nlist = reinterpret_cast<int32_t*>(base + head);
  // pass 2: place every entry at its row's next slot, in input order
  std::vector<int64_t> pos(deg.begin(), deg.end() - 1);
  int64_t nodes2 = 0;
  err = mtx_detail::for_each_entry(in, &nodes2, [&](int32_t u, int32_t v) { nlist[pos[(size_t)u]++] = v; });
  if (err.empty() && msync(map, bytes, MS_SYNC) != 0) err = std::string("cannot write ") + out + ": " + strerror(errno);
  munmap(map, bytes);
  return err;
}

}  // namespace scout
