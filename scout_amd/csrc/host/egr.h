// egr.h: the .egr binary CSR file, read through a read-only mmap and written by a
// durable temp-file + fsync + rename.
//
// Layout (little-endian):  int32 nodes;  int32 edges | -1;  [uint64 edges if -1];
//                          nindex[nodes+1] (int32, or uint64 when the -1 sentinel is used);
//                          nlist[edges] (int32).
// Follows core/FlashGraph.h:130-192 (header parse, nindex/nlist byte offsets, MADV_RANDOM
// unless SCOUT_NO_MADV_RANDOM) and core/flash_compactor.h:47-118 (always-wide output,
// atomic replace, directory fsync).
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
#include <sys/stat.h>
#include <unistd.h>

namespace scout {

class EgrFile {
 public:
  int32_t nodes = 0;
  uint64_t edges = 0;
  bool wide = false;

  EgrFile() = default;
  EgrFile(const EgrFile&) = delete;
  EgrFile& operator=(const EgrFile&) = delete;
  ~EgrFile() { close_map(); }

  // Returns an empty string on success, else an error message.
  std::string open(const char* path) {
    Map m;
    std::string err = map_file(path, m);
    if (!err.empty()) return err;
    close_map();
    adopt(m);
    return "";
  }

  // Map the new file first, then drop the old mapping (FlashGraph::remap semantics).
  std::string remap(const char* path) { return open(path); }

  inline int64_t nindex_at(int64_t v) const {
    return wide ? (int64_t)((const uint64_t*)nindex_)[v] : (int64_t)((const int32_t*)nindex_)[v];
  }
  inline int64_t degree(int32_t v) const { return nindex_at(v + 1) - nindex_at(v); }
  inline const int32_t* row(int32_t v) const { return nlist_ + nindex_at(v); }
  const int32_t* nlist() const { return nlist_; }
  const void* nindex_raw() const { return nindex_; }

  // Copies the offsets into a host int64 vector (n+1 entries).
  void offsets64(std::vector<int64_t>& out) const {
    out.resize((size_t)nodes + 1);
    for (int64_t v = 0; v <= nodes; v++) out[(size_t)v] = nindex_at(v);
  }

 private:
  struct Map {
    int fd = -1;
    void* base = nullptr;
    size_t size = 0;
    int32_t nodes = 0;
    uint64_t edges = 0;
    bool wide = false;
    size_t nindex_off = 0, nlist_off = 0;
  };
  int fd_ = -1;
  void* base_ = nullptr;
  size_t size_ = 0;
  const char* nindex_ = nullptr;
  const int32_t* nlist_ = nullptr;

  void adopt(const Map& m) {
    fd_ = m.fd;
    base_ = m.base;
    size_ = m.size;
    nodes = m.nodes;
    edges = m.edges;
    wide = m.wide;
    nindex_ = (const char*)base_ + m.nindex_off;
    nlist_ = (const int32_t*)((const char*)base_ + m.nlist_off);
  }
  void close_map() {
    if (base_ && base_ != MAP_FAILED && size_) munmap(base_, size_);
    if (fd_ >= 0) ::close(fd_);
    base_ = nullptr;
    fd_ = -1;
    size_ = 0;
  }

  static std::string map_file(const char* path, Map& m) {
    int fd = ::open(path, O_RDONLY);
    if (fd < 0) return std::string("could not open ") + path;
    struct stat st;
    if (fstat(fd, &st) != 0) { ::close(fd); return std::string("fstat failed for ") + path; }
    size_t sz = (size_t)st.st_size;
    if (sz < 8) { ::close(fd); return std::string("truncated header in ") + path; }
    int32_t hdr[2];
    if (pread(fd, hdr, 8, 0) != 8) { ::close(fd); return std::string("failed to read header of ") + path; }
    size_t header_end = 8;
    if (hdr[1] == -1) {
      uint64_t e64 = 0;
      if (sz < 16 || pread(fd, &e64, 8, 8) != 8) { ::close(fd); return std::string("failed to read 64-bit edge count from ") + path; }
      m.edges = e64;
      m.wide = true;
      header_end = 16;
    } else {
      if (hdr[0] < 1 || hdr[1] < 0) { ::close(fd); return std::string("invalid header in ") + path; }
      m.edges = (uint64_t)hdr[1];
      m.wide = false;
    }
    m.nodes = hdr[0];
    if (m.nodes < 0) { ::close(fd); return std::string("invalid node count in ") + path; }
    size_t stride = m.wide ? 8 : 4;
    m.nindex_off = header_end;
    m.nlist_off = header_end + ((size_t)m.nodes + 1) * stride;
    size_t need = m.nlist_off + (size_t)m.edges * 4;
    if (need > sz) { ::close(fd); return std::string("file shorter than its header promises: ") + path; }
    void* p = mmap(nullptr, sz, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) { ::close(fd); return std::string("mmap failed for ") + path; }
    if (!std::getenv("SCOUT_NO_MADV_RANDOM")) madvise(p, sz, MADV_RANDOM);
    m.fd = fd;
    m.base = p;
    m.size = sz;
    return "";
  }
};

// Writes a wide .egr (sentinel -1 + uint64 count + uint64 nindex) durably: temp file in the
// same directory, fflush + fsync, rename over `path`, then fsync of the directory.
// `row_writer(v, FILE*)` is not used here: callers pass a flat CSR.
inline std::string egr_write_wide(const std::string& path, int32_t nodes, const int64_t* off,
                                  const int32_t* col) {
  const std::string tmp = path + ".tmp." + std::to_string(getpid());
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return "could not open " + tmp + " for writing";
  const int32_t sentinel = -1;
  const uint64_t edges = (uint64_t)off[nodes];
  bool ok = fwrite(&nodes, 4, 1, f) == 1 && fwrite(&sentinel, 4, 1, f) == 1 && fwrite(&edges, 8, 1, f) == 1;
  // int64 and uint64 offsets share a bit pattern for non-negative values.
  ok = ok && fwrite(off, 8, (size_t)nodes + 1, f) == (size_t)nodes + 1;
  ok = ok && (edges == 0 || fwrite(col, 4, (size_t)edges, f) == (size_t)edges);
  ok = ok && fflush(f) == 0 && fsync(fileno(f)) == 0;
  fclose(f);
  if (!ok) { unlink(tmp.c_str()); return "failed writing " + tmp; }
  if (rename(tmp.c_str(), path.c_str()) != 0) { unlink(tmp.c_str()); return "rename to " + path + " failed"; }
  std::string dir = ".";
  size_t slash = path.find_last_of('/');
  if (slash != std::string::npos) dir = path.substr(0, slash);
  int dfd = ::open(dir.c_str(), O_RDONLY);
  if (dfd >= 0) { fsync(dfd); ::close(dfd); }
  return "";
}

// Narrow writer (int32 header count and int32 nindex), for graphs with < 2^31 entries.
inline std::string egr_write_narrow(const std::string& path, int32_t nodes, const int64_t* off,
                                    const int32_t* col) {
  if (off[nodes] > INT32_MAX) return "narrow .egr cannot hold more than 2^31-1 entries";
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return "could not open " + path + " for writing";
  int32_t e32 = (int32_t)off[nodes];
  bool ok = fwrite(&nodes, 4, 1, f) == 1 && fwrite(&e32, 4, 1, f) == 1;
  std::vector<int32_t> n32((size_t)nodes + 1);
  for (size_t i = 0; i <= (size_t)nodes; i++) n32[i] = (int32_t)off[i];
  ok = ok && fwrite(n32.data(), 4, n32.size(), f) == n32.size();
  ok = ok && (e32 == 0 || fwrite(col, 4, (size_t)e32, f) == (size_t)e32);
  ok = ok && fclose(f) == 0;
  return ok ? "" : "failed writing " + path;
}

}  // namespace scout
