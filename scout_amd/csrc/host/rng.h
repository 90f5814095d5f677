// rng.h: counter-based random numbers shared by the host generators and the GPU
// generators, so both produce the same graph and stream for the same seed.
//
// draw(seed, stream, i) = splitmix64(splitmix64(seed ^ stream * K) + i), where splitmix64 is
// Steele, Lea and Flood's SplitMix64 finaliser (OOPSLA 2014). Uniform integers in [0, n) use
// the multiply-shift map (x >> 32) * n >> 32 of Lemire (2019); its bias is < n / 2^32.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SCOUT_HD __host__ __device__ __forceinline__
#else
#define SCOUT_HD inline
#endif

namespace scout {

SCOUT_HD uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

SCOUT_HD uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return splitmix64(seed ^ (stream * 0xd1342543de82ef95ULL));
}

SCOUT_HD uint64_t draw(uint64_t key, uint64_t i) { return splitmix64(key + i); }

// Uniform in [0, n) for n < 2^32.
SCOUT_HD uint32_t uniform_below(uint64_t x, uint32_t n) { return (uint32_t)(((x >> 32) * (uint64_t)n) >> 32); }

// Named streams, so different uses of one seed never share draws.
enum : uint64_t {
  RNG_RMAT = 1,
  RNG_ER = 2,
  RNG_BA = 3,
  RNG_STREAM_DEL = 4,
  RNG_STREAM_INS = 5,
  RNG_STREAM_TREE = 6,
  RNG_STREAM_SIZE = 7,
  RNG_MLP_INIT = 8,
  RNG_ARRIVAL = 9,
  RNG_MULTI = 10,
};

// R-MAT: one 32-bit uniform per recursion level picks a quadrant with probabilities
// (a, b, c, d); thresholds are fixed-point so host and device agree bit for bit.
struct RmatParams {
  uint32_t ta, tab, tabc;  // cumulative thresholds scaled by 2^32
  int scale;
};

SCOUT_HD RmatParams rmat_params(double a, double b, double c, int scale) {
  RmatParams p;
  p.ta = (uint32_t)(a * 4294967296.0);
  p.tab = (uint32_t)((a + b) * 4294967296.0);
  p.tabc = (uint32_t)((a + b + c) * 4294967296.0);
  p.scale = scale;
  return p;
}

// Edge i of the R-MAT stream with key `key`: returns (u, v) packed as u << 32 | v.
SCOUT_HD uint64_t rmat_edge(const RmatParams& p, uint64_t key, uint64_t i) {
  uint32_t u = 0, v = 0;
  for (int lvl = 0; lvl < p.scale; lvl++) {
    uint64_t x = draw(key, i * (uint64_t)p.scale + (uint64_t)lvl);
    uint32_t r = (uint32_t)(x >> 32);
    uint32_t bu = 0, bv = 0;
    if (r < p.ta) {
    } else if (r < p.tab) {
      bv = 1;
    } else if (r < p.tabc) {
      bu = 1;
    } else {
      bu = 1;
      bv = 1;
    }
    u = (u << 1) | bu;
    v = (v << 1) | bv;
  }
  return ((uint64_t)u << 32) | (uint64_t)v;
}

}  // namespace scout
