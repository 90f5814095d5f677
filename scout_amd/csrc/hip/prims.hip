// prims.hip: device-wide scans and sorts (hipCUB), kept in one translation unit.
#include <hipcub/hipcub.hpp>

#include "internal.h"

namespace scout {

namespace {
template <class Fn>
int with_temp(Ctx& c, Fn fn) {
  size_t bytes = 0;
  SCOUT_TRY(fn(nullptr, bytes));
  SCOUT_CHECK(c.tmp.reserve((int64_t)bytes + 256));
  bytes = (size_t)c.tmp.cap;
  SCOUT_TRY(fn((void*)c.tmp.p, bytes));
  return SCOUT_OK;
}
}  // namespace

int scan_exclusive_i64(Ctx& c, const int64_t* in, int64_t* out, int64_t count) {
  if (count <= 0) return SCOUT_OK;
  return with_temp(c, [&](void* t, size_t& b) {
    return hipcub::DeviceScan::ExclusiveSum(t, b, in, out, count, c.s);
  });
}

int scan_exclusive_i32_to_i64(Ctx& c, const int32_t* in, int64_t* out, int64_t count) {
  if (count <= 0) return SCOUT_OK;
  // Widen on the fly so offsets above 2^31 are exact.
  hipcub::TransformInputIterator<int64_t, hipcub::CastOp<int64_t>, const int32_t*> it(in, hipcub::CastOp<int64_t>());
  return with_temp(c, [&](void* t, size_t& b) {
    return hipcub::DeviceScan::ExclusiveSum(t, b, it, out, count, c.s);
  });
}

namespace {
struct RowLen {
  __host__ __device__ int64_t operator()(const AHead& h) const { return h.len; }
};
}  // namespace

int scan_row_lens(Ctx& c, const AHead* in, int64_t* out, int64_t count) {
  if (count <= 0) return SCOUT_OK;
  hipcub::TransformInputIterator<int64_t, RowLen, const AHead*> it(in, RowLen());
  return with_temp(c, [&](void* t, size_t& b) {
    return hipcub::DeviceScan::ExclusiveSum(t, b, it, out, count, c.s);
  });
}

int sort_pairs_u64(Ctx& c, const uint64_t* kin, uint64_t* kout, const int32_t* vin, int32_t* vout, int64_t count,
                   int end_bit) {
  if (count <= 0) return SCOUT_OK;
  return with_temp(c, [&](void* t, size_t& b) {
    return hipcub::DeviceRadixSort::SortPairs(t, b, kin, kout, vin, vout, (int)count, 0, end_bit, c.s);
  });
}

int sort_keys_u64(Ctx& c, const uint64_t* kin, uint64_t* kout, int64_t count, int begin_bit, int end_bit) {
  if (count <= 0) return SCOUT_OK;
  // hipCUB sorts at most 2^31-1 items per call; larger sorts are split by the caller.
  if (count > (int64_t)0x7fffffff) {
    set_error("sort_keys_u64: more than 2^31-1 keys");
    return SCOUT_ERR_ARG;
  }
  return with_temp(c, [&](void* t, size_t& b) {
    return hipcub::DeviceRadixSort::SortKeys(t, b, kin, kout, (int)count, begin_bit, end_bit, c.s);
  });
}

int reduce_max_i32(Ctx& c, const int32_t* in, int64_t count, int32_t* d_out) {
  if (count <= 0) return SCOUT_OK;
  return with_temp(c, [&](void* t, size_t& b) { return hipcub::DeviceReduce::Max(t, b, in, d_out, count, c.s); });
}

int reduce_sum_i64(Ctx& c, const int64_t* in, int64_t count, int64_t* d_out) {
  if (count <= 0) return SCOUT_OK;
  return with_temp(c, [&](void* t, size_t& b) { return hipcub::DeviceReduce::Sum(t, b, in, d_out, count, c.s); });
}

}  // namespace scout
