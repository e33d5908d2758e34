// Device layout of the HBM join table, shared by the plain join kernels and
// the batched executor: 64-byte buckets of 8 int64 keys (kEmpty = free slot)
// with a parallel int32 row-id array; the bucket of a key is the multiply-high
// of a 64-bit mix of the key and the bucket count; insertion claims the first
// free slot in bucket order with a 64-bit atomicCAS, moving to the next bucket
// when one is full. Replaces runtime::HashTable (runtime/hash_table.cc:13-67).
#pragma once

#include "kg_ops.h"

namespace kg {

constexpr int64_t kEmpty = INT64_MIN;
constexpr int kSlots = 8;         // slots per 64-byte bucket
constexpr int kProbeLanes = 4;    // lanes cooperating on one bucket
constexpr int kJBlock = 256;
constexpr int64_t kSideCap = 1 << 20;

struct Acc {
  unsigned long long sum;
  unsigned long long count;
};

struct TableView {
  int64_t* keys;
  int32_t* rows;
  uint64_t nbuckets;
  int32_t* side_rows;
  unsigned long long* side_count;
};

__device__ __forceinline__ uint64_t bucket_of(int64_t key, uint64_t nbuckets) {
  return __umul64hi(mix64(static_cast<uint64_t>(key)), nbuckets);
}

// Streamed probe-side key: a non-temporal load, so the key stream does not
// displace the table from the caches.
__device__ __forceinline__ bool load_key_nt(const kg_col& c, int64_t r, int64_t& key) {
  if (c.nulls && c.nulls[r]) return false;
  if (c.type == KG_BIGINT) {
    key = __builtin_nontemporal_load(static_cast<const int64_t*>(c.data) + r);
    return true;
  }
  if (c.type == KG_INT || c.type == KG_DATE || c.type == KG_ENUM) {
    key = __builtin_nontemporal_load(static_cast<const int32_t*>(c.data) + r);
    return true;
  }
  if (c.type == KG_REAL) {
    const double d = __builtin_nontemporal_load(static_cast<const double*>(c.data) + r);
    if (d != d) return false;
    key = d == 0.0 ? 0 : __double_as_longlong(d);
    return true;
  }
  key = load_int(c.type, c.data, r);
  return true;
}

// Join key of row r as 64 bits. REAL keys compare as doubles in the reference
// (Equal, compile/proxy/evaluate.cc:449-453): -0.0 is folded onto +0.0 and a
// NaN key never matches, so it is treated like NULL.
__device__ __forceinline__ bool load_key(const kg_col& c, int64_t r, int64_t& key) {
  if (c.nulls && c.nulls[r]) return false;
  if (c.type == KG_REAL) {
    const double d = static_cast<const double*>(c.data)[r];
    if (d != d) return false;
    key = d == 0.0 ? 0 : __double_as_longlong(d);
    return true;
  }
  key = load_int(c.type, c.data, r);
  return true;
}

}  // namespace kg
