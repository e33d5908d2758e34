// Device layout of the HBM join table, shared by the plain join kernels and
// the batched executor: 64-byte buckets of 8 int64 keys (kEmpty = free slot)
// with a parallel int32 row-id array; the bucket of a key is the multiply-high
// of a 64-bit mix of the key and the bucket count; insertion claims the first
// free slot in bucket order with a 64-bit atomicCAS, moving to the next bucket
// when one is full. Replaces runtime::HashTable (runtime/hash_table.cc:13-67).
//
// Narrow form: when every key of the table lies in [base, base + 2^32 - 2],
// the same key buffer holds uint32 offsets key - base in 64-byte buckets of 16
// slots (kEmpty32 = free), half as many buckets for the same slot count. The
// table then takes half the bytes (a 16M-key table at load factor 0.5 is
// 128 MB, inside the 256 MB Infinity Cache), and a probe key outside the range
// misses without touching memory. Row ids stay per slot.
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

constexpr uint32_t kEmpty32 = 0xFFFFFFFFu;
constexpr int kSlots32 = 16;      // slots per 64-byte narrow bucket

struct TableView {
  int64_t* keys;
  int32_t* rows;
  uint64_t nbuckets;              // buckets of 8 wide slots (narrow: nbuckets / 2 buckets of 16)
  int32_t* side_rows;
  unsigned long long* side_count;
  int32_t narrow;
  int32_t reserved;
  int64_t base;                   // narrow: key = base + stored offset
};

__host__ __device__ __forceinline__ uint32_t* keys32(const TableView& t) {
  return reinterpret_cast<uint32_t*>(t.keys);
}

__host__ __device__ __forceinline__ uint64_t nbuckets32(const TableView& t) { return t.nbuckets / 2; }

// Stored offset of k in a narrow table; false when k cannot be in it.
__device__ __forceinline__ bool narrow_key(const TableView& t, int64_t k, uint32_t& k32) {
  const uint64_t d = static_cast<uint64_t>(k) - static_cast<uint64_t>(t.base);
  k32 = static_cast<uint32_t>(d);
  return d < static_cast<uint64_t>(kEmpty32);
}

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

// Insert (k, row) into a wide table; false when every bucket is full. Keys equal
// to kEmpty go to the side list.
__device__ __forceinline__ bool insert_wide(const TableView& t, int64_t k, int32_t row, int* overflow) {
  if (k == kEmpty) {
    const unsigned long long pos = atomicAdd(t.side_count, 1ull);
    if (pos < static_cast<unsigned long long>(kSideCap)) t.side_rows[pos] = row;
    else atomicOr(overflow, 1);
    return true;
  }
  uint64_t b = bucket_of(k, t.nbuckets);
  for (uint64_t probes = 0; probes < t.nbuckets; probes++) {
    for (int s = 0; s < kSlots; s++) {
      if (t.keys[b * kSlots + s] != kEmpty) continue;
      const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(t.keys + b * kSlots + s),
                                                static_cast<unsigned long long>(kEmpty),
                                                static_cast<unsigned long long>(k));
      if (prev == static_cast<unsigned long long>(kEmpty)) {
        t.rows[b * kSlots + s] = row;
        return true;
      }
    }
    b = (b + 1 == t.nbuckets) ? 0 : b + 1;
  }
  atomicOr(overflow, 1);
  return false;
}

// Insert (k, row) into a narrow table (k must be in range; else overflow).
__device__ __forceinline__ bool insert_narrow(const TableView& t, int64_t k, int32_t row, int* overflow) {
  uint32_t k32;
  if (!narrow_key(t, k, k32)) {
    atomicOr(overflow, 1);
    return false;
  }
  uint32_t* keys = keys32(t);
  const uint64_t nb = nbuckets32(t);
  uint64_t b = bucket_of(k, nb);
  for (uint64_t probes = 0; probes < nb; probes++) {
    for (int s = 0; s < kSlots32; s++) {
      if (keys[b * kSlots32 + s] != kEmpty32) continue;
      if (atomicCAS(keys + b * kSlots32 + s, kEmpty32, k32) == kEmpty32) {
        t.rows[b * kSlots32 + s] = row;
        return true;
      }
    }
    b = (b + 1 == nb) ? 0 : b + 1;
  }
  atomicOr(overflow, 1);
  return false;
}

// Calls f(build row) for every slot whose key equals k, in bucket order
// (side-list rows for kEmpty in a wide table first).
template <typename F>
__device__ __forceinline__ void table_matches(const TableView& t, int64_t k, F&& f) {
  if (t.narrow) {
    uint32_t k32;
    if (!narrow_key(t, k, k32)) return;
    const uint32_t* keys = keys32(t);
    const uint64_t nb = nbuckets32(t);
    uint64_t b = bucket_of(k, nb);
    for (uint64_t scanned = 0; scanned < nb; scanned++) {
      const uint4* bk = reinterpret_cast<const uint4*>(keys + b * kSlots32);
      bool empty = false;
#pragma unroll
      for (int q = 0; q < kSlots32 / 4; q++) {
        const uint4 x = bk[q];
        const int64_t s0 = static_cast<int64_t>(b) * kSlots32 + 4 * q;
        if (x.x == k32) f(t.rows[s0]);
        if (x.y == k32) f(t.rows[s0 + 1]);
        if (x.z == k32) f(t.rows[s0 + 2]);
        if (x.w == k32) f(t.rows[s0 + 3]);
        empty |= x.x == kEmpty32 || x.y == kEmpty32 || x.z == kEmpty32 || x.w == kEmpty32;
      }
      if (empty) break;
      b = (b + 1 == nb) ? 0 : b + 1;
    }
    return;
  }
  if (k == kEmpty) {
    unsigned long long ns = *t.side_count;
    if (ns > static_cast<unsigned long long>(kSideCap)) ns = kSideCap;
    for (unsigned long long j = 0; j < ns; j++) f(t.side_rows[j]);
    return;
  }
  uint64_t b = bucket_of(k, t.nbuckets);
  for (uint64_t scanned = 0; scanned < t.nbuckets; scanned++) {
    const longlong2* bk = reinterpret_cast<const longlong2*>(t.keys + b * kSlots);
    bool empty = false;
#pragma unroll
    for (int q = 0; q < kSlots / 2; q++) {
      const longlong2 kv = bk[q];
      if (kv.x == k) f(t.rows[b * kSlots + 2 * q]);
      if (kv.y == k) f(t.rows[b * kSlots + 2 * q + 1]);
      empty |= kv.x == kEmpty || kv.y == kEmpty;
    }
    if (empty) break;
    b = (b + 1 == t.nbuckets) ? 0 : b + 1;
  }
}

}  // namespace kg
