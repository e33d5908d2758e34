// Column-index keys and directory lookups shared by the index kernels
// (kg_index.hip) and the index-nested-loop expansion (kg_text.hip).
//
// An index keys its tuple ids by a 64-bit value, encoded one of three ways:
//  * I32: 8/16/32-bit integer columns (BOOLEAN, SMALLINT, INT, DATE, ENUM);
//    sorted on 32 bits;
//  * I64: BIGINT; sorted on 64 bits;
//  * F64: REAL, or any integer column joined with a REAL one: the value as a
//    double (integers converted), -0.0 folded onto +0.0, NaN never indexed and
//    never found — the reference compares such keys as doubles
//    (compile/proxy/evaluate.cc:449-453).
// A NULL key is neither indexed nor found (permutable_skinner_join_translator
// .cc:972-994 indexes only non-NULL keys; NULL = x is not TRUE).
#pragma once

#include "kg_ops.h"

namespace kg {

enum IndexEnc { kEncI32 = 0, kEncI64 = 1, kEncF64 = 2 };

constexpr int kDirSlots = 8;
constexpr int64_t kDirEmpty = INT64_MIN;

// Directory from key to its run of tuple ids: dense (dspan > 0: offset/count
// indexed by key - dmin, I32/I64 only) or hashed (8-slot buckets of keys, free
// = kDirEmpty; the run of the key INT64_MIN itself, which cannot live in a
// slot, is min_off/min_cnt).
struct IndexDir {
  const int64_t* keys;
  const int64_t* offset;
  const int32_t* count;
  uint64_t nbuckets;
  int64_t dmin;
  uint64_t dspan;
  int64_t min_off;
  int32_t min_cnt;
  int32_t enc;
};

__host__ __device__ inline int index_enc_of(int32_t type) {
  return type == KG_REAL ? kEncF64 : (type == KG_BIGINT ? kEncI64 : kEncI32);
}

__device__ __forceinline__ int64_t f64_key(double x) {
  return __double_as_longlong(x == 0.0 ? 0.0 : x);
}

// Index key of row r of column c under encoding enc; false when the row has
// no key (NULL, NaN, or a REAL value that no integer key can equal).
__device__ __forceinline__ bool index_key(const kg_col& c, int64_t r, int enc, int64_t& k) {
  if (c.nulls && c.nulls[r]) return false;
  if (enc == kEncF64) {
    const double x = c.type == KG_REAL ? static_cast<const double*>(c.data)[r]
                                       : static_cast<double>(load_int(c.type, c.data, r));
    if (x != x) return false;
    k = f64_key(x);
    return true;
  }
  if (c.type == KG_REAL) {  // a REAL probe into an integer index: only integral values can match
    const double x = static_cast<const double*>(c.data)[r];
    if (!(x >= -9.2233720368547758e18 && x < 9.2233720368547758e18)) return false;  // NaN, out of range
    k = static_cast<int64_t>(x);
    return static_cast<double>(k) == x;
  }
  k = load_int(c.type, c.data, r);
  return true;
}

__device__ __forceinline__ uint64_t dir_bucket(int64_t key, uint64_t nbuckets) {
  return __umul64hi(mix64(static_cast<uint64_t>(key)), nbuckets);
}

// The run (offset into the tid array, count) of key k; count 0 when absent.
__device__ __forceinline__ void index_find(const IndexDir& d, int64_t k, int64_t& off, int32_t& cnt) {
  off = 0;
  cnt = 0;
  if (d.dspan) {
    const uint64_t u = static_cast<uint64_t>(k) - static_cast<uint64_t>(d.dmin);
    if (k >= d.dmin && u < d.dspan) {
      cnt = d.count[u];
      if (cnt) off = d.offset[u];
    }
    return;
  }
  if (k == kDirEmpty) {
    off = d.min_off;
    cnt = d.min_cnt;
    return;
  }
  uint64_t b = dir_bucket(k, d.nbuckets);
  for (uint64_t probes = 0; probes < d.nbuckets; probes++) {
    const longlong2* bk = reinterpret_cast<const longlong2*>(d.keys + b * kDirSlots);
    bool empty = false;
#pragma unroll
    for (int q = 0; q < kDirSlots / 2; q++) {
      const longlong2 kv = bk[q];
      if (kv.x == k) {
        off = d.offset[b * kDirSlots + 2 * q];
        cnt = d.count[b * kDirSlots + 2 * q];
        return;
      }
      if (kv.y == k) {
        off = d.offset[b * kDirSlots + 2 * q + 1];
        cnt = d.count[b * kDirSlots + 2 * q + 1];
        return;
      }
      empty |= kv.x == kDirEmpty || kv.y == kDirEmpty;
    }
    if (empty) return;
    b = (b + 1 == d.nbuckets) ? 0 : b + 1;
  }
}

// Builds an index over the non-NULL keys of rows d_rows[0..n) (or 0..n-1) of
// `key` with an explicit encoding (kg_index_build chooses it from the type).
int index_build(kg_ctx* ctx, const kg_col* key, const int32_t* d_rows, int64_t n, int enc, kg_index** out);

// The directory and tid array of an index.
void index_view(const kg_index* idx, IndexDir* dir, const int32_t** tids);

// The encoding an index was built with.
int index_enc(const kg_index* idx);

}  // namespace kg
