// ORACLE PIN — test infrastructure. A thin extern "C" harness over the
// reference's own runtime sources (compiled in place from the reference
// checkout by oracle/Makefile into oracle/_ref/libkushref.so). Only tests/
// load it, to check the oracle restatement (kush_oracle.cc) against the
// reference's compiled code where that code builds standalone.
//
// The harness drives:
//   runtime/hash_table.cc          HashCombine, Create/Insert/GetBucket/Free
//   runtime/vector.cc              bucket vectors; Sort (the ORDER BY merge
//                                  sort) over row positions
//   runtime/aggregate_hash_table.cc Init/Resize/AllocateNewPage/ConstructEntry/
//                                  GetEntry/GetPayload/Size/Free
//   runtime/column_index_bucket.cc FastForwardBucket, sorted intersection
//   runtime/tuple_idx_table/*      ART Insert / iteration
//   runtime/date.cc                BuildDate / SplitDate
// The generated-code halves (the probe loop of
// compile/proxy/aggregate_hash_table.cc:187-243 and the bucket loop of
// compile/translators/hash_join_translator.cc:105-139) exist only as IR
// builder code; the harness drives the runtime calls in the order that IR
// emits them.
#include <cstdint>
#include <cstring>
#include <vector>

#include "runtime/aggregate_hash_table.h"
#include "runtime/column_index_bucket.h"
#include "runtime/date.h"
#include "runtime/hash_table.h"
#include "runtime/tuple_idx_table/tuple_idx_table.h"
#include "runtime/vector.h"

using namespace kush::runtime;

extern "C" {

void ref_hash_combine(int32_t* hash, int64_t v) { HashTable::HashCombine(hash, v); }

int32_t ref_build_date(int32_t y, int32_t m, int32_t d) {
  return static_cast<int32_t>(Date::DateBuilder(y, m, d).Build());
}

void ref_split_date(int32_t jd, int32_t* y, int32_t* m, int32_t* d) { Date::SplitDate(jd, y, m, d); }

int32_t ref_fast_forward(const int32_t* data, int32_t size, int32_t prev) {
  ColumnIndexBucket b{const_cast<int32_t*>(data), size, 0, nullptr};
  return FastForwardBucket(&b, prev);
}

// Full drain of the k-way intersection in chunks, as the Skinner table
// functions do: Init at next_tuple, then PopulateResult until it returns 0.
int64_t ref_intersect_all(int32_t nlists, const int32_t* const* lists, const int32_t* sizes,
                          int32_t next_tuple, int32_t chunk, int32_t* out, int64_t out_cap) {
  std::vector<ColumnIndexBucket> b(nlists);
  for (int32_t i = 0; i < nlists; i++) b[i] = {const_cast<int32_t*>(lists[i]), sizes[i], 0, nullptr};
  std::vector<int32_t> state(nlists > 0 ? nlists : 1);
  BucketListSortedIntersectionInit(b.data(), nlists, state.data(), next_tuple);
  std::vector<int32_t> buf(chunk);
  int64_t total = 0;
  while (true) {
    int32_t got = BucketListSortedIntersectionPopulateResult(b.data(), nlists, state.data(), buf.data(), chunk);
    if (got == 0) break;
    for (int32_t i = 0; i < got; i++) {
      if (total < out_cap) out[total] = buf[i];
      total++;
    }
  }
  return total;
}

int64_t ref_tuple_dedup(int64_t n, int32_t len, const int32_t* tuples, uint8_t* is_new,
                        int32_t* sorted_out) {
  auto* t = TupleIdxTable::Create();
  for (int64_t i = 0; i < n; i++) {
    bool fresh = TupleIdxTable::Insert(t, tuples + i * len, len);
    if (is_new) is_new[i] = fresh;
  }
  auto* it = TupleIdxTable::CreateIt();
  int64_t count = 0;
  if (TupleIdxTable::Begin(t, it)) {
    do {
      if (sorted_out) std::memcpy(sorted_out + count * len, TupleIdxTable::Get(it), len * 4);
      count++;
    } while (TupleIdxTable::IteratorNext(t, it));
  }
  TupleIdxTable::FreeIt(it);
  TupleIdxTable::Free(t);
  return count;
}

// Join on one BIGINT key: build rows are int64 keys packed into the bucket
// vectors (element = {int64 key, int32 row}), probe walks GetBucket and
// compares keys; emits (probe row, build row) pairs.
int64_t ref_hash_join_i64(int64_t nb, const int64_t* bkeys, int64_t np, const int64_t* pkeys,
                          int64_t out_cap, int32_t* out_probe, int32_t* out_build) {
  struct Elem {
    int64_t key;
    int32_t row;
  };
  HashTable::HashTable ht;
  HashTable::Create(&ht, sizeof(Elem));
  for (int64_t i = 0; i < nb; i++) {
    int32_t h = 0;
    HashTable::HashCombine(&h, bkeys[i]);
    auto* slot = reinterpret_cast<Elem*>(HashTable::Insert(&ht, h));
    slot->key = bkeys[i];
    slot->row = static_cast<int32_t>(i);
  }
  int64_t out = 0;
  for (int64_t j = 0; j < np; j++) {
    int32_t h = 0;
    HashTable::HashCombine(&h, pkeys[j]);
    Vector::Vector* bucket = HashTable::GetBucket(&ht, h);
    for (int32_t k = 0; k < Vector::Size(bucket); k++) {
      auto* e = reinterpret_cast<Elem*>(Vector::Get(bucket, k));
      if (e->key != pkeys[j]) continue;
      if (out < out_cap) {
        out_probe[out] = static_cast<int32_t>(j);
        out_build[out] = e->row;
      }
      out++;
    }
  }
  HashTable::Free(&ht);
  return out;
}

// GroupBy(one BIGINT key; SUM(BIGINT), COUNT(*)) through the reference's
// AggregateHashTable, with the generated UpdateOrInsert loop driven here.
// Payload = {int64 hash; int64 key; int64 sum; int64 count} (32 B, hash at 0).
// Emits groups by walking payload blocks in order (ForEach).
int64_t ref_group_by_i64(int64_t n, const int64_t* keys, const int64_t* vals, int64_t max_groups,
                         int64_t* out_key, int64_t* out_sum, int64_t* out_cnt) {
  struct Payload {
    int64_t hash;
    int64_t key;
    int64_t sum;
    int64_t count;
  };
  AggregateHashTable::AggregateHashTable ht;
  AggregateHashTable::Init(&ht, sizeof(Payload), 0);
  const uint64_t m = 0xc6a4a7935bd1e995ull;
  for (int64_t i = 0; i < n; i++) {
    if (ht.size == ht.capacity ||
        static_cast<double>(ht.size) > static_cast<double>(ht.capacity) / AggregateHashTable::LOAD_FACTOR)
      AggregateHashTable::Resize(&ht);
    uint64_t k = static_cast<uint64_t>(keys[i]);
    k *= m;
    k ^= k >> 47;
    k *= m;
    uint64_t hash = 0;
    hash ^= k;
    hash *= m;
    hash += 0xe6546b64ull;
    const uint16_t salt = static_cast<uint16_t>(hash >> 48);
    uint64_t idx = hash & ht.mask;
    while (true) {
      auto* entry = static_cast<uint64_t*>(AggregateHashTable::GetEntry(&ht, static_cast<uint32_t>(idx)));
      const uint32_t block_idx = static_cast<uint32_t>(*entry & 0xFFFFFFFFull);
      const uint16_t block_offset = static_cast<uint16_t>((*entry >> 32) & 0xFFFF);
      const uint16_t esalt = static_cast<uint16_t>(*entry >> 48);
      if (block_idx == 0) {
        ht.size++;
        if (ht.last_payload_offset + ht.payload_size > AggregateHashTable::BLOCK_SIZE)
          AggregateHashTable::AllocateNewPage(&ht);
        const uint32_t bidx = ht.payload_block_size - 1;
        const uint16_t off = ht.last_payload_offset;
        ht.last_payload_offset = off + ht.payload_size;
        *entry = AggregateHashTable::ConstructEntry(salt, off, bidx);
        auto* p = static_cast<Payload*>(AggregateHashTable::GetPayload(&ht, bidx, off));
        p->hash = static_cast<int64_t>(hash);
        p->key = keys[i];
        p->sum = vals[i];
        p->count = 1;
        break;
      }
      if (esalt == salt) {
        auto* p = static_cast<Payload*>(AggregateHashTable::GetPayload(&ht, block_idx, block_offset));
        if (p->key == keys[i]) {
          p->sum = static_cast<int64_t>(static_cast<uint64_t>(p->sum) + static_cast<uint64_t>(vals[i]));
          p->count += 1;
          break;
        }
      }
      idx = (idx + 1) & ht.mask;
    }
  }
  const int32_t ng = AggregateHashTable::Size(&ht);
  int64_t g = 0;
  for (uint32_t b = 1; b < ht.payload_block_size; b++) {
    const uint32_t end = (b == ht.payload_block_size - 1) ? ht.last_payload_offset : AggregateHashTable::BLOCK_SIZE;
    for (uint32_t off = 0; off + ht.payload_size <= end; off += ht.payload_size) {
      auto* p = static_cast<Payload*>(AggregateHashTable::GetPayload(&ht, b, off));
      if (g < max_groups) {
        out_key[g] = p->key;
        out_sum[g] = p->sum;
        out_cnt[g] = p->count;
      }
      g++;
    }
  }
  AggregateHashTable::Free(&ht);
  return ng == g ? ng : -1;
}

// ORDER BY's sort (compile/translators/order_by_translator.cc:50-80): the
// reference's Vector::Sort over 4-byte row positions with a comparison
// function that walks the keys as the generated comparator does, NULLs
// ordered by compile/proxy/evaluate.cc:339-395 (NULL never less, non-NULL
// less than NULL). Keys are typed integer / REAL columns (no strings here;
// TEXT and ENUM ordering is pinned through the fixtures).
struct RefSortKey {
  int32_t type;  // 0 SMALLINT, 1 INT, 2 BIGINT, 3 REAL, 4 DATE
  int32_t asc;
  const void* data;
  const int8_t* nulls;
};
static const RefSortKey* g_sort_keys;
static int32_t g_sort_nkeys;

static bool ref_key_less(const RefSortKey& k, int32_t a, int32_t b) {
  if (k.nulls && k.nulls[a]) return false;
  if (k.nulls && k.nulls[b]) return true;
  switch (k.type) {
    case 0: return static_cast<const int16_t*>(k.data)[a] < static_cast<const int16_t*>(k.data)[b];
    case 2: return static_cast<const int64_t*>(k.data)[a] < static_cast<const int64_t*>(k.data)[b];
    case 3: return static_cast<const double*>(k.data)[a] < static_cast<const double*>(k.data)[b];
    default: return static_cast<const int32_t*>(k.data)[a] < static_cast<const int32_t*>(k.data)[b];
  }
}

static bool ref_sort_comp(int8_t* x, int8_t* y) {
  int32_t a, b;
  std::memcpy(&a, x, 4);
  std::memcpy(&b, y, 4);
  for (int32_t i = 0; i < g_sort_nkeys; i++) {
    if (ref_key_less(g_sort_keys[i], a, b)) return g_sort_keys[i].asc != 0;
    if (ref_key_less(g_sort_keys[i], b, a)) return g_sort_keys[i].asc == 0;
  }
  return false;
}

void ref_sort_rows(int32_t n, const RefSortKey* keys, int32_t nkeys, int32_t* out) {
  g_sort_keys = keys;
  g_sort_nkeys = nkeys;
  Vector::Vector vec;
  Vector::Create(&vec, 4, n > 0 ? n : 1);
  for (int32_t i = 0; i < n; i++) std::memcpy(Vector::PushBack(&vec), &i, 4);
  Vector::Sort(&vec, ref_sort_comp);
  for (int32_t i = 0; i < n; i++) std::memcpy(&out[i], Vector::Get(&vec, i), 4);
  Vector::Free(&vec);
}

}  // extern "C"
