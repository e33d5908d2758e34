// Column indexes, batched bucket lookups, batched k-way sorted intersection
// and the tuple-id dedup set — the device side of Skinner's per-episode work.
//
// Reference behaviour:
//  * MemoryColumnIndex (runtime/memory_column_index.cc:17-121,
//    compile/proxy/memory_column_index.cc:282-428): key -> bucket of tuple ids,
//    appended in scan order (ascending); only non-NULL keys are indexed
//    (permutable_skinner_join_translator.cc:972-994); the tuple id is the
//    row's position in the materialized child (its selection position).
//  * GetBucket returns runtime::EMPTY for a missing key.
//  * BucketListSortedIntersection* (runtime/column_index_bucket.cc:59-182):
//    the sorted intersection of k buckets from next_tuple on.
//  * TupleIdxTable (runtime/tuple_idx_table/tuple_idx_table.cc:19-99): a set of
//    int32[len] tuples; Insert returns false for a tuple already present.
//
// Layout:
//  * index: tids[] grouped by key, ascending within a key (a stable radix sort
//    of (key, tid) pairs), plus an open-addressing directory key -> (offset,
//    count) in 64-byte buckets of eight 8-byte keys.
//  * tuple set: open-addressing slots of {int64 state, int32 tuple[len]};
//    a slot is claimed with a 64-bit atomicCAS on its state word.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <new>

#include "kg_index.h"

namespace kg {

constexpr int kIBlock = 256;
constexpr int kMaxTupleLen = 8;
constexpr int kMaxIntersect = 8;

// Sortable form of an index key: I32 keys as a 32-bit order-preserving word
// (sorted on 32 bits), I64 keys with the sign bit flipped, F64 keys as their
// bits (equal keys only need to be adjacent).
__device__ __forceinline__ uint64_t enc_key(int64_t k, int enc) {
  if (enc == kEncI32) return static_cast<uint64_t>(static_cast<uint32_t>(static_cast<int32_t>(k)) ^ 0x80000000u);
  if (enc == kEncI64) return static_cast<uint64_t>(k) ^ (1ull << 63);
  return static_cast<uint64_t>(k);
}

__host__ __device__ __forceinline__ int64_t dec_key(uint64_t u, int enc) {
  if (enc == kEncI32) return static_cast<int32_t>(static_cast<uint32_t>(u) ^ 0x80000000u);
  if (enc == kEncI64) return static_cast<int64_t>(u ^ (1ull << 63));
  return static_cast<int64_t>(u);
}

// Sortable keys of the selected rows plus their tuple ids (selection
// positions); rows without a key get valid = 0.
__global__ void k_index_gather(kg_col key, const int32_t* sel, int64_t n, int enc, uint64_t* okey, int32_t* otid,
                               int32_t* valid) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t r = sel ? sel[i] : i;
    int64_t k = 0;
    const bool ok = index_key(key, r, enc, k);
    okey[i] = enc_key(ok ? k : 0, enc);
    otid[i] = static_cast<int32_t>(i);
    valid[i] = ok ? 1 : 0;
  }
}

// Owning view of a directory (see IndexDir).
struct Dir {
  int64_t* keys = nullptr;     // [nbuckets * 8] (hashed)
  int64_t* offset = nullptr;   // [nbuckets * 8] or [dspan]
  int32_t* count = nullptr;    // [nbuckets * 8] or [dspan]
  uint64_t nbuckets = 0;
  int64_t dmin = 0;
  uint64_t dspan = 0;
  int64_t min_off = 0;
  int32_t min_cnt = 0;
  int32_t enc = kEncI32;
  IndexDir view() const { return IndexDir{keys, offset, count, nbuckets, dmin, dspan, min_off, min_cnt, enc}; }
};

// Dense directory: one thread per run, runs in ascending key order.
__global__ void k_dir_dense(Dir d, const uint64_t* ukeys, const int32_t* counts, const int64_t* offsets,
                            const int32_t* nruns) {
  const int64_t n = *nruns;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t u = static_cast<uint64_t>(dec_key(ukeys[i], d.enc)) - static_cast<uint64_t>(d.dmin);
    d.offset[u] = offsets[i];
    d.count[u] = counts[i];
  }
}

// Sets *unsorted when some key is greater than the next one.
__global__ void k_keys_sorted(const uint64_t* k, int64_t n, int* unsorted) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  bool bad = false;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i + 1 < n; i += stride)
    bad |= k[i] > k[i + 1];
  if (__any(bad) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(unsorted, 1);
}

// One thread per distinct key (run): claim a directory slot. The run of the
// key kDirEmpty (never a slot key) is kept by the host in min_off / min_cnt.
__global__ void k_dir_insert(Dir d, const uint64_t* ukeys, const int32_t* counts, const int64_t* offsets,
                             const int32_t* nruns, int* overflow) {
  const int64_t n = *nruns;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t k = dec_key(ukeys[i], d.enc);
    if (k == kDirEmpty) continue;
    uint64_t b = dir_bucket(k, d.nbuckets);
    bool done = false;
    for (uint64_t probes = 0; probes < d.nbuckets && !done; probes++) {
      for (int s = 0; s < kDirSlots && !done; s++) {
        int64_t* p = d.keys + b * kDirSlots + s;
        if (*p != kDirEmpty) continue;
        const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(p),
                                                  static_cast<unsigned long long>(kDirEmpty),
                                                  static_cast<unsigned long long>(k));
        if (prev == static_cast<unsigned long long>(kDirEmpty)) {
          d.offset[b * kDirSlots + s] = offsets[i];
          d.count[b * kDirSlots + s] = counts[i];
          done = true;
        }
      }
      b = (b + 1 == d.nbuckets) ? 0 : b + 1;
    }
    if (!done) atomicOr(overflow, 1);
  }
}

// Batched GetBucket: bucket (pointer into tids, size) of each probe key; a
// missing or NULL key yields (tids, 0), the EMPTY bucket.
__global__ void k_dir_lookup(IndexDir d, const int32_t* tids, kg_col key, const int32_t* sel, int64_t n,
                             const int32_t** out_ptr, int32_t* out_cnt) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t r = sel ? sel[i] : i;
    int64_t k = 0, off = 0;
    int32_t cnt = 0;
    if (index_key(key, r, d.enc, k)) index_find(d, k, off, cnt);
    out_ptr[i] = tids + (cnt ? off : 0);
    out_cnt[i] = cnt;
  }
}

__global__ void k_fill_i32x(int32_t* p, int32_t v, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_i64x(int64_t* p, int64_t v, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// lower_bound of x in a[0, n) (FastForwardBucket without the INT32_MAX case).
__device__ __forceinline__ int32_t lower_bound(const int32_t* a, int32_t n, int32_t x) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One wave per task: the k-way sorted intersection of lists[t*k .. t*k+k-1]
// from next_tuple on. The shortest list drives; 64 candidates at a time are
// checked by binary search in the others. COUNT pass writes counts[t]; WRITE
// pass writes the ids at out + offsets[t] (ascending).
template <bool WRITE>
__global__ __launch_bounds__(kIBlock) void k_intersect(int64_t ntasks, int32_t k, const int32_t* const* lists,
                                                        const int32_t* sizes, const int32_t* next_tuple,
                                                        int64_t* counts, const int64_t* offsets, int32_t* out,
                                                        int64_t cap) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kIBlock + threadIdx.x) / kWave;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * kIBlock / kWave;
  for (int64_t t = wave; t < ntasks; t += nwaves) {
    const int32_t* const* L = lists + t * k;
    const int32_t* S = sizes + t * k;
    const int32_t nt = next_tuple ? next_tuple[t] : INT32_MIN;
    int drv = 0;
    for (int j = 1; j < k; j++)
      if (S[j] < S[drv]) drv = j;
    const int32_t* D = L[drv];
    const int32_t dn = S[drv];
    int32_t start = (k > 0 && dn > 0) ? lower_bound(D, dn, nt) : 0;
    int64_t written = 0;
    for (int32_t base = start; base < dn && k > 0; base += kWave) {
      const int32_t i = base + lane;
      bool keep = i < dn;
      int32_t c = keep ? D[i] : 0;
      for (int j = 0; j < k && keep; j++) {
        if (j == drv) continue;
        const int32_t p = lower_bound(L[j], S[j], c);
        keep = p < S[j] && L[j][p] == c;
      }
      const uint64_t m = __ballot(keep);
      if (WRITE && keep) {
        const int64_t pos = offsets[t] + written + __popcll(m & ((1ull << lane) - 1ull));
        if (pos < cap) out[pos] = c;
      }
      written += __popcll(m);
    }
    if (!WRITE && lane == 0) counts[t] = written;
  }
}

// --- tuple-id set -----------------------------------------------------------
// Slot: state (int64) then len int32s. state: 0 free, 1 being written,
// 2 | (tag << 2) complete, where tag is a 32-bit fingerprint of the tuple.
struct TupleSetView {
  unsigned long long* state;  // [nslots]
  int32_t* data;              // [nslots * len]
  uint64_t nslots;
  int32_t len;
};

__device__ __forceinline__ uint64_t tuple_hash(const int32_t* t, int len) {
  uint64_t h = 0x9E3779B97F4A7C15ull ^ static_cast<uint64_t>(len);
  for (int i = 0; i < len; i++) h = mix64(h ^ static_cast<uint32_t>(t[i]) ^ (static_cast<uint64_t>(i) << 40));
  return h;
}

// Insert each tuple; is_new[i] = 1 when this call added it, and ties between
// equal tuples inside one batch go to the lowest input position (the first to
// be inserted in the reference's sequential order): the claimant records
// min(position) and a second pass compares.
__global__ __launch_bounds__(kIBlock) void k_tuple_insert(TupleSetView s, const int32_t* tuples, int64_t n,
                                                           int64_t* slot_of, int32_t* first_pos, uint8_t* claimed,
                                                           unsigned long long* fresh_count, int* overflow,
                                                           int32_t spin_limit) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kIBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kIBlock + threadIdx.x; i < n; i += stride) {
    int32_t t[kMaxTupleLen];
    for (int q = 0; q < s.len; q++) t[q] = tuples[i * s.len + q];
    const uint64_t h = tuple_hash(t, s.len);
    const unsigned long long ready = 2ull | (static_cast<unsigned long long>(static_cast<uint32_t>(h >> 32)) << 2);
    uint64_t slot = __umul64hi(h, s.nslots);
    int64_t found = -1;
    int32_t spins = 0;
    for (uint64_t probes = 0; probes < s.nslots && found < 0;) {
      unsigned long long st = __hip_atomic_load(&s.state[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      if (st == 0) {
        const unsigned long long prev = atomicCAS(&s.state[slot], 0ull, 1ull);
        if (prev == 0ull) {
          for (int q = 0; q < s.len; q++)
            __hip_atomic_store(&s.data[slot * s.len + q], t[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(&s.state[slot], ready, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          found = static_cast<int64_t>(slot);
          first_pos[slot] = static_cast<int32_t>(i);  // lowered below by equal tuples of this batch
          claimed[slot] = 1;
          atomicAdd(fresh_count, 1ull);
          break;
        }
        continue;  // lost the race: re-read this slot
      }
      if (st == 1ull) {
        // Another lane is writing this slot: re-read it (bounded).
        if (++spins > spin_limit) {
          atomicOr(overflow, 2);
          break;
        }
        continue;
      }
      if (st == ready) {
        bool eq = true;
        for (int q = 0; q < s.len; q++)
          eq &= __hip_atomic_load(&s.data[slot * s.len + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == t[q];
        if (eq) {
          found = static_cast<int64_t>(slot);
          break;
        }
      }
      slot = (slot + 1 == s.nslots) ? 0 : slot + 1;
      probes++;
    }
    if (found < 0 && spins <= spin_limit) atomicOr(overflow, 1);
    slot_of[i] = found;
  }
}

// Second pass: the lowest input position per claimed-in-this-batch slot wins.
__global__ void k_tuple_min_pos(const int64_t* slot_of, int64_t n, int32_t* first_pos, const uint8_t* claimed) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t sl = slot_of[i];
    if (sl >= 0 && claimed[sl]) atomicMin(&first_pos[sl], static_cast<int32_t>(i));
  }
}


__global__ void k_tuple_is_new(const int64_t* slot_of, int64_t n, const int32_t* first_pos, const uint8_t* claimed,
                               uint8_t* is_new) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t sl = slot_of[i];
    is_new[i] = (sl >= 0 && claimed[sl] && first_pos[sl] == static_cast<int32_t>(i)) ? 1 : 0;
  }
}

__global__ void k_tuple_clear_claimed(const int64_t* slot_of, int64_t n, uint8_t* claimed) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t sl = slot_of[i];
    if (sl >= 0) claimed[sl] = 0;
  }
}

// Compaction of complete slots into out[k * len].
__global__ void k_tuple_extract(TupleSetView s, unsigned long long* counter, int32_t* out, int64_t cap) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < static_cast<int64_t>(s.nslots);
       i += stride) {
    if ((s.state[i] & 3ull) != 2ull) continue;
    const unsigned long long pos = atomicAdd(counter, 1ull);
    if (pos < static_cast<unsigned long long>(cap))
      for (int q = 0; q < s.len; q++) out[pos * s.len + q] = s.data[i * s.len + q];
  }
}

inline int igrid(kg_ctx* ctx, int64_t work) {
  int64_t g = ceil_div(std::max<int64_t>(work, 1), kIBlock * 4);
  return static_cast<int>(std::min<int64_t>(g, static_cast<int64_t>(ctx->num_cus) * 16));
}

}  // namespace kg

using namespace kg;

struct kg_index {
  int32_t* tids = nullptr;  // [ntids]
  int64_t ntids = 0;
  int64_t nkeys = 0;
  Dir dir{};
};

struct kg_tuple_set {
  unsigned long long* state = nullptr;
  int32_t* data = nullptr;
  uint64_t nslots = 0;
  int32_t len = 0;
  int64_t capacity = 0;
  int64_t size = 0;
  uint8_t* claimed = nullptr;      // [nslots] scratch per insert batch
  int32_t* first_pos = nullptr;    // [nslots]
  unsigned long long* counters = nullptr;
  TupleSetView view() const { return TupleSetView{state, data, nslots, len}; }
};

namespace {

// Index and tuple-set memory comes from the context's block cache: a plan
// builds and frees several of them per query.
int ialloc(kg_ctx* ctx, void** p, size_t bytes) { return kg_dev_alloc(ctx, p, bytes ? bytes : 8); }

// Returns a block to ctx's cache (stream-ordered reuse), or frees it when
// there is no context.
void ifree(kg_ctx* ctx, void* p) {
  if (!p) return;
  if (ctx)
    kg_dev_release(ctx, p);
  else
    (void)hipFree(p);
}

int read_u64(kg_ctx* ctx, const void* d, void* h, size_t bytes) {
  KG_HIP_CHECK(ctx, hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return KG_OK;
}

int check_index_key(kg_ctx* ctx, const kg_col* key, int64_t n) {
  KG_REQUIRE(ctx, key && (n == 0 || key->data), "key column is NULL");
  if (type_width(key->type) == 0)
    return kg_fail(ctx, KG_ERR_UNSUPPORTED, "index keys must be fixed-width (type %d; TEXT keys index their "
                   "kg_text_hash codes)", key->type);
  return KG_OK;
}

}  // namespace

extern "C" int kg_index_free(kg_ctx* ctx, kg_index* idx) {
  if (!idx) return KG_OK;
  if (ctx) (void)hipSetDevice(ctx->device);
  ifree(ctx, idx->tids);
  ifree(ctx, idx->dir.keys);
  ifree(ctx, idx->dir.offset);
  ifree(ctx, idx->dir.count);
  delete idx;
  return KG_OK;
}

int kg::index_build(kg_ctx* ctx, const kg_col* key, const int32_t* d_rows, int64_t nrows, int enc,
                    kg_index** out) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, out, "out is NULL");
  KG_REQUIRE(ctx, nrows >= 0 && nrows < INT32_MAX, "row count out of range");
  KG_REQUIRE(ctx, enc == kEncI32 || enc == kEncI64 || enc == kEncF64, "bad index encoding %d", enc);
  int st = check_index_key(ctx, key, nrows);
  if (st != KG_OK) return st;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  kg_index* idx = new (std::nothrow) kg_index();
  if (!idx) return kg_fail(ctx, KG_ERR_OOM, "host allocation failed");
  idx->dir.enc = enc;
  const int64_t n = nrows;
  const int end_bit = enc == kEncI32 ? 32 : 64;
  // Rows without a key (NULL; NaN under F64) are dropped first.
  const bool select = key->nulls || enc == kEncF64 || key->type == KG_REAL;
  // Workspace: keys/tids in and out, valid flags, run outputs, cub temp.
  size_t cub_sort = 0, cub_sel = 0, cub_rle = 0, cub_scan = 0;
  uint64_t* nk = nullptr;
  int32_t* ni = nullptr;
  const int nmax = static_cast<int>(std::max<int64_t>(n, 1));
  hipcub::DeviceRadixSort::SortPairs(nullptr, cub_sort, nk, nk, ni, ni, nmax, 0, end_bit, ctx->stream);
  hipcub::DeviceSelect::Flagged(nullptr, cub_sel, nk, ni, nk, ni, nmax, ctx->stream);
  hipcub::DeviceRunLengthEncode::Encode(nullptr, cub_rle, nk, nk, ni, ni, nmax, ctx->stream);
  hipcub::DeviceScan::ExclusiveSum(nullptr, cub_scan, ni, reinterpret_cast<int64_t*>(ni), nmax, ctx->stream);
  const size_t cub_bytes = std::max(std::max(cub_sort, cub_sel), std::max(cub_rle, cub_scan));
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t nn = static_cast<size_t>(nmax);
  const size_t o_k0 = carve(nn * 8), o_t0 = carve(nn * 4), o_k1 = carve(nn * 8), o_t1 = carve(nn * 4),
               o_valid = carve(nn * 4), o_cnt = carve(nn * 4), o_off = carve(nn * 8), o_nsel = carve(16),
               o_nruns = carve(16), o_flag = carve(16), o_cub = carve(cub_bytes);
  if ((st = kg_work_reserve(ctx, off)) != KG_OK) {
    delete idx;
    return st;
  }
  char* w = static_cast<char*>(ctx->work);
  uint64_t* k0 = reinterpret_cast<uint64_t*>(w + o_k0);
  int32_t* t0 = reinterpret_cast<int32_t*>(w + o_t0);
  uint64_t* k1 = reinterpret_cast<uint64_t*>(w + o_k1);
  int32_t* t1 = reinterpret_cast<int32_t*>(w + o_t1);
  int32_t* valid = reinterpret_cast<int32_t*>(w + o_valid);
  int32_t* cnt = reinterpret_cast<int32_t*>(w + o_cnt);
  int64_t* offs = reinterpret_cast<int64_t*>(w + o_off);
  int32_t* nsel = reinterpret_cast<int32_t*>(w + o_nsel);
  int32_t* nruns = reinterpret_cast<int32_t*>(w + o_nruns);
  int* flag = reinterpret_cast<int*>(w + o_flag);
  void* cub = w + o_cub;
  size_t cb = cub_bytes;

  if (n > 0) {
    kg_time_begin(ctx);
    hipLaunchKernelGGL(k_index_gather, dim3(igrid(ctx, n)), dim3(kIBlock), 0, ctx->stream, *key, d_rows, n, enc, k0,
                       t0, valid);
    KG_LAUNCH_CHECK(ctx);
    // Drop rows without a key (keeping order), then a stable sort by key keeps
    // tuple ids ascending per key. Otherwise every key stays in place.
    if (select) {
      cb = cub_bytes;
      KG_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(cub, cb, k0, valid, k1, nsel, static_cast<int>(n), ctx->stream));
      cb = cub_bytes;
      KG_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(cub, cb, t0, valid, t1, nsel, static_cast<int>(n), ctx->stream));
    }
  }
  int32_t hsel = static_cast<int32_t>(n);
  if (n > 0 && select && (st = read_u64(ctx, nsel, &hsel, 4)) != KG_OK) {
    delete idx;
    return st;
  }
  if (n > 0 && !select) {
    std::swap(k0, k1);
    std::swap(t0, t1);
  }
  idx->ntids = hsel;
  if ((st = ialloc(ctx, reinterpret_cast<void**>(&idx->tids), static_cast<size_t>(hsel) * 4)) != KG_OK) {
    kg_index_free(ctx, idx);
    return st;
  }
  int32_t hruns = 0;
  if (hsel > 0) {
    // Keys already in order (an id column over an ascending selection) need no
    // sort: the stable sort of sorted keys is the identity.
    hipLaunchKernelGGL(k_fill_i32x, dim3(1), dim3(64), 0, ctx->stream, flag, 0, 4);
    hipLaunchKernelGGL(k_keys_sorted, dim3(igrid(ctx, hsel)), dim3(kIBlock), 0, ctx->stream, k1, hsel, flag);
    KG_LAUNCH_CHECK(ctx);
    int unsorted = 0;
    if ((st = read_u64(ctx, flag, &unsorted, 4)) != KG_OK) {
      kg_index_free(ctx, idx);
      return st;
    }
    cb = cub_bytes;
    if (unsorted) {
      KG_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(cub, cb, k1, k0, t1, idx->tids, hsel, 0, end_bit,
                                                           ctx->stream));
    } else {
      KG_HIP_CHECK(ctx, hipMemcpyAsync(k0, k1, static_cast<size_t>(hsel) * 8, hipMemcpyDeviceToDevice, ctx->stream));
      KG_HIP_CHECK(ctx, hipMemcpyAsync(idx->tids, t1, static_cast<size_t>(hsel) * 4, hipMemcpyDeviceToDevice,
                                       ctx->stream));
    }
    cb = cub_bytes;
    KG_HIP_CHECK(ctx, hipcub::DeviceRunLengthEncode::Encode(cub, cb, k0, k1, cnt, nruns, hsel, ctx->stream));
    if ((st = read_u64(ctx, nruns, &hruns, 4)) != KG_OK) {
      kg_index_free(ctx, idx);
      return st;
    }
    cb = cub_bytes;
    KG_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveSum(cub, cb, cnt, offs, hruns, ctx->stream));
  }
  idx->nkeys = hruns;
  if (hruns > 0) {
    // Key range of the runs (ascending in I32 / I64): first and last run key.
    uint64_t ends[2] = {0, 0};
    KG_HIP_CHECK(ctx, hipMemcpyAsync(&ends[0], k1, 8, hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP_CHECK(ctx, hipMemcpyAsync(&ends[1], k1 + hruns - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t kmin = dec_key(ends[0], enc), kmax = dec_key(ends[1], enc);
    const uint64_t span1 = static_cast<uint64_t>(kmax) - static_cast<uint64_t>(kmin);  // span - 1
    if (enc != kEncF64 && span1 < 3 * static_cast<uint64_t>(hruns) + 1024) {
      const int64_t span = static_cast<int64_t>(span1) + 1;
      idx->dir.dmin = kmin;
      idx->dir.dspan = static_cast<uint64_t>(span);
      if ((st = ialloc(ctx, reinterpret_cast<void**>(&idx->dir.offset), span * 8)) != KG_OK ||
          (st = ialloc(ctx, reinterpret_cast<void**>(&idx->dir.count), span * 4)) != KG_OK) {
        kg_index_free(ctx, idx);
        return st;
      }
      hipLaunchKernelGGL(k_fill_i32x, dim3(igrid(ctx, span)), dim3(kIBlock), 0, ctx->stream, idx->dir.count, 0, span);
      hipLaunchKernelGGL(k_dir_dense, dim3(igrid(ctx, hruns)), dim3(kIBlock), 0, ctx->stream, idx->dir, k1, cnt, offs,
                         nruns);
      KG_LAUNCH_CHECK(ctx);
      kg_time_end(ctx);
      *out = idx;
      return KG_OK;
    }
    if (enc == kEncI64 && kmin == kDirEmpty) {  // the run of INT64_MIN is the first run
      int32_t c0 = 0;
      if ((st = read_u64(ctx, cnt, &c0, 4)) != KG_OK) {
        kg_index_free(ctx, idx);
        return st;
      }
      idx->dir.min_off = 0;
      idx->dir.min_cnt = c0;
    }
  }
  idx->dir.nbuckets = static_cast<uint64_t>(std::max<int64_t>(ceil_div(std::max<int64_t>(hruns, 1) * 2, kDirSlots), 1));
  const size_t ns = idx->dir.nbuckets * kDirSlots;
  if ((st = ialloc(ctx, reinterpret_cast<void**>(&idx->dir.keys), ns * 8)) != KG_OK ||
      (st = ialloc(ctx, reinterpret_cast<void**>(&idx->dir.offset), ns * 8)) != KG_OK ||
      (st = ialloc(ctx, reinterpret_cast<void**>(&idx->dir.count), ns * 4)) != KG_OK) {
    kg_index_free(ctx, idx);
    return st;
  }
  hipLaunchKernelGGL(k_fill_i64x, dim3(igrid(ctx, ns)), dim3(kIBlock), 0, ctx->stream, idx->dir.keys, kDirEmpty,
                     static_cast<int64_t>(ns));
  KG_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_fill_i32x, dim3(1), dim3(64), 0, ctx->stream, flag, 0, 4);
  if (hruns > 0) {
    hipLaunchKernelGGL(k_dir_insert, dim3(igrid(ctx, hruns)), dim3(kIBlock), 0, ctx->stream, idx->dir, k1, cnt, offs,
                       nruns, flag);
    KG_LAUNCH_CHECK(ctx);
  }
  if (n > 0) kg_time_end(ctx);
  int hflag = 0;
  if ((st = read_u64(ctx, flag, &hflag, 4)) != KG_OK) {
    kg_index_free(ctx, idx);
    return st;
  }
  if (hflag) {
    kg_index_free(ctx, idx);
    return kg_fail(ctx, KG_ERR_CAPACITY, "index directory overflow");
  }
  *out = idx;
  return KG_OK;
}

extern "C" int kg_index_build(kg_ctx* ctx, const kg_col* key, const int32_t* d_rows, int64_t nrows,
                              kg_index** out) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, key, "key column is NULL");
  return index_build(ctx, key, d_rows, nrows, index_enc_of(key->type), out);
}

void kg::index_view(const kg_index* idx, IndexDir* dir, const int32_t** tids) {
  *dir = idx->dir.view();
  *tids = idx->tids;
}

int kg::index_enc(const kg_index* idx) { return idx->dir.enc; }

extern "C" int kg_index_info(kg_ctx* ctx, const kg_index* idx, const int32_t** d_tids, int64_t* ntids,
                             int64_t* nkeys) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, idx, "index is NULL");
  if (d_tids) *d_tids = idx->tids;
  if (ntids) *ntids = idx->ntids;
  if (nkeys) *nkeys = idx->nkeys;
  return KG_OK;
}

extern "C" int kg_index_lookup(kg_ctx* ctx, const kg_index* idx, const kg_col* key, const int32_t* d_rows,
                               int64_t n, const int32_t** d_out_ptr, int32_t* d_out_count) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, idx, "index is NULL");
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count out of range");
  KG_REQUIRE(ctx, n == 0 || (d_out_ptr && d_out_count), "output arrays are NULL");
  int st = check_index_key(ctx, key, n);
  if (st != KG_OK) return st;
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_dir_lookup, dim3(igrid(ctx, n)), dim3(kIBlock), 0, ctx->stream, idx->dir.view(), idx->tids,
                     *key, d_rows, n, d_out_ptr, d_out_count);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  return KG_OK;
}

extern "C" int kg_intersect_batch(kg_ctx* ctx, int64_t ntasks, int32_t k, const int32_t* const* d_lists,
                                  const int32_t* d_sizes, const int32_t* d_next_tuple, int64_t* d_out_offsets,
                                  int32_t* d_out, int64_t out_cap, int64_t* out_total) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, ntasks >= 0 && ntasks < INT32_MAX, "task count out of range");
  KG_REQUIRE(ctx, k >= 1 && k <= kMaxIntersect, "1..%d lists per task (got %d)", kMaxIntersect, k);
  KG_REQUIRE(ctx, out_total && d_out_offsets, "null output pointer");
  KG_REQUIRE(ctx, ntasks == 0 || (d_lists && d_sizes), "list arrays are NULL");
  KG_REQUIRE(ctx, out_cap == 0 || d_out, "d_out is NULL");
  *out_total = 0;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (ntasks == 0) {
    KG_HIP_CHECK(ctx, hipMemsetAsync(d_out_offsets, 0, 8, ctx->stream));
    return KG_OK;
  }
  size_t cub_scan = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, cub_scan, static_cast<int64_t*>(nullptr), static_cast<int64_t*>(nullptr),
                                   static_cast<int>(ntasks + 1), ctx->stream);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_cnt = carve((ntasks + 1) * 8), o_cub = carve(cub_scan);
  int st = kg_work_reserve(ctx, off);
  if (st != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  int64_t* counts = reinterpret_cast<int64_t*>(w + o_cnt);
  const int grid = static_cast<int>(std::min<int64_t>(ceil_div(ntasks, kIBlock / kWave), ctx->num_cus * 16));
  KG_HIP_CHECK(ctx, hipMemsetAsync(counts + ntasks, 0, 8, ctx->stream));
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_intersect<false>, dim3(grid), dim3(kIBlock), 0, ctx->stream, ntasks, k, d_lists, d_sizes,
                     d_next_tuple, counts, nullptr, nullptr, 0);
  KG_LAUNCH_CHECK(ctx);
  size_t cb = cub_scan;
  KG_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveSum(w + o_cub, cb, counts, d_out_offsets, static_cast<int>(ntasks + 1),
                                                     ctx->stream));
  hipLaunchKernelGGL(k_intersect<true>, dim3(grid), dim3(kIBlock), 0, ctx->stream, ntasks, k, d_lists, d_sizes,
                     d_next_tuple, counts, d_out_offsets, d_out, out_cap);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  int64_t total = 0;
  if ((st = read_u64(ctx, d_out_offsets + ntasks, &total, 8)) != KG_OK) return st;
  *out_total = total;
  if (total > out_cap)
    return kg_fail(ctx, KG_ERR_CAPACITY, "%lld intersection ids exceed out_cap %lld", (long long)total,
                   (long long)out_cap);
  return KG_OK;
}

extern "C" int kg_tuple_set_free(kg_ctx* ctx, kg_tuple_set* s) {
  if (!s) return KG_OK;
  if (ctx) (void)hipSetDevice(ctx->device);
  ifree(ctx, s->state);
  ifree(ctx, s->data);
  ifree(ctx, s->claimed);
  ifree(ctx, s->first_pos);
  ifree(ctx, s->counters);
  delete s;
  return KG_OK;
}

extern "C" int kg_tuple_set_create(kg_ctx* ctx, int32_t len, int64_t capacity, kg_tuple_set** out) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, out, "out is NULL");
  KG_REQUIRE(ctx, len >= 1 && len <= kMaxTupleLen, "tuple length 1..%d (got %d)", kMaxTupleLen, len);
  KG_REQUIRE(ctx, capacity >= 0 && capacity < (1ll << 31), "capacity out of range");
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  kg_tuple_set* s = new (std::nothrow) kg_tuple_set();
  if (!s) return kg_fail(ctx, KG_ERR_OOM, "host allocation failed");
  s->len = len;
  s->capacity = capacity;
  s->nslots = static_cast<uint64_t>(std::max<int64_t>(capacity * 2, 64));
  int st;
  if ((st = ialloc(ctx, reinterpret_cast<void**>(&s->state), s->nslots * 8)) != KG_OK ||
      (st = ialloc(ctx, reinterpret_cast<void**>(&s->data), s->nslots * len * 4)) != KG_OK ||
      (st = ialloc(ctx, reinterpret_cast<void**>(&s->claimed), s->nslots)) != KG_OK ||
      (st = ialloc(ctx, reinterpret_cast<void**>(&s->first_pos), s->nslots * 4)) != KG_OK ||
      (st = ialloc(ctx, reinterpret_cast<void**>(&s->counters), 64)) != KG_OK) {
    kg_tuple_set_free(ctx, s);
    return st;
  }
  KG_HIP_CHECK(ctx, hipMemsetAsync(s->state, 0, s->nslots * 8, ctx->stream));
  KG_HIP_CHECK(ctx, hipMemsetAsync(s->claimed, 0, s->nslots, ctx->stream));
  KG_HIP_CHECK(ctx, hipMemsetAsync(s->counters, 0, 64, ctx->stream));
  *out = s;
  return KG_OK;
}

extern "C" int kg_tuple_set_insert(kg_ctx* ctx, kg_tuple_set* s, const int32_t* d_tuples, int64_t n,
                                   uint8_t* d_is_new, int64_t* out_new) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, s && out_new, "null pointer argument");
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "tuple count out of range");
  KG_REQUIRE(ctx, n == 0 || d_tuples, "d_tuples is NULL");
  *out_new = 0;
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_slot = carve(n * 8), o_new = carve(n);
  int st = kg_work_reserve(ctx, off);
  if (st != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  int64_t* slot_of = reinterpret_cast<int64_t*>(w + o_slot);
  uint8_t* is_new = d_is_new ? d_is_new : reinterpret_cast<uint8_t*>(w + o_new);
  unsigned long long* fresh = s->counters;
  int* overflow = reinterpret_cast<int*>(s->counters + 1);
  KG_HIP_CHECK(ctx, hipMemsetAsync(s->counters, 0, 16, ctx->stream));
  const int grid = igrid(ctx, n);
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_tuple_insert, dim3(grid), dim3(kIBlock), 0, ctx->stream, s->view(), d_tuples, n, slot_of,
                     s->first_pos, s->claimed, fresh, overflow, 1 << 20);
  KG_LAUNCH_CHECK(ctx);
  // The claimant marked its slot and wrote its position; every equal tuple of
  // the batch lowers that position, then is_new is decided.
  hipLaunchKernelGGL(k_tuple_min_pos, dim3(grid), dim3(kIBlock), 0, ctx->stream, slot_of, n, s->first_pos,
                     s->claimed);
  hipLaunchKernelGGL(k_tuple_is_new, dim3(grid), dim3(kIBlock), 0, ctx->stream, slot_of, n, s->first_pos, s->claimed,
                     is_new);
  hipLaunchKernelGGL(k_tuple_clear_claimed, dim3(grid), dim3(kIBlock), 0, ctx->stream, slot_of, n, s->claimed);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  unsigned long long h[2] = {0, 0};
  if ((st = read_u64(ctx, s->counters, h, 16)) != KG_OK) return st;
  if (h[1] & 1) return kg_fail(ctx, KG_ERR_CAPACITY, "tuple set is full");
  if (h[1] & 2) return kg_fail(ctx, KG_ERR_HIP, "tuple set slot wait exceeded its bound");
  s->size += static_cast<int64_t>(h[0]);
  *out_new = static_cast<int64_t>(h[0]);
  if (s->size > s->capacity)
    return kg_fail(ctx, KG_ERR_CAPACITY, "tuple set holds %lld tuples, above its capacity %lld", (long long)s->size,
                   (long long)s->capacity);
  return KG_OK;
}

extern "C" int kg_tuple_set_extract(kg_ctx* ctx, const kg_tuple_set* s, int32_t* d_out, int64_t cap,
                                    int64_t* out_n) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, s && out_n, "null pointer argument");
  KG_REQUIRE(ctx, cap == 0 || d_out, "d_out is NULL");
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  unsigned long long* cnt = s->counters + 2;
  KG_HIP_CHECK(ctx, hipMemsetAsync(cnt, 0, 8, ctx->stream));
  hipLaunchKernelGGL(k_tuple_extract, dim3(igrid(ctx, s->nslots)), dim3(kIBlock), 0, ctx->stream, s->view(), cnt,
                     d_out, cap);
  KG_LAUNCH_CHECK(ctx);
  unsigned long long h = 0;
  int st = read_u64(ctx, cnt, &h, 8);
  if (st != KG_OK) return st;
  *out_n = static_cast<int64_t>(h);
  if (*out_n > cap) return kg_fail(ctx, KG_ERR_CAPACITY, "%lld tuples exceed cap %lld", (long long)*out_n, (long long)cap);
  return KG_OK;
}
