// TEXT predicates, MIN/MAX over TEXT, and the index-nested-loop expansion
// step of a multi-way join.
//
// Reference behaviour:
//  * runtime/string.cc:14-82: StartsWith (rfind(p, 0) == 0), EndsWith,
//    Contains, Equals and the comparisons are byte-wise string_view
//    operations; Like is an anchored regex full match built by the planner
//    (plan/planner.cc:284-323): '%' -> ".*", '_' -> ".", every other pattern
//    byte literal. (The planner rewrites LIKE '%x' to STARTS_WITH and 'x%' to
//    ENDS_WITH, planner.cc:249-269; callers pass the rewritten operator.)
//  * MIN/MAX over TEXT compare with string_view's operator< (unsigned bytes,
//    a proper prefix first) and keep the first value among equals
//    (compile/proxy/aggregator.cc:155-160).
//  * The Skinner table functions extend a partial tuple by the bucket of the
//    next table's index for the join key (permutable_skinner_join_translator.cc:359-433).
//
// TEXT layout on the device: offsets[n+1] (int64) into a byte heap, plus an
// optional int8 NULL array — the slot table of the .kdb TEXT format
// (runtime/column_data.cc:175-229) flattened to prefix offsets.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "kg_ops.h"
#include "kg_index.h"
#include "kg_strops.h"

namespace kg {

constexpr int kTBlock = 256;
constexpr int kMaxPattern = 240;

struct Pattern {
  int32_t op;
  int32_t len;
  uint8_t p[kMaxPattern];
};

// `s op p` with a LIKE pattern already checked (like_pattern_ok) on the host.
__device__ __forceinline__ bool text_pred(int32_t op, const uint8_t* s, int64_t n, const uint8_t* p, int32_t m) {
  return op == KG_TEXT_LIKE ? like_match(s, n, p, m) : text_match(op, s, n, p, m);
}

// Heap bytes a workgroup stages in LDS for its 256 rows: the bytes of
// consecutive rows are one contiguous range of the heap.
constexpr int kTextTileBytes = 32 * 1024;

// One row per thread. Without a selection (and with a 4-byte aligned heap), a
// workgroup's rows are consecutive: their heap range is copied to LDS with
// coalesced dword loads
// (when it fits kTextTileBytes) and matched there, so the per-character loads
// of the matching loop hit LDS instead of waiting on global memory. Rows of a
// selection, and tiles whose strings exceed the LDS range, match from global
// memory. The pattern is read from LDS.
__global__ __launch_bounds__(kTBlock) void k_text_flags(kg_text_col col, Pattern P, const int32_t* sel, int64_t n,
                                                         int32_t* ids, uint8_t* flags) {
  __shared__ uint32_t tile[kTextTileBytes / 4];
  __shared__ uint8_t pat[kMaxPattern];
  for (int j = threadIdx.x; j < P.len; j += kTBlock) pat[j] = P.p[j];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kTBlock;
  for (int64_t t0 = static_cast<int64_t>(blockIdx.x) * kTBlock; t0 < n; t0 += stride) {
    const int64_t i = t0 + threadIdx.x;
    const int64_t tn = n - t0 < kTBlock ? n - t0 : kTBlock;
    int64_t base = 0;
    bool staged = false;
    if (!sel && (reinterpret_cast<uintptr_t>(col.bytes) & 3) == 0) {
      const int64_t lo = col.offsets[t0], hi = col.offsets[t0 + tn];
      base = lo & ~int64_t{3};
      if (hi - base <= kTextTileBytes) {
        staged = true;
        const int64_t nfull = (hi - base) / 4;
        const uint32_t* src = reinterpret_cast<const uint32_t*>(col.bytes + base);
        for (int64_t w = threadIdx.x; w < nfull; w += kTBlock) tile[w] = __builtin_nontemporal_load(src + w);
        // The last partial word byte by byte: nothing past the heap's end is read.
        const int64_t rem = (hi - base) - nfull * 4;
        if (threadIdx.x < rem)
          reinterpret_cast<uint8_t*>(tile)[nfull * 4 + threadIdx.x] = col.bytes[base + nfull * 4 + threadIdx.x];
      }
    }
    __syncthreads();
    if (i < n) {
      const int64_t r = sel ? sel[i] : i;
      bool ok = false;
      if (!(col.nulls && col.nulls[r])) {
        const int64_t a = col.offsets[r], b = col.offsets[r + 1];
        const uint8_t* s = staged ? reinterpret_cast<const uint8_t*>(tile) + (a - base) : col.bytes + a;
        ok = text_pred(P.op, s, b - a, pat, P.len);
      }
      ids[i] = static_cast<int32_t>(r);
      flags[i] = ok ? 1 : 0;
    }
    __syncthreads();
  }
}

// Position (in the input order) of the MIN (or MAX) non-NULL string; among
// equal strings the earliest position wins. One candidate per thread, then
// a tree over the workgroup; per-workgroup winners go to `best`.
__device__ __forceinline__ bool better(const kg_text_col& c, const int32_t* sel, int64_t a, int64_t b, bool is_max) {
  // true when position a should replace position b (b < 0 = none yet)
  if (a < 0) return false;
  if (b < 0) return true;
  const int64_t ra = sel ? sel[a] : a, rb = sel ? sel[b] : b;
  const int cmp = str_cmp(c.bytes + c.offsets[ra], c.offsets[ra + 1] - c.offsets[ra], c.bytes + c.offsets[rb],
                          c.offsets[rb + 1] - c.offsets[rb]);
  if (cmp != 0) return is_max ? cmp > 0 : cmp < 0;
  return a < b;
}

__global__ __launch_bounds__(kTBlock) void k_text_best(kg_text_col c, const int32_t* sel, int64_t n, int is_max,
                                                        int64_t* best) {
  __shared__ int64_t cand[kTBlock];
  int64_t mine = -1;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kTBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kTBlock + threadIdx.x; i < n; i += stride) {
    const int64_t r = sel ? sel[i] : i;
    if (c.nulls && c.nulls[r]) continue;
    if (better(c, sel, i, mine, is_max)) mine = i;
  }
  cand[threadIdx.x] = mine;
  __syncthreads();
  for (int s = kTBlock / 2; s >= 1; s >>= 1) {
    if (threadIdx.x < s) {
      const int64_t o = cand[threadIdx.x + s];
      if (better(c, sel, o, cand[threadIdx.x], is_max)) cand[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) best[blockIdx.x] = cand[0];
}

__global__ void k_text_best_final(kg_text_col c, const int32_t* sel, const int64_t* best, int nb, int is_max,
                                  int64_t* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t b = -1;
  for (int i = 0; i < nb; i++)
    if (better(c, sel, best[i], b, is_max)) b = best[i];
  *out = b;
}

// --- index-nested-loop expansion --------------------------------------------

struct Expand {
  int32_t width;          // partial tuple width
  int32_t src_pos;        // tuple position holding the probing table's tuple id
  kg_col key;             // key column of that table (indexed by row id)
  const int32_t* src_rows;  // that table's selection: tuple id -> row id (NULL = identity)
  const int32_t* tids;    // the index's tid array
};

__global__ void k_zero_i64(int64_t* p, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Pass 1: bucket (offset, size) per partial tuple.
__global__ void k_expand_count(Expand e, IndexDir d, const int32_t* tuples, int64_t n, int64_t* boff,
                               int64_t* bcnt) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t tid = tuples ? tuples[i * e.width + e.src_pos] : static_cast<int32_t>(i);
    const int64_t r = e.src_rows ? e.src_rows[tid] : tid;
    int64_t k = 0, off = 0;
    int32_t cnt = 0;
    if (index_key(e.key, r, d.enc, k)) index_find(d, k, off, cnt);
    boff[i] = off;
    bcnt[i] = cnt;
  }
}

// Pass 2: one thread per output tuple; its source tuple by binary search over
// the exclusive-scanned counts (balanced under skewed bucket sizes).
__global__ void k_expand_write(Expand e, const int32_t* tuples, int64_t n, const int64_t* boff,
                               const int64_t* starts, int64_t total, int32_t* out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int w1 = e.width + 1;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < total; j += stride) {
    int64_t lo = 0, hi = n;  // last i with starts[i] <= j
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (starts[mid] <= j) lo = mid;
      else hi = mid;
    }
    const int64_t i = lo;
    if (tuples) {
      for (int q = 0; q < e.width; q++) out[j * w1 + q] = tuples[i * e.width + q];
    } else {
      out[j * w1] = static_cast<int32_t>(i);
    }
    out[j * w1 + e.width] = e.tids[boff[i] + (j - starts[i])];
  }
}

// Row ids of one position of a tuple batch: out[i] = sel[t[i*width+pos]].
__global__ void k_tuple_rows(const int32_t* t, int32_t width, int32_t pos, const int32_t* sel, int64_t n,
                             int32_t* out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t v = t[i * width + pos];
    out[i] = sel ? sel[v] : v;
  }
}

// Materialize rows[i] of a fixed-width column (and its NULL flag).
__global__ void k_gather(const uint8_t* src, const int8_t* src_nulls, int w, const int32_t* rows, int64_t n,
                         uint8_t* dst, int8_t* dst_nulls) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t r = rows[i];
    switch (w) {
      case 1: dst[i] = src[r]; break;
      case 2: reinterpret_cast<int16_t*>(dst)[i] = reinterpret_cast<const int16_t*>(src)[r]; break;
      case 4: reinterpret_cast<int32_t*>(dst)[i] = reinterpret_cast<const int32_t*>(src)[r]; break;
      default: reinterpret_cast<int64_t*>(dst)[i] = reinterpret_cast<const int64_t*>(src)[r]; break;
    }
    if (dst_nulls) dst_nulls[i] = src_nulls ? src_nulls[r] : 0;
  }
}

// One item per thread up to 16 workgroups per CU (grid-stride beyond): these
// kernels are bound by dependent loads, so more waves hide more latency.
inline int tgrid(kg_ctx* ctx, int64_t work) {
  int64_t g = ceil_div(std::max<int64_t>(work, 1), kTBlock);
  return static_cast<int>(std::min<int64_t>(g, static_cast<int64_t>(ctx->num_cus) * 16));
}

}  // namespace kg

using namespace kg;

namespace {

int check_text(kg_ctx* ctx, const kg_text_col* c) {
  KG_REQUIRE(ctx, c && c->offsets, "text column is NULL");
  return KG_OK;
}

// 64-bit FNV-1a over a TEXT value's bytes, finished with mix64 (fmix64).
// Equal strings hash equal; placement only, never compared as a result.
__device__ __forceinline__ uint64_t text_hash(const kg_text_col& c, int64_t r) {
  const int64_t b = c.offsets[r], e = c.offsets[r + 1];
  uint64_t h = 0xcbf29ce484222325ull;
  for (int64_t i = b; i < e; i++) h = (h ^ c.bytes[i]) * 0x100000001b3ull;
  return mix64(h ^ static_cast<uint64_t>(e - b));
}

__global__ void k_text_hash(kg_text_col col, const int32_t* rows, int64_t n, int64_t* out, int8_t* out_nulls) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t r = rows ? rows[i] : i;
    const bool null = col.nulls && col.nulls[r];
    out[i] = null ? 0 : static_cast<int64_t>(text_hash(col, r));
    if (out_nulls) out_nulls[i] = null ? 1 : 0;
  }
}

// Keep flag of pair i: a[ra[i]] == b[rb[i]] byte for byte, neither NULL.
__global__ void k_text_pairs_eq(kg_text_col a, const int32_t* ra, kg_text_col b, const int32_t* rb, int64_t n,
                                int32_t* ids, uint8_t* fl) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t x = ra ? ra[i] : i, y = rb ? rb[i] : i;
    bool eq = !(a.nulls && a.nulls[x]) && !(b.nulls && b.nulls[y]);
    if (eq) {
      const int64_t ab = a.offsets[x], al = a.offsets[x + 1] - ab;
      const int64_t bb = b.offsets[y], bl = b.offsets[y + 1] - bb;
      eq = al == bl;
      for (int64_t k = 0; eq && k < al; k++) eq = a.bytes[ab + k] == b.bytes[bb + k];
    }
    ids[i] = static_cast<int32_t>(i);
    fl[i] = eq ? 1 : 0;
  }
}

// Exactness check of TEXT hash codes (kg_text_hash_verify): an open-addressing
// table of codes (kCodeFree = free; a code equal to it is taken as
// kCodeFree + 1 in both passes) with the smallest row of each code, then every
// row compared byte for byte with its code's smallest row.
constexpr int64_t kCodeFree = INT64_MIN;

__device__ __forceinline__ int64_t code_of(const int64_t* hash, int64_t i) {
  const int64_t h = hash[i];
  return h == kCodeFree ? kCodeFree + 1 : h;
}

__global__ void k_code_fill(int64_t* keys, int32_t* reps, int64_t nslots, unsigned long long* count) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nslots; i += stride) {
    keys[i] = kCodeFree;
    reps[i] = INT32_MAX;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0;
}

__global__ void k_code_claim(kg_text_col col, const int32_t* rows, const int64_t* hash, int64_t n, int64_t* keys,
                             int32_t* reps, uint64_t mask) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t r = rows ? rows[i] : i;
    if (col.nulls && col.nulls[r]) continue;
    const int64_t h = code_of(hash, i);
    uint64_t s = mix64(static_cast<uint64_t>(h)) & mask;
    for (uint64_t probes = 0; probes <= mask; probes++) {
      int64_t cur = keys[s];
      if (cur == kCodeFree)
        cur = static_cast<int64_t>(atomicCAS(reinterpret_cast<unsigned long long*>(keys + s),
                                             static_cast<unsigned long long>(kCodeFree),
                                             static_cast<unsigned long long>(h)));
      if (cur == kCodeFree || cur == h) {
        atomicMin(reps + s, static_cast<int32_t>(r));
        break;
      }
      s = (s + 1) & mask;
    }
  }
}

__global__ void k_code_check(kg_text_col col, const int32_t* rows, const int64_t* hash, int64_t n,
                             const int64_t* keys, const int32_t* reps, uint64_t mask, unsigned long long* count) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t r = rows ? rows[i] : i;
    if (col.nulls && col.nulls[r]) continue;
    const int64_t h = code_of(hash, i);
    uint64_t s = mix64(static_cast<uint64_t>(h)) & mask;
    for (uint64_t probes = 0; probes <= mask && keys[s] != h; probes++) s = (s + 1) & mask;
    if (keys[s] != h) continue;  // not reached: every code was claimed by the first pass
    const int64_t q = reps[s];
    const int64_t ab = col.offsets[r], al = col.offsets[r + 1] - ab;
    const int64_t bb = col.offsets[q], bl = col.offsets[q + 1] - bb;
    bool eq = al == bl;
    for (int64_t k = 0; eq && k < al; k++) eq = col.bytes[ab + k] == col.bytes[bb + k];
    if (!eq) atomicAdd(count, 1ull);
  }
}

}  // namespace

extern "C" int kg_text_select(kg_ctx* ctx, const kg_text_col* col, int32_t op, const char* pattern, int32_t plen,
                              const int32_t* d_rows_in, int64_t n, int32_t* d_rows_out, int64_t* out_count) {
  if (!ctx) return KG_ERR_INVALID;
  int st = check_text(ctx, col);
  if (st != KG_OK) return st;
  KG_REQUIRE(ctx, op >= KG_TEXT_EQ && op <= KG_TEXT_LIKE, "bad text operator %d", op);
  KG_REQUIRE(ctx, plen >= 0 && plen <= kMaxPattern, "pattern length %d out of 0..%d", plen, kMaxPattern);
  KG_REQUIRE(ctx, plen == 0 || pattern, "pattern is NULL");
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count out of range");
  KG_REQUIRE(ctx, out_count && (n == 0 || d_rows_out), "null output pointer");
  *out_count = 0;
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // A LIKE pattern re2 would reject (not UTF-8) matches nothing.
  if (op == KG_TEXT_LIKE && !like_pattern_ok(reinterpret_cast<const uint8_t*>(pattern), plen)) return KG_OK;
  Pattern P{};
  P.op = op;
  P.len = plen;
  if (plen) std::memcpy(P.p, pattern, plen);
  size_t cub_sel = 0;
  hipcub::DeviceSelect::Flagged(nullptr, cub_sel, static_cast<int32_t*>(nullptr), static_cast<uint8_t*>(nullptr),
                                static_cast<int32_t*>(nullptr), static_cast<int64_t*>(nullptr), static_cast<int>(n),
                                ctx->stream);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_ids = carve(n * 4), o_fl = carve(n), o_cnt = carve(16), o_cub = carve(cub_sel);
  if ((st = kg_work_reserve(ctx, off)) != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  int32_t* ids = reinterpret_cast<int32_t*>(w + o_ids);
  uint8_t* fl = reinterpret_cast<uint8_t*>(w + o_fl);
  int64_t* cnt = reinterpret_cast<int64_t*>(w + o_cnt);
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_text_flags, dim3(tgrid(ctx, n)), dim3(kTBlock), 0, ctx->stream, *col, P, d_rows_in, n, ids, fl);
  KG_LAUNCH_CHECK(ctx);
  size_t cb = cub_sel;
  KG_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(w + o_cub, cb, ids, fl, d_rows_out, cnt, static_cast<int>(n),
                                                  ctx->stream));
  kg_time_end(ctx);
  int64_t h = 0;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *out_count = h;
  return KG_OK;
}

extern "C" int kg_text_minmax(kg_ctx* ctx, const kg_text_col* col, const int32_t* d_rows, int64_t n, int32_t is_max,
                              int64_t* out_pos, char* out_buf, int64_t buf_cap, int64_t* out_len) {
  if (!ctx) return KG_ERR_INVALID;
  int st = check_text(ctx, col);
  if (st != KG_OK) return st;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count out of range");
  KG_REQUIRE(ctx, out_pos && out_len, "null output pointer");
  *out_pos = -1;
  *out_len = 0;
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const int nb = tgrid(ctx, n);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_best = carve(nb * 8), o_out = carve(16);
  if ((st = kg_work_reserve(ctx, off)) != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  int64_t* best = reinterpret_cast<int64_t*>(w + o_best);
  int64_t* dout = reinterpret_cast<int64_t*>(w + o_out);
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_text_best, dim3(nb), dim3(kTBlock), 0, ctx->stream, *col, d_rows, n, is_max ? 1 : 0, best);
  hipLaunchKernelGGL(k_text_best_final, dim3(1), dim3(64), 0, ctx->stream, *col, d_rows, best, nb, is_max ? 1 : 0,
                     dout);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  int64_t pos = -1;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(&pos, dout, 8, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *out_pos = pos;
  if (pos < 0) return KG_OK;
  // Read-backs on the context's stream (a null-stream copy would also wait
  // for every other stream's work on the device).
  int32_t row = static_cast<int32_t>(pos);
  if (d_rows) {
    KG_HIP_CHECK(ctx, hipMemcpyAsync(&row, d_rows + pos, 4, hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  }
  int64_t ab[2];
  KG_HIP_CHECK(ctx, hipMemcpyAsync(ab, col->offsets + row, 16, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *out_len = ab[1] - ab[0];
  if (out_buf && buf_cap > 0) {
    const int64_t m = std::min(buf_cap, *out_len);
    if (m > 0) {
      KG_HIP_CHECK(ctx, hipMemcpyAsync(out_buf, col->bytes + ab[0], m, hipMemcpyDeviceToHost, ctx->stream));
      KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
  }
  return KG_OK;
}

extern "C" int kg_index_join_expand(kg_ctx* ctx, int64_t n, int32_t width, const int32_t* d_tuples, int32_t src_pos,
                                    const kg_col* src_key, const int32_t* d_src_rows, const kg_index* idx,
                                    int64_t out_cap, int32_t* d_out, int64_t* out_n) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, idx && src_key && (n == 0 || src_key->data) && out_n, "null pointer argument");
  KG_REQUIRE(ctx, width >= 1 && width <= 15 && src_pos >= 0 && src_pos < width, "bad width / position");
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "tuple count out of range");
  KG_REQUIRE(ctx, d_tuples || width == 1, "d_tuples is NULL with width %d (the identity tuples have width 1)", width);
  KG_REQUIRE(ctx, out_cap == 0 || d_out, "d_out is NULL");
  if (type_width(src_key->type) == 0)
    return kg_fail(ctx, KG_ERR_UNSUPPORTED, "join keys must be fixed-width (type %d)", src_key->type);
  *out_n = 0;
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  IndexDir d;
  const int32_t* tids = nullptr;
  index_view(idx, &d, &tids);
  Expand e{width, src_pos, *src_key, d_src_rows, tids};
  size_t cub_scan = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, cub_scan, static_cast<int64_t*>(nullptr), static_cast<int64_t*>(nullptr),
                                   static_cast<int>(n + 1), ctx->stream);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_boff = carve(n * 8), o_bcnt = carve((n + 1) * 8), o_start = carve((n + 1) * 8),
               o_cub = carve(cub_scan);
  int st = kg_work_reserve(ctx, off);
  if (st != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  int64_t* boff = reinterpret_cast<int64_t*>(w + o_boff);
  int64_t* bcnt = reinterpret_cast<int64_t*>(w + o_bcnt);
  int64_t* starts = reinterpret_cast<int64_t*>(w + o_start);
  hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, ctx->stream, bcnt + n, 1);
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_expand_count, dim3(tgrid(ctx, n)), dim3(kTBlock), 0, ctx->stream, e, d, d_tuples, n, boff, bcnt);
  KG_LAUNCH_CHECK(ctx);
  size_t cb = cub_scan;
  KG_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveSum(w + o_cub, cb, bcnt, starts, static_cast<int>(n + 1), ctx->stream));
  int64_t total = 0;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(&total, starts + n, 8, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *out_n = total;
  if (total > out_cap) {
    kg_time_end(ctx);
    return kg_fail(ctx, KG_ERR_CAPACITY, "%lld joined tuples exceed out_cap %lld", (long long)total,
                   (long long)out_cap);
  }
  if (total > 0)
    hipLaunchKernelGGL(k_expand_write, dim3(tgrid(ctx, total)), dim3(kTBlock), 0, ctx->stream, e, d_tuples, n, boff,
                       starts, total, d_out);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  return KG_OK;
}

extern "C" int kg_tuple_rows(kg_ctx* ctx, const int32_t* d_tuples, int32_t width, int32_t pos, const int32_t* d_sel,
                             int64_t n, int32_t* d_out) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, width >= 1 && pos >= 0 && pos < width, "bad width / position");
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "tuple count out of range");
  KG_REQUIRE(ctx, n == 0 || (d_tuples && d_out), "null pointer argument");
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(k_tuple_rows, dim3(tgrid(ctx, n)), dim3(kTBlock), 0, ctx->stream, d_tuples, width, pos, d_sel, n,
                     d_out);
  KG_LAUNCH_CHECK(ctx);
  return KG_OK;
}

extern "C" int kg_gather(kg_ctx* ctx, const kg_col* src, const int32_t* d_rows, int64_t n, void* d_out,
                         int8_t* d_out_nulls) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, src && src->data, "source column is NULL");
  const int w = type_width(src->type);
  if (w == 0) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "gather of type %d", src->type);
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count out of range");
  KG_REQUIRE(ctx, n == 0 || (d_rows && d_out), "null pointer argument");
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(k_gather, dim3(tgrid(ctx, n)), dim3(kTBlock), 0, ctx->stream,
                     static_cast<const uint8_t*>(src->data), src->nulls, w, d_rows, n, static_cast<uint8_t*>(d_out),
                     d_out_nulls);
  KG_LAUNCH_CHECK(ctx);
  return KG_OK;
}

extern "C" int kg_text_hash(kg_ctx* ctx, const kg_text_col* col, const int32_t* d_rows, int64_t n, int64_t* d_out,
                            int8_t* d_out_nulls) {
  if (!ctx) return KG_ERR_INVALID;
  int st = check_text(ctx, col);
  if (st != KG_OK) return st;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count out of range");
  KG_REQUIRE(ctx, n == 0 || d_out, "d_out is NULL");
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(k_text_hash, dim3(tgrid(ctx, n)), dim3(kTBlock), 0, ctx->stream, *col, d_rows, n, d_out,
                     d_out_nulls);
  KG_LAUNCH_CHECK(ctx);
  return KG_OK;
}

extern "C" int kg_text_pairs_equal(kg_ctx* ctx, const kg_text_col* a, const int32_t* d_rows_a, const kg_text_col* b,
                                   const int32_t* d_rows_b, int64_t n, int32_t* d_out_idx, int64_t* out_count) {
  if (!ctx) return KG_ERR_INVALID;
  int st = check_text(ctx, a);
  if (st != KG_OK || (st = check_text(ctx, b)) != KG_OK) return st;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "pair count out of range");
  KG_REQUIRE(ctx, out_count && (n == 0 || d_out_idx), "null output pointer");
  *out_count = 0;
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  size_t cub_sel = 0;
  KG_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(nullptr, cub_sel, static_cast<int32_t*>(nullptr),
                                                  static_cast<uint8_t*>(nullptr), static_cast<int32_t*>(nullptr),
                                                  static_cast<int64_t*>(nullptr), static_cast<int>(n), ctx->stream));
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_ids = carve(n * 4), o_fl = carve(n), o_cnt = carve(16), o_cub = carve(cub_sel);
  if ((st = kg_work_reserve(ctx, off)) != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  int32_t* ids = reinterpret_cast<int32_t*>(w + o_ids);
  uint8_t* fl = reinterpret_cast<uint8_t*>(w + o_fl);
  int64_t* cnt = reinterpret_cast<int64_t*>(w + o_cnt);
  hipLaunchKernelGGL(k_text_pairs_eq, dim3(tgrid(ctx, n)), dim3(kTBlock), 0, ctx->stream, *a, d_rows_a, *b, d_rows_b,
                     n, ids, fl);
  KG_LAUNCH_CHECK(ctx);
  size_t cb = cub_sel;
  KG_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(w + o_cub, cb, ids, fl, d_out_idx, cnt, static_cast<int>(n),
                                                  ctx->stream));
  int64_t h = 0;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *out_count = h;
  return KG_OK;
}

extern "C" int kg_text_hash_verify(kg_ctx* ctx, const kg_text_col* col, const int32_t* d_rows, const int64_t* d_hash,
                                   int64_t n, int64_t* out_mismatches) {
  if (!ctx) return KG_ERR_INVALID;
  int st = check_text(ctx, col);
  if (st != KG_OK) return st;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count out of range");
  KG_REQUIRE(ctx, out_mismatches && (n == 0 || d_hash), "null pointer argument");
  *out_mismatches = 0;
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  uint64_t nslots = 64;
  while (nslots < 2 * static_cast<uint64_t>(n)) nslots <<= 1;
  char* buf = nullptr;
  const size_t kb = nslots * 8, rb = (nslots * 4 + 255) / 256 * 256;
  if ((st = kg_dev_alloc(ctx, reinterpret_cast<void**>(&buf), kb + rb + 256)) != KG_OK) return st;
  int64_t* keys = reinterpret_cast<int64_t*>(buf);
  int32_t* reps = reinterpret_cast<int32_t*>(buf + kb);
  unsigned long long* count = reinterpret_cast<unsigned long long*>(buf + kb + rb);
  const int grid = tgrid(ctx, static_cast<int64_t>(std::max<uint64_t>(nslots, static_cast<uint64_t>(n))));
  hipLaunchKernelGGL(k_code_fill, dim3(grid), dim3(kTBlock), 0, ctx->stream, keys, reps,
                     static_cast<int64_t>(nslots), count);
  hipLaunchKernelGGL(k_code_claim, dim3(tgrid(ctx, n)), dim3(kTBlock), 0, ctx->stream, *col, d_rows, d_hash, n, keys,
                     reps, nslots - 1);
  hipLaunchKernelGGL(k_code_check, dim3(tgrid(ctx, n)), dim3(kTBlock), 0, ctx->stream, *col, d_rows, d_hash, n,
                     keys, reps, nslots - 1, count);
  hipError_t e = hipGetLastError();
  unsigned long long h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, count, 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  kg_dev_release(ctx, buf);
  if (e != hipSuccess) return kg_fail(ctx, KG_ERR_HIP, "text code check failed: %s", hipGetErrorString(e));
  *out_mismatches = static_cast<int64_t>(h);
  return KG_OK;
}
