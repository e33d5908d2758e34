// Internal definitions shared by the kushgpu HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "kushgpu.h"

struct kg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int deterministic = 0;  // kg_ctx_set_deterministic
  // Region count for partitioned join probes: -1 chooses by table size, 0 probes
  // directly, otherwise a multiple of 8 (kg_ctx_set_probe_regions).
  int probe_regions = -1;
  int probe_wg_per_cu = 0;    // probe workgroups per CU; 0 = 32
  // Join builds: -1 chooses (region-partitioned for large builds into tables
  // beyond the caches), 0 always inserts directly, > 0 at least that many
  // regions (kg_ctx_set_build_regions).
  int build_regions = -1;
  // Probe + GROUP BY matched key (kg_ht_probe_agg): -1 chooses (probe rows
  // partitioned by table region and aggregated per region in LDS for large
  // probes into tables beyond the caches), 0 always uses per-slot device
  // atomics, > 0 partitions with at least that many regions
  // (kg_ctx_set_agg_regions).
  int agg_regions = -1;
  // Join tuning choices per (build rows, key cardinality, skew) class.
  struct JoinTune {
    uint64_t cls;
    double load_factor;
    int32_t wg_per_cu, regions;
    double probe_ms;
  };
  std::vector<JoinTune> join_tunes;
  // Rows per lane step of the last fixed-order aggregation (1 or 4): with
  // kg_agg_blocks_for(n) it defines that aggregation's FP64 addition order.
  int last_agg_vec = 1;
  // The GROUP BY path of the last aggregation call (kg_ctx_last_path).
  const char* last_path = "";
  int num_cus = 256;
  int64_t max_lds = 64 * 1024;  // LDS bytes one workgroup may use (160 KB on gfx950)
  std::string err;
  // Grow-only device workspace (block partials, counters) and a small pinned
  // host buffer for scalar read-backs.
  void* work = nullptr;
  size_t work_bytes = 0;
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  // Dominant-kernel timing: event pairs recorded since the last stats reset.
  int timing = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
  // Columns loaded from the reference's on-disk files, resident by path until
  // released (kg_file.hip); freed with the context.
  struct FileCol {
    std::string path;
    int32_t type = 0;
    int64_t n = 0;
    void* data = nullptr;       // typed array, or TEXT bytes
    int8_t* nulls = nullptr;
    int64_t* offsets = nullptr; // TEXT: n + 1 offsets into data
  };
  std::vector<FileCol> files;
  // Device address of this device's static scratch array (kg_context.hip),
  // used by the entry points that allocate nothing.
  void* scratch = nullptr;
  // Device blocks handed out by kg_dev_alloc (pointer -> bytes), and released
  // blocks kept for reuse (bytes -> pointer), at most kg::kBlockCacheBytes in
  // all. Reuse is ordered on ctx->stream, where every user of these blocks
  // enqueues its work.
  std::multimap<size_t, void*> block_cache;
  std::unordered_map<void*, size_t> block_live;
  size_t block_cache_bytes = 0;
  // Contexts of the batched serving loop's plain lanes (kg_serve): queries
  // the batches do not run execute there, each lane on its own stream and
  // host thread. Created on first use, destroyed with this context.
  std::vector<kg_ctx*> serve_lanes;
};

// Device memory for per-query structures (join tables, their accumulators):
// a released block of at least `bytes` (and at most about twice that) from
// the context's cache, else hipMalloc. kg_dev_release returns a block to the
// cache (or frees it when the cache is full or the block is not the
// context's). A plain plan creates and frees a join table per query; on the
// device a hipMalloc / hipFree pair costs more than a small build and probe.
int kg_dev_alloc(kg_ctx* ctx, void** p, size_t bytes);
void kg_dev_release(kg_ctx* ctx, void* p);

// Frees the device memory of a loaded file column.
void kg_file_col_free(kg_ctx::FileCol& f);

// Records a timing start/stop on ctx->stream when timing is on (no-op otherwise).
// Usage: kg_time_begin(ctx); launch...; kg_time_end(ctx);
void kg_time_begin(kg_ctx* ctx);
void kg_time_end(kg_ctx* ctx);

// Lets `kernel` launch with up to `bytes` of dynamic LDS (beyond the 64 KB
// default; at most ctx->max_lds).
int kg_allow_lds(kg_ctx* ctx, const void* kernel, size_t bytes);

// Sets ctx->err and returns the status.
int kg_fail(kg_ctx* ctx, int status, const char* fmt, ...);

// Ensures ctx->work holds at least `bytes`; returns KG_OK or an error status.
int kg_work_reserve(kg_ctx* ctx, size_t bytes);
int kg_pinned_reserve(kg_ctx* ctx, size_t bytes);

#define KG_HIP_CHECK(ctx, expr)                                                       \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return kg_fail((ctx), KG_ERR_HIP, "%s failed: %s (%s:%d)", #expr,              \
                     hipGetErrorString(e_), __FILE__, __LINE__);                      \
  } while (0)

#define KG_LAUNCH_CHECK(ctx)                                                          \
  do {                                                                                \
    hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess)                                                             \
      return kg_fail((ctx), KG_ERR_HIP, "kernel launch failed: %s (%s:%d)",           \
                     hipGetErrorString(e_), __FILE__, __LINE__);                      \
  } while (0)

#define KG_REQUIRE(ctx, cond, ...)                                                    \
  do {                                                                                \
    if (!(cond)) return kg_fail((ctx), KG_ERR_INVALID, __VA_ARGS__);                  \
  } while (0)

namespace kg {

constexpr int kWave = 64;

// Join-table memory reuse (kg_join.hip): whether t's key / row memory holds a
// table of capacity_rows at load_factor, and resetting it to such an empty
// table (free markers, counters, form) without allocating.
bool ht_fits(const kg_hash_table* t, int64_t capacity_rows, double load_factor);
int ht_reset(kg_ctx* ctx, kg_hash_table* t, int64_t capacity_rows, double load_factor);
// The load factor kg_join_tune chose on this context for builds of about
// build_rows rows (same power-of-two class, unique keys preferred), or 0.5.
double tuned_load_factor(const kg_ctx* ctx, int64_t build_rows);

// Batch footprint parts (kg_batch.hip), for admission: a query's own result
// area and accumulators, and its build structures, which the queries of one
// batch that share a build source (same build key column, rows and group
// codes; see kg_batch_submit) hold once.
struct BatchSource {
  const void* key;
  const int8_t* nulls;
  const void* codes;
  int64_t bn;
  int32_t type, G;
  bool operator==(const BatchSource& o) const {
    return key == o.key && nulls == o.nulls && codes == o.codes && bn == o.bn && type == o.type && G == o.G;
  }
};
size_t batch_own_bytes(const kg_query& q);
// Whether kg_batch_submit orders the batch after the context stream's prior
// work (default on; the serving loop turns it off, see kg_batch).
void batch_order_after_ctx(kg_batch* b, bool on);
size_t batch_source_bytes(const kg_query& q);
bool batch_source_id(const kg_query& q, BatchSource* id);

// The static device scratch of the no-allocation entry points: tile counts
// (int32) and tile offsets (int64) of kg_scan_select_noalloc for up to
// kNoAllocMaxTiles tiles of 4096 rows, then the selected-row total.
constexpr int64_t kNoAllocMaxTiles = 131072;
constexpr int64_t kNoAllocScratchBytes = kNoAllocMaxTiles * 12 + 256;

// Bytes of released device blocks a context keeps for reuse (kg_dev_alloc).
constexpr size_t kBlockCacheBytes = size_t{2} << 30;

// Width in bytes of a column's values for the SQL type ids of kushgpu.h.
inline int type_width(int32_t type) {
  switch (type) {
    case KG_SMALLINT: return 2;
    case KG_INT:
    case KG_DATE:
    case KG_ENUM: return 4;
    case KG_BIGINT:
    case KG_REAL: return 8;
    case KG_BOOLEAN: return 1;
    default: return 0;
  }
}

// Column access on the device: integers sign-extended to int64, REAL as double.
__device__ __forceinline__ int64_t load_int(int32_t type, const void* p, int64_t i) {
  switch (type) {
    case KG_SMALLINT: return static_cast<const int16_t*>(p)[i];
    case KG_BIGINT: return static_cast<const int64_t*>(p)[i];
    case KG_BOOLEAN: return static_cast<const int8_t*>(p)[i];
    default: return static_cast<const int32_t*>(p)[i];
  }
}

__device__ __forceinline__ bool cmp_i64(int32_t op, int64_t x, int64_t l) {
  switch (op) {
    case KG_EQ: return x == l;
    case KG_NEQ: return x != l;
    case KG_LT: return x < l;
    case KG_LEQ: return x <= l;
    case KG_GT: return x > l;
    default: return x >= l;
  }
}

__device__ __forceinline__ bool cmp_f64(int32_t op, double x, double l) {
  switch (op) {
    case KG_EQ: return x == l;
    case KG_NEQ: return x != l;
    case KG_LT: return x < l;
    case KG_LEQ: return x <= l;
    case KG_GT: return x > l;
    default: return x >= l;
  }
}

// 64-bit finalizer (MurmurHash3 fmix64) used to place keys in the device
// tables. Placement is not visible in results, so it need not match the
// reference's 32-bit HashCombine bucket choice.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace kg
