// Batched execution: many queries' pipelines in one persistent kernel fed by
// device work queues, captured once as a HIP graph per batch slot.
//
// Two pipeline shapes are batched, the ones a host plan hands to the plain
// entry points one query at a time:
//   SCAN_AGG  ScanSelect -> GroupByAggregate over dense code keys
//             (scan_select_translator.cc:92-156 ->
//             group_by_aggregate_translator.cc:36-111), as kg_scan_group_agg;
//   JOIN_AGG  ScanSelect(build) -> HashJoin build, then ScanSelect(probe) ->
//             HashJoin probe -> GroupByAggregate on a build-side code column or
//             the ungrouped Aggregate (hash_join_translator.cc:85-139), as
//             kg_ht_build + kg_scan_join_group_agg.
// A batch runs as one captured graph: fills (accumulators, table slots), the
// build phase (work items = (query, build tile)), the scan/probe kernels, and
// a finalize for queries without scan/probe items. A query's result area lives
// in pinned host memory; the wave that finishes its last item folds the
// accumulator copies into it and publishes a completion word.
//
// Scans with dense code keys and non-NULL columns (the register path) run as
// shared scans: the queries over one table form groups whose tiles (512 or
// 1024 rows) are dealt round robin to the workgroups; the group's columns are
// staged once into LDS and every query of the group reads the tile from
// there. Other scan/probe work items are (query, tile of 256 to kBTile rows,
// one size per queue, smaller when the queue's queries hold few rows in all);
// queries over the same table share its tiles through L2: the items of tile k
// of a table are adjacent and sit in queue k mod 8, one queue per XCD (the
// waves of workgroup b take the items of queue b mod 8, the XCD the
// dispatcher places it on: each wave's first item by its index, the rest in
// queue order from the queue's counter). There is no inter-workgroup waiting,
// so any grid of at least eight workgroups is correct.
//
// Register-path partials live in lane registers, reduced across the wave by a
// butterfly; other queries use a wave-private LDS table updated with LDS
// atomics (up to kBLdsWords 8-byte words per workgroup). Partials go to one of
// kBSlots accumulator copies per query with global atomics; the wave that
// finishes a query's last item folds the copies. FP64 sums are therefore
// added in a timing-dependent order: a query's result in a mix equals its
// result alone within FP64 rounding (integers, counts and group order exactly).
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "kg_table.h"

namespace kg {

constexpr int kBTile = 16384;         // most rows per scan/probe work item
constexpr int kBTileMin = 256;        // fewest rows per scan/probe work item (one row group per lane)
constexpr int64_t kBItemTarget = 1024;  // fewest scan/probe items a queue aims for (before kBTileMin)
constexpr int kBBuildTile = 512;      // rows per build work item
constexpr int kBThreads = 256;        // threads per workgroup
constexpr int kBWaves = kBThreads / kWave;
constexpr int kBQueues = 8;           // scan/probe work queues (one per XCD)
constexpr int kBHeadStride = 32;      // 4-byte words between the queues' item counters
constexpr int kBMaxIvl = 8;           // key intervals of a shared build source's predicate union
constexpr int kRangeCopies = 16;      // copies of a direct map's key range, one 128-byte line each
constexpr int kRangeStride = 16;      // int64 words between copies
constexpr size_t kRangeBytes = sizeof(int64_t) * kRangeCopies * kRangeStride;
constexpr int kBSlots = 64;           // accumulator copies per query (8 per XCD)
constexpr int kBMaxQueries = 256;     // queries per batch
constexpr int kBRegG = 8;             // register path: at most this many groups
constexpr int kBLdsWords = 2048;      // LDS path: 8-byte words of the item table
constexpr int kBFastPreds = 4;
constexpr int kBMaxAggs = 16;
constexpr int kBMaxFills = 8 * kBMaxQueries + 8;
constexpr int kBMapMax = 1 << 21;     // direct-map joins: at most this many key offsets
constexpr int kBBigMapMax = 1 << 25;  // map-only joins (build sides beyond kBMapMax rows): map entries
constexpr uint32_t kMapEmpty = 0xFFFFFFFFu;

// kFlagNoMap (the direct map of a join cannot be used) is reported only for a
// map-only build source, as KG_ERR_UNSUPPORTED.
enum BFlag : uint32_t { kFlagEmptyKey = 1u, kFlagBadCode = 2u, kFlagOverflow = 4u, kFlagNoMap = 8u };

// Register-path view of a scan: 4-byte non-NULL predicate, key and INT
// measure columns, non-NULL REAL measures, all 16-byte aligned (vector loads).
struct BFast {
  int32_t npred;
  int32_t op[kBFastPreds];
  int64_t lit[kBFastPreds];
  const int32_t* pcol[kBFastPreds];
  int32_t nk;
  int32_t card1;
  const int32_t* key[2];
  int32_t real[kMaxMeasures];
  const void* mcol[kMaxMeasures];
  // Columns in the staged tile of the query's shared-scan group.
  int8_t pcx[kBFastPreds];
  int8_t kcx[2];
  int8_t mcx[kMaxMeasures];
};

// Typed view of a join's probe side: 4-byte non-NULL predicate columns, a
// non-NULL 4- or 8-byte integer probe key, at most two non-NULL measures
// (REAL or 4-byte integer, optionally times a REAL / 4-byte integer factor),
// all 16-byte aligned (vector loads).
struct BJoinFast {
  int32_t npred;
  int32_t op[kBFastPreds];
  int64_t lit[kBFastPreds];
  const int32_t* pcol[kBFastPreds];
  const void* key;
  int32_t key64;
  int32_t mkind[2];     // 1 REAL, 2 4-byte integer
  int32_t fkind[2];     // 0 none, 1 REAL, 2 4-byte integer
  const void* mcol[2];
  const void* fcol[2];
};

// Device descriptor of one query.
struct BQuery {
  int32_t kind;         // KG_Q_SCAN_AGG / KG_Q_JOIN_AGG
  int32_t reg;          // 1: register path (BFast valid)
  int32_t jfast;        // 1: typed join path (J valid)
  BFast F;
  BJoinFast J;
  int32_t G;            // dense groups
  int32_t nm;           // measures
  int64_t n;            // scanned / probing rows
  PredSet P;            // scan / probe side predicates
  KeySet K;             // SCAN_AGG group keys
  MeasureSet M;
  kg_col pkey;          // JOIN_AGG probe key
  const int32_t* codes; // JOIN_AGG build row -> group code (nullptr: group 0)
  // Direct map of a typed join with a small build side: entry key - min for
  // keys in [min, min + msize) = group code << 24 | build row (kMapEmpty: no
  // row), usable when the qualifying build keys are unique and span at most
  // msize values. krange = {-min, max} of the qualifying build keys, held in
  // kRangeCopies copies kRangeStride words apart (see k_batch_range).
  uint32_t* map;
  int64_t* krange;
  int32_t msize;
  // Build source: typed joins over the same build key column, row count and
  // group codes in one batch share one map (and table) built over the build
  // rows that pass any of their build predicates (build_all, on the source's
  // first query); each then re-checks its own
  // build predicates at a match (recheck): on the probe key when every one is
  // a conjunct on the build key column (kp_*), else on the build row. A build
  // side beyond kBMapMax rows has the map only (map_only): if it cannot be used
  // the query reports KG_ERR_UNSUPPORTED. sflag: the source's build-phase flags.
  uint32_t* sflag;
  int32_t map_only, build_all, recheck, nkp, kp_other;
  int32_t bmem_lo, bmem_hi;   // build_all: the source's queries, BHeader::bmem[bmem_lo .. bmem_hi)
  // The query that builds (alone, or build_all for a source), when every
  // query's build predicates compare the build key with literals in one
  // conjunction: the union of their key intervals (nivl merged, ascending,
  // inclusive bounds), tested in place of the predicates; nivl < 0 otherwise.
  int32_t nivl;
  int64_t ivl_lo[kBMaxIvl], ivl_hi[kBMaxIvl];
  int32_t kp_op[kBFastPreds];
  int64_t kp_lit[kBFastPreds];
  TableView t;          // JOIN_AGG table (slots in the arena)
  int64_t bn;           // JOIN_AGG build rows
  PredSet PB;           // build side predicates
  kg_col bkey;          // build key
  // Accumulators in the arena: sum/nn [kBSlots][G][nm], rows [kBSlots][G],
  // first [kBSlots][G] (min over copies), flag (BFlag bits).
  uint64_t* sum;
  uint64_t* nn;
  uint64_t* rows;
  uint64_t* first;
  uint32_t* flag;
  char* res;            // result area (pinned host memory, written by the fold)
  unsigned long long* rem;  // scan/probe items not yet finished
  int32_t* done;        // host-visible completion word, set to `seq` after the fold
  int32_t seq;          // batch sequence number
  int32_t items;        // scan/probe items of the query (0: folded by k_batch_final)
  int32_t naggs;
  int32_t first_shift;  // first >> first_shift = the scanned / probe row
  const int8_t* cnull[kBMaxAggs];  // COUNT(col): NULL array of the column
};

struct BFill {
  uint64_t* p;
  int64_t n;
  uint64_t v;
};

// Scan/probe work of one kernel: tables (queries sharing table_id and n) and
// one queue per XCD. Queue x holds, table by table, the tiles k = x mod 8,
// each followed by all of that table's queries.
// Shared-scan work: groups (SGroup) in order, and the prefix of their tile
// counts; tile t of the sequence is taken by workgroup t mod gridDim.x.
struct STiles {
  int32_t ng;
  int32_t tpre[kBMaxQueries + 1];
};

struct BQueues {
  int32_t nt;
  int32_t tile;                               // rows per item (a power of two, kBTileMin..kBTile)
  int32_t qpre[kBQueues][kBMaxQueries + 1];   // per queue, item prefix over tables
  int32_t toff[kBMaxQueries + 1];             // table t's queries: tq[toff[t] .. toff[t+1])
  int32_t tq[kBMaxQueries];
  int64_t tn[kBMaxQueries];                   // rows of table t
  // Per queue, the count of items taken past the first W (see drain_wave),
  // each on its own 128-byte line.
  unsigned int next[kBQueues * kBHeadStride];
};

// Batch header, copied to the device with the query descriptors.
struct BHeader {
  int32_t nq, nfills, nbq, pad;
  int32_t bpre[kBMaxQueries + 1];             // build phase: item prefix over bq entries
  int32_t bq[kBMaxQueries];                   // build phase: query of entry j
  int32_t bmem[kBMaxQueries];                 // queries of each shared build source
  int32_t nrq;                                // key-range phase (direct-map joins only):
  int32_t rpre[kBMaxQueries + 1];             //   item prefix over rq entries
  int32_t rq[kBMaxQueries];                   //   query of entry j
  STiles shared;                              // register-path scans, by group
  BQueues join;                               // joins with a typed view
  BQueues gen;                                // the others
  BFill fills[kBMaxFills];
};

// Result area of one query: flag, then AggFinal[G*max(nm,1)], rows[G],
// first[G], fnull[naggs][G] (padded to 8 bytes).
struct ResLayout {
  size_t fin, rows, first, fnull, bytes;
};

__host__ __device__ inline ResLayout res_layout(int G, int nm, int naggs) {
  ResLayout L;
  L.fin = 16;
  L.rows = L.fin + static_cast<size_t>(G) * (nm > 0 ? nm : 1) * sizeof(AggFinal);
  L.first = L.rows + static_cast<size_t>(G) * 8;
  L.fnull = L.first + static_cast<size_t>(G) * 8;
  L.bytes = (L.fnull + static_cast<size_t>(naggs) * G + 255) / 256 * 256;
  return L;
}

__device__ __forceinline__ uint64_t acc_add(bool real, uint64_t a, uint64_t b) {
  return real ? fadd_bits(a, b) : a + b;
}

__device__ __forceinline__ int find_prefix(const int32_t* pre, int n, int32_t it) {
  // Last j in [0, n) with pre[j] <= it (pre ascending, pre[0] = 0).
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= it) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kBThreads) void k_batch_clear(const BHeader* h) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBThreads;
  const int nf = h->nfills;
  for (int f = 0; f < nf; f++) {
    const BFill F = h->fills[f];
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBThreads + threadIdx.x; i < F.n; i += stride) F.p[i] = F.v;
  }
}

__device__ __forceinline__ uint64_t coherent_load(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Whether build row r goes into query d's map / table (its build predicates,
// or for a shared source (build_all) those of any of the source's queries,
// tested as key intervals where they have that form), with its key in *k. A
// row with a NULL key goes nowhere.
__device__ __forceinline__ bool build_row_key(const BHeader* h, const BQuery* qs, const BQuery& d, int64_t r,
                                              int64_t& k) {
  if (d.nivl >= 0) {
    if (!load_key(d.bkey, r, k)) return false;
    for (int i = 0; i < d.nivl; i++)
      if (k >= d.ivl_lo[i] && k <= d.ivl_hi[i]) return true;
    return false;
  }
  if (!d.build_all) {
    if (!eval_preds(d.PB, r)) return false;
  } else {
    bool any = false;
    for (int m = d.bmem_lo; m < d.bmem_hi && !any; m++) any = eval_preds(qs[h->bmem[m]].PB, r);
    if (!any) return false;
  }
  return load_key(d.bkey, r, k);
}

// Key range of a direct-map join: min and max of the qualifying build keys,
// over the build items (before the build phase). Items are dealt by a grid
// stride (no shared counter), so a workgroup's items ascend and those of one
// query are consecutive: the workgroup accumulates its range over them and
// reduces it once per query, when the query changes and at the end. It then
// touches the words of copy blockIdx mod kRangeCopies, each copy on its own
// cache line, only when it widens them (every workgroup finds the initial
// range, so with one copy all their atomics would queue on one line).
__global__ __launch_bounds__(kBThreads) void k_batch_range(BHeader* h, const BQuery* qs) {
  __shared__ int64_t s_red[2][kBWaves];
  constexpr int R = kBBuildTile / kBThreads;  // rows per thread and item
  static_assert(R * kBThreads == kBBuildTile, "build items are whole rows per thread");
  const int32_t total = h->rpre[h->nrq];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  int64_t nmin = INT64_MIN, mx = INT64_MIN;  // max of -key, max of key
  // Uniform over the workgroup (it and j depend on blockIdx only).
  auto flush = [&](const BQuery& d) {
    for (int s = kWave / 2; s >= 1; s >>= 1) {
      const int64_t x = static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(nmin), s));
      const int64_t y = static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(mx), s));
      nmin = x > nmin ? x : nmin;
      mx = y > mx ? y : mx;
    }
    if (lane == 0) {
      s_red[0][wave] = nmin;
      s_red[1][wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < kBWaves; w++) {
        nmin = s_red[0][w] > nmin ? s_red[0][w] : nmin;
        mx = s_red[1][w] > mx ? s_red[1][w] : mx;
      }
      if (mx != INT64_MIN) {
        int64_t* kr = d.krange + static_cast<int64_t>(blockIdx.x % kRangeCopies) * kRangeStride;
        if (nmin > static_cast<int64_t>(coherent_load(reinterpret_cast<const uint64_t*>(&kr[0]))))
          atomicMax(reinterpret_cast<long long*>(&kr[0]), static_cast<long long>(nmin));
        if (mx > static_cast<int64_t>(coherent_load(reinterpret_cast<const uint64_t*>(&kr[1]))))
          atomicMax(reinterpret_cast<long long*>(&kr[1]), static_cast<long long>(mx));
      }
    }
    __syncthreads();
    nmin = mx = INT64_MIN;
  };
  int j = -1;
  for (int32_t it = blockIdx.x; it < total; it += gridDim.x) {
    int jj = j < 0 ? find_prefix(h->rpre, h->nrq, it) : j;
    while (h->rpre[jj + 1] <= it) jj++;
    if (jj != j) {
      if (j >= 0) flush(qs[h->rq[j]]);
      j = jj;
    }
    const BQuery& d = qs[h->rq[j]];
    const int64_t lo = static_cast<int64_t>(it - h->rpre[j]) * kBBuildTile;
    const int64_t hi = lo + kBBuildTile < d.bn ? lo + kBBuildTile : d.bn;
#pragma unroll
    for (int e = 0; e < R; e++) {
      const int64_t r = lo + threadIdx.x + static_cast<int64_t>(e) * kBThreads;
      int64_t k;
      if (r >= hi || !build_row_key(h, qs, d, r, k) || k == kEmpty) continue;
      nmin = -k > nmin ? -k : nmin;
      mx = k > mx ? k : mx;
    }
  }
  if (j >= 0) flush(qs[h->rq[j]]);
}

// Whether a direct-map join can use its map (range known, within the map, no
// repeated key); *mn and *span receive the map's base key and span. Called by
// all 64 lanes of a wave (lane c reads range copy c).
__device__ __forceinline__ bool map_usable(const BQuery& d, int64_t* mn, int64_t* span) {
  if (!d.map) return false;
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t a = static_cast<uint64_t>(INT64_MIN), b = static_cast<uint64_t>(INT64_MIN);
  if (lane < kRangeCopies) {
    a = coherent_load(reinterpret_cast<const uint64_t*>(d.krange + static_cast<int64_t>(lane) * kRangeStride));
    b = coherent_load(reinterpret_cast<const uint64_t*>(d.krange + static_cast<int64_t>(lane) * kRangeStride + 1));
  }
  for (int st = kRangeCopies / 2; st >= 1; st >>= 1) {
    const int64_t x = static_cast<int64_t>(shfl_xor_u64(a, st));
    const int64_t y = static_cast<int64_t>(shfl_xor_u64(b, st));
    a = static_cast<uint64_t>(x > static_cast<int64_t>(a) ? x : static_cast<int64_t>(a));
    b = static_cast<uint64_t>(y > static_cast<int64_t>(b) ? y : static_cast<int64_t>(b));
  }
  const int64_t nmin = static_cast<int64_t>(__shfl(static_cast<unsigned long long>(a), 0));
  const int64_t mx = static_cast<int64_t>(__shfl(static_cast<unsigned long long>(b), 0));
  const uint32_t fl = __hip_atomic_load(d.sflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (fl & kFlagNoMap) return false;
  if (mx == INT64_MIN) {  // no qualifying build row: every probe misses
    *mn = 0;
    *span = 0;
    return true;
  }
  *mn = -nmin;
  const uint64_t sp = static_cast<uint64_t>(mx) - static_cast<uint64_t>(*mn) + 1;
  *span = static_cast<int64_t>(sp);
  return sp <= static_cast<uint64_t>(d.msize);
}

// Build phase: items (query, build tile); same insertion as k_ht_build, and
// for a direct-map join the entry of the key in the map (a repeated key, or a
// group code outside [0, G), turns the map off). Items are dealt by a grid
// stride; a workgroup checks the map once per query (its items of one query
// are consecutive), and the map entries of a thread's rows of an item are
// claimed together.
__global__ __launch_bounds__(kBThreads) void k_batch_build(BHeader* h, const BQuery* qs) {
  constexpr int R = kBBuildTile / kBThreads;  // rows per thread and item
  const int32_t total = h->bpre[h->nbq];
  int j = -1;
  int64_t mn = 0, span = 0;
  bool use_map = false;
  for (int32_t it = blockIdx.x; it < total; it += gridDim.x) {
    int jj = j < 0 ? find_prefix(h->bpre, h->nbq, it) : j;
    while (h->bpre[jj + 1] <= it) jj++;
    const BQuery& d = qs[h->bq[jj]];
    if (jj != j) {
      j = jj;
      use_map = map_usable(d, &mn, &span);
    }
    const int64_t lo = static_cast<int64_t>(it - h->bpre[j]) * kBBuildTile;
    const int64_t hi = lo + kBBuildTile < d.bn ? lo + kBBuildTile : d.bn;
    const TableView t = d.t;
    if (d.map_only && !use_map) {  // no table to fall back to: the queries report it
      if (threadIdx.x == 0) atomicOr(d.sflag, kFlagNoMap);
      continue;
    }
    int64_t key[R];
    bool ok[R];
#pragma unroll
    for (int e = 0; e < R; e++) {
      const int64_t r = lo + threadIdx.x + static_cast<int64_t>(e) * kBThreads;
      key[e] = 0;
      ok[e] = r < hi && build_row_key(h, qs, d, r, key[e]);
      if (ok[e] && key[e] == kEmpty) {
        atomicOr(d.sflag, kFlagEmptyKey);
        ok[e] = false;
      }
    }
    if (use_map) {
      uint32_t prev[R];
#pragma unroll
      for (int e = 0; e < R; e++) {
        const int64_t r = lo + threadIdx.x + static_cast<int64_t>(e) * kBThreads;
        prev[e] = kMapEmpty;
        if (!ok[e]) continue;
        const uint64_t off = static_cast<uint64_t>(key[e]) - static_cast<uint64_t>(mn);
        const int g = d.codes ? d.codes[r] : 0;
        if (off >= static_cast<uint64_t>(span) || g < 0 || g >= d.G || r >= (int64_t{1} << 24)) {
          atomicOr(d.sflag, kFlagNoMap);
        } else {
          const uint32_t ent = (static_cast<uint32_t>(g) << 24) | static_cast<uint32_t>(r);
          prev[e] = atomicCAS(&d.map[off], kMapEmpty, ent);
        }
      }
#pragma unroll
      for (int e = 0; e < R; e++)
        if (prev[e] != kMapEmpty) atomicOr(d.sflag, kFlagNoMap);
    }
    if (d.map_only) continue;
#pragma unroll 1
    for (int e = 0; e < R; e++) {
      if (!ok[e]) continue;
      const int64_t r = lo + threadIdx.x + static_cast<int64_t>(e) * kBThreads;
      const int64_t k = key[e];
      uint64_t b = bucket_of(k, t.nbuckets);
      bool done = false;
      for (uint64_t probes = 0; probes < t.nbuckets && !done; probes++) {
        for (int s = 0; s < kSlots && !done; s++) {
          if (t.keys[b * kSlots + s] != kEmpty) continue;
          const unsigned long long p =
              atomicCAS(reinterpret_cast<unsigned long long*>(t.keys + b * kSlots + s),
                        static_cast<unsigned long long>(kEmpty), static_cast<unsigned long long>(k));
          if (p == static_cast<unsigned long long>(kEmpty)) {
            t.rows[b * kSlots + s] = static_cast<int32_t>(r);
            done = true;
          }
        }
        b = (b + 1 == t.nbuckets) ? 0 : b + 1;
      }
      if (!done) atomicOr(d.sflag, kFlagOverflow);
    }
  }
}

// Whether build row brow (key k, equal to the probe key) passes query d's own
// build predicates, for a query whose map or table holds its source's rows.
__device__ __forceinline__ bool build_recheck(const BQuery& d, int64_t k, int32_t brow) {
  if (!d.recheck) return true;
  if (d.kp_other) return eval_preds(d.PB, brow);
  bool ok = true;
#pragma unroll
  for (int p = 0; p < kBFastPreds; p++)
    if (p < d.nkp) ok = ok && cmp_i64(d.kp_op[p], k, d.kp_lit[p]);
  return ok;
}

// Calls f(brow) for every build row whose key equals k (bucket order).
template <typename F>
__device__ __forceinline__ void for_matches(const TableView& t, int64_t k, F&& f) {
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

// Visits, for the calling wave, the qualifying output rows of rows [lo, hi) of
// query d: for a scan, emit(group, first-key, row) per qualifying row; for a
// join, once per match.
template <typename Emit>
__device__ __forceinline__ void visit_rows(const BQuery& d, int64_t lo, int64_t hi, Emit&& emit) {
  for (int64_t r = lo + (threadIdx.x & (kWave - 1)); r < hi; r += kWave) {
    if (!eval_preds(d.P, r)) continue;
    if (d.kind == KG_Q_SCAN_AGG) {
      const int g = group_of(d.K, r);
      if (static_cast<unsigned>(g) >= static_cast<unsigned>(d.G)) {
        atomicOr(d.flag, kFlagBadCode);
        continue;
      }
      emit(g, static_cast<uint64_t>(r), r);
    } else {
      int64_t k;
      if (!load_key(d.pkey, r, k) || k == kEmpty) continue;
      for_matches(d.t, k, [&](int32_t brow) {
        const int g = d.codes ? d.codes[brow] : 0;
        if (static_cast<unsigned>(g) >= static_cast<unsigned>(d.G)) {
          atomicOr(d.flag, kFlagBadCode);
          return;
        }
        emit(g, (static_cast<uint64_t>(r) << 32) | static_cast<uint32_t>(brow), r);
      });
    }
  }
}

// Vector-load helpers of the typed join path: lane l owns the row groups
// lo + 4 l + 256 k (k < item rows / 256) of 4 consecutive rows of an item,
// read with 16-byte vector loads (FULL) or element by element at the end of a
// table.

template <bool FULL>
__device__ __forceinline__ void ld4_i32(const int32_t* col, int64_t base, int nv, int32_t (&o)[4]) {
  if (FULL) {
    const int4 x = *reinterpret_cast<const int4*>(col + base);
    o[0] = x.x;
    o[1] = x.y;
    o[2] = x.z;
    o[3] = x.w;
  } else {
#pragma unroll
    for (int e = 0; e < 4; e++) o[e] = e < nv ? col[base + e] : 0;
  }
}

// K = 1: REAL (double bits), K = 2: 4-byte integer (sign-extended).
template <bool FULL, int K>
__device__ __forceinline__ void ld4_measure(const void* col, int64_t base, int nv, uint64_t (&o)[4]) {
  if (K == 1) {
    const double* c = static_cast<const double*>(col);
    if (FULL) {
      const double2 a = *reinterpret_cast<const double2*>(c + base);
      const double2 b = *reinterpret_cast<const double2*>(c + base + 2);
      o[0] = static_cast<uint64_t>(__double_as_longlong(a.x));
      o[1] = static_cast<uint64_t>(__double_as_longlong(a.y));
      o[2] = static_cast<uint64_t>(__double_as_longlong(b.x));
      o[3] = static_cast<uint64_t>(__double_as_longlong(b.y));
    } else {
#pragma unroll
      for (int e = 0; e < 4; e++) o[e] = e < nv ? static_cast<uint64_t>(__double_as_longlong(c[base + e])) : 0;
    }
  } else {
    int32_t t[4];
    ld4_i32<FULL>(static_cast<const int32_t*>(col), base, nv, t);
#pragma unroll
    for (int e = 0; e < 4; e++) o[e] = static_cast<uint64_t>(static_cast<int64_t>(t[e]));
  }
}

template <int K>
__device__ __forceinline__ uint64_t kadd(uint64_t a, uint64_t b) {
  return K == 1 ? fadd_bits(a, b) : a + b;
}

template <int K>
__device__ __forceinline__ uint64_t wave_sum(uint64_t a) {
  for (int s = kWave / 2; s >= 1; s >>= 1) a = kadd<K>(a, shfl_xor_u64(a, s));
  return a;
}

template <int K>
__device__ __forceinline__ void atomic_acc(uint64_t* p, uint64_t v) {
  if (K == 1) atomicAdd(reinterpret_cast<double*>(p), __longlong_as_double(v));
  else atomicAdd(reinterpret_cast<unsigned long long*>(p), v);
}

__device__ __forceinline__ int rows_left(int64_t hi, int64_t base) {
  const int64_t left = hi - base;
  return static_cast<int>(left < 0 ? 0 : (left > 4 ? 4 : left));
}

// Shared scans (SCAN_AGG with a BFast view, at most kBRegG groups). The
// register-path queries over one table form groups whose columns, together at
// most kSRowBytes bytes per row, are staged tile by tile (512 or 1024 rows)
// into the workgroup's LDS with direct global-to-LDS loads; every query of the
// group then reads the tile from LDS, so the group reads the table from HBM
// once, and the queries' per-row work runs at LDS latency. The work of a group
// is split into at most kSWaves units, one per wave: a unit is (query, pair of
// its measures, half of its groups when it has more than four, slice of the
// tile's rows). Lane l handles row groups of 4
// consecutive rows (LDS vector reads), evaluates the predicates and the group
// code, and adds the unit's measures into lane registers with select-adds.
// The partials stay in registers across all tiles of the group the workgroup
// takes, and are reduced (butterfly) and added to accumulator copy `slot` once
// per stay.
constexpr int kSThreads = 512;        // threads per shared-scan workgroup (two per CU)
constexpr int kSWaves = kSThreads / kWave;
constexpr int kSBufBytes = 32768;    // LDS of a staged tile (two per workgroup)
constexpr int kSMaxTile = 1024;       // rows per tile when a row takes <= 32 bytes (else 512)
constexpr int kSRowBytes = kSBufBytes / (kSMaxTile / 2);
constexpr int kSMaxCols = 16;
constexpr int kSMaxQ = kSWaves;       // queries per group
constexpr int kSUnitM = 4;            // measures per unit
constexpr int kSUnitG = 4;            // groups per unit
constexpr int kSMaxRG = kSMaxTile / (kWave * 4);   // row groups per lane in a whole-tile unit

// A group of register-path queries over one table, and its units.
struct SGroup {
  int64_t n;                          // table rows
  int32_t nq, ncol, rows, nunits;     // rows: tile rows
  int32_t tile_bytes;                 // staged bytes of a tile (columns' tiles back to back, 16-byte multiple)
  int32_t q[kSMaxQ];
  const void* col[kSMaxCols];
  int32_t width[kSMaxCols];           // bytes per value (4 or 8)
  int32_t off[kSMaxCols];             // LDS byte offset of the column's tile
  // Unit w: query q[uq[w]], measures um[w] and um[w] + 1 (those below nm),
  // groups ug[w] .. ug[w] + 3 (those below G), tile rows [ur[w], ur[w] + 256
  // urg[w]); ucnt[w]: it also counts rows and first rows of its groups (one
  // unit per query, group and row does).
  int8_t uq[kSWaves], um[kSWaves], ug[kSWaves], urg[kSWaves], ucnt[kSWaves];
  int16_t ur[kSWaves];
};

// s_waitcnt immediate (gfx9 encoding: vmcnt in bits 3:0 and 15:14, expcnt in
// 6:4, lgkmcnt in 11:8) that waits for LDS operations only: lgkmcnt(0), with
// vmcnt and expcnt at their maximum.
constexpr int kWaitLgkm0 = 0xC07F;

// LDS pointers (address space 3), so that tile reads compile to ds_read.
typedef __attribute__((address_space(3))) char lds_char;
typedef int v4i __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) const v4i lds_v4i;
typedef __attribute__((address_space(3))) const v2d lds_v2d;

__device__ __forceinline__ void lds4_i32(const lds_char* col, int row, int32_t (&o)[4]) {
  const v4i x = *reinterpret_cast<lds_v4i*>(col + row * 4);
  o[0] = x[0];
  o[1] = x[1];
  o[2] = x[2];
  o[3] = x[3];
}

// K = 1: REAL (double bits), K = 2: 4-byte integer (sign-extended).
template <int K>
__device__ __forceinline__ void lds4_measure(const lds_char* col, int row, uint64_t (&o)[4]) {
  if (K == 1) {
    const v2d a = *reinterpret_cast<lds_v2d*>(col + row * 8);
    const v2d b = *reinterpret_cast<lds_v2d*>(col + row * 8 + 16);
    o[0] = static_cast<uint64_t>(__double_as_longlong(a[0]));
    o[1] = static_cast<uint64_t>(__double_as_longlong(a[1]));
    o[2] = static_cast<uint64_t>(__double_as_longlong(b[0]));
    o[3] = static_cast<uint64_t>(__double_as_longlong(b[1]));
  } else {
    int32_t t[4];
    lds4_i32(col, row, t);
#pragma unroll
    for (int e = 0; e < 4; e++) o[e] = static_cast<uint64_t>(static_cast<int64_t>(t[e]));
  }
}

// Partials of one unit, kept in lane registers while the workgroup stays on
// the unit's group: per measure and group the sum (double bits or wrapping
// int64), per group the row count and first row.
struct UnitAcc {
  uint64_t s[kSUnitM][kSUnitG];
  uint32_t cnt[kSUnitG];
  uint32_t first[kSUnitG];  // table rows are below 2^31
};

__device__ __forceinline__ void unit_reset(UnitAcc& A) {
#pragma unroll
  for (int g = 0; g < kSUnitG; g++) {
#pragma unroll
    for (int m = 0; m < kSUnitM; m++) A.s[m][g] = 0;
    A.cnt[g] = 0;
    A.first[g] = ~0u;
  }
}

// What a wave reads of the descriptors, loaded once per stay into
// wave-uniform registers (the descriptors live in global memory that the
// kernel also writes, so the compiler would otherwise re-read them with
// vector loads inside the row loops). CodeDesc: the query whose row codes the
// wave computes (phase A); UnitDesc: the wave's unit (phase B).
struct CodeDesc {
  int q, npred, nk, card1, G, rg0, rgs;   // local query; row groups rg0, rg0 + rgs, ...
  int op[kBFastPreds];
  int64_t lit[kBFastPreds];
  int poff[kBFastPreds];          // LDS byte offsets of the columns in the tile
  int koff[2];
};

struct UnitDesc {
  int q, G, gb, gc, nm, nmu, m0, counts, r0, rg;
  uint32_t real;                  // bit i: measure m0 + i is REAL
  int moff[kSUnitM];
};

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int64_t uni64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(static_cast<uint64_t>(v) >> 32));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Phase A of wave w: local query w mod nq, row groups w / nq + k * (waves per
// query) of the tile.
__device__ __forceinline__ void code_desc(const BQuery* qs, const SGroup& S, int w, CodeDesc& c) {
  const int nq = uni(S.nq);
  c.q = w % nq;
  c.rg0 = w / nq;
  c.rgs = kSWaves / nq;
  const BQuery& d = qs[uni(S.q[c.q])];
  const BFast& F = d.F;
  c.npred = uni(F.npred);
  c.nk = uni(F.nk);
  c.card1 = uni(F.card1);
  c.G = uni(d.G);
#pragma unroll
  for (int p = 0; p < kBFastPreds; p++) {
    const bool on = p < c.npred;
    c.op[p] = on ? uni(F.op[p]) : 0;
    c.lit[p] = on ? uni64(F.lit[p]) : 0;
    c.poff[p] = on ? uni(S.off[F.pcx[p]]) : 0;
  }
#pragma unroll
  for (int k = 0; k < 2; k++) c.koff[k] = k < c.nk ? uni(S.off[F.kcx[k]]) : 0;
}

// Unit w of group S, whose query is d.
__device__ __forceinline__ void unit_desc(const BQuery& d, const SGroup& S, int w, UnitDesc& u) {
  const BFast& F = d.F;
  u.q = uni(S.uq[w]);
  u.G = uni(d.G);
  u.gb = uni(S.ug[w]);
  u.gc = min(u.G - u.gb, kSUnitG);
  u.nm = uni(d.nm);
  u.m0 = uni(S.um[w]);
  u.counts = uni(S.ucnt[w]);
  u.r0 = uni(S.ur[w]);
  u.rg = uni(S.urg[w]);
  uint32_t real = 0;
  int nmu = 0;
#pragma unroll
  for (int i = 0; i < kSUnitM; i++) {
    const int m = u.m0 + i;
    const bool on = m < u.nm;
    u.moff[i] = on ? uni(S.off[F.mcx[m]]) : 0;
    real |= (on && F.real[m]) ? 1u << i : 0u;
    nmu += on ? 1 : 0;
  }
  u.nmu = uni(nmu);
  u.real = uni(static_cast<int>(real));
}

// Phase A: the group code of each of the tile's rows for query c.q (-1: the
// row fails or is past the vr valid rows), one byte per row in codes[c.q].
// Returns whether a qualifying row had a code outside [0, G).
__device__ __forceinline__ bool tile_codes(const CodeDesc& c, const lds_char* sh, int rows, int vr,
                                           __attribute__((address_space(3))) uint32_t* codes) {
  const int lane = threadIdx.x & (kWave - 1);
  bool bad = false;
  for (int j = c.rg0; j < rows / (kWave * 4); j += c.rgs) {
    const int row = 4 * (j * kWave + lane);
    bool ok[4];
#pragma unroll
    for (int e = 0; e < 4; e++) ok[e] = row + e < vr;
#pragma unroll
    for (int p = 0; p < kBFastPreds; p++) {
      if (p >= c.npred) break;
      int32_t x[4];
      lds4_i32(sh + c.poff[p], row, x);
#pragma unroll
      for (int e = 0; e < 4; e++) ok[e] = ok[e] && cmp_i64(c.op[p], x[e], c.lit[p]);
    }
    int32_t g[4] = {0, 0, 0, 0};
    if (c.nk >= 1) {
      int32_t a[4];
      lds4_i32(sh + c.koff[0], row, a);
      if (c.nk == 2) {
        int32_t b[4];
        lds4_i32(sh + c.koff[1], row, b);
#pragma unroll
        for (int e = 0; e < 4; e++) g[e] = a[e] * c.card1 + b[e];
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++) g[e] = a[e];
      }
    }
    uint32_t word = 0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      if (ok[e] && static_cast<unsigned>(g[e]) >= static_cast<unsigned>(c.G)) {
        bad = true;
        ok[e] = false;
      }
      word |= (ok[e] ? static_cast<uint32_t>(g[e]) : 0xFFu) << (8 * e);
    }
    codes[c.q * (kSMaxTile / 4) + row / 4] = word;
  }
  return bad;
}

// Phase B: adds the unit's rows of the staged tile (first table row lo) to A,
// for GC groups gb .. gb + GC - 1 (a compile-time count, so the select-adds
// carry no per-group branches), taking each row's group from the phase A
// codes. Rows that fail, or belong to other groups, get a local group outside
// [0, GC), so a select-add tests the local group alone; INT measures add
// x * (g == gg) with one 32x32+64 multiply-add, REAL measures a select and an
// add (a multiply by 0.0 would turn an infinite value into NaN in the other
// groups). Adding +0.0 or 0 leaves a sum unchanged: sums start at +0.0 / 0 and
// never become -0.0. Every accumulator index is static: a dynamically indexed
// update would move A to scratch.
template <int GC>
__device__ __forceinline__ void unit_scan(const UnitDesc& u, const lds_char* sh, int64_t lo,
                                          const __attribute__((address_space(3))) uint32_t* codes, UnitAcc& A) {
  const int nmu = u.nmu;
  const int lane = threadIdx.x & (kWave - 1);
  // Row groups one at a time (no unrolling: copies of the measure loop per
  // row group cost registers and spill).
#pragma unroll 1
  for (int j = 0; j < u.rg; j++) {
    const int row = u.r0 + 4 * (j * kWave + lane);
    const uint32_t word = codes[u.q * (kSMaxTile / 4) + row / 4];
    if (word == 0xFFFFFFFFu) continue;  // no qualifying row
    int32_t g[4];
#pragma unroll
    for (int e = 0; e < 4; e++) g[e] = static_cast<int32_t>(static_cast<int8_t>(word >> (8 * e))) - u.gb;
    if (u.counts) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const uint32_t r = static_cast<uint32_t>(lo + row + e);
#pragma unroll
        for (int gg = 0; gg < GC; gg++) {
          const bool hit = gg == g[e];
          A.cnt[gg] += hit ? 1u : 0u;
          A.first[gg] = min(A.first[gg], hit ? r : ~0u);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kSUnitM; i++) {
      if (i >= nmu) break;
      const lds_char* col = sh + u.moff[i];
      if ((u.real >> i) & 1u) {
        uint64_t v[4];
        lds4_measure<1>(col, row, v);
#pragma unroll
        for (int e = 0; e < 4; e++)
#pragma unroll
          for (int gg = 0; gg < GC; gg++) A.s[i][gg] = fadd_bits(A.s[i][gg], gg == g[e] ? v[e] : 0);
      } else {
        int32_t x[4];
        lds4_i32(col, row, x);
#pragma unroll
        for (int e = 0; e < 4; e++)
#pragma unroll
          for (int gg = 0; gg < GC; gg++)
            A.s[i][gg] += static_cast<uint64_t>(static_cast<int64_t>(x[e]) * static_cast<int64_t>(gg == g[e]));
      }
    }
  }
}

// Reduces a unit's partials across the wave and adds them to query d's
// accumulator copy `slot`, lane l for group gb + l.
__device__ __forceinline__ void unit_flush(const BQuery& d, const UnitDesc& u, const UnitAcc& A, int slot) {
  const int G = u.G, nm = u.nm, gc = u.gc;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t base = static_cast<int64_t>(slot) * G + u.gb + lane;   // (slot, group) of lane < gc
  if (u.counts) {
    uint64_t myc = 0, myf = ~0ull;
#pragma unroll
    for (int gg = 0; gg < kSUnitG; gg++) {
      if (gg >= gc) break;
      uint32_t c = A.cnt[gg];
      uint32_t f = A.first[gg];
      for (int s = kWave / 2; s >= 1; s >>= 1) {
        c += __shfl_xor(c, s);
        f = min(f, static_cast<uint32_t>(__shfl_xor(static_cast<int>(f), s)));
      }
      if (lane == gg) {
        myc = c;
        myf = f == ~0u ? ~0ull : f;
      }
    }
    if (lane < gc && myc) {
      atomicAdd(reinterpret_cast<unsigned long long*>(&d.rows[base]), myc);
      atomicMin(reinterpret_cast<unsigned long long*>(&d.first[base]), myf);
      for (int m = 0; m < nm; m++) atomicAdd(reinterpret_cast<unsigned long long*>(&d.nn[base * nm + m]), myc);
    }
  }
#pragma unroll
  for (int i = 0; i < kSUnitM; i++) {
    if (i >= u.nmu) break;
    const bool real = (u.real >> i) & 1u;
    uint64_t mine = 0;
#pragma unroll
    for (int gg = 0; gg < kSUnitG; gg++) {
      if (gg >= gc) break;
      const uint64_t a = real ? wave_sum<1>(A.s[i][gg]) : wave_sum<2>(A.s[i][gg]);
      if (lane == gg) mine = a;
    }
    if (lane < gc && mine) {
      if (real) atomic_acc<1>(&d.sum[base * nm + u.m0 + i], mine);
      else atomic_acc<2>(&d.sum[base * nm + u.m0 + i], mine);
    }
  }
}

// LDS path, one wave per item: a wave-private LDS table of G*nm sums and
// non-NULL counts, G row counts and G first keys (at most kBWaveWords words),
// updated with LDS atomics. Each lane keeps the partial of the last group it
// emitted in registers and flushes it to the table when its group changes, so
// runs of one group (an ungrouped join, a clustered key) cost one LDS update
// per run. The wave's LDS operations complete in order, so no barrier is
// needed between the updates and the final read of the table.
constexpr int kBWaveWords = kBLdsWords / kBWaves;

__device__ __forceinline__ void item_lds(const BQuery& d, int64_t lo, int64_t hi, int slot, uint64_t* lds) {
  const int G = d.G, nm = d.nm;
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t* lsum = lds;
  uint64_t* lnn = lsum + G * nm;
  uint64_t* lrows = lnn + G * nm;
  uint64_t* lfirst = lrows + G;
  const int words = G * nm * 2 + G * 2;
  for (int i = lane; i < words; i += kWave) lds[i] = i < G * nm * 2 + G ? 0ull : ~0ull;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  int cg = -1;
  uint32_t crows = 0;
  uint64_t cfirst = ~0ull;
  uint64_t csum[kMaxMeasures];
  uint32_t cnn[kMaxMeasures];
  auto flush = [&]() {
    if (cg < 0) return;
    atomicAdd(reinterpret_cast<unsigned long long*>(&lrows[cg]), static_cast<unsigned long long>(crows));
    atomicMin(reinterpret_cast<unsigned long long*>(&lfirst[cg]), cfirst);
#pragma unroll
    for (int m = 0; m < kMaxMeasures; m++) {
      if (m >= nm || cnn[m] == 0) continue;
      atomicAdd(reinterpret_cast<unsigned long long*>(&lnn[cg * nm + m]), static_cast<unsigned long long>(cnn[m]));
      if (d.M.is_real[m]) atomicAdd(reinterpret_cast<double*>(&lsum[cg * nm + m]), __longlong_as_double(csum[m]));
      else atomicAdd(reinterpret_cast<unsigned long long*>(&lsum[cg * nm + m]), csum[m]);
    }
  };
  visit_rows(d, lo, hi, [&](int g, uint64_t f, int64_t r) {
    if (g != cg) {
      flush();
      cg = g;
      crows = 0;
      cfirst = ~0ull;
#pragma unroll
      for (int m = 0; m < kMaxMeasures; m++) {
        csum[m] = 0;
        cnn[m] = 0;
      }
    }
    crows++;
    cfirst = f < cfirst ? f : cfirst;
#pragma unroll
    for (int m = 0; m < kMaxMeasures; m++) {
      if (m >= nm) continue;
      bool null;
      const uint64_t v = load_measure(d.M, m, r, null);
      if (null) continue;
      cnn[m]++;
      csum[m] = acc_add(d.M.is_real[m], csum[m], v);
    }
  });
  flush();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int i = lane; i < G; i += kWave) {
    if (lrows[i] == 0) continue;
    atomicAdd(reinterpret_cast<unsigned long long*>(&d.rows[static_cast<int64_t>(slot) * G + i]), lrows[i]);
    atomicMin(reinterpret_cast<unsigned long long*>(&d.first[static_cast<int64_t>(slot) * G + i]), lfirst[i]);
  }
  for (int i = lane; i < G * nm; i += kWave) {
    if (lnn[i] == 0) continue;
    const int64_t o = static_cast<int64_t>(slot) * G * nm + i;
    atomicAdd(reinterpret_cast<unsigned long long*>(&d.nn[o]), lnn[i]);
    if (d.M.is_real[i % nm]) atomicAdd(reinterpret_cast<double*>(&d.sum[o]), __longlong_as_double(lsum[i]));
    else atomicAdd(reinterpret_cast<unsigned long long*>(&d.sum[o]), lsum[i]);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Join items (LDS path, one wave per item), four probe rows per lane at a
// time: rows lo + lane + 64 (j + e), e < 4. Predicates, keys and measures of
// the four rows are loaded together, then their first buckets; a bucket's
// matches are consumed directly, and only a probe whose first bucket is full
// (no free slot, so the chain continues) walks the following buckets.
// Aggregation into the wave's LDS table as in item_lds; at most NM measures.
template <int NM>
__device__ __forceinline__ void item_join(const BQuery& d, int64_t lo, int64_t hi, int slot, uint64_t* lds) {
  constexpr int R = 4;
  const int G = d.G, nm = d.nm;
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t* lsum = lds;
  uint64_t* lnn = lsum + G * nm;
  uint64_t* lrows = lnn + G * nm;
  uint64_t* lfirst = lrows + G;
  const int words = G * nm * 2 + G * 2;
  for (int i = lane; i < words; i += kWave) lds[i] = i < G * nm * 2 + G ? 0ull : ~0ull;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const TableView& t = d.t;
  const PredSet& P = d.P;
  for (int64_t r0 = lo + lane; r0 < hi; r0 += kWave * R) {
    int64_t r[R];
    bool ok[R];
#pragma unroll
    for (int e = 0; e < R; e++) {
      r[e] = r0 + static_cast<int64_t>(e) * kWave;
      ok[e] = r[e] < hi;
    }
    // Predicates, four rows per column.
    for (int i = 0; i < P.n; i++) {
#pragma unroll
      for (int e = 0; e < R; e++) {
        if (!ok[e]) continue;
        bool v;
        if (P.type[i] == KG_REAL) v = cmp_f64(P.op[i], static_cast<const double*>(P.data[i])[r[e]], P.flit[i]);
        else v = cmp_i64(P.op[i], load_int(P.type[i], P.data[i], r[e]), P.ilit[i]);
        if (P.nulls[i] && P.nulls[i][r[e]]) v = false;
        ok[e] = v;
      }
    }
    int64_t k[R];
    uint64_t v[R][NM];
    bool vnull[R][NM];
#pragma unroll
    for (int e = 0; e < R; e++) {
      k[e] = 0;
      if (ok[e]) ok[e] = load_key(d.pkey, r[e], k[e]) && k[e] != kEmpty;
#pragma unroll
      for (int m = 0; m < NM; m++) {
        v[e][m] = 0;
        vnull[e][m] = true;
        if (m < nm && ok[e]) v[e][m] = load_measure(d.M, m, r[e], vnull[e][m]);
      }
    }
    // First buckets of the four probes.
    uint64_t b[R];
    longlong2 kv[R][kSlots / 2];
#pragma unroll
    for (int e = 0; e < R; e++) {
      b[e] = ok[e] ? bucket_of(k[e], t.nbuckets) : 0;
#pragma unroll
      for (int q = 0; q < kSlots / 2; q++)
        kv[e][q] = ok[e] ? reinterpret_cast<const longlong2*>(t.keys + b[e] * kSlots)[q]
                         : make_longlong2(kEmpty, kEmpty);
    }
#pragma unroll
    for (int e = 0; e < R; e++) {
      if (!ok[e]) continue;
      uint32_t hit = 0;
      bool empty = false;
#pragma unroll
      for (int q = 0; q < kSlots / 2; q++) {
        hit |= (kv[e][q].x == k[e] ? 1u : 0u) << (2 * q);
        hit |= (kv[e][q].y == k[e] ? 1u : 0u) << (2 * q + 1);
        empty |= kv[e][q].x == kEmpty || kv[e][q].y == kEmpty;
      }
      auto consume = [&](int32_t brow) {
        const int g = d.codes ? d.codes[brow] : 0;
        if (static_cast<unsigned>(g) >= static_cast<unsigned>(G)) {
          atomicOr(d.flag, kFlagBadCode);
          return;
        }
        atomicAdd(reinterpret_cast<unsigned long long*>(&lrows[g]), 1ull);
        atomicMin(reinterpret_cast<unsigned long long*>(&lfirst[g]),
                  (static_cast<unsigned long long>(r[e]) << 32) | static_cast<uint32_t>(brow));
#pragma unroll
        for (int m = 0; m < NM; m++) {
          if (m >= nm || vnull[e][m]) continue;
          atomicAdd(reinterpret_cast<unsigned long long*>(&lnn[g * nm + m]), 1ull);
          if (d.M.is_real[m]) atomicAdd(reinterpret_cast<double*>(&lsum[g * nm + m]), __longlong_as_double(v[e][m]));
          else atomicAdd(reinterpret_cast<unsigned long long*>(&lsum[g * nm + m]), v[e][m]);
        }
      };
      while (hit) {
        const int sidx = __builtin_ctz(hit);
        hit &= hit - 1;
        consume(t.rows[b[e] * kSlots + sidx]);
      }
      if (!empty) {
        // The chain continues past a full first bucket.
        uint64_t bb = (b[e] + 1 == t.nbuckets) ? 0 : b[e] + 1;
        for (uint64_t scanned = 1; scanned < t.nbuckets; scanned++) {
          const longlong2* bk = reinterpret_cast<const longlong2*>(t.keys + bb * kSlots);
          bool emp = false;
#pragma unroll
          for (int q = 0; q < kSlots / 2; q++) {
            const longlong2 x = bk[q];
            if (x.x == k[e]) consume(t.rows[bb * kSlots + 2 * q]);
            if (x.y == k[e]) consume(t.rows[bb * kSlots + 2 * q + 1]);
            emp |= x.x == kEmpty || x.y == kEmpty;
          }
          if (emp) break;
          bb = (bb + 1 == t.nbuckets) ? 0 : bb + 1;
        }
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int i = lane; i < G; i += kWave) {
    if (lrows[i] == 0) continue;
    atomicAdd(reinterpret_cast<unsigned long long*>(&d.rows[static_cast<int64_t>(slot) * G + i]), lrows[i]);
    atomicMin(reinterpret_cast<unsigned long long*>(&d.first[static_cast<int64_t>(slot) * G + i]), lfirst[i]);
  }
  for (int i = lane; i < G * nm; i += kWave) {
    if (lnn[i] == 0) continue;
    const int64_t o = static_cast<int64_t>(slot) * G * nm + i;
    atomicAdd(reinterpret_cast<unsigned long long*>(&d.nn[o]), lnn[i]);
    if (d.M.is_real[i % nm]) atomicAdd(reinterpret_cast<double*>(&d.sum[o]), __longlong_as_double(lsum[i]));
    else atomicAdd(reinterpret_cast<unsigned long long*>(&d.sum[o]), lsum[i]);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Typed join items (one wave per item): lane l owns the row groups
// lo + 4 l + 256 k (k < ngroups, the item's rows / 256) of 4 consecutive rows.
// Per row group, the
// predicate, key and measure columns are read with 16-byte vector loads, then
// the first buckets of the passing rows are loaded together, then their
// matches' row ids and group codes. Aggregation into the wave's LDS table: row
// count, sums, and the first output row of each group (a lane issues that
// atomicMin once per group and item, since its rows ascend); non-NULL counts
// equal row counts for these non-NULL measures.
template <bool FULL>
__device__ __forceinline__ void ld4_i64(const int64_t* col, int64_t base, int nv, int64_t (&o)[4]) {
  if (FULL) {
    const longlong2 a = *reinterpret_cast<const longlong2*>(col + base);
    const longlong2 b = *reinterpret_cast<const longlong2*>(col + base + 2);
    o[0] = a.x;
    o[1] = a.y;
    o[2] = b.x;
    o[3] = b.y;
  } else {
#pragma unroll
    for (int e = 0; e < 4; e++) o[e] = e < nv ? col[base + e] : 0;
  }
}

// Measure value of 4 rows as double bits (REAL or product) or int64.
template <bool FULL>
__device__ __forceinline__ void ld4_jmeasure(int kind, int fkind, const void* col, const void* fcol, int64_t base,
                                             int nv, uint64_t (&o)[4]) {
  if (kind == 1) ld4_measure<FULL, 1>(col, base, nv, o);
  else ld4_measure<FULL, 2>(col, base, nv, o);
  if (fkind == 0) return;
  uint64_t f[4];
  if (fkind == 1) ld4_measure<FULL, 1>(fcol, base, nv, f);
  else ld4_measure<FULL, 2>(fcol, base, nv, f);
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const double x = kind == 1 ? __longlong_as_double(o[e]) : static_cast<double>(static_cast<int64_t>(o[e]));
    const double y = fkind == 1 ? __longlong_as_double(f[e]) : static_cast<double>(static_cast<int64_t>(f[e]));
    o[e] = static_cast<uint64_t>(__double_as_longlong(x * y));
  }
}

template <bool FULL, bool ONEG>
__device__ __forceinline__ void item_join_fast(const BQuery& d, int64_t lo, int64_t hi, int slot, uint64_t* lds,
                                               int ngroups) {
  const BJoinFast& J = d.J;
  const int G = d.G, nm = d.nm;
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t* lsum = lds;
  uint64_t* lrows = lsum + G * nm;
  uint64_t* lfirst = lrows + G;
  const int words = G * nm + G * 2;
  if (!ONEG) {
    for (int i = lane; i < words; i += kWave) lds[i] = i < G * nm + G ? 0ull : ~0ull;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  // One group (an ungrouped join): the lane's row count, sums and first output
  // row stay in registers until the item ends (every match would otherwise
  // add to the same LDS words).
  uint64_t rc = 0, rs0 = 0, rs1 = 0, rfirst = ~0ull;
  const TableView& t = d.t;
  int64_t mn = 0, span = 0;
  const bool use_map = map_usable(d, &mn, &span);
  // Product measures are REAL; others keep their kind.
  const bool real0 = J.mkind[0] == 1 || J.fkind[0] != 0;
  const bool real1 = J.mkind[1] == 1 || J.fkind[1] != 0;
  uint64_t seen = 0;  // groups (< 64) whose first row this lane has reported
  // A row group's streamed values: 4 consecutive rows per lane.
  struct JRows {
    int64_t base;
    bool pass[4];
    int64_t key[4];
    uint64_t v0[4], v1[4];
  };
  // False when no lane of the wave has rows left (tail items).
  auto load_rows = [&](int k, JRows& R) -> bool {
    R.base = lo + 4 * lane + static_cast<int64_t>(k) * kWave * 4;
    const int64_t base = R.base;
    const int nv = FULL ? 4 : rows_left(hi, base);
    if (!FULL && __builtin_amdgcn_readfirstlane(__ballot(nv > 0) != 0 ? 1 : 0) == 0) return false;
#pragma unroll
    for (int e = 0; e < 4; e++) R.pass[e] = e < nv;
#pragma unroll
    for (int p = 0; p < kBFastPreds; p++) {
      if (p >= J.npred) break;
      int32_t x[4];
      ld4_i32<FULL>(J.pcol[p], base, nv, x);
#pragma unroll
      for (int e = 0; e < 4; e++) R.pass[e] = R.pass[e] && cmp_i64(J.op[p], x[e], J.lit[p]);
    }
    if (J.key64) {
      ld4_i64<FULL>(static_cast<const int64_t*>(J.key), base, nv, R.key);
    } else {
      int32_t k32[4];
      ld4_i32<FULL>(static_cast<const int32_t*>(J.key), base, nv, k32);
#pragma unroll
      for (int e = 0; e < 4; e++) R.key[e] = k32[e];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) R.v0[e] = R.v1[e] = 0;
    if (nm > 0) ld4_jmeasure<FULL>(J.mkind[0], J.fkind[0], J.mcol[0], J.fcol[0], base, nv, R.v0);
    if (nm > 1) ld4_jmeasure<FULL>(J.mkind[1], J.fkind[1], J.mcol[1], J.fcol[1], base, nv, R.v1);
#pragma unroll
    for (int e = 0; e < 4; e++) R.pass[e] = R.pass[e] && R.key[e] != kEmpty;
    return true;
  };
  // Adds probe row `row` matched with build row brow of group g.
  auto consume_g = [&](int64_t row, int32_t brow, int g, uint64_t a0, uint64_t a1) {
    if (ONEG) {
      rc++;
      const uint64_t fr = (static_cast<uint64_t>(row) << 32) | static_cast<uint32_t>(brow);
      rfirst = fr < rfirst ? fr : rfirst;
      if (nm > 0) rs0 = real0 ? fadd_bits(rs0, a0) : rs0 + a0;
      if (nm > 1) rs1 = real1 ? fadd_bits(rs1, a1) : rs1 + a1;
      return;
    }
    atomicAdd(reinterpret_cast<unsigned long long*>(&lrows[g]), 1ull);
    if (g >= 64 || !((seen >> g) & 1ull)) {
      atomicMin(reinterpret_cast<unsigned long long*>(&lfirst[g]),
                (static_cast<unsigned long long>(row) << 32) | static_cast<uint32_t>(brow));
      if (g < 64) seen |= 1ull << g;
    }
    if (nm > 0) {
      if (real0) atomicAdd(reinterpret_cast<double*>(&lsum[g * nm]), __longlong_as_double(a0));
      else atomicAdd(reinterpret_cast<unsigned long long*>(&lsum[g * nm]), a0);
    }
    if (nm > 1) {
      if (real1) atomicAdd(reinterpret_cast<double*>(&lsum[g * nm + 1]), __longlong_as_double(a1));
      else atomicAdd(reinterpret_cast<unsigned long long*>(&lsum[g * nm + 1]), a1);
    }
  };
  if (use_map) {
    // Direct map: one 4-byte entry per probe, the four loaded together. The
    // next row group's streamed values are loaded before this one's map
    // entries, so the streamed loads overlap the lookups (the table path below
    // holds too many registers to do the same).
    JRows cur, nxt;
    bool have = load_rows(0, cur);
#pragma unroll 1
    for (int k = 0; have; k++) {
      const bool have_next = k + 1 < ngroups && load_rows(k + 1, nxt);
      uint32_t me[4];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const uint64_t off = static_cast<uint64_t>(cur.key[e]) - static_cast<uint64_t>(mn);
        me[e] = (cur.pass[e] && off < static_cast<uint64_t>(span)) ? d.map[off] : kMapEmpty;
      }
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (me[e] != kMapEmpty && build_recheck(d, cur.key[e], static_cast<int32_t>(me[e] & 0xFFFFFFu)))
          consume_g(cur.base + e, static_cast<int32_t>(me[e] & 0xFFFFFFu), static_cast<int>(me[e] >> 24),
                    cur.v0[e], cur.v1[e]);
      cur = nxt;
      have = have_next;
    }
  } else if (!d.map_only) {  // map_only without a usable map: the query reports it
#pragma unroll 1
    for (int k = 0; k < ngroups; k++) {
      JRows cur;
      if (!load_rows(k, cur)) break;
      const int64_t base = cur.base;
      const bool* pass = cur.pass;
      const int64_t* key = cur.key;
      const uint64_t* v0 = cur.v0;
      const uint64_t* v1 = cur.v1;
      // First buckets of the passing rows, two rows at a time (a bucket is 16
      // registers; four in flight would raise the register budget of every
      // join item, map path included, and cut the waves per SIMD).
#pragma unroll
      for (int h2 = 0; h2 < 4; h2 += 2) {
        uint64_t b[2];
        longlong2 kv[2][kSlots / 2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
          b[e] = pass[h2 + e] ? bucket_of(key[h2 + e], t.nbuckets) : 0;
#pragma unroll
          for (int q = 0; q < kSlots / 2; q++)
            kv[e][q] = pass[h2 + e] ? reinterpret_cast<const longlong2*>(t.keys + b[e] * kSlots)[q]
                                    : make_longlong2(kEmpty, kEmpty);
        }
        // Matching slots of the first buckets (bit kSlots e + s: row h2 + e,
        // slot s) and the rows whose first bucket is full, so their chain
        // goes on.
        static_assert(2 * kSlots <= 32, "two rows' slot bits fit one word");
        uint32_t hits = 0, more = 0;
#pragma unroll
        for (int e = 0; e < 2; e++) {
          uint32_t h = 0;
          bool empty = false;
#pragma unroll
          for (int q = 0; q < kSlots / 2; q++) {
            h |= (kv[e][q].x == key[h2 + e] ? 1u : 0u) << (2 * q);
            h |= (kv[e][q].y == key[h2 + e] ? 1u : 0u) << (2 * q + 1);
            empty |= kv[e][q].x == kEmpty || kv[e][q].y == kEmpty;
          }
          hits |= (pass[h2 + e] ? h : 0u) << (kSlots * e);
          more |= (pass[h2 + e] && !empty ? 1u : 0u) << e;
        }
        // Rows in ascending order, each one's first-bucket matches before its
        // chain; the row's values are selected, not indexed, so they stay in
        // registers.
#pragma unroll 1
        for (int e = 0; e < 2; e++) {
          uint32_t h = (hits >> (kSlots * e)) & ((1u << kSlots) - 1u);
          const bool chain = (more >> e) & 1u;
          if (h == 0 && !chain) continue;
          const int64_t row = base + h2 + e;
          const int64_t ke = e == 0 ? key[h2] : key[h2 + 1];
          const uint64_t be = e == 0 ? b[0] : b[1];
          const uint64_t a0 = e == 0 ? v0[h2] : v0[h2 + 1];
          const uint64_t a1 = e == 0 ? v1[h2] : v1[h2 + 1];
          auto consume = [&](int32_t brow) {
            if (!build_recheck(d, ke, brow)) return;
            const int g = d.codes ? d.codes[brow] : 0;
            if (static_cast<unsigned>(g) >= static_cast<unsigned>(G)) {
              atomicOr(d.flag, kFlagBadCode);
              return;
            }
            consume_g(row, brow, g, a0, a1);
          };
          while (h) {
            const int sidx = __builtin_ctz(h);
            h &= h - 1;
            consume(t.rows[be * kSlots + sidx]);
          }
          if (chain) {
            uint64_t bb = (be + 1 == t.nbuckets) ? 0 : be + 1;
            for (uint64_t scanned = 1; scanned < t.nbuckets; scanned++) {
              const longlong2* bk = reinterpret_cast<const longlong2*>(t.keys + bb * kSlots);
              bool emp = false;
#pragma unroll
              for (int q = 0; q < kSlots / 2; q++) {
                const longlong2 x = bk[q];
                if (x.x == ke) consume(t.rows[bb * kSlots + 2 * q]);
                if (x.y == ke) consume(t.rows[bb * kSlots + 2 * q + 1]);
                emp |= x.x == kEmpty || x.y == kEmpty;
              }
              if (emp) break;
              bb = (bb + 1 == t.nbuckets) ? 0 : bb + 1;
            }
          }
        }
      }
    }
  }
  if (ONEG) {
    for (int st = kWave / 2; st >= 1; st >>= 1) {
      rc += shfl_xor_u64(rc, st);
      const uint64_t of = shfl_xor_u64(rfirst, st);
      rfirst = of < rfirst ? of : rfirst;
    }
    if (nm > 0) rs0 = real0 ? wave_sum<1>(rs0) : wave_sum<2>(rs0);
    if (nm > 1) rs1 = real1 ? wave_sum<1>(rs1) : wave_sum<2>(rs1);
    if (lane == 0 && rc) {
      atomicAdd(reinterpret_cast<unsigned long long*>(&d.rows[slot]), rc);
      atomicMin(reinterpret_cast<unsigned long long*>(&d.first[slot]), rfirst);
      for (int m = 0; m < nm; m++) {
        const int64_t o = static_cast<int64_t>(slot) * nm + m;
        const uint64_t v = m == 0 ? rs0 : rs1;
        atomicAdd(reinterpret_cast<unsigned long long*>(&d.nn[o]), rc);
        if (d.M.is_real[m]) atomicAdd(reinterpret_cast<double*>(&d.sum[o]), __longlong_as_double(v));
        else atomicAdd(reinterpret_cast<unsigned long long*>(&d.sum[o]), v);
      }
    }
    return;
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int i = lane; i < G; i += kWave) {
    const uint64_t c = lrows[i];
    if (c == 0) continue;
    atomicAdd(reinterpret_cast<unsigned long long*>(&d.rows[static_cast<int64_t>(slot) * G + i]), c);
    atomicMin(reinterpret_cast<unsigned long long*>(&d.first[static_cast<int64_t>(slot) * G + i]), lfirst[i]);
    for (int m = 0; m < nm; m++) {
      const int64_t o = (static_cast<int64_t>(slot) * G + i) * nm + m;
      atomicAdd(reinterpret_cast<unsigned long long*>(&d.nn[o]), c);
      if (d.M.is_real[m]) atomicAdd(reinterpret_cast<double*>(&d.sum[o]), __longlong_as_double(lsum[i * nm + m]));
      else atomicAdd(reinterpret_cast<unsigned long long*>(&d.sum[o]), lsum[i * nm + m]);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Folds query d's accumulator copies into its result area (pinned host
// memory) and publishes the completion word. Called by one wave, after an
// acquire, once every item of the query has added its partials. The copies are
// read with device-scope atomic loads (they were written by atomics from every
// XCD).
__device__ void fold_query(const BQuery& d) {
  const int lane = threadIdx.x & (kWave - 1);
  const int G = d.G, nm = d.nm;
  const ResLayout L = res_layout(G, nm, d.naggs);
  AggFinal* fin = reinterpret_cast<AggFinal*>(d.res + L.fin);
  uint64_t* rows = reinterpret_cast<uint64_t*>(d.res + L.rows);
  uint64_t* first = reinterpret_cast<uint64_t*>(d.res + L.first);
  int8_t* fnull = reinterpret_cast<int8_t*>(d.res + L.fnull);
  if (lane == 0)
    *reinterpret_cast<uint32_t*>(d.res) =
        __hip_atomic_load(d.flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
        (d.sflag ? __hip_atomic_load(d.sflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u);
  for (int i = lane; i < G * nm; i += kWave) {
    const bool real = d.M.is_real[i % nm];
    uint64_t s = 0, c = 0;
    for (int sl = 0; sl < kBSlots; sl++) {
      s = acc_add(real, s, coherent_load(&d.sum[static_cast<int64_t>(sl) * G * nm + i]));
      c += coherent_load(&d.nn[static_cast<int64_t>(sl) * G * nm + i]);
    }
    fin[i] = AggFinal{s, 0, 0, c};
  }
  for (int g = lane; g < G; g += kWave) {
    uint64_t c = 0;
    for (int sl = 0; sl < kBSlots; sl++) c += coherent_load(&d.rows[static_cast<int64_t>(sl) * G + g]);
    rows[g] = c;
    uint64_t f = ~0ull;
    for (int sl = 0; sl < kBSlots; sl++) {
      const uint64_t x = coherent_load(&d.first[static_cast<int64_t>(sl) * G + g]);
      f = x < f ? x : f;
    }
    first[g] = f;
    for (int a = 0; a < d.naggs; a++)
      fnull[static_cast<int64_t>(a) * G + g] =
          (d.cnull[a] && c > 0 && f != ~0ull) ? d.cnull[a][f >> d.first_shift] : 0;
  }
  // Results before the completion word, for the host polling it.
  __threadfence_system();
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) __hip_atomic_store(d.done, d.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Ends one scan/probe item of query d: the wave that finishes the query's last
// item folds it. The item's partials are device-scope atomics; waiting until
// the wave's outstanding memory operations are acknowledged orders them before
// the count, and the fold reads the copies with device-scope loads. (A device
// release fence would also write back and invalidate the XCD's L2 on every
// item, discarding the tiles the queries of a table share.)
__device__ __forceinline__ void item_done(const BQuery& d) {
  const int lane = threadIdx.x & (kWave - 1);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  unsigned long long old = 0;
  if (lane == 0) old = atomicAdd(d.rem, ~0ull);  // - 1
  old = __shfl(old, 0);
  if (old == 1) fold_query(d);
}

// Accumulator copy of the calling wave: copies are spread over the waves of
// an XCD (slot mod 8 = the XCD of workgroup b, which is b mod 8) so that
// concurrent flushes of one query's partials rarely hit the same address.
__device__ __forceinline__ int acc_slot() {
  const int wave = static_cast<int>(blockIdx.x / kBQueues) * kBWaves + static_cast<int>(threadIdx.x / kWave);
  return static_cast<int>(blockIdx.x % kBQueues) + kBQueues * (wave % (kBSlots / kBQueues));
}

// Scan/probe items of one queue set, one wave per item. The waves of XCD x
// (workgroups b with b mod 8 = x) take the items of queue x: wave j of the XCD
// (W waves on it) starts with item j, with no atomic, and takes each further
// item from the queue's counter, W + count. Items are thus taken in queue
// order, so the items of one tile (one per query of its table) run at about
// the same time and share the tile in L2; starting every wave through the
// counter would queue the atomics of the whole grid on its line at once.
// Every item is taken once. Needs a grid of at least kBQueues workgroups,
// which the host checks.
template <typename Run>
__device__ __forceinline__ void drain_wave(BQueues& Q, const BQuery* qs, Run&& run) {
  const int nt = Q.nt;
  const int lane = threadIdx.x & (kWave - 1);
  const int x = static_cast<int>(blockIdx.x % kBQueues);
  const int32_t total = Q.qpre[x][nt];
  const int32_t nwg = (static_cast<int32_t>(gridDim.x) - x + kBQueues - 1) / kBQueues;
  const int32_t W = nwg * kBWaves;
  int t = 0;
  int32_t it = static_cast<int32_t>(blockIdx.x / kBQueues) * kBWaves + uni(static_cast<int>(threadIdx.x / kWave));
  while (it < total) {
    while (Q.qpre[x][t + 1] <= it) t++;  // a wave's items ascend
    const int32_t local = it - Q.qpre[x][t];
    const int32_t m = Q.toff[t + 1] - Q.toff[t];
    const int32_t tile = x + kBQueues * (local / m);
    const BQuery& d = qs[Q.tq[Q.toff[t] + local % m]];
    const int64_t lo = static_cast<int64_t>(tile) * Q.tile;
    const int64_t hi = lo + Q.tile < d.n ? lo + Q.tile : d.n;
    run(d, lo, hi);
    item_done(d);
    if (total <= W) break;  // the first items were all: no counter traffic
    uint32_t v = 0;
    if (lane == 0) v = atomicAdd(&Q.next[x * kBHeadStride], 1u);
    it = W + static_cast<int32_t>(__builtin_amdgcn_readfirstlane(v));
  }
}

// Ends a workgroup's stay on group S after `ntiles` of its tiles: the units'
// partials go to the accumulator copies, and each query's remaining-tile count
// drops by ntiles once every wave's partials are acknowledged; the wave that
// takes a count to zero folds that query.
__device__ __forceinline__ void group_leave(const SGroup& S, const BQuery* qs, bool active, const UnitDesc& u,
                                            const UnitAcc& A, const CodeDesc& c, bool cbad, int slot,
                                            int32_t ntiles) {
  const int wave = uni(static_cast<int>(threadIdx.x / kWave));
  const int lane = threadIdx.x & (kWave - 1);
  if (active) unit_flush(qs[uni(S.q[u.q])], u, A, slot);
  if (__builtin_amdgcn_readfirstlane(__ballot(cbad) != 0 ? 1 : 0) && lane == 0)
    atomicOr(qs[uni(S.q[c.q])].flag, kFlagBadCode);
  // The partials' atomics are acknowledged (and not moved past the wait)
  // before the barrier that precedes the count.
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __syncthreads();
  if (wave < S.nq) {
    const BQuery& d = qs[uni(S.q[wave])];
    unsigned long long old = 0;
    if (lane == 0) old = atomicAdd(d.rem, static_cast<unsigned long long>(-static_cast<long long>(ntiles)));
    old = __shfl(old, 0);
    if (old == static_cast<unsigned long long>(ntiles)) fold_query(d);
  }
}

// Scan/probe phase, shared scans. The groups' tiles form one sequence
// (groups in order, smallest tables first) dealt round robin: workgroup b takes
// tiles b, b + B, b + 2B, ... (B = gridDim.x), so no work queue is needed and
// each workgroup knows its next tile; that tile is staged into the second LDS
// buffer while the current one is scanned. Per tile, phase A writes every
// query's row codes to LDS and phase B runs the units: wave w runs unit w of
// the group and keeps its partials in registers for as long as the workgroup
// takes tiles of that group; they are flushed when the workgroup's next tile
// belongs to another group, and at the end. Two workgroups of 8 waves per CU.
__global__ __launch_bounds__(kSThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_shared(
    BHeader* h, const BQuery* qs, const SGroup* gs) {
  // Tiles are staged into `stg` with direct global-to-LDS loads and copied to
  // `cur_tile` (LDS to LDS) before they are computed. The compiler can tell
  // that the computation's LDS reads (of cur_tile) do not alias the
  // global-to-LDS loads in flight (into stg), so it does not make them wait
  // for the next tile's loads: staging overlaps computation. (With one array
  // of two buffers indexed by parity, every LDS read waited for the whole next
  // tile.)
  __shared__ __attribute__((aligned(16))) char stg[kSBufBytes];
  __shared__ __attribute__((aligned(16))) char cur_tile[kSBufBytes];
  // Tile prefix, rows and tile rows of every group, read once: finding a
  // tile's group and rows then takes LDS reads only.
  __shared__ int32_t s_tpre[kBMaxQueries + 1];
  __shared__ int64_t s_n[kBMaxQueries];
  __shared__ int32_t s_rows[kBMaxQueries];
  // Phase A output: per query of the group, one group-code byte per tile row.
  __shared__ uint32_t s_codes[kSMaxQ * kSMaxTile / 4];
  const int wave = uni(static_cast<int>(threadIdx.x / kWave));
  const int lane = threadIdx.x & (kWave - 1);
  // Accumulator copy: slot mod 8 = the workgroup's XCD (b mod 8), spread over
  // the XCD's waves.
  const int slot = static_cast<int>(blockIdx.x % kBQueues) +
                   kBQueues * static_cast<int>((blockIdx.x / kBQueues * kSWaves + wave) % (kBSlots / kBQueues));
  const STiles& T = h->shared;
  const int ng = T.ng;
  for (int i = threadIdx.x; i <= ng; i += kSThreads) s_tpre[i] = T.tpre[i];
  for (int i = threadIdx.x; i < ng; i += kSThreads) {
    s_n[i] = gs[i].n;
    s_rows[i] = gs[i].rows;
  }
  __syncthreads();
  const int32_t total = s_tpre[ng];
  const int32_t step = static_cast<int32_t>(gridDim.x);
  UnitAcc A;
  unit_reset(A);
  UnitDesc u{};
  CodeDesc cd{};
  int cur = -1;
  bool active = false, cbad = false;
  int32_t ntiles = 0;
  // Tile t of the sequence: group g (searched upward from the previous one:
  // a workgroup's tiles ascend), first row lo, valid rows vr.
  auto decode = [&](int32_t t, int& g, int64_t& lo, int& vr) {
    while (g + 1 < ng && s_tpre[g + 1] <= t) g++;
    const int rows = s_rows[g];
    lo = static_cast<int64_t>(t - s_tpre[g]) * rows;
    const int64_t n = s_n[g];
    vr = static_cast<int>(n - lo < rows ? n - lo : rows);
  };
  // Staging: 16-byte chunks with direct global-to-LDS loads (one 1 KB wave
  // instruction covers 64 chunks; the tile's instructions are dealt round
  // robin over the waves, across columns). The last chunk of a tile's column
  // may run past the table's last row: the column starts 16-byte aligned
  // (register-path columns are), so that chunk starts at a 16-byte boundary
  // and lies in the same 4 KB page as the valid bytes it holds; its extra
  // bytes land in LDS past the valid rows, inside the column's tile, and are
  // never read. The staged group's column table lives in LDS (s_col, s_cw,
  // s_coff, s_ncol), rewritten when the staged group changes: reading it
  // from global memory per column would make each column wait for the loads
  // of the previous one, and holding it in vector registers across the loop
  // spills.
  __shared__ uint64_t s_col[kSMaxCols];
  __shared__ int32_t s_cw[kSMaxCols], s_coff[kSMaxCols];
  __shared__ int32_t s_ncol;
  int staged_g = -1;
  auto stage = [&](int g, int64_t lo, int vr, char* buf) {
    if (g != staged_g) {  // uniform: every wave stages the same tiles
      // No wave reads the table here: its last reads were in the previous
      // stage call, before this iteration's first barrier.
      const SGroup& S = gs[g];
      if (threadIdx.x < kSMaxCols) {
        s_col[threadIdx.x] = reinterpret_cast<uint64_t>(S.col[threadIdx.x]);
        s_cw[threadIdx.x] = S.width[threadIdx.x];
        s_coff[threadIdx.x] = S.off[threadIdx.x];
      }
      if (threadIdx.x == 0) s_ncol = S.ncol;
      __syncthreads();
      staged_g = g;
    }
    const int ncol = uni(s_ncol);
    const int rows = s_rows[g];
    int base = 0;
    for (int c = 0; c < ncol; c++) {
      const uint64_t pa = s_col[c];
      const uint32_t plo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(pa));
      const uint32_t phi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(pa >> 32));
      const int w = uni(s_cw[c]);
      const int off = uni(s_coff[c]);
      const char* src = reinterpret_cast<const char*>((static_cast<uint64_t>(phi) << 32) | plo) + lo * w;
      char* dst = buf + off;
      const int chunks = (vr * w + 15) >> 4;
      const int ninst = (rows * w) >> 10;
      for (int i = (wave - base % kSWaves + kSWaves) % kSWaves; i < ninst; i += kSWaves) {
        const int chunk = i * kWave + lane;
        if (chunk < chunks) __builtin_amdgcn_global_load_lds(src + chunk * 16, dst + i * 1024, 16, 0, 0);
      }
      base += ninst;
    }
  };
  int32_t t = static_cast<int32_t>(blockIdx.x);
  int g = 0, vr = 0;
  int64_t lo = 0;
  if (t < total) {
    decode(t, g, lo, vr);
    stage(g, lo, vr, stg);
  }
  const lds_char* buf = (const lds_char*)cur_tile;
  while (t < total) {
    // This tile's staging is complete; move it to the compute buffer, which no
    // wave still reads.
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    {
      const int nv = uni(gs[uni(g)].tile_bytes) >> 4;
      const v4i* src = reinterpret_cast<const v4i*>(stg);
      v4i* dst = reinterpret_cast<v4i*>(cur_tile);
      for (int i = threadIdx.x; i < nv; i += kSThreads) dst[i] = src[i];
    }
    __syncthreads();
    if (g != cur) {
      if (cur >= 0) {
        group_leave(gs[cur], qs, active, u, A, cd, cbad, slot, ntiles);
        unit_reset(A);
        cbad = false;
      }
      cur = g;
      ntiles = 0;
      const SGroup& S = gs[g];
      active = wave < uni(S.nunits);
      if (active) unit_desc(qs[uni(S.q[S.uq[wave]])], S, wave, u);
      code_desc(qs, S, wave, cd);
    }
    const int32_t tn = t + step;
    int gn = g, vrn = 0;
    int64_t lon = 0;
    if (tn < total) {
      decode(tn, gn, lon, vrn);
      stage(gn, lon, vrn, stg);
    }
    __attribute__((address_space(3))) uint32_t* codes = (__attribute__((address_space(3))) uint32_t*)(s_codes);
    cbad |= tile_codes(cd, buf, s_rows[g], vr, codes);
    // The codes are complete: an LDS-only barrier (a full __syncthreads would
    // also wait for the next tile's global-to-LDS loads, which must stay in
    // flight during the units' work).
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (active) {
      switch (u.gc) {
        case 1: unit_scan<1>(u, buf, lo, codes, A); break;
        case 2: unit_scan<2>(u, buf, lo, codes, A); break;
        case 3: unit_scan<3>(u, buf, lo, codes, A); break;
        default: unit_scan<4>(u, buf, lo, codes, A); break;
      }
    }
    ntiles++;
    t = tn;
    g = gn;
    lo = lon;
    vr = vrn;
  }
  if (cur >= 0) group_leave(gs[cur], qs, active, u, A, cd, cbad, slot, ntiles);
}

// Scan/probe phase, joins with a typed view (one wave per item). Held to 4
// waves per SIMD (128 VGPRs): the pipelined map loop wants about 157, and at
// 3 waves per SIMD large batches ran slower, while at 4 the few registers
// spilled cost less than the waves gained (served mix at the plain p99: 4.2x
// plain against 3.6x at 3 waves per SIMD).
__global__ __launch_bounds__(kBThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_join(
    BHeader* h, const BQuery* qs) {
  __shared__ uint64_t lds[kBLdsWords];
  const int slot = acc_slot();
  uint64_t* mine = lds + (threadIdx.x / kWave) * kBWaveWords;
  const int tile = h->join.tile;
  const int ngroups = tile / (kWave * 4);
  drain_wave(h->join, qs, [&](const BQuery& d, int64_t lo, int64_t hi) {
    if (d.G == 1) {
      if (hi - lo == tile) item_join_fast<true, true>(d, lo, hi, slot, mine, ngroups);
      else item_join_fast<false, true>(d, lo, hi, slot, mine, ngroups);
    } else {
      if (hi - lo == tile) item_join_fast<true, false>(d, lo, hi, slot, mine, ngroups);
      else item_join_fast<false, false>(d, lo, hi, slot, mine, ngroups);
    }
  });
}

// Scan/probe phase, the other queries (generic LDS path, one wave per item).
__global__ __launch_bounds__(kBThreads) void k_batch_gen(BHeader* h, const BQuery* qs) {
  __shared__ uint64_t lds[kBLdsWords];
  const int slot = acc_slot();
  uint64_t* mine = lds + (threadIdx.x / kWave) * kBWaveWords;
  drain_wave(h->gen, qs, [&](const BQuery& d, int64_t lo, int64_t hi) {
    if (d.kind == KG_Q_JOIN_AGG && d.nm <= 2) item_join<2>(d, lo, hi, slot, mine);
    else item_lds(d, lo, hi, slot, mine);
  });
}

// Queries without scan/probe items (no rows) are folded here, after the
// scan/probe kernels.
__global__ __launch_bounds__(kWave) void k_batch_final(const BHeader* h, const BQuery* qs) {
  if (static_cast<int>(blockIdx.x) >= h->nq) return;
  const BQuery& d = qs[blockIdx.x];
  if (d.G == 0 || !d.res || d.items != 0) return;  // not planned, or folded by its last item
  fold_query(d);
}

}  // namespace kg

using namespace kg;

namespace {

// Host state of one planned query.
struct BPlan {
  int32_t status = KG_OK;
  std::string err;
  int32_t G = 0, nkeys_out = 0, card1 = 0;
  MeasureSet M;
  std::vector<AggPlan> plan;
  size_t res_off = 0;
  ResLayout L{};
  bool join = false;
  bool map_only = false;     // its build source has the map only
  bool build_entry = true;   // it adds the build phase's items of its source
  int64_t bn = 0;
};

int check_join_key(kg_ctx* ctx, const kg_col& c, int64_t n, const char* what) {
  KG_REQUIRE(ctx, n == 0 || c.data, "%s column has no data", what);
  switch (c.type) {
    case KG_SMALLINT:
    case KG_INT:
    case KG_BIGINT:
    case KG_DATE:
    case KG_ENUM:
    case KG_REAL:
    case KG_BOOLEAN:
      return KG_OK;
    default:
      return kg_fail(ctx, KG_ERR_UNSUPPORTED, "%s of type %d is not supported", what, c.type);
  }
}

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

uint64_t table_buckets(int64_t build_rows, double lf) {
  const double slots = std::max<double>(static_cast<double>(build_rows) / lf, kSlots);
  return static_cast<uint64_t>(ceil_div(static_cast<int64_t>(slots), kSlots));
}

double query_lf(const kg_query& q) { return q.load_factor > 0 ? q.load_factor : 0.5; }

// Groups of a query as planned (0 when it cannot be planned cheaply).
int64_t query_groups(const kg_query& q) {
  if (q.kind == KG_Q_JOIN_AGG) return q.build_code_col >= 0 ? q.group_card : 1;
  int64_t G = 1;
  for (int k = 0; k < q.nkeys && q.key_card; k++) G *= std::max(q.key_card[k], 1);
  return G;
}

// Device bytes one query takes in a batch arena: result area, accumulator
// copies, and the join table's key and row slots.
// Direct-map entries of a join (0: no map): typed joins whose build side has
// at most kBMapMax rows and at most 254 groups; twice the rows, rounded up to
// a power of two, capped at kBMapMax.
int32_t map_entries(bool typed_join, int64_t bn, int64_t G) {
  if (!typed_join || bn <= 0 || bn > kBMapMax || G > 254) return 0;
  int32_t m = 2;
  while (m < 2 * bn && m < kBMapMax) m *= 2;
  return m;
}

// Map entries of a shared build source of bn rows (0: none): twice the rows,
// rounded up to a power of two, at most kBMapMax for build sides of at most
// kBMapMax rows (which also get a table) and kBBigMapMax beyond.
int32_t source_map_entries(int64_t bn, int64_t G) {
  if (bn <= 0 || G > 254 || bn > (int64_t{1} << 24)) return 0;  // map entries hold 24-bit build rows
  const int64_t cap = bn > kBMapMax ? kBBigMapMax : kBMapMax;
  int64_t m = 2;
  while (m < 2 * bn && m < cap) m *= 2;
  return m >= bn ? static_cast<int32_t>(m) : 0;
}

// Bytes of a query's own result area and accumulators in a batch.
size_t query_own_bytes(const kg_query& q) {
  const int64_t G = std::min<int64_t>(query_groups(q), kMaxLdsGroups);
  const int nm = std::min(std::max(q.naggs, 1), kMaxMeasures);
  size_t b = res_layout(static_cast<int>(G), nm, std::min(q.naggs, kBMaxAggs)).bytes;
  b += round256(static_cast<size_t>(kBSlots) * G * (2 * nm + 1) * 8 + 16);
  b += round256(static_cast<size_t>(kBSlots) * G * 8);
  return b;
}

// Bytes of a join's build structures: map-only (build sides beyond kBMapMax
// rows that a map can take), or table plus the map of a small build side.
size_t join_table_bytes(const kg_query& q) {
  if (q.kind != KG_Q_JOIN_AGG) return 0;
  const int64_t G = std::min<int64_t>(query_groups(q), kMaxLdsGroups);
  const int64_t bn = std::max<int64_t>(q.build_n, 0);
  const int32_t sm = source_map_entries(bn, G);
  if (bn > kBMapMax && sm) return round256(static_cast<size_t>(sm) * 4) + 256 + kRangeBytes;
  const uint64_t nb = table_buckets(bn, query_lf(q));
  size_t b = round256(nb * kSlots * 8) + round256(nb * kSlots * 4) + 256 + kRangeBytes;
  const int32_t m = map_entries(true, bn, G);
  if (m) b += round256(static_cast<size_t>(m) * 4) + kRangeBytes;
  return b;
}

size_t query_footprint(const kg_query& q) { return query_own_bytes(q) + join_table_bytes(q); }

}  // namespace

namespace kg {

size_t batch_own_bytes(const kg_query& q) { return query_own_bytes(q); }

size_t batch_source_bytes(const kg_query& q) { return join_table_bytes(q); }

bool batch_source_id(const kg_query& q, BatchSource* id) {
  if (q.kind != KG_Q_JOIN_AGG || !q.build_cols || q.build_key_col < 0 || q.build_key_col >= q.build_ncols) return false;
  const int64_t G = query_groups(q);
  if (!source_map_entries(q.build_n, G)) return false;
  const kg_col& k = q.build_cols[q.build_key_col];
  id->key = k.data;
  id->nulls = k.nulls;
  id->codes = q.build_code_col >= 0 && q.build_code_col < q.build_ncols ? q.build_cols[q.build_code_col].data
                                                                       : nullptr;
  id->bn = q.build_n;
  id->type = k.type;
  id->G = static_cast<int32_t>(G);
  return true;
}

}  // namespace kg

namespace {

}  // namespace

struct kg_batch {
  kg_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hipEvent_t ready = nullptr;  // the context stream's work before a submit
  // Whether a submit orders the batch after the context stream's work (its
  // columns may have been written there). The serving loop, which reads only
  // resident columns and synchronizes the device before its clock starts,
  // turns it off: a batch must not wait for work another stream shares a
  // hardware queue with.
  bool after_ctx = true;
  int grid = 0;
  int grid_gen = 0;
  int grid_shared = 0;
  char* d_desc = nullptr;      // BHeader, then BQuery[kBMaxQueries]
  char* h_desc = nullptr;      // pinned staging of the same
  char* arena = nullptr;
  size_t arena_bytes = 0;
  char* h_res = nullptr;       // result areas: coherent pinned host memory written by the device
  char* d_res = nullptr;       // its device address
  int32_t* h_done = nullptr;   // per-query completion words (coherent pinned host memory)
  int32_t* d_done = nullptr;
  int32_t seq = 0;             // sequence number of the batch in flight
  size_t res_cap = 0;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  // in flight
  bool pending = false;
  int32_t nq = 0;
  size_t res_used = 0;
  std::vector<BPlan> plans;
  BHeader* hdr() { return reinterpret_cast<BHeader*>(h_desc); }
  BQuery* hq() { return reinterpret_cast<BQuery*>(h_desc + round256(sizeof(BHeader))); }
  BHeader* dhdr() { return reinterpret_cast<BHeader*>(d_desc); }
  BQuery* dq() { return reinterpret_cast<BQuery*>(d_desc + round256(sizeof(BHeader))); }
  size_t groups_off() const { return round256(sizeof(BHeader)) + round256(sizeof(BQuery) * kBMaxQueries); }
  SGroup* hg() { return reinterpret_cast<SGroup*>(h_desc + groups_off()); }
  SGroup* dg() { return reinterpret_cast<SGroup*>(d_desc + groups_off()); }
};

namespace {

size_t desc_bytes() {
  return round256(sizeof(BHeader)) + round256(sizeof(BQuery) * kBMaxQueries) + round256(sizeof(SGroup) * kBMaxQueries);
}

int capture(kg_batch* b) {
  kg_ctx* ctx = b->ctx;
  KG_HIP_CHECK(ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(k_batch_clear, dim3(b->grid), dim3(kBThreads), 0, b->stream, b->dhdr());
  hipLaunchKernelGGL(k_batch_range, dim3(b->grid_gen), dim3(kBThreads), 0, b->stream, b->dhdr(), b->dq());
  hipLaunchKernelGGL(k_batch_build, dim3(b->grid_gen), dim3(kBThreads), 0, b->stream, b->dhdr(), b->dq());
  hipLaunchKernelGGL(k_batch_shared, dim3(b->grid_shared), dim3(kSThreads), 0, b->stream, b->dhdr(), b->dq(),
                     b->dg());
  hipLaunchKernelGGL(k_batch_join, dim3(b->grid_gen), dim3(kBThreads), 0, b->stream, b->dhdr(), b->dq());
  hipLaunchKernelGGL(k_batch_gen, dim3(b->grid_gen), dim3(kBThreads), 0, b->stream, b->dhdr(), b->dq());
  hipLaunchKernelGGL(k_batch_final, dim3(kBMaxQueries), dim3(kWave), 0, b->stream, b->dhdr(), b->dq());
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess) return kg_fail(ctx, KG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  b->graph = g;
  KG_HIP_CHECK(ctx, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return KG_OK;
}

bool fits_group(const BQuery& D);

// For the query that builds a map or table (alone, or the leader of a shared
// build source with members mem): when every member's build predicates are
// one conjunction of comparisons of the build key column (not REAL) with
// literals, the union of the members' key intervals, merged, into lead.nivl /
// ivl_lo / ivl_hi (at most kBMaxIvl intervals, else nivl stays -1).
// A row passes some member's predicates exactly when its non-NULL key lies in
// the union; a member without predicates passes every row.
void key_interval_union(BQuery& lead, const int32_t* mem, int nmem, const BQuery* Q) {
  std::vector<std::pair<int64_t, int64_t>> iv;
  for (int m = 0; m < nmem; m++) {
    const BQuery& M = Q[mem[m]];
    if (M.PB.n > 0 && !M.PB.conj) return;
    int64_t lo = INT64_MIN, hi = INT64_MAX;
    for (int p = 0; p < M.PB.n; p++) {
      if (M.PB.data[p] != M.bkey.data || M.PB.type[p] != M.bkey.type || M.PB.type[p] == KG_REAL ||
          M.PB.nulls[p] != M.bkey.nulls)
        return;
      const int64_t v = M.PB.ilit[p];
      switch (M.PB.op[p]) {
        case KG_EQ: lo = std::max(lo, v); hi = std::min(hi, v); break;
        case KG_LEQ: hi = std::min(hi, v); break;
        case KG_GEQ: lo = std::max(lo, v); break;
        case KG_LT:  // key < INT64_MIN: empty
          if (v == INT64_MIN) {
            lo = 1;
            hi = 0;
          } else {
            hi = std::min(hi, v - 1);
          }
          break;
        case KG_GT:  // key > INT64_MAX: empty
          if (v == INT64_MAX) {
            lo = 1;
            hi = 0;
          } else {
            lo = std::max(lo, v + 1);
          }
          break;
        default: return;  // <> splits the key range
      }
    }
    if (lo <= hi) iv.push_back({lo, hi});
  }
  std::sort(iv.begin(), iv.end());
  std::vector<std::pair<int64_t, int64_t>> u;
  for (const auto& x : iv) {
    if (!u.empty() && (u.back().second == INT64_MAX || x.first <= u.back().second + 1))
      u.back().second = std::max(u.back().second, x.second);
    else
      u.push_back(x);
  }
  if (u.size() > static_cast<size_t>(kBMaxIvl)) return;
  lead.nivl = static_cast<int32_t>(u.size());
  for (size_t i = 0; i < u.size(); i++) {
    lead.ivl_lo[i] = u[i].first;
    lead.ivl_hi[i] = u[i].second;
  }
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// The register-path view of a scan, when its columns allow vector loads.
bool fast_view(const PredSet& P, const KeySet& K, const MeasureSet& M, BFast& F) {
  std::memset(&F, 0, sizeof(F));
  auto is_i32 = [](int32_t t) { return t == KG_INT || t == KG_DATE || t == KG_ENUM; };
  if (!P.conj || P.n > kBFastPreds) return false;
  F.npred = P.n;
  for (int i = 0; i < P.n; i++) {
    if (!is_i32(P.type[i]) || P.nulls[i] || !aligned16(P.data[i])) return false;
    F.op[i] = P.op[i];
    F.lit[i] = P.ilit[i];
    F.pcol[i] = static_cast<const int32_t*>(P.data[i]);
  }
  F.nk = K.nk;
  F.card1 = K.card1;
  for (int k = 0; k < K.nk; k++) {
    if (!aligned16(K.data[k])) return false;
    F.key[k] = K.data[k];
  }
  for (int m = 0; m < M.nm; m++) {
    if (M.nulls[m] || M.data2[m] || !aligned16(M.data[m])) return false;
    if (M.type[m] == KG_REAL) F.real[m] = 1;
    else if (is_i32(M.type[m])) F.real[m] = 0;
    else return false;
    F.mcol[m] = M.data[m];
  }
  return true;
}

// The typed join view, when the probe side's columns allow vector loads.
bool join_view(const kg_query& q, const PredSet& P, const MeasureSet& M, BJoinFast& J) {
  std::memset(&J, 0, sizeof(J));
  auto is_i32 = [](int32_t t) { return t == KG_INT || t == KG_DATE || t == KG_ENUM; };
  if (!P.conj || P.n > kBFastPreds || M.nm > 2) return false;
  J.npred = P.n;
  for (int i = 0; i < P.n; i++) {
    if (!is_i32(P.type[i]) || P.nulls[i] || !aligned16(P.data[i])) return false;
    J.op[i] = P.op[i];
    J.lit[i] = P.ilit[i];
    J.pcol[i] = static_cast<const int32_t*>(P.data[i]);
  }
  const kg_col& k = q.cols[q.probe_key_col];
  if (k.nulls || !aligned16(k.data)) return false;
  if (k.type == KG_BIGINT) J.key64 = 1;
  else if (!is_i32(k.type)) return false;
  J.key = k.data;
  for (int m = 0; m < M.nm; m++) {
    if (M.nulls[m] || !aligned16(M.data[m])) return false;
    if (M.type[m] == KG_REAL) J.mkind[m] = 1;
    else if (is_i32(M.type[m])) J.mkind[m] = 2;
    else return false;
    J.mcol[m] = M.data[m];
    if (M.data2[m]) {
      if (M.nulls2[m] || !aligned16(M.data2[m])) return false;
      if (M.type2[m] == KG_REAL) J.fkind[m] = 1;
      else if (is_i32(M.type2[m])) J.fkind[m] = 2;
      else return false;
      J.fcol[m] = M.data2[m];
    }
  }
  return true;
}

// Validates and plans query q into the device descriptor D and host plan P.
int plan_query(kg_ctx* ctx, const kg_query& q, BQuery& D, BPlan& P) {
  std::memset(&D, 0, sizeof(D));
  D.nivl = -1;
  if (q.kind == KG_Q_CHAIN) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "chain queries are not batched");
  KG_REQUIRE(ctx, q.kind == KG_Q_SCAN_AGG || q.kind == KG_Q_JOIN_AGG, "bad query kind %d", q.kind);
  KG_REQUIRE(ctx, q.n >= 0 && q.n < INT32_MAX, "row count %lld out of range", (long long)q.n);
  KG_REQUIRE(ctx, q.naggs >= 0 && q.naggs <= kBMaxAggs, "at most %d aggregates", kBMaxAggs);
  KG_REQUIRE(ctx, q.ncols > 0 && q.cols, "no columns");
  int st;
  if ((st = build_preds(ctx, q.cols, q.ncols, q.preds, q.npreds, 1, q.n, D.P)) != KG_OK) return st;
  bool minmax = false;
  if ((st = build_measures(ctx, q.cols, q.ncols, q.aggs, q.naggs, q.n, P.M, P.plan, minmax)) != KG_OK) return st;
  if (minmax) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "MIN/MAX are not batched");
  // Batched REAL sums add in a timing-dependent order; in the deterministic
  // mode such a query is not taken (the serving loop runs it on the plain path).
  if (ctx->deterministic)
    for (int m = 0; m < P.M.nm; m++)
      if (P.M.is_real[m])
        return kg_fail(ctx, KG_ERR_UNSUPPORTED, "REAL sums are not batched in the deterministic mode");
  D.M = P.M;
  D.kind = q.kind;
  D.n = q.n;
  D.nm = P.M.nm;
  D.naggs = q.naggs;
  int64_t G = 1;
  if (q.kind == KG_Q_SCAN_AGG) {
    KG_REQUIRE(ctx, q.nkeys >= 0 && q.nkeys <= 2, "0 to 2 dense key columns (got %d)", q.nkeys);
    KG_REQUIRE(ctx, q.nkeys == 0 || (q.key_cols && q.key_card), "key arrays are NULL");
    D.K.nk = q.nkeys;
    for (int k = 0; k < q.nkeys; k++) {
      KG_REQUIRE(ctx, q.key_cols[k] >= 0 && q.key_cols[k] < q.ncols, "bad key column %d", q.key_cols[k]);
      const kg_col& c = q.cols[q.key_cols[k]];
      if (c.type != KG_INT && c.type != KG_ENUM && c.type != KG_DATE)
        return kg_fail(ctx, KG_ERR_UNSUPPORTED, "dense group keys must be 32-bit codes (type %d)", c.type);
      if (c.nulls) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "nullable dense group keys are not supported");
      KG_REQUIRE(ctx, q.key_card[k] > 0, "key cardinality must be positive");
      D.K.data[k] = static_cast<const int32_t*>(c.data);
      G *= q.key_card[k];
      if (G > kMaxLdsGroups) break;
    }
    if (q.nkeys == 2) D.K.card1 = q.key_card[1];
    P.nkeys_out = q.nkeys;
    P.card1 = D.K.card1;
    D.first_shift = 0;
  } else {
    KG_REQUIRE(ctx, q.probe_key_col >= 0 && q.probe_key_col < q.ncols, "bad probe key column %d", q.probe_key_col);
    if ((st = check_join_key(ctx, q.cols[q.probe_key_col], q.n, "probe key")) != KG_OK) return st;
    KG_REQUIRE(ctx, q.build_n >= 0 && q.build_n < INT32_MAX, "build row count out of range");
    KG_REQUIRE(ctx, q.build_ncols > 0 && q.build_cols, "no build columns");
    KG_REQUIRE(ctx, q.build_key_col >= 0 && q.build_key_col < q.build_ncols, "bad build key column %d",
               q.build_key_col);
    if ((st = check_join_key(ctx, q.build_cols[q.build_key_col], q.build_n, "build key")) != KG_OK) return st;
    if ((st = build_preds(ctx, q.build_cols, q.build_ncols, q.build_preds, q.build_npreds, 1, q.build_n, D.PB)) !=
        KG_OK)
      return st;
    D.pkey = q.cols[q.probe_key_col];
    D.bkey = q.build_cols[q.build_key_col];
    D.bn = q.build_n;
    if (q.build_code_col >= 0) {
      KG_REQUIRE(ctx, q.build_code_col < q.build_ncols, "bad build code column %d", q.build_code_col);
      const kg_col& c = q.build_cols[q.build_code_col];
      if (c.type != KG_INT && c.type != KG_ENUM)
        return kg_fail(ctx, KG_ERR_UNSUPPORTED, "build group codes must be INT/ENUM codes");
      if (c.nulls) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "nullable build group codes");
      KG_REQUIRE(ctx, q.build_n == 0 || c.data, "build code column has no data");
      KG_REQUIRE(ctx, q.group_card > 0, "group cardinality must be positive");
      D.codes = static_cast<const int32_t*>(c.data);
      G = q.group_card;
      P.nkeys_out = 1;
    } else {
      G = 1;
      P.nkeys_out = 0;
    }
    const double lf = query_lf(q);
    KG_REQUIRE(ctx, lf > 0 && lf <= 0.9, "load factor %g out of (0, 0.9]", lf);
    D.first_shift = 32;
    P.join = true;
    P.bn = q.build_n;
  }
  if (G > kMaxLdsGroups)
    return kg_fail(ctx, KG_ERR_UNSUPPORTED, "%lld dense groups exceed the on-chip limit %d", (long long)G,
                   kMaxLdsGroups);
  const int nm = P.M.nm;
  bool nonnull = true;
  for (int m = 0; m < nm; m++)
    if (P.M.nulls[m] || (P.M.data2[m] && P.M.nulls2[m])) nonnull = false;
  D.reg = (q.kind == KG_Q_SCAN_AGG && G <= kBRegG && nonnull && fast_view(D.P, D.K, P.M, D.F)) ? 1 : 0;
  if (D.reg && !fits_group(D)) D.reg = 0;  // its columns alone exceed a staged tile
  D.jfast = (q.kind == KG_Q_JOIN_AGG && join_view(q, D.P, P.M, D.J)) ? 1 : 0;
  if (!D.reg && G * nm * 2 + G * 2 > kBWaveWords)
    return kg_fail(ctx, KG_ERR_UNSUPPORTED, "%lld groups x %d measures exceed the batched LDS table", (long long)G,
                   nm);
  D.G = static_cast<int32_t>(G);
  P.G = D.G;
  for (int a = 0; a < q.naggs; a++) D.cnull[a] = P.plan[a].kind == KG_AGG_COUNT ? P.plan[a].count_nulls : nullptr;
  return KG_OK;
}

// Rows per item of a queue over the queries `ids`: the largest power of two
// up to kBTile that still gives kBItemTarget items over all their rows (about
// a quarter of the waves the chip holds at the join kernel's occupancy: more,
// smaller items measured slower on 1M- to 50M-row probes), so a
// batch of a few queries over small or mid-size tables spreads over the waves
// instead of running one long item per wave (an item's rows are walked by one
// wave, one dependent map or bucket lookup per row group). Larger items keep
// the per-item partial flushes (global atomics) few.
int queue_tile(const std::vector<int32_t>& ids, const kg_query* qs) {
  int64_t rows = 0;
  for (int i : ids) rows += qs[i].n;
  int tile = kBTile;
  while (tile > kBTileMin && rows / tile < kBItemTarget) tile /= 2;
  return tile;
}

// Tables of the queries `ids` (equal table_id and n share tiles), in order of
// first appearance, and the item prefix of each XCD queue.
bool layout_queues(BQueues& Q, const std::vector<int32_t>& ids, const kg_query* qs) {
  std::vector<std::pair<std::pair<int32_t, int64_t>, std::vector<int32_t>>> tables;
  for (int i : ids) {
    if (qs[i].n == 0) continue;
    const std::pair<int32_t, int64_t> key{qs[i].table_id, qs[i].n};
    auto it = std::find_if(tables.begin(), tables.end(), [&](const auto& e) { return e.first == key; });
    if (it == tables.end()) {
      tables.push_back({key, {}});
      it = tables.end() - 1;
    }
    it->second.push_back(i);
  }
  // Smallest tables first: their queries complete (and are published) early.
  std::stable_sort(tables.begin(), tables.end(),
                   [](const auto& a, const auto& b) { return a.first.second < b.first.second; });
  const int nt = static_cast<int>(tables.size());
  Q.nt = nt;
  Q.toff[0] = 0;
  for (int t = 0; t < nt; t++) {
    Q.tn[t] = tables[t].first.second;
    for (size_t j = 0; j < tables[t].second.size(); j++) Q.tq[Q.toff[t] + j] = tables[t].second[j];
    Q.toff[t + 1] = Q.toff[t] + static_cast<int32_t>(tables[t].second.size());
  }
  std::memset(Q.next, 0, sizeof(Q.next));
  for (int x = 0; x < kBQueues; x++) {
    Q.qpre[x][0] = 0;
    for (int t = 0; t < nt; t++) {
      const int64_t tiles = ceil_div(Q.tn[t], Q.tile);
      const int64_t mine = tiles > x ? (tiles - x + kBQueues - 1) / kBQueues : 0;
      const int64_t items = mine * (Q.toff[t + 1] - Q.toff[t]);
      if (Q.qpre[x][t] + items >= INT32_MAX) return false;
      Q.qpre[x][t + 1] = Q.qpre[x][t] + static_cast<int32_t>(items);
    }
  }
  return true;
}

// Columns (pointer, bytes per value) a register-path query reads.
int fast_columns(const BQuery& D, std::pair<const void*, int> (&cols)[kSMaxCols]) {
  int n = 0;
  auto add = [&](const void* p, int w) {
    for (int c = 0; c < n; c++)
      if (cols[c].first == p) return;
    if (n < kSMaxCols) cols[n] = {p, w};
    n++;
  };
  for (int p = 0; p < D.F.npred; p++) add(D.F.pcol[p], 4);
  for (int k = 0; k < D.F.nk; k++) add(D.F.key[k], 4);
  for (int m = 0; m < D.nm; m++) add(D.F.mcol[m], D.F.real[m] ? 8 : 4);
  return n;
}

int fast_row_bytes(const BQuery& D) {
  std::pair<const void*, int> cols[kSMaxCols];
  const int n = fast_columns(D, cols);
  if (n > kSMaxCols) return kSBufBytes;
  int bytes = 0;
  for (int c = 0; c < n; c++) bytes += cols[c].second;
  return bytes;
}

// Whether a register-path query fits one shared-scan group on its own.
bool fits_group(const BQuery& D) { return fast_row_bytes(D) <= kSRowBytes; }

// Groups the register-path queries by table (table_id, n), smallest tables
// first; a group holds at most kSMaxQ queries whose columns together fit a
// staged tile (a query that does not fit starts the next group of its table).
// Fills the queries' column indices and item counts and the tile sequence.
bool layout_groups(STiles& ST, SGroup* gs, const std::vector<int32_t>& ids, const kg_query* qs, BQuery* Q) {
  std::vector<std::pair<std::pair<int32_t, std::pair<int64_t, int>>, std::vector<int32_t>>> tables;
  for (int i : ids) {
    if (qs[i].n == 0) continue;
    const std::pair<int32_t, std::pair<int64_t, int>> key{qs[i].table_id, {qs[i].n, 0}};
    auto it = std::find_if(tables.begin(), tables.end(), [&](const auto& e) { return e.first == key; });
    if (it == tables.end()) {
      tables.push_back({key, {}});
      it = tables.end() - 1;
    }
    it->second.push_back(i);
  }
  std::stable_sort(tables.begin(), tables.end(),
                   [](const auto& a, const auto& b) { return a.first.second.first < b.first.second.first; });
  int ng = 0;
  for (const auto& t : tables) {
    // A group's tiles take 1024 rows while its columns take at most 32 bytes
    // per row, else 512 (at most kSRowBytes bytes per row).
    const int max_bytes = kSRowBytes;
    SGroup* cur = nullptr;
    int bytes = 0, units = 0;
    for (int i : t.second) {
      BQuery& D = Q[i];
      const int mu = std::max(1, (D.nm + kSUnitM - 1) / kSUnitM) * ((D.G + kSUnitG - 1) / kSUnitG);  // its units
      std::pair<const void*, int> cols[kSMaxCols];
      const int nc = fast_columns(D, cols);
      auto find = [](const SGroup* g, const void* p) {
        for (int c = 0; c < g->ncol; c++)
          if (g->col[c] == p) return c;
        return -1;
      };
      int add_cols = 0, add_bytes = 0;
      for (int c = 0; c < nc && cur; c++)
        if (find(cur, cols[c].first) < 0) {
          add_cols++;
          add_bytes += cols[c].second;
        }
      if (!cur || cur->nq == kSMaxQ || units + mu > kSWaves || cur->ncol + add_cols > kSMaxCols ||
          bytes + add_bytes > max_bytes) {
        if (ng == kBMaxQueries) return false;
        cur = &gs[ng++];
        std::memset(cur, 0, sizeof(SGroup));
        cur->n = t.first.second.first;
        bytes = 0;
        units = 0;
      }
      units += mu;
      auto index = [&](const void* p, int w) {
        int c = find(cur, p);
        if (c < 0) {
          c = cur->ncol++;
          cur->col[c] = p;
          cur->width[c] = w;
          bytes += w;
        }
        return static_cast<int8_t>(c);
      };
      for (int p = 0; p < D.F.npred; p++) D.F.pcx[p] = index(D.F.pcol[p], 4);
      for (int k = 0; k < D.F.nk; k++) D.F.kcx[k] = index(D.F.key[k], 4);
      for (int m = 0; m < D.nm; m++) D.F.mcx[m] = index(D.F.mcol[m], D.F.real[m] ? 8 : 4);
      cur->q[cur->nq++] = i;
    }
  }
  ST.ng = ng;
  ST.tpre[0] = 0;
  for (int g = 0; g < ng; g++) {
    SGroup& S = gs[g];
    int rb = 0;
    for (int c = 0; c < S.ncol; c++) rb += S.width[c];
    S.rows = rb * kSMaxTile <= kSBufBytes ? kSMaxTile : kSMaxTile / 2;
    for (int q = 0; q < S.nq; q++) Q[S.q[q]].items = static_cast<int32_t>(ceil_div(S.n, S.rows));
    int off = 0;
    for (int c = 0; c < S.ncol; c++) {
      S.off[c] = off;
      off += S.rows * S.width[c];
    }
    S.tile_bytes = (off + 15) / 16 * 16;  // <= kSBufBytes: a group's row bytes are bounded (kSRowBytes)
    // Units: each query's measures in pairs, then every pair over rs row
    // slices (whole row groups of 4 rows per lane), at most one unit per wave.
    auto pairs = [&](int q) { return std::max(1, (Q[S.q[q]].nm + kSUnitM - 1) / kSUnitM); };
    auto halves = [&](int q) { return (Q[S.q[q]].G + kSUnitG - 1) / kSUnitG; };
    int mu = 0;
    for (int q = 0; q < S.nq; q++) mu += pairs(q) * halves(q);
    int rs = 1;
    while (mu * rs * 2 <= kSWaves && S.rows / (rs * 2) >= kWave * 4) rs *= 2;
    int w = 0;
    for (int q = 0; q < S.nq; q++)
      for (int k = 0; k < pairs(q); k++)
        for (int hf = 0; hf < halves(q); hf++)
          for (int r = 0; r < rs; r++, w++) {
            S.uq[w] = static_cast<int8_t>(q);
            S.um[w] = static_cast<int8_t>(k * kSUnitM);
            S.ug[w] = static_cast<int8_t>(hf * kSUnitG);
            S.ur[w] = static_cast<int16_t>(r * (S.rows / rs));
            S.urg[w] = static_cast<int8_t>(S.rows / rs / (kWave * 4));
            S.ucnt[w] = k == 0 ? 1 : 0;
          }
    S.nunits = w;
    const int64_t tiles = ceil_div(S.n, S.rows);
    if (ST.tpre[g] + tiles >= INT32_MAX) return false;
    ST.tpre[g + 1] = ST.tpre[g] + static_cast<int32_t>(tiles);
  }
  return true;
}

}  // namespace

extern "C" int kg_batch_create(kg_ctx* ctx, const kg_batch_opts* opts, kg_batch** out) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, out, "out is NULL");
  *out = nullptr;
  const int64_t arena = (opts && opts->arena_bytes > 0) ? opts->arena_bytes : (int64_t{1} << 30);
  KG_REQUIRE(ctx, arena >= (1 << 20), "arena of %lld bytes is too small", (long long)arena);
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  kg_batch* b = new (std::nothrow) kg_batch();
  if (!b) return kg_fail(ctx, KG_ERR_OOM, "host allocation failed");
  b->ctx = ctx;
  b->arena_bytes = static_cast<size_t>(arena);
  b->res_cap = std::min<size_t>(b->arena_bytes, size_t{64} << 20);
  b->grid = ctx->num_cus * 3;
  b->grid_gen = std::max(ctx->num_cus * 8, kBQueues);  // drain_wave needs a workgroup per XCD queue
  b->grid_shared = ctx->num_cus * 2;
  auto fail = [&](int st) {
    kg_batch_free(b);
    return st;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ready, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    return fail(kg_fail(ctx, KG_ERR_HIP, "stream/event creation failed"));
  }
  if (hipMalloc(reinterpret_cast<void**>(&b->d_desc), desc_bytes()) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->arena), b->arena_bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(kg_fail(ctx, KG_ERR_OOM, "device allocation of the batch arena (%zu bytes) failed", b->arena_bytes));
  }
  const unsigned coherent = hipHostMallocCoherent | hipHostMallocMapped;
  if (hipHostMalloc(reinterpret_cast<void**>(&b->h_desc), desc_bytes(), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&b->h_res), b->res_cap, coherent) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&b->h_done), kBMaxQueries * sizeof(int32_t), coherent) != hipSuccess ||
      hipHostGetDevicePointer(reinterpret_cast<void**>(&b->d_res), b->h_res, 0) != hipSuccess ||
      hipHostGetDevicePointer(reinterpret_cast<void**>(&b->d_done), b->h_done, 0) != hipSuccess) {
    (void)hipGetLastError();
    return fail(kg_fail(ctx, KG_ERR_OOM, "pinned host allocation failed"));
  }
  std::memset(b->h_done, 0, kBMaxQueries * sizeof(int32_t));
  std::memset(b->h_desc, 0, desc_bytes());
  int st = capture(b);
  if (st != KG_OK) return fail(st);
  *out = b;
  return KG_OK;
}

extern "C" int kg_batch_free(kg_batch* b) {
  if (!b) return KG_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->arena) (void)hipFree(b->arena);
  if (b->h_desc) (void)hipHostFree(b->h_desc);
  if (b->h_res) (void)hipHostFree(b->h_res);
  if (b->h_done) (void)hipHostFree(b->h_done);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->ready) (void)hipEventDestroy(b->ready);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  (void)hipGetLastError();
  delete b;
  return KG_OK;
}

void kg::batch_order_after_ctx(kg_batch* b, bool on) { b->after_ctx = on; }

extern "C" int kg_batch_footprint(const kg_query* q, int64_t* bytes) {
  if (!q || !bytes) return KG_ERR_INVALID;
  *bytes = static_cast<int64_t>(query_footprint(*q));
  return KG_OK;
}

extern "C" int64_t kg_batch_arena_bytes(const kg_batch* b) { return b ? static_cast<int64_t>(b->arena_bytes) : 0; }

extern "C" int kg_batch_submit(kg_batch* b, const kg_query* qs, int32_t nq) {
  if (!b) return KG_ERR_INVALID;
  kg_ctx* ctx = b->ctx;
  KG_REQUIRE(ctx, !b->pending, "the previous batch has not been collected");
  KG_REQUIRE(ctx, nq >= 0 && nq <= kBMaxQueries, "1 to %d queries per batch (got %d)", kBMaxQueries, nq);
  KG_REQUIRE(ctx, nq == 0 || qs, "queries are NULL");
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // Join queries without a load factor take the context's tuned choice for
  // their build size (kg_join_tune), else 0.5.
  std::vector<kg_query> resolved(qs, qs + std::max(nq, 0));
  for (kg_query& q : resolved)
    if (q.kind == KG_Q_JOIN_AGG && !(q.load_factor > 0)) q.load_factor = kg::tuned_load_factor(ctx, q.build_n);
  qs = resolved.data();
  BHeader* H = b->hdr();
  BQuery* Q = b->hq();
  std::memset(H, 0, sizeof(BHeader));
  b->plans.assign(nq, BPlan{});
  // Plan; result areas go to the pinned result buffer, accumulators, counters
  // and tables to the device arena.
  b->seq++;
  std::vector<int32_t> ok;
  size_t off = 0;
  for (int i = 0; i < nq; i++) {
    BPlan& P = b->plans[i];
    const std::string saved = ctx->err;
    P.status = plan_query(ctx, qs[i], Q[i], P);
    if (P.status != KG_OK) {
      P.err = ctx->err;
      ctx->err = saved;
      continue;
    }
    P.L = res_layout(P.G, P.M.nm, qs[i].naggs);
    if (off + P.L.bytes > b->res_cap) {
      P.status = KG_ERR_CAPACITY;
      P.err = "result areas exceed the batch's result buffer";
      continue;
    }
    P.res_off = off;
    off += P.L.bytes;
    ok.push_back(i);
  }
  b->res_used = off;
  // Build sources: typed joins that a direct map may serve, by build key
  // column, build rows and group codes. A source with several queries, or with
  // a build side beyond kBMapMax rows, is built once for the batch.
  struct Source {
    const void* key;
    int32_t type;
    const int8_t* nulls;
    int64_t bn;
    const int32_t* codes;
    int32_t G;
    double lf;
    std::vector<int32_t> members;
    int32_t msize = 0;
    bool map_only = false, ok = true;
  };
  std::vector<Source> srcs;
  std::vector<int> src_of(nq, -1);
  for (int i : ok) {
    const BPlan& P = b->plans[i];
    const BQuery& D = Q[i];
    if (!P.join || !D.jfast || !source_map_entries(P.bn, P.G)) continue;
    auto it = std::find_if(srcs.begin(), srcs.end(), [&](const Source& x) {
      return x.key == D.bkey.data && x.type == D.bkey.type && x.nulls == D.bkey.nulls && x.bn == P.bn &&
             x.codes == D.codes && x.G == P.G;
    });
    if (it == srcs.end()) {
      srcs.push_back(Source{D.bkey.data, D.bkey.type, D.bkey.nulls, P.bn, D.codes, P.G, query_lf(qs[i]), {}});
      it = srcs.end() - 1;
    }
    it->members.push_back(i);
    src_of[i] = static_cast<int>(it - srcs.begin());
  }
  for (size_t si = 0; si < srcs.size(); si++) {
    Source& S = srcs[si];
    if (S.members.size() < 2 && S.bn <= kBMapMax) {  // one query, small build side: its own map and table
      src_of[S.members[0]] = -1;
      S.members.clear();
      continue;
    }
    S.map_only = S.bn > kBMapMax;
    S.msize = source_map_entries(S.bn, S.G);
  }
  off = 0;
  int nf = 0;
  std::vector<int> rem_fill(nq, -1);  // fill entry of each query's remaining-item count
  // Arena: each query's accumulators (and the table of a join that is not in
  // a shared source), then the shared sources. A query or source that does not
  // fit is left to the plain path (KG_ERR_UNSUPPORTED).
  for (int i : ok) {
    BPlan& P = b->plans[i];
    BQuery& D = Q[i];
    const int G = P.G, nm = P.M.nm;
    const size_t acc_bytes = round256(static_cast<size_t>(kBSlots) * G * (2 * std::max(nm, 1) + 1) * 8 + 16);
    const size_t first_bytes = round256(static_cast<size_t>(kBSlots) * G * 8);
    size_t need = acc_bytes + first_bytes;
    uint64_t nbuckets = 0;
    const bool own_table = P.join && src_of[i] < 0;
    const int32_t msize = own_table ? map_entries(D.jfast != 0, P.bn, G) : 0;
    if (own_table) {
      const int32_t self = i;
      key_interval_union(D, &self, 1, Q);
      nbuckets = table_buckets(P.bn, query_lf(qs[i]));
      need += round256(nbuckets * kSlots * 8) + round256(nbuckets * kSlots * 4);
      if (msize) need += round256(static_cast<size_t>(msize) * 4) + kRangeBytes;
    }
    if (off + need > b->arena_bytes) {
      P.status = KG_ERR_UNSUPPORTED;
      P.err = "the batch arena cannot hold this query's tables";
      continue;
    }
    char* base = b->arena + off;
    D.sum = reinterpret_cast<uint64_t*>(base);
    D.nn = D.sum + static_cast<size_t>(kBSlots) * G * std::max(nm, 1);
    D.rows = D.nn + static_cast<size_t>(kBSlots) * G * std::max(nm, 1);
    D.flag = reinterpret_cast<uint32_t*>(D.rows + static_cast<size_t>(kBSlots) * G);
    D.rem = reinterpret_cast<unsigned long long*>(D.rows + static_cast<size_t>(kBSlots) * G + 1);
    D.sflag = P.join ? D.flag : nullptr;
    // Fill ranges must not overlap: the fill kernel writes them in no fixed order.
    H->fills[nf++] = BFill{D.sum, static_cast<int64_t>(kBSlots) * G * (2 * std::max(nm, 1) + 1) + 1, 0};
    D.items = 0;  // set with the layout of the query's queue or shared-scan group
    rem_fill[i] = nf;
    H->fills[nf++] = BFill{reinterpret_cast<uint64_t*>(D.rem), 1, static_cast<uint64_t>(D.items)};
    D.first = reinterpret_cast<uint64_t*>(base + acc_bytes);
    H->fills[nf++] = BFill{D.first, static_cast<int64_t>(kBSlots) * G, ~0ull};
    D.map = nullptr;
    D.krange = nullptr;
    D.msize = 0;
    if (own_table) {
      char* tb = base + acc_bytes + first_bytes;
      D.t.keys = reinterpret_cast<int64_t*>(tb);
      D.t.rows = reinterpret_cast<int32_t*>(tb + round256(nbuckets * kSlots * 8));
      D.t.nbuckets = nbuckets;
      D.t.side_rows = nullptr;
      D.t.side_count = nullptr;
      H->fills[nf++] = BFill{reinterpret_cast<uint64_t*>(D.t.keys), static_cast<int64_t>(nbuckets * kSlots),
                             static_cast<uint64_t>(kEmpty)};
      if (msize) {
        char* mb = tb + round256(nbuckets * kSlots * 8) + round256(nbuckets * kSlots * 4);
        D.map = reinterpret_cast<uint32_t*>(mb);
        D.krange = reinterpret_cast<int64_t*>(mb + round256(static_cast<size_t>(msize) * 4));
        D.msize = msize;
        H->fills[nf++] = BFill{reinterpret_cast<uint64_t*>(D.map), msize / 2, ~0ull};
        H->fills[nf++] = BFill{reinterpret_cast<uint64_t*>(D.krange), kRangeCopies * kRangeStride,
                               static_cast<uint64_t>(INT64_MIN)};
      }
    }
    D.res = b->d_res + P.res_off;
    D.done = b->d_done + i;
    D.seq = b->seq;
    off += need;
  }
  int32_t nbmem = 0;
  for (Source& S : srcs) {
    if (S.members.empty()) continue;
    // flag word, key range copies, map, then (small build sides) the table
    const uint64_t nbuckets = S.map_only ? 0 : table_buckets(S.bn, S.lf);
    const size_t map_b = round256(static_cast<size_t>(S.msize) * 4);
    const size_t keys_b = round256(nbuckets * kSlots * 8), rows_b = round256(nbuckets * kSlots * 4);
    const size_t need = 256 + kRangeBytes + map_b + keys_b + rows_b;
    bool live = false;
    for (int i : S.members) live |= b->plans[i].status == KG_OK;
    if (!live) continue;
    if (off + need > b->arena_bytes) {
      for (int i : S.members)
        if (b->plans[i].status == KG_OK) {
          b->plans[i].status = KG_ERR_UNSUPPORTED;
          b->plans[i].err = "the batch arena cannot hold this query's tables";
        }
      continue;
    }
    char* base = b->arena + off;
    uint32_t* sflag = reinterpret_cast<uint32_t*>(base);
    int64_t* krange = reinterpret_cast<int64_t*>(base + 256);
    uint32_t* map = reinterpret_cast<uint32_t*>(base + 256 + kRangeBytes);
    H->fills[nf++] = BFill{reinterpret_cast<uint64_t*>(sflag), 1, 0};
    H->fills[nf++] = BFill{reinterpret_cast<uint64_t*>(krange), kRangeCopies * kRangeStride,
                           static_cast<uint64_t>(INT64_MIN)};
    H->fills[nf++] = BFill{reinterpret_cast<uint64_t*>(map), S.msize / 2, ~0ull};
    TableView tv{};
    if (!S.map_only) {
      tv.keys = reinterpret_cast<int64_t*>(base + 256 + kRangeBytes + map_b);
      tv.rows = reinterpret_cast<int32_t*>(base + 256 + kRangeBytes + map_b + keys_b);
      tv.nbuckets = nbuckets;
      H->fills[nf++] = BFill{reinterpret_cast<uint64_t*>(tv.keys), static_cast<int64_t>(nbuckets * kSlots),
                             static_cast<uint64_t>(kEmpty)};
    }
    off += need;
    const bool several = S.members.size() > 1;
    bool leader = true;
    int lead = -1;
    const int32_t blo = nbmem;
    for (int i : S.members)
      if (b->plans[i].status == KG_OK) H->bmem[nbmem++] = i;
    for (int i : S.members) {
      BPlan& P = b->plans[i];
      if (P.status != KG_OK) continue;
      BQuery& D = Q[i];
      P.map_only = S.map_only;
      D.map = map;
      D.krange = krange;
      D.msize = S.msize;
      D.t = tv;
      D.sflag = sflag;
      D.map_only = S.map_only ? 1 : 0;
      D.build_all = (leader && several) ? 1 : 0;  // the leader builds the source (the others add no build items)
      D.nivl = -1;
      if (leader) lead = i;
      D.bmem_lo = blo;
      D.bmem_hi = nbmem;
      D.recheck = several ? 1 : 0;
      P.build_entry = leader;
      leader = false;
      // Build predicates that are conjuncts on the build key column are
      // checked on the (equal) probe key; any other form on the build row.
      D.nkp = 0;
      D.kp_other = 0;
      if (D.PB.n > 0 && (!D.PB.conj || D.PB.n > kBFastPreds)) D.kp_other = 1;
      for (int p = 0; p < D.PB.n && !D.kp_other; p++) {
        if (D.PB.data[p] == D.bkey.data && D.PB.type[p] == D.bkey.type && D.PB.type[p] != KG_REAL) {
          D.kp_op[D.nkp] = D.PB.op[p];
          D.kp_lit[D.nkp] = D.PB.ilit[p];
          D.nkp++;
        } else {
          D.kp_other = 1;
        }
      }
    }
    if (lead >= 0) key_interval_union(Q[lead], H->bmem + blo, nbmem - blo, Q);
  }
  H->nfills = nf;
  // Build phase entries; tables (queries sharing table_id and n) for the scan
  // phase, in order of first appearance.
  std::vector<int32_t> live;
  for (int i : ok)
    if (b->plans[i].status == KG_OK) live.push_back(i);
  H->nq = nq;
  int nbq = 0;
  H->bpre[0] = 0;
  for (int i : live) {
    if (!b->plans[i].join || b->plans[i].bn == 0 || !b->plans[i].build_entry) continue;
    H->bq[nbq] = i;
    H->bpre[nbq + 1] = H->bpre[nbq] + static_cast<int32_t>(ceil_div(b->plans[i].bn, kBBuildTile));
    nbq++;
  }
  H->nbq = nbq;
  int nrq = 0;
  H->rpre[0] = 0;
  for (int j = 0; j < nbq; j++) {
    if (!Q[H->bq[j]].map) continue;
    H->rq[nrq] = H->bq[j];
    H->rpre[nrq + 1] = H->rpre[nrq] + (H->bpre[j + 1] - H->bpre[j]);
    nrq++;
  }
  H->nrq = nrq;
  std::vector<int32_t> fast_ids, join_ids, gen_ids;
  for (int i : live) {
    if (Q[i].reg) fast_ids.push_back(i);
    else if (Q[i].jfast) join_ids.push_back(i);
    else gen_ids.push_back(i);
  }
  H->join.tile = queue_tile(join_ids, qs);
  H->gen.tile = queue_tile(gen_ids, qs);
  if (!layout_groups(H->shared, b->hg(), fast_ids, qs, Q) || !layout_queues(H->join, join_ids, qs) ||
      !layout_queues(H->gen, gen_ids, qs))
    return kg_fail(ctx, KG_ERR_INVALID, "batch too large");
  for (int i : join_ids) Q[i].items = static_cast<int32_t>(ceil_div(qs[i].n, H->join.tile));
  for (int i : gen_ids) Q[i].items = static_cast<int32_t>(ceil_div(qs[i].n, H->gen.tile));
  for (int i : live) H->fills[rem_fill[i]].v = static_cast<uint64_t>(Q[i].items);
  // Queries that failed planning keep a zero descriptor: G = 0, nothing to fold.
  for (int i = 0; i < nq; i++)
    if (b->plans[i].status != KG_OK) {
      std::memset(&Q[i], 0, sizeof(BQuery));
    }
  const size_t bytes = round256(sizeof(BHeader)) + sizeof(BQuery) * static_cast<size_t>(nq);
  // Columns written on the context's stream before this call are visible to the batch.
  if (b->after_ctx) {
    KG_HIP_CHECK(ctx, hipEventRecord(b->ready, ctx->stream));
    KG_HIP_CHECK(ctx, hipStreamWaitEvent(b->stream, b->ready, 0));
  }
  KG_HIP_CHECK(ctx, hipMemcpyAsync(b->d_desc, b->h_desc, bytes, hipMemcpyHostToDevice, b->stream));
  if (H->shared.ng)
    KG_HIP_CHECK(ctx, hipMemcpyAsync(b->dg(), b->hg(), sizeof(SGroup) * H->shared.ng, hipMemcpyHostToDevice,
                                     b->stream));
  KG_HIP_CHECK(ctx, hipGraphLaunch(b->exec, b->stream));
  KG_HIP_CHECK(ctx, hipEventRecord(b->done, b->stream));
  b->nq = nq;
  b->pending = true;
  return KG_OK;
}

extern "C" int kg_batch_poll(kg_batch* b, int32_t* done) {
  if (!b || !done) return KG_ERR_INVALID;
  if (!b->pending) {
    *done = 1;
    return KG_OK;
  }
  const hipError_t e = hipEventQuery(b->done);
  if (e == hipSuccess) {
    *done = 1;
    return KG_OK;
  }
  if (e == hipErrorNotReady) {
    *done = 0;
    return KG_OK;
  }
  return kg_fail(b->ctx, KG_ERR_HIP, "batch event query failed: %s", hipGetErrorString(e));
}

namespace {

// Writes query i's result from its (completed) result area.
void emit_query(kg_batch* b, int i, kg_query_result& R) {
  kg_ctx* ctx = b->ctx;
  BPlan& P = b->plans[i];
  R.ngroups = 0;
  R.status = P.status;
  if (P.status != KG_OK) return;
  const char* base = b->h_res + P.res_off;
  const uint32_t flag = *reinterpret_cast<const volatile uint32_t*>(base);
  if ((flag & kFlagEmptyKey) || ((flag & kFlagNoMap) && P.map_only)) {
    R.status = KG_ERR_UNSUPPORTED;
    return;
  }
  if (flag & kFlagBadCode) {
    R.status = KG_ERR_INVALID;
    return;
  }
  if (flag & kFlagOverflow) {
    R.status = KG_ERR_CAPACITY;
    return;
  }
  const int G = P.G, nm = P.M.nm, naggs = static_cast<int>(P.plan.size());
  const AggFinal* f = reinterpret_cast<const AggFinal*>(base + P.L.fin);
  std::vector<AggFinal> fin(f, f + static_cast<size_t>(G) * std::max(nm, 1));
  const uint64_t* r = reinterpret_cast<const uint64_t*>(base + P.L.rows);
  std::vector<uint64_t> rowsv(r, r + G);
  const uint64_t* fr = reinterpret_cast<const uint64_t*>(base + P.L.first);
  std::vector<int64_t> firstv(G);
  for (int g = 0; g < G; g++) firstv[g] = static_cast<int64_t>(fr[g]);
  std::vector<std::vector<int8_t>> first_null(naggs);
  const int8_t* fn = reinterpret_cast<const int8_t*>(base + P.L.fnull);
  for (int a = 0; a < naggs; a++)
    if (P.plan[a].kind == KG_AGG_COUNT && P.plan[a].count_nulls)
      first_null[a].assign(fn + static_cast<size_t>(a) * G, fn + static_cast<size_t>(a + 1) * G);
  if ((naggs > 0 && (!R.vals || !R.nulls)) || (P.nkeys_out > 0 && !R.keys)) {
    R.status = KG_ERR_INVALID;
    return;
  }
  const std::string saved = ctx->err;
  R.status = emit_groups(ctx, G, P.nkeys_out, P.card1, P.M, P.plan, fin, rowsv, firstv, first_null, R.max_groups,
                         &R.ngroups, R.keys, R.vals, R.nulls);
  ctx->err = saved;
}

bool query_ready(const kg_batch* b, int i) {
  if (b->plans[i].status != KG_OK) return true;
  return __atomic_load_n(&b->h_done[i], __ATOMIC_ACQUIRE) == b->seq;
}

}  // namespace

extern "C" int kg_batch_query_ready(kg_batch* b, int32_t i, int32_t* ready) {
  if (!b || !ready) return KG_ERR_INVALID;
  KG_REQUIRE(b->ctx, b->pending && i >= 0 && i < b->nq, "no query %d in flight", i);
  *ready = query_ready(b, i) ? 1 : 0;
  return KG_OK;
}

extern "C" int kg_batch_collect_query(kg_batch* b, int32_t i, kg_query_result* res) {
  if (!b || !res) return KG_ERR_INVALID;
  KG_REQUIRE(b->ctx, b->pending && i >= 0 && i < b->nq, "no query %d in flight", i);
  KG_REQUIRE(b->ctx, query_ready(b, i), "query %d has not completed", i);
  emit_query(b, i, *res);
  return KG_OK;
}

extern "C" int kg_batch_release(kg_batch* b) {
  if (!b) return KG_ERR_INVALID;
  kg_ctx* ctx = b->ctx;
  KG_REQUIRE(ctx, b->pending, "no batch in flight");
  b->pending = false;
  KG_HIP_CHECK(ctx, hipEventSynchronize(b->done));
  return KG_OK;
}

extern "C" int kg_batch_collect(kg_batch* b, kg_query_result* res) {
  if (!b) return KG_ERR_INVALID;
  kg_ctx* ctx = b->ctx;
  KG_REQUIRE(ctx, b->pending, "no batch in flight");
  KG_REQUIRE(ctx, b->nq == 0 || res, "results are NULL");
  b->pending = false;
  KG_HIP_CHECK(ctx, hipEventSynchronize(b->done));
  for (int i = 0; i < b->nq; i++) {
    if (!query_ready(b, i)) {
      res[i].ngroups = 0;
      res[i].status = kg_fail(ctx, KG_ERR_HIP, "query %d was not published by the device", i);
      continue;
    }
    emit_query(b, i, res[i]);
  }
  return KG_OK;
}

extern "C" int kg_batch_run(kg_batch* b, const kg_query* qs, int32_t nq, kg_query_result* res) {
  int st = kg_batch_submit(b, qs, nq);
  if (st != KG_OK) return st;
  return kg_batch_collect(b, res);
}
