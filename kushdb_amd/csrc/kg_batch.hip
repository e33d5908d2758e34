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
// A batch runs four kernels: fills (accumulators, table slots), the build
// phase (work items = (query, build tile)), the scan/probe phase, and a
// finalize that folds each query's accumulator copies into its result area,
// which is copied back in one transfer.
//
// Scan/probe work items are (query, tile of kBTile rows). Queries over the same
// table share its tiles: the items of tile k of a table are adjacent and sit in
// queue k mod 8, one queue per XCD (workgroup b starts on queue b mod 8, the
// XCD the dispatcher places it on), so the queries reading tile k run back to
// back on one XCD and all but the first read it from that XCD's L2. A
// workgroup drains its own queue, then the others, and exits when all eight are
// empty; there is no inter-workgroup waiting, so any grid size is correct.
//
// Per item, a query with at most 8 groups and non-NULL measures accumulates in
// lane-private registers (the register path), reduced across the wave by a
// butterfly and across the workgroup in LDS; other queries use a
// workgroup-private LDS table updated with LDS atomics (up to kBLdsWords 8-byte
// words). Item partials go to one of eight accumulator copies per query with
// global atomics; the finalize adds the eight copies. FP64 sums are therefore
// added in a timing-dependent order: a query's result in a mix equals its
// result alone within FP64 rounding (integers, counts and group order exactly).
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "kg_table.h"

namespace kg {

constexpr int kBTile = 4096;          // rows per scan/probe work item
constexpr int kBBuildTile = 512;      // rows per build work item
constexpr int kBThreads = 256;        // threads per workgroup
constexpr int kBWaves = kBThreads / kWave;
constexpr int kBQueues = 8;           // scan/probe work queues (one per XCD)
constexpr int kBSlots = 64;           // accumulator copies per query (8 per XCD)
constexpr int kBMaxQueries = 256;     // queries per batch
constexpr int kBRegG = 8;             // register path: at most this many groups
constexpr int kBLdsWords = 2048;      // LDS path: 8-byte words of the item table
constexpr int kBFastPreds = 4;
constexpr int kBMaxAggs = 16;
constexpr int kBMaxFills = 4 * kBMaxQueries + 8;

enum BFlag : uint32_t { kFlagEmptyKey = 1u, kFlagBadCode = 2u, kFlagOverflow = 4u };

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
struct BQueues {
  int32_t nt;
  int32_t qpre[kBQueues][kBMaxQueries + 1];   // per queue, item prefix over tables
  int32_t toff[kBMaxQueries + 1];             // table t's queries: tq[toff[t] .. toff[t+1])
  int32_t tq[kBMaxQueries];
  int64_t tn[kBMaxQueries];                   // rows of table t
  unsigned int heads[kBQueues];
};

// Batch header, copied to the device with the query descriptors.
struct BHeader {
  int32_t nq, nfills, nbq, pad;
  int32_t bpre[kBMaxQueries + 1];             // build phase: item prefix over bq entries
  int32_t bq[kBMaxQueries];                   // build phase: query of entry j
  unsigned int bhead;
  BQueues fast;                               // register-path scans
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

// Takes the next item (of `total`) of a queue for the whole workgroup.
__device__ __forceinline__ int32_t take_item(unsigned int* head, int32_t total, int32_t* s_it) {
  if (threadIdx.x == 0) {
    // An exhausted queue is detected without an atomic (heads only grow).
    uint32_t v = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (static_cast<int32_t>(v) < total) v = atomicAdd(head, 1u);
    *s_it = static_cast<int32_t>(v);
  }
  __syncthreads();
  // Uniform across the workgroup: keeps descriptor reads on the scalar path.
  const int32_t it = __builtin_amdgcn_readfirstlane(*s_it);
  __syncthreads();
  return it;
}

// Build phase: items (query, build tile); same insertion as k_ht_build.
__global__ __launch_bounds__(kBThreads) void k_batch_build(BHeader* h, const BQuery* qs) {
  __shared__ int32_t s_it;
  const int32_t total = h->bpre[h->nbq];
  while (true) {
    const int32_t it = take_item(&h->bhead, total, &s_it);
    if (it >= total) break;
    const int j = find_prefix(h->bpre, h->nbq, it);
    const BQuery& d = qs[h->bq[j]];
    const int64_t lo = static_cast<int64_t>(it - h->bpre[j]) * kBBuildTile;
    const int64_t hi = lo + kBBuildTile < d.bn ? lo + kBBuildTile : d.bn;
    const TableView t = d.t;
    for (int64_t r = lo + threadIdx.x; r < hi; r += kBThreads) {
      if (!eval_preds(d.PB, r)) continue;
      int64_t k;
      if (!load_key(d.bkey, r, k)) continue;  // NULL build keys are skipped
      if (k == kEmpty) {
        atomicOr(d.flag, kFlagEmptyKey);
        continue;
      }
      uint64_t b = bucket_of(k, t.nbuckets);
      bool done = false;
      for (uint64_t probes = 0; probes < t.nbuckets && !done; probes++) {
        for (int s = 0; s < kSlots && !done; s++) {
          if (t.keys[b * kSlots + s] != kEmpty) continue;
          const unsigned long long prev =
              atomicCAS(reinterpret_cast<unsigned long long*>(t.keys + b * kSlots + s),
                        static_cast<unsigned long long>(kEmpty), static_cast<unsigned long long>(k));
          if (prev == static_cast<unsigned long long>(kEmpty)) {
            t.rows[b * kSlots + s] = static_cast<int32_t>(r);
            done = true;
          }
        }
        b = (b + 1 == t.nbuckets) ? 0 : b + 1;
      }
      if (!done) atomicOr(d.flag, kFlagOverflow);
    }
  }
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

// Register path (SCAN_AGG with a BFast view, at most kBRegG groups), one wave
// per item: lane l owns the row groups lo + 4 l + 256 k (k < 16) of 4
// consecutive rows, read with 16-byte vector loads, four row groups' loads in
// flight at once. A first pass turns predicates and group codes into a 64-bit
// row mask and packed 4-bit codes and counts rows and first rows per group; the
// measures are then summed two at a time into lane-private registers, so each
// column is read once. Partials are reduced across the wave by a butterfly and
// added to accumulator copy `slot` by one atomic instruction per value kind.
// Waves work independently (no LDS, no barriers).
constexpr int kWRowGroups = kBTile / (kWave * 4);   // row groups per lane
constexpr int kWBatch = 2;                          // row groups loaded together
constexpr int kWBatches = kWRowGroups / kWBatch;
constexpr int kWBatchRows = 4 * kWBatch;            // rows per lane per batch
static_assert(kWBatchRows * 4 * kWBatches == 256, "4-bit codes of all rows fill codes[4]");

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

// Packed 4-bit codes of all rows of a lane: batch kb's kWBatchRows codes are
// bits [32 kb, 32 kb + 32) of the 256-bit value (w0, w1, w2, w3).
struct Codes {
  uint64_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
  __device__ __forceinline__ uint32_t get(int kb) const {
    const int word = kb >> 1;
    const uint64_t x = word == 0 ? w0 : (word == 1 ? w1 : (word == 2 ? w2 : w3));
    return static_cast<uint32_t>(x >> (32 * (kb & 1)));
  }
  __device__ __forceinline__ void put(int kb, uint32_t c) {
    const uint64_t part = static_cast<uint64_t>(c) << (32 * (kb & 1));
    const int word = kb >> 1;
    w0 |= word == 0 ? part : 0;
    w1 |= word == 1 ? part : 0;
    w2 |= word == 2 ? part : 0;
    w3 |= word == 3 ? part : 0;
  }
};
static_assert(kWBatchRows * 4 == 32, "one 32-bit code word per batch");

// Sums measures m0 (kind K0) and m0 + 1 (kind K1, 0 = absent) over the rows in
// `mask` and adds the wave's totals to accumulator copy `slot`.
template <bool FULL, int K0, int K1>
__device__ __forceinline__ void wave_chunk(const BQuery& d, int64_t lo, int64_t hi, int m0, uint64_t mask,
                           const Codes& codes, int slot) {
  const BFast& F = d.F;
  const int G = d.G, nm = d.nm;
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t acc[2][kBRegG];
#pragma unroll
  for (int gg = 0; gg < kBRegG; gg++) acc[0][gg] = acc[1][gg] = 0;
  // Two load buffers: batch kb + 1 is in flight while batch kb is summed.
  static_assert(kWBatches % 2 == 0, "batches are processed in pairs");
  uint64_t va[kWBatch][2][4], vb[kWBatch][2][4];
  auto load = [&](int kb, uint64_t (&dst)[kWBatch][2][4]) {
#pragma unroll
    for (int kk = 0; kk < kWBatch; kk++) {
      const int64_t base = lo + 4 * lane + static_cast<int64_t>(kb * kWBatch + kk) * kWave * 4;
      const int nv = FULL ? 4 : rows_left(hi, base);
      ld4_measure<FULL, K0>(F.mcol[m0], base, nv, dst[kk][0]);
      if (K1) ld4_measure<FULL, K1>(F.mcol[m0 + 1], base, nv, dst[kk][1]);
    }
  };
  auto sum = [&](int kb, const uint64_t (&cur)[kWBatch][2][4]) {
    const uint32_t mw = static_cast<uint32_t>(mask >> (kWBatchRows * kb));
    const uint32_t cw = codes.get(kb);
#pragma unroll
    for (int kk = 0; kk < kWBatch; kk++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int bit = 4 * kk + e;
        if (!((mw >> bit) & 1u)) continue;
        const int g = static_cast<int>((cw >> (4 * bit)) & 15u);
        // Branch-free select-add (a dynamically indexed update would move acc
        // to scratch). Adding +0.0 or 0 leaves an accumulator unchanged: they
        // start at +0.0 / 0 and never become -0.0.
#pragma unroll
        for (int gg = 0; gg < kBRegG; gg++) {
          if (gg >= G) break;
          acc[0][gg] = kadd<K0>(acc[0][gg], gg == g ? cur[kk][0][e] : 0);
          if (K1) acc[1][gg] = kadd<K1>(acc[1][gg], gg == g ? cur[kk][1][e] : 0);
        }
      }
  };
  load(0, va);
#pragma unroll 1
  for (int kb = 0; kb < kWBatches; kb += 2) {
    load(kb + 1, vb);
    sum(kb, va);
    if (kb + 2 < kWBatches) load(kb + 2, va);
    sum(kb + 1, vb);
  }
  uint64_t mine0 = 0, mine1 = 0;
#pragma unroll
  for (int gg = 0; gg < kBRegG; gg++) {
    if (gg >= G) break;
    const uint64_t a0 = wave_sum<K0>(acc[0][gg]);
    if (lane == gg) mine0 = a0;
    if (K1) {
      const uint64_t a1 = wave_sum<K1>(acc[1][gg]);
      if (lane == gg) mine1 = a1;
    }
  }
  if (lane < G) {
    const int64_t o = (static_cast<int64_t>(slot) * G + lane) * nm + m0;
    if (mine0) atomic_acc<K0>(&d.sum[o], mine0);
    if (K1 && mine1) atomic_acc<K1>(&d.sum[o + 1], mine1);
  }
}

template <bool FULL>
__device__ __forceinline__ void item_fast(const BQuery& d, int64_t lo, int64_t hi, int slot) {
  const BFast& F = d.F;
  const int G = d.G, nm = d.nm;
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t mask = 0;
  Codes codes;
  uint32_t cnt[kBRegG];
  uint64_t first[kBRegG];
#pragma unroll
  for (int g = 0; g < kBRegG; g++) {
    cnt[g] = 0;
    first[g] = ~0ull;
  }
  bool bad = false;
#pragma unroll 1
  for (int kb = 0; kb < kWBatches; kb++) {
    uint32_t mw = 0;
    uint32_t cw = 0;
#pragma unroll
    for (int kk = 0; kk < kWBatch; kk++) {
      const int64_t base = lo + 4 * lane + static_cast<int64_t>(kb * kWBatch + kk) * kWave * 4;
      const int nv = FULL ? 4 : rows_left(hi, base);
      bool pass[4];
#pragma unroll
      for (int e = 0; e < 4; e++) pass[e] = e < nv;
#pragma unroll
      for (int p = 0; p < kBFastPreds; p++) {
        if (p >= F.npred) break;
        int32_t x[4];
        ld4_i32<FULL>(F.pcol[p], base, nv, x);
#pragma unroll
        for (int e = 0; e < 4; e++) pass[e] = pass[e] && cmp_i64(F.op[p], x[e], F.lit[p]);
      }
      int32_t g[4] = {0, 0, 0, 0};
      if (F.nk >= 1) {
        int32_t a[4];
        ld4_i32<FULL>(F.key[0], base, nv, a);
        if (F.nk == 2) {
          int32_t b[4];
          ld4_i32<FULL>(F.key[1], base, nv, b);
#pragma unroll
          for (int e = 0; e < 4; e++) g[e] = a[e] * F.card1 + b[e];
        } else {
#pragma unroll
          for (int e = 0; e < 4; e++) g[e] = a[e];
        }
      }
#pragma unroll
      for (int e = 0; e < 4; e++) {
        if (!pass[e]) continue;
        if (static_cast<unsigned>(g[e]) >= static_cast<unsigned>(G)) {
          bad = true;
          continue;
        }
        const int bit = 4 * kk + e;
        mw |= 1u << bit;
        cw |= static_cast<uint32_t>(g[e]) << (4 * bit);
        const uint64_t row = static_cast<uint64_t>(base + e);
#pragma unroll
        for (int gg = 0; gg < kBRegG; gg++) {
          if (gg >= G) break;
          cnt[gg] += gg == g[e] ? 1u : 0u;
          first[gg] = (gg == g[e] && row < first[gg]) ? row : first[gg];
        }
      }
    }
    mask |= static_cast<uint64_t>(mw) << (kWBatchRows * kb);
    codes.put(kb, cw);
  }
  if (bad) atomicOr(d.flag, kFlagBadCode);
  // Rows, first rows and each measure's non-NULL count (= rows: the measures
  // are non-NULL), one lane per group.
  uint64_t myc = 0, myf = ~0ull;
#pragma unroll
  for (int gg = 0; gg < kBRegG; gg++) {
    if (gg >= G) break;
    uint32_t c = cnt[gg];
    uint64_t f = first[gg];
    for (int s = kWave / 2; s >= 1; s >>= 1) {
      c += __shfl_xor(c, s);
      const uint64_t o = shfl_xor_u64(f, s);
      f = o < f ? o : f;
    }
    if (lane == gg) {
      myc = c;
      myf = f;
    }
  }
  if (lane < G && myc) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&d.rows[static_cast<int64_t>(slot) * G + lane]), myc);
    atomicMin(reinterpret_cast<unsigned long long*>(&d.first[static_cast<int64_t>(slot) * G + lane]), myf);
    for (int m = 0; m < nm; m++)
      atomicAdd(reinterpret_cast<unsigned long long*>(&d.nn[(static_cast<int64_t>(slot) * G + lane) * nm + m]), myc);
  }
  if (__builtin_amdgcn_readfirstlane(__ballot(mask != 0) != 0 ? 1 : 0) == 0) return;  // no qualifying rows
  for (int m0 = 0; m0 < nm; m0 += 2) {
    const int k0 = F.real[m0] ? 1 : 2;
    const int k1 = m0 + 1 < nm ? (F.real[m0 + 1] ? 1 : 2) : 0;
    switch (k0 * 3 + k1) {
      case 3: wave_chunk<FULL, 1, 0>(d, lo, hi, m0, mask, codes, slot); break;
      case 4: wave_chunk<FULL, 1, 1>(d, lo, hi, m0, mask, codes, slot); break;
      case 5: wave_chunk<FULL, 1, 2>(d, lo, hi, m0, mask, codes, slot); break;
      case 6: wave_chunk<FULL, 2, 0>(d, lo, hi, m0, mask, codes, slot); break;
      case 7: wave_chunk<FULL, 2, 1>(d, lo, hi, m0, mask, codes, slot); break;
      default: wave_chunk<FULL, 2, 2>(d, lo, hi, m0, mask, codes, slot); break;
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
// lo + 4 l + 256 k (k < 16) of 4 consecutive rows. Per row group, the
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

template <bool FULL>
__device__ __forceinline__ void item_join_fast(const BQuery& d, int64_t lo, int64_t hi, int slot, uint64_t* lds) {
  const BJoinFast& J = d.J;
  const int G = d.G, nm = d.nm;
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t* lsum = lds;
  uint64_t* lrows = lsum + G * nm;
  uint64_t* lfirst = lrows + G;
  const int words = G * nm + G * 2;
  for (int i = lane; i < words; i += kWave) lds[i] = i < G * nm + G ? 0ull : ~0ull;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const TableView& t = d.t;
  // Product measures are REAL; others keep their kind.
  const bool real0 = J.mkind[0] == 1 || J.fkind[0] != 0;
  const bool real1 = J.mkind[1] == 1 || J.fkind[1] != 0;
  uint64_t seen = 0;  // groups (< 64) whose first row this lane has reported
#pragma unroll 1
  for (int k = 0; k < kWRowGroups; k++) {
    const int64_t base = lo + 4 * lane + static_cast<int64_t>(k) * kWave * 4;
    const int nv = FULL ? 4 : rows_left(hi, base);
    if (!FULL && __builtin_amdgcn_readfirstlane(__ballot(nv > 0) != 0 ? 1 : 0) == 0) break;
    bool pass[4];
#pragma unroll
    for (int e = 0; e < 4; e++) pass[e] = e < nv;
#pragma unroll
    for (int p = 0; p < kBFastPreds; p++) {
      if (p >= J.npred) break;
      int32_t x[4];
      ld4_i32<FULL>(J.pcol[p], base, nv, x);
#pragma unroll
      for (int e = 0; e < 4; e++) pass[e] = pass[e] && cmp_i64(J.op[p], x[e], J.lit[p]);
    }
    int64_t key[4];
    if (J.key64) {
      ld4_i64<FULL>(static_cast<const int64_t*>(J.key), base, nv, key);
    } else {
      int32_t k32[4];
      ld4_i32<FULL>(static_cast<const int32_t*>(J.key), base, nv, k32);
#pragma unroll
      for (int e = 0; e < 4; e++) key[e] = k32[e];
    }
    uint64_t v0[4] = {0, 0, 0, 0}, v1[4] = {0, 0, 0, 0};
    if (nm > 0) ld4_jmeasure<FULL>(J.mkind[0], J.fkind[0], J.mcol[0], J.fcol[0], base, nv, v0);
    if (nm > 1) ld4_jmeasure<FULL>(J.mkind[1], J.fkind[1], J.mcol[1], J.fcol[1], base, nv, v1);
    // First buckets of the passing rows, loaded together.
    uint64_t b[4];
    longlong2 kv[4][kSlots / 2];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      pass[e] = pass[e] && key[e] != kEmpty;
      b[e] = pass[e] ? bucket_of(key[e], t.nbuckets) : 0;
#pragma unroll
      for (int q = 0; q < kSlots / 2; q++)
        kv[e][q] = pass[e] ? reinterpret_cast<const longlong2*>(t.keys + b[e] * kSlots)[q]
                           : make_longlong2(kEmpty, kEmpty);
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      if (!pass[e]) continue;
      const int64_t row = base + e;
      auto consume = [&](int32_t brow) {
        const int g = d.codes ? d.codes[brow] : 0;
        if (static_cast<unsigned>(g) >= static_cast<unsigned>(G)) {
          atomicOr(d.flag, kFlagBadCode);
          return;
        }
        atomicAdd(reinterpret_cast<unsigned long long*>(&lrows[g]), 1ull);
        if (g >= 64 || !((seen >> g) & 1ull)) {
          atomicMin(reinterpret_cast<unsigned long long*>(&lfirst[g]),
                    (static_cast<unsigned long long>(row) << 32) | static_cast<uint32_t>(brow));
          if (g < 64) seen |= 1ull << g;
        }
        if (nm > 0) {
          if (real0) atomicAdd(reinterpret_cast<double*>(&lsum[g * nm]), __longlong_as_double(v0[e]));
          else atomicAdd(reinterpret_cast<unsigned long long*>(&lsum[g * nm]), v0[e]);
        }
        if (nm > 1) {
          if (real1) atomicAdd(reinterpret_cast<double*>(&lsum[g * nm + 1]), __longlong_as_double(v1[e]));
          else atomicAdd(reinterpret_cast<unsigned long long*>(&lsum[g * nm + 1]), v1[e]);
        }
      };
      uint32_t hit = 0;
      bool empty = false;
#pragma unroll
      for (int q = 0; q < kSlots / 2; q++) {
        hit |= (kv[e][q].x == key[e] ? 1u : 0u) << (2 * q);
        hit |= (kv[e][q].y == key[e] ? 1u : 0u) << (2 * q + 1);
        empty |= kv[e][q].x == kEmpty || kv[e][q].y == kEmpty;
      }
      while (hit) {
        const int sidx = __builtin_ctz(hit);
        hit &= hit - 1;
        consume(t.rows[b[e] * kSlots + sidx]);
      }
      if (!empty) {
        uint64_t bb = (b[e] + 1 == t.nbuckets) ? 0 : b[e] + 1;
        for (uint64_t scanned = 1; scanned < t.nbuckets; scanned++) {
          const longlong2* bk = reinterpret_cast<const longlong2*>(t.keys + bb * kSlots);
          bool emp = false;
#pragma unroll
          for (int q = 0; q < kSlots / 2; q++) {
            const longlong2 x = bk[q];
            if (x.x == key[e]) consume(t.rows[bb * kSlots + 2 * q]);
            if (x.y == key[e]) consume(t.rows[bb * kSlots + 2 * q + 1]);
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
__device__ __forceinline__ uint64_t coherent_load(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ void fold_query(const BQuery& d) {
  const int lane = threadIdx.x & (kWave - 1);
  const int G = d.G, nm = d.nm;
  const ResLayout L = res_layout(G, nm, d.naggs);
  AggFinal* fin = reinterpret_cast<AggFinal*>(d.res + L.fin);
  uint64_t* rows = reinterpret_cast<uint64_t*>(d.res + L.rows);
  uint64_t* first = reinterpret_cast<uint64_t*>(d.res + L.first);
  int8_t* fnull = reinterpret_cast<int8_t*>(d.res + L.fnull);
  if (lane == 0) *reinterpret_cast<uint32_t*>(d.res) = __hip_atomic_load(d.flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// Wave-level form of drain(): each wave takes items on its own.
template <typename Run>
__device__ __forceinline__ void drain_wave(BQueues& Q, const BQuery* qs, Run&& run) {
  const int nt = Q.nt;
  const int lane = threadIdx.x & (kWave - 1);
  int x = blockIdx.x % kBQueues;
  for (int exhausted = 0; exhausted < kBQueues;) {
    const int32_t total = Q.qpre[x][nt];
    // Heads only grow: a queue seen at or past its end stays empty, and is
    // skipped without an atomic (all waves finish by probing every queue).
    uint32_t v = 0;
    if (lane == 0) {
      v = __hip_atomic_load(&Q.heads[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (static_cast<int32_t>(v) < total) v = atomicAdd(&Q.heads[x], 1u);
    }
    const int32_t it = static_cast<int32_t>(__builtin_amdgcn_readfirstlane(v));
    if (it >= total) {
      x = (x + 1) % kBQueues;
      exhausted++;
      continue;
    }
    const int t = find_prefix(Q.qpre[x], nt, it);
    const int32_t local = it - Q.qpre[x][t];
    const int32_t m = Q.toff[t + 1] - Q.toff[t];
    const int32_t tile = x + kBQueues * (local / m);
    const BQuery& d = qs[Q.tq[Q.toff[t] + local % m]];
    const int64_t lo = static_cast<int64_t>(tile) * kBTile;
    const int64_t hi = lo + kBTile < d.n ? lo + kBTile : d.n;
    run(d, lo, hi);
    item_done(d);
  }
}

// Scan/probe phase, register-path scans (one wave per item).
__global__ __launch_bounds__(kBThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_batch_fast(BHeader* h, const BQuery* qs) {
  const int slot = acc_slot();
  drain_wave(h->fast, qs, [&](const BQuery& d, int64_t lo, int64_t hi) {
    if (hi - lo == kBTile) item_fast<true>(d, lo, hi, slot);
    else item_fast<false>(d, lo, hi, slot);
  });
}

// Scan/probe phase, joins with a typed view (one wave per item).
__global__ __launch_bounds__(kBThreads) void k_batch_join(BHeader* h, const BQuery* qs) {
  __shared__ uint64_t lds[kBLdsWords];
  const int slot = acc_slot();
  uint64_t* mine = lds + (threadIdx.x / kWave) * kBWaveWords;
  drain_wave(h->join, qs, [&](const BQuery& d, int64_t lo, int64_t hi) {
    if (hi - lo == kBTile) item_join_fast<true>(d, lo, hi, slot, mine);
    else item_join_fast<false>(d, lo, hi, slot, mine);
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
  int64_t table = -1;  // index of the shared table entry
  bool join = false;
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
size_t query_footprint(const kg_query& q) {
  const int64_t G = std::min<int64_t>(query_groups(q), kMaxLdsGroups);
  const int nm = std::min(std::max(q.naggs, 1), kMaxMeasures);
  size_t b = res_layout(static_cast<int>(G), nm, std::min(q.naggs, kBMaxAggs)).bytes;
  b += round256(static_cast<size_t>(kBSlots) * G * (2 * nm + 1) * 8 + 16);
  b += round256(static_cast<size_t>(kBSlots) * G * 8);
  if (q.kind == KG_Q_JOIN_AGG) {
    const uint64_t nb = table_buckets(std::max<int64_t>(q.build_n, 0), query_lf(q));
    b += round256(nb * kSlots * 8) + round256(nb * kSlots * 4);
  }
  return b;
}

}  // namespace

struct kg_batch {
  kg_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hipEvent_t ready = nullptr;  // the context stream's work before a submit
  int grid = 0;
  int grid_gen = 0;
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
};

namespace {

size_t desc_bytes() { return round256(sizeof(BHeader)) + round256(sizeof(BQuery) * kBMaxQueries); }

int capture(kg_batch* b) {
  kg_ctx* ctx = b->ctx;
  KG_HIP_CHECK(ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(k_batch_clear, dim3(b->grid), dim3(kBThreads), 0, b->stream, b->dhdr());
  hipLaunchKernelGGL(k_batch_build, dim3(b->grid), dim3(kBThreads), 0, b->stream, b->dhdr(), b->dq());
  hipLaunchKernelGGL(k_batch_fast, dim3(b->grid), dim3(kBThreads), 0, b->stream, b->dhdr(), b->dq());
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
  KG_REQUIRE(ctx, q.kind == KG_Q_SCAN_AGG || q.kind == KG_Q_JOIN_AGG, "bad query kind %d", q.kind);
  KG_REQUIRE(ctx, q.n >= 0 && q.n < INT32_MAX, "row count %lld out of range", (long long)q.n);
  KG_REQUIRE(ctx, q.naggs >= 0 && q.naggs <= kBMaxAggs, "at most %d aggregates", kBMaxAggs);
  KG_REQUIRE(ctx, q.ncols > 0 && q.cols, "no columns");
  int st;
  if ((st = build_preds(ctx, q.cols, q.ncols, q.preds, q.npreds, 1, q.n, D.P)) != KG_OK) return st;
  bool minmax = false;
  if ((st = build_measures(ctx, q.cols, q.ncols, q.aggs, q.naggs, q.n, P.M, P.plan, minmax)) != KG_OK) return st;
  if (minmax) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "MIN/MAX are not batched");
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
  D.jfast = (q.kind == KG_Q_JOIN_AGG && join_view(q, D.P, P.M, D.J)) ? 1 : 0;
  if (!D.reg && G * nm * 2 + G * 2 > kBWaveWords)
    return kg_fail(ctx, KG_ERR_UNSUPPORTED, "%lld groups x %d measures exceed the batched LDS table", (long long)G,
                   nm);
  D.G = static_cast<int32_t>(G);
  P.G = D.G;
  for (int a = 0; a < q.naggs; a++) D.cnull[a] = P.plan[a].kind == KG_AGG_COUNT ? P.plan[a].count_nulls : nullptr;
  return KG_OK;
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
  for (int x = 0; x < kBQueues; x++) {
    Q.qpre[x][0] = 0;
    for (int t = 0; t < nt; t++) {
      const int64_t tiles = ceil_div(Q.tn[t], kBTile);
      const int64_t mine = tiles > x ? (tiles - x + kBQueues - 1) / kBQueues : 0;
      const int64_t items = mine * (Q.toff[t + 1] - Q.toff[t]);
      if (Q.qpre[x][t] + items >= INT32_MAX) return false;
      Q.qpre[x][t + 1] = Q.qpre[x][t] + static_cast<int32_t>(items);
    }
    Q.heads[x] = 0;
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
  b->grid_gen = ctx->num_cus * 8;
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
  off = 0;
  int nf = 0;
  for (int i : ok) {
    BPlan& P = b->plans[i];
    BQuery& D = Q[i];
    const int G = P.G, nm = P.M.nm;
    const size_t acc_bytes = round256(static_cast<size_t>(kBSlots) * G * (2 * std::max(nm, 1) + 1) * 8 + 16);
    const size_t first_bytes = round256(static_cast<size_t>(kBSlots) * G * 8);
    size_t need = acc_bytes + first_bytes;
    uint64_t nbuckets = 0;
    if (P.join) {
      nbuckets = table_buckets(P.bn, query_lf(qs[i]));
      need += round256(nbuckets * kSlots * 8) + round256(nbuckets * kSlots * 4);
    }
    if (off + need > b->arena_bytes) {
      P.status = KG_ERR_CAPACITY;
      P.err = "the batch arena cannot hold this query's tables";
      continue;
    }
    char* base = b->arena + off;
    D.sum = reinterpret_cast<uint64_t*>(base);
    D.nn = D.sum + static_cast<size_t>(kBSlots) * G * std::max(nm, 1);
    D.rows = D.nn + static_cast<size_t>(kBSlots) * G * std::max(nm, 1);
    D.flag = reinterpret_cast<uint32_t*>(D.rows + static_cast<size_t>(kBSlots) * G);
    D.rem = reinterpret_cast<unsigned long long*>(D.rows + static_cast<size_t>(kBSlots) * G + 1);
    // Fill ranges must not overlap: the fill kernel writes them in no fixed order.
    H->fills[nf++] = BFill{D.sum, static_cast<int64_t>(kBSlots) * G * (2 * std::max(nm, 1) + 1) + 1, 0};
    D.items = static_cast<int32_t>(ceil_div(qs[i].n, kBTile));
    H->fills[nf++] = BFill{reinterpret_cast<uint64_t*>(D.rem), 1, static_cast<uint64_t>(D.items)};
    D.first = reinterpret_cast<uint64_t*>(base + acc_bytes);
    H->fills[nf++] = BFill{D.first, static_cast<int64_t>(kBSlots) * G, ~0ull};
    if (P.join) {
      char* tb = base + acc_bytes + first_bytes;
      D.t.keys = reinterpret_cast<int64_t*>(tb);
      D.t.rows = reinterpret_cast<int32_t*>(tb + round256(nbuckets * kSlots * 8));
      D.t.nbuckets = nbuckets;
      D.t.side_rows = nullptr;
      D.t.side_count = nullptr;
      H->fills[nf++] = BFill{reinterpret_cast<uint64_t*>(D.t.keys), static_cast<int64_t>(nbuckets * kSlots),
                             static_cast<uint64_t>(kEmpty)};
    }
    D.res = b->d_res + P.res_off;
    D.done = b->d_done + i;
    D.seq = b->seq;
    off += need;
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
    if (!b->plans[i].join || b->plans[i].bn == 0) continue;
    H->bq[nbq] = i;
    H->bpre[nbq + 1] = H->bpre[nbq] + static_cast<int32_t>(ceil_div(b->plans[i].bn, kBBuildTile));
    nbq++;
  }
  H->nbq = nbq;
  std::vector<int32_t> fast_ids, join_ids, gen_ids;
  for (int i : live) {
    if (Q[i].reg) fast_ids.push_back(i);
    else if (Q[i].jfast) join_ids.push_back(i);
    else gen_ids.push_back(i);
  }
  if (!layout_queues(H->fast, fast_ids, qs) || !layout_queues(H->join, join_ids, qs) ||
      !layout_queues(H->gen, gen_ids, qs))
    return kg_fail(ctx, KG_ERR_INVALID, "batch too large");
  // Queries that failed planning keep a zero descriptor: G = 0, nothing to fold.
  for (int i = 0; i < nq; i++)
    if (b->plans[i].status != KG_OK) {
      std::memset(&Q[i], 0, sizeof(BQuery));
    }
  const size_t bytes = round256(sizeof(BHeader)) + sizeof(BQuery) * static_cast<size_t>(nq);
  // Columns written on the context's stream before this call are visible to the batch.
  KG_HIP_CHECK(ctx, hipEventRecord(b->ready, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamWaitEvent(b->stream, b->ready, 0));
  KG_HIP_CHECK(ctx, hipMemcpyAsync(b->d_desc, b->h_desc, bytes, hipMemcpyHostToDevice, b->stream));
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
  if (flag & kFlagEmptyKey) {
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
