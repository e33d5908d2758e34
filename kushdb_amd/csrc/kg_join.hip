// HBM-resident join hash table and high-cardinality group-by.
//
// Reference behaviour: runtime::HashTable (runtime/hash_table.cc:13-67) used by
// HashJoinTranslator (compile/translators/hash_join_translator.cc:34-139), and
// runtime::AggregateHashTable (runtime/aggregate_hash_table.cc:9-135) with the
// generated UpdateOrInsert loop (compile/proxy/aggregate_hash_table.cc:187-243).
//
// Layout (join table):
//   keys[nbuckets * 8]   int64, 64-byte aligned buckets of 8 slots, kEmpty = free
//   rows[nbuckets * 8]   int32 build row id of each slot
//   acc [nbuckets * 8]   {int64 sum, int64 count} per slot, for the fused
//                        probe + group-by on the matched build key
// A key equal to kEmpty cannot live in a slot; such build rows go to a side
// list that probes with that key scan.
// Insert: hash -> bucket (multiply-high of a 64-bit mix), first free slot in
// bucket order claimed with a 64-bit atomicCAS, next bucket when full.
// Probe: four lanes read one 64-byte bucket together (16 bytes each), compare
// two keys per lane, and ballot; a bucket with a free slot ends the scan.
#include <algorithm>
#include <cstring>
#include <new>

#include "kg_table.h"

namespace kg {

__global__ __launch_bounds__(kJBlock) void k_ht_build(TableView t, kg_col key, const int32_t* sel, int64_t n,
                                                       unsigned long long* inserted, int* overflow) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kJBlock;
  unsigned long long local = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kJBlock + threadIdx.x; i < n; i += stride) {
    const int64_t r = sel ? sel[i] : i;
    int64_t k;
    if (!load_key(key, r, k)) continue;  // NULL build keys are skipped
    local++;
    if (k == kEmpty) {
      const unsigned long long pos = atomicAdd(t.side_count, 1ull);
      if (pos < static_cast<unsigned long long>(kSideCap)) t.side_rows[pos] = static_cast<int32_t>(r);
      else atomicOr(overflow, 1);
      continue;
    }
    uint64_t b = bucket_of(k, t.nbuckets);
    bool done = false;
    for (uint64_t probes = 0; probes < t.nbuckets && !done; probes++) {
      const int64_t* bk = t.keys + b * kSlots;
      for (int s = 0; s < kSlots && !done; s++) {
        if (bk[s] != kEmpty) continue;
        const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(t.keys + b * kSlots + s),
                                                  static_cast<unsigned long long>(kEmpty),
                                                  static_cast<unsigned long long>(k));
        if (prev == static_cast<unsigned long long>(kEmpty)) {
          t.rows[b * kSlots + s] = static_cast<int32_t>(r);
          done = true;
        }
      }
      b = (b + 1 == t.nbuckets) ? 0 : b + 1;
    }
    if (!done) atomicOr(overflow, 1);
  }
  // One atomic per wave for the row count.
  for (int s = kWave / 2; s >= 1; s >>= 1) local += __shfl_xor(local, s);
  if ((threadIdx.x & (kWave - 1)) == 0 && local) atomicAdd(inserted, local);
}

enum ProbeMode { kCount = 0, kEmit = 1, kAgg = 2 };

struct ProbeOut {
  unsigned long long* count;   // kCount / kEmit: total pairs
  int32_t* out_probe;          // kEmit
  int32_t* out_build;          // kEmit
  int64_t cap;                 // kEmit
  Acc* acc;                    // kAgg: per-slot accumulators (+1 side slot at the end)
  int32_t* first;              // kAgg: per-slot first probe row (optional)
  const int64_t* vals;         // kAgg: probe-side values
};

// Wave-aggregated append of `want` items per lane; returns this lane's base.
__device__ __forceinline__ unsigned long long wave_append(unsigned long long* counter, unsigned want) {
  const int lane = threadIdx.x & (kWave - 1);
  unsigned incl = want;
  for (int s = 1; s < kWave; s <<= 1) {
    const unsigned o = __shfl_up(incl, s);
    if (lane >= s) incl += o;
  }
  const unsigned total = __shfl(incl, kWave - 1);
  unsigned long long base = 0;
  if (lane == 0 && total) base = atomicAdd(counter, static_cast<unsigned long long>(total));
  base = __shfl(base, 0);
  return base + (incl - want);
}

// One probe (held by a 4-lane group) from its first bucket on. kv0 is this
// lane's 16 bytes of the first bucket, already loaded by the caller so that
// several probes' first loads are in flight together. Must be called by the
// whole wave (ballots).
template <int MODE, bool FIRST>
__device__ __forceinline__ void probe_one(const TableView& t, const ProbeOut& o, int64_t r, int64_t k, bool active,
                                          uint64_t b, longlong2 kv0, int sub, int grp,
                                          unsigned long long& local) {
  unsigned nmatch = 0;
  int64_t first_slot = -1;
  // The side list holds build rows whose key is the free-slot marker.
  if (active && k == kEmpty) {
    active = false;
    if (sub == 0) {
      unsigned long long ns = *t.side_count;
      if (ns > static_cast<unsigned long long>(kSideCap)) ns = kSideCap;  // overflowed build (reported there)
      if (MODE == kCount) local += ns;
      if (MODE == kAgg && ns) {
        nmatch = static_cast<unsigned>(ns);
        first_slot = static_cast<int64_t>(t.nbuckets) * kSlots;
      }
      if (MODE == kEmit) {
        const unsigned long long pos = atomicAdd(o.count, ns);
        for (unsigned long long j = 0; j < ns; j++)
          if (pos + j < static_cast<unsigned long long>(o.cap)) {
            o.out_probe[pos + j] = static_cast<int32_t>(r);
            o.out_build[pos + j] = t.side_rows[j];
          }
      }
    }
  }
  uint64_t scanned = 0;
  bool first_bucket = true;
  while (__any(active)) {
    bool m0 = false, m1 = false, e = false;
    int64_t slot0 = 0;
    if (active) {
      slot0 = static_cast<int64_t>(b) * kSlots + sub * 2;
      const longlong2 kv = first_bucket ? kv0 : *reinterpret_cast<const longlong2*>(t.keys + slot0);
      m0 = kv.x == k;
      m1 = kv.y == k;
      e = kv.x == kEmpty || kv.y == kEmpty;
    }
    first_bucket = false;
    const uint64_t bm0 = __ballot(m0), bm1 = __ballot(m1), be = __ballot(e);
    const unsigned g0 = (bm0 >> (grp * kProbeLanes)) & 0xF;
    const unsigned g1 = (bm1 >> (grp * kProbeLanes)) & 0xF;
    const unsigned ge = (be >> (grp * kProbeLanes)) & 0xF;
    if (MODE == kCount) {
      if (active && sub == 0) local += __popc(g0) + __popc(g1);
    } else if (MODE == kEmit) {
      const unsigned want = (m0 ? 1u : 0u) + (m1 ? 1u : 0u);
      const unsigned long long pos = wave_append(o.count, active ? want : 0u);
      if (m0 && pos < static_cast<unsigned long long>(o.cap)) {
        o.out_probe[pos] = static_cast<int32_t>(r);
        o.out_build[pos] = t.rows[slot0];
      }
      if (m1 && pos + (m0 ? 1 : 0) < static_cast<unsigned long long>(o.cap)) {
        o.out_probe[pos + (m0 ? 1 : 0)] = static_cast<int32_t>(r);
        o.out_build[pos + (m0 ? 1 : 0)] = t.rows[slot0 + 1];
      }
    } else {
      if (active) {
        nmatch += __popc(g0) + __popc(g1);
        if (first_slot < 0 && (g0 | g1)) {
          // Slot order within the bucket: lane sub holds slots 2*sub, 2*sub+1.
          unsigned bits = 0;
          for (int q = 0; q < kProbeLanes; q++) bits |= (((g0 >> q) & 1u) << (2 * q)) | (((g1 >> q) & 1u) << (2 * q + 1));
          first_slot = static_cast<int64_t>(b) * kSlots + (__ffs(bits) - 1);
        }
      }
    }
    if (active) {
      scanned++;
      if (ge || scanned >= t.nbuckets) active = false;
      b = (b + 1 == t.nbuckets) ? 0 : b + 1;
    }
  }
  if (MODE == kAgg && sub == 0 && nmatch) {
    Acc* a = &o.acc[first_slot];
    atomicAdd(&a->sum, static_cast<unsigned long long>(o.vals[r]) * nmatch);
    atomicAdd(&a->count, static_cast<unsigned long long>(nmatch));
    if (FIRST) atomicMin(&o.first[first_slot], static_cast<int32_t>(r));
  }
}

// Probe kernel: every 4-lane group handles kUnroll probes per wave step; the
// first-bucket loads of all of them are issued before any is examined, so a
// wave keeps 16 * kUnroll random 64-byte reads in flight.
constexpr int kUnroll = 4;

template <int MODE, bool FIRST>
__global__ __launch_bounds__(kJBlock) void k_ht_probe(TableView t, kg_col key, const int32_t* sel, int64_t n,
                                                       ProbeOut o) {
  const int lane = threadIdx.x & (kWave - 1);
  const int sub = lane & (kProbeLanes - 1);
  const int grp = lane / kProbeLanes;
  constexpr int kGroups = kWave / kProbeLanes;
  constexpr int kStep = kGroups * kUnroll;
  const int64_t gwave = (static_cast<int64_t>(blockIdx.x) * kJBlock + threadIdx.x) / kWave;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * kJBlock / kWave;
  unsigned long long local = 0;
  for (int64_t base = gwave * kStep; base < n; base += nwaves * kStep) {
    int64_t r[kUnroll], k[kUnroll];
    bool active[kUnroll];
    uint64_t b[kUnroll];
    longlong2 kv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const int64_t i = base + u * kGroups + grp;
      r[u] = -1;
      k[u] = 0;
      active[u] = false;
      if (i < n) {
        r[u] = sel ? sel[i] : i;
        active[u] = load_key_nt(key, r[u], k[u]);
      }
      b[u] = active[u] ? bucket_of(k[u], t.nbuckets) : 0;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      kv[u] = make_longlong2(kEmpty, kEmpty);
      if (active[u] && k[u] != kEmpty)
        kv[u] = *reinterpret_cast<const longlong2*>(t.keys + static_cast<int64_t>(b[u]) * kSlots + sub * 2);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++)
      probe_one<MODE, FIRST>(t, o, r[u], k[u], active[u], b[u], kv[u], sub, grp, local);
  }
  if (MODE == kCount) {
    for (int s = kWave / 2; s >= 1; s >>= 1) local += __shfl_xor(local, s);
    if (lane == 0 && local) atomicAdd(o.count, local);
  }
}

__global__ void k_fill_i64(int64_t* p, int64_t v, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_i32j(int32_t* p, int32_t v, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Compaction of the occupied accumulator slots into (key, sum, count, first).
// Each workgroup owns a contiguous range of slots: it counts its live slots,
// takes one range of output positions with a single atomic, then writes them
// in slot order (wave ballot + popcount prefix, wave offsets through LDS).
__global__ __launch_bounds__(kJBlock) void k_collect(const int64_t* keys, const Acc* acc, const int32_t* first,
                                                      int64_t nslots, int64_t side_key_slot,
                                                      unsigned long long* count, int64_t cap,
                                                      int64_t* out_key, int64_t* out_sum, int64_t* out_cnt,
                                                      int32_t* out_first) {
  const int64_t total = nslots + 1;
  const int64_t chunk = (total + gridDim.x - 1) / gridDim.x;
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * chunk;
  const int64_t hi = lo + chunk < total ? lo + chunk : total;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  constexpr int NW = kJBlock / kWave;
  __shared__ unsigned int wsum[NW];
  __shared__ unsigned long long base;
  unsigned int mine = 0;
  for (int64_t s = lo + threadIdx.x; s < hi; s += kJBlock) mine += acc[s].count > 0 ? 1u : 0u;
  for (int st = kWave / 2; st >= 1; st >>= 1) mine += __shfl_xor(mine, st);
  if (lane == 0) wsum[wave] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < NW; w++) t += wsum[w];
    base = t ? atomicAdd(count, t) : 0;
  }
  __syncthreads();
  unsigned long long pos = base;
  __shared__ unsigned int wc[2][NW];
  int it = 0;
  for (int64_t s0 = lo; s0 < hi; s0 += kJBlock, it++) {
    const int64_t s = s0 + threadIdx.x;
    const bool live = s < hi && acc[s].count > 0;
    const uint64_t m = __ballot(live);
    if (lane == 0) wc[it & 1][wave] = __popcll(m);
    __syncthreads();
    unsigned int before = 0, tot = 0;
    for (int w = 0; w < NW; w++) {
      const unsigned int c = wc[it & 1][w];
      before += w < wave ? c : 0;
      tot += c;
    }
    if (live) {
      const unsigned long long p = pos + before + __popcll(m & ((1ull << lane) - 1ull));
      if (p < static_cast<unsigned long long>(cap)) {
        out_key[p] = s == side_key_slot ? kEmpty : keys[s];
        out_sum[p] = static_cast<int64_t>(acc[s].sum);
        out_cnt[p] = static_cast<int64_t>(acc[s].count);
        if (out_first) out_first[p] = first[s];
      }
    }
    pos += tot;
  }
}

// High-cardinality group-by: insert-or-update one (key, value) per row into an
// open-addressing table of keys + accumulators. A row scans its bucket's eight
// keys; a match updates; a free slot is claimed with atomicCAS (if the CAS
// loses to the same key, that slot is updated).
template <bool FIRST>
__global__ __launch_bounds__(kJBlock) void k_group_by_i64(int64_t* keys, Acc* acc, int32_t* first, uint64_t nbuckets,
                                                           const int64_t* in_keys, const int64_t* in_vals, int64_t n,
                                                           int* overflow) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kJBlock;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kJBlock + threadIdx.x; r < n; r += stride) {
    const int64_t k = in_keys[r];
    const unsigned long long v = static_cast<unsigned long long>(in_vals[r]);
    int64_t slot = -1;
    if (k == kEmpty) {
      slot = static_cast<int64_t>(nbuckets) * kSlots;  // side slot
    } else {
      uint64_t b = bucket_of(k, nbuckets);
      for (uint64_t probes = 0; probes < nbuckets && slot < 0; probes++) {
        int64_t* bk = keys + b * kSlots;
        int s = 0;
        while (s < kSlots) {
          const int64_t cur = bk[s];
          if (cur == k) {
            slot = static_cast<int64_t>(b) * kSlots + s;
            break;
          }
          if (cur == kEmpty) {
            const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(bk + s),
                                                      static_cast<unsigned long long>(kEmpty),
                                                      static_cast<unsigned long long>(k));
            if (prev == static_cast<unsigned long long>(kEmpty) || static_cast<int64_t>(prev) == k) {
              slot = static_cast<int64_t>(b) * kSlots + s;
              break;
            }
            continue;  // lost to another key: re-read this slot (now occupied)
          }
          s++;
        }
        b = (b + 1 == nbuckets) ? 0 : b + 1;
      }
    }
    if (slot < 0) {
      atomicOr(overflow, 1);
      continue;
    }
    atomicAdd(&acc[slot].sum, v);
    atomicAdd(&acc[slot].count, 1ull);
    if (FIRST) atomicMin(&first[slot], static_cast<int32_t>(r));
  }
}

// Star pipeline in one kernel: ScanSelect on the probe (fact) side, probe of
// the join table, group code of the matched build row, aggregation into a
// workgroup-private LDS table, merged into a global table with atomics.
// Reference chain: ScanSelectTranslator -> HashJoinTranslator::Consume(right)
// (hash_join_translator.cc:105-139) -> GroupByAggregateTranslator::Consume
// (group_by_aggregate_translator.cc:102-111). `first` records the first join
// output row of each group as (probe row << 32 | build row), the order in
// which the reference's aggregate table inserts groups.
template <bool MINMAX>
__global__ __launch_bounds__(kJBlock) void k_join_group_agg(PredSet P, kg_col key, TableView t,
                                                             const int32_t* codes, int G, MeasureSet M, int64_t n,
                                                             AggFinal* gfin, unsigned long long* grows,
                                                             unsigned long long* gfirst) {
  extern __shared__ unsigned long long lds[];
  const int nm = M.nm;
  unsigned long long* lsum = lds;
  unsigned long long* lnn = lsum + G * nm;
  unsigned long long* lrows = lnn + G * nm;
  unsigned long long* lfirst = lrows + G;
  unsigned long long* lmn = lfirst + G;
  unsigned long long* lmx = lmn + (MINMAX ? G * nm : 0);
  const int words = G * nm * 2 + G * 2 + (MINMAX ? 2 * G * nm : 0);
  for (int i = threadIdx.x; i < words; i += kJBlock) lds[i] = 0;
  for (int i = threadIdx.x; i < G; i += kJBlock) lfirst[i] = ~0ull;
  if (MINMAX)
    for (int i = threadIdx.x; i < G * nm; i += kJBlock) {
      const bool real = M.is_real[i % nm];
      lmn[i] = real ? static_cast<unsigned long long>(__double_as_longlong(__builtin_inf()))
                    : static_cast<unsigned long long>(INT64_MAX);
      lmx[i] = real ? static_cast<unsigned long long>(__double_as_longlong(-__builtin_inf()))
                    : static_cast<unsigned long long>(INT64_MIN);
    }
  __syncthreads();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kJBlock;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kJBlock + threadIdx.x; r < n; r += stride) {
    if (!eval_preds(P, r)) continue;
    int64_t k;
    if (!load_key(key, r, k)) continue;
    uint64_t v[kMaxMeasures];
    bool vnull[kMaxMeasures];
#pragma unroll
    for (int m = 0; m < kMaxMeasures; m++)
      if (m < nm) v[m] = load_measure(M, m, r, vnull[m]);
    auto consume = [&](int32_t brow) {
      const int g = codes ? codes[brow] : 0;
      if (g < 0 || g >= G) return;
      atomicAdd(&lrows[g], 1ull);
      atomicMin(&lfirst[g], (static_cast<unsigned long long>(r) << 32) | static_cast<uint32_t>(brow));
#pragma unroll
      for (int m = 0; m < kMaxMeasures; m++) {
        if (m >= nm || vnull[m]) continue;
        atomicAdd(&lnn[g * nm + m], 1ull);
        if (M.is_real[m]) {
          atomicAdd(reinterpret_cast<double*>(&lsum[g * nm + m]), __longlong_as_double(v[m]));
          if (MINMAX) {
            atomicMin(reinterpret_cast<double*>(&lmn[g * nm + m]), __longlong_as_double(v[m]));
            atomicMax(reinterpret_cast<double*>(&lmx[g * nm + m]), __longlong_as_double(v[m]));
          }
        } else {
          atomicAdd(&lsum[g * nm + m], static_cast<unsigned long long>(v[m]));
          if (MINMAX) {
            atomicMin(reinterpret_cast<long long*>(&lmn[g * nm + m]), static_cast<long long>(v[m]));
            atomicMax(reinterpret_cast<long long*>(&lmx[g * nm + m]), static_cast<long long>(v[m]));
          }
        }
      }
    };
    if (k == kEmpty) {
      unsigned long long ns = *t.side_count;
      if (ns > static_cast<unsigned long long>(kSideCap)) ns = kSideCap;
      for (unsigned long long j = 0; j < ns; j++) consume(t.side_rows[j]);
      continue;
    }
    uint64_t b = bucket_of(k, t.nbuckets);
    for (uint64_t scanned = 0; scanned < t.nbuckets; scanned++) {
      const longlong2* bk = reinterpret_cast<const longlong2*>(t.keys + b * kSlots);
      bool empty = false;
#pragma unroll
      for (int q = 0; q < kSlots / 2; q++) {
        const longlong2 kv = bk[q];
        if (kv.x == k) consume(t.rows[b * kSlots + 2 * q]);
        if (kv.y == k) consume(t.rows[b * kSlots + 2 * q + 1]);
        empty |= kv.x == kEmpty || kv.y == kEmpty;
      }
      if (empty) break;
      b = (b + 1 == t.nbuckets) ? 0 : b + 1;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < G; i += kJBlock) {
    if (lrows[i] == 0) continue;
    atomicAdd(&grows[i], lrows[i]);
    atomicMin(&gfirst[i], lfirst[i]);
  }
  for (int i = threadIdx.x; i < G * nm; i += kJBlock) {
    if (lnn[i] == 0) continue;
    const int m = i % nm;
    AggFinal* f = &gfin[i];
    atomicAdd(reinterpret_cast<unsigned long long*>(&f->nn), lnn[i]);
    if (M.is_real[m]) {
      atomicAdd(reinterpret_cast<double*>(&f->sum), __longlong_as_double(lsum[i]));
      if (MINMAX) {
        atomicMin(reinterpret_cast<double*>(&f->mn), __longlong_as_double(lmn[i]));
        atomicMax(reinterpret_cast<double*>(&f->mx), __longlong_as_double(lmx[i]));
      }
    } else {
      atomicAdd(reinterpret_cast<unsigned long long*>(&f->sum), lsum[i]);
      if (MINMAX) {
        atomicMin(reinterpret_cast<long long*>(&f->mn), static_cast<long long>(lmn[i]));
        atomicMax(reinterpret_cast<long long*>(&f->mx), static_cast<long long>(lmx[i]));
      }
    }
  }
}

__global__ void k_init_join_final(AggFinal* fin, int G, MeasureSet M, unsigned long long* rows,
                                  unsigned long long* first) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nm = M.nm;
  if (i < G * nm) {
    const bool real = M.is_real[i % nm];
    fin[i].sum = 0;
    fin[i].nn = 0;
    fin[i].mn = real ? static_cast<uint64_t>(__double_as_longlong(__builtin_inf())) : static_cast<uint64_t>(INT64_MAX);
    fin[i].mx = real ? static_cast<uint64_t>(__double_as_longlong(-__builtin_inf())) : static_cast<uint64_t>(INT64_MIN);
  }
  if (i < G) {
    rows[i] = 0;
    first[i] = ~0ull;
  }
}

// COUNT(col) over the probe side: whether each group's first join row has a
// NULL in `nulls` (indexed by probe row = first >> 32).
__global__ void k_first_null_join(const unsigned long long* first, int G, const int8_t* nulls, int8_t* out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G) out[g] = (first[g] != ~0ull && nulls) ? nulls[first[g] >> 32] : 0;
}

inline int grid_for(kg_ctx* ctx, int64_t work, int per_thread) {
  int64_t g = ceil_div(std::max<int64_t>(work, 1), static_cast<int64_t>(kJBlock) * per_thread);
  return static_cast<int>(std::min<int64_t>(g, static_cast<int64_t>(ctx->num_cus) * 32));
}

}  // namespace kg

using namespace kg;

struct kg_hash_table {
  uint64_t nbuckets = 0;
  int64_t capacity = 0;
  int64_t* keys = nullptr;
  int32_t* rows = nullptr;
  Acc* acc = nullptr;          // lazily allocated for probe_agg
  int32_t* first = nullptr;    // lazily allocated for probe_agg with first rows
  int32_t* side_rows = nullptr;
  unsigned long long* counters = nullptr;  // [0] inserted rows, [1] side count, [2] overflow flag, [3] scratch
  int device = 0;
  TableView view() const { return TableView{keys, rows, nbuckets, side_rows, counters + 1}; }
};

namespace {

int alloc(kg_ctx* ctx, void** p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return kg_fail(ctx, KG_ERR_OOM, "device allocation of %zu bytes failed", bytes);
  }
  return KG_OK;
}

// A column's data may be NULL only when it has no rows (an empty device tensor).
int check_key(kg_ctx* ctx, const kg_col* key, int64_t n) {
  KG_REQUIRE(ctx, key && (n == 0 || key->data), "key column is NULL");
  switch (key->type) {
    case KG_SMALLINT:
    case KG_INT:
    case KG_BIGINT:
    case KG_DATE:
    case KG_ENUM:
    case KG_REAL:
    case KG_BOOLEAN:
      return KG_OK;
    default:
      return kg_fail(ctx, KG_ERR_UNSUPPORTED, "join keys of type %d are not supported", key->type);
  }
}

int read_counter(kg_ctx* ctx, const unsigned long long* d, unsigned long long* h) {
  int st = kg_pinned_reserve(ctx, 64);
  if (st != KG_OK) return st;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(ctx->pinned, d, 8, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *h = *static_cast<unsigned long long*>(ctx->pinned);
  return KG_OK;
}

}  // namespace

extern "C" int kg_ht_create(kg_ctx* ctx, int64_t capacity_rows, double load_factor, kg_hash_table** out) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, out, "out is NULL");
  KG_REQUIRE(ctx, capacity_rows >= 0 && capacity_rows < INT32_MAX, "capacity %lld out of range",
             (long long)capacity_rows);
  if (load_factor == 0) load_factor = 0.5;
  KG_REQUIRE(ctx, load_factor > 0 && load_factor <= 0.9, "load factor %g out of (0, 0.9]", load_factor);
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  kg_hash_table* t = new (std::nothrow) kg_hash_table();
  if (!t) return kg_fail(ctx, KG_ERR_OOM, "host allocation failed");
  t->device = ctx->device;
  t->capacity = capacity_rows;
  const double slots = std::max<double>(static_cast<double>(capacity_rows) / load_factor, kSlots);
  t->nbuckets = static_cast<uint64_t>(ceil_div(static_cast<int64_t>(slots), kSlots));
  const size_t nslots = t->nbuckets * kSlots;
  int st;
  if ((st = alloc(ctx, reinterpret_cast<void**>(&t->keys), nslots * 8)) != KG_OK ||
      (st = alloc(ctx, reinterpret_cast<void**>(&t->rows), nslots * 4)) != KG_OK ||
      (st = alloc(ctx, reinterpret_cast<void**>(&t->side_rows), kSideCap * 4)) != KG_OK ||
      (st = alloc(ctx, reinterpret_cast<void**>(&t->counters), 64)) != KG_OK) {
    kg_ht_free(ctx, t);
    return st;
  }
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(ctx, nslots, 4)), dim3(kJBlock), 0, ctx->stream, t->keys, kEmpty,
                     static_cast<int64_t>(nslots));
  KG_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, t->counters, 0ull, 8);
  KG_LAUNCH_CHECK(ctx);
  *out = t;
  return KG_OK;
}

extern "C" int kg_ht_build(kg_ctx* ctx, kg_hash_table* ht, const kg_col* key, const int32_t* d_rows,
                           int64_t nrows) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, ht, "table is NULL");
  KG_REQUIRE(ctx, nrows >= 0 && nrows < INT32_MAX, "row count out of range");
  int st = check_key(ctx, key, nrows);
  if (st != KG_OK) return st;
  if (nrows == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  unsigned long long inserted = 0;
  if ((st = read_counter(ctx, ht->counters, &inserted)) != KG_OK) return st;
  if (static_cast<int64_t>(inserted) + nrows > ht->capacity)
    return kg_fail(ctx, KG_ERR_CAPACITY, "table capacity %lld rows exceeded", (long long)ht->capacity);
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_ht_build, dim3(grid_for(ctx, nrows, 4)), dim3(kJBlock), 0, ctx->stream, ht->view(), *key,
                     d_rows, nrows, ht->counters, reinterpret_cast<int*>(ht->counters + 2));
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  unsigned long long overflow = 0;
  if ((st = read_counter(ctx, ht->counters + 2, &overflow)) != KG_OK) return st;
  if (overflow) return kg_fail(ctx, KG_ERR_CAPACITY, "join table overflow (side list or buckets full)");
  return KG_OK;
}

namespace {

int probe_common(kg_ctx* ctx, const kg_hash_table* ht, const kg_col* key, int64_t nrows) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, ht, "table is NULL");
  KG_REQUIRE(ctx, nrows >= 0 && nrows < INT32_MAX, "row count out of range");
  int st = check_key(ctx, key, nrows);
  if (st != KG_OK) return st;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  return KG_OK;
}

}  // namespace

extern "C" int kg_ht_probe_count(kg_ctx* ctx, const kg_hash_table* ht, const kg_col* key,
                                 const int32_t* d_rows, int64_t nrows, int64_t* out_count) {
  int st = probe_common(ctx, ht, key, nrows);
  if (st != KG_OK) return st;
  KG_REQUIRE(ctx, out_count, "out_count is NULL");
  *out_count = 0;
  if (nrows == 0) return KG_OK;
  unsigned long long* cnt = ht->counters + 3;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, cnt, 0ull, 1);
  KG_LAUNCH_CHECK(ctx);
  ProbeOut o{};
  o.count = cnt;
  kg_time_begin(ctx);
  hipLaunchKernelGGL((k_ht_probe<kCount, false>), dim3(grid_for(ctx, nrows, 64)), dim3(kJBlock), 0, ctx->stream,
                     ht->view(), *key, d_rows, nrows, o);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  unsigned long long h = 0;
  if ((st = read_counter(ctx, cnt, &h)) != KG_OK) return st;
  *out_count = static_cast<int64_t>(h);
  return KG_OK;
}

extern "C" int kg_ht_probe(kg_ctx* ctx, const kg_hash_table* ht, const kg_col* key, const int32_t* d_rows,
                           int64_t nrows, int32_t* d_out_probe, int32_t* d_out_build, int64_t out_cap,
                           int64_t* out_count) {
  int st = probe_common(ctx, ht, key, nrows);
  if (st != KG_OK) return st;
  KG_REQUIRE(ctx, out_count, "out_count is NULL");
  KG_REQUIRE(ctx, out_cap >= 0, "negative out_cap");
  KG_REQUIRE(ctx, out_cap == 0 || (d_out_probe && d_out_build), "output arrays are NULL");
  *out_count = 0;
  if (nrows == 0) return KG_OK;
  unsigned long long* cnt = ht->counters + 3;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, cnt, 0ull, 1);
  KG_LAUNCH_CHECK(ctx);
  ProbeOut o{};
  o.count = cnt;
  o.out_probe = d_out_probe;
  o.out_build = d_out_build;
  o.cap = out_cap;
  kg_time_begin(ctx);
  hipLaunchKernelGGL((k_ht_probe<kEmit, false>), dim3(grid_for(ctx, nrows, 64)), dim3(kJBlock), 0, ctx->stream,
                     ht->view(), *key, d_rows, nrows, o);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  unsigned long long h = 0;
  if ((st = read_counter(ctx, cnt, &h)) != KG_OK) return st;
  *out_count = static_cast<int64_t>(h);
  if (*out_count > out_cap)
    return kg_fail(ctx, KG_ERR_CAPACITY, "%lld join pairs exceed out_cap %lld", (long long)*out_count,
                   (long long)out_cap);
  return KG_OK;
}

extern "C" int kg_ht_probe_agg(kg_ctx* ctx, const kg_hash_table* ht, const kg_col* key, const int64_t* d_vals,
                               int64_t nrows, int64_t max_groups, int64_t* d_out_key, int64_t* d_out_sum,
                               int64_t* d_out_count, int32_t* d_out_first, int64_t* out_ngroups) {
  int st = probe_common(ctx, ht, key, nrows);
  if (st != KG_OK) return st;
  KG_REQUIRE(ctx, out_ngroups && d_vals, "null pointer argument");
  KG_REQUIRE(ctx, max_groups >= 0, "negative max_groups");
  KG_REQUIRE(ctx, max_groups == 0 || (d_out_key && d_out_sum && d_out_count), "output arrays are NULL");
  *out_ngroups = 0;
  kg_hash_table* t = const_cast<kg_hash_table*>(ht);
  const int64_t nslots = static_cast<int64_t>(t->nbuckets) * kSlots;
  if (!t->acc && (st = alloc(ctx, reinterpret_cast<void**>(&t->acc), (nslots + 1) * sizeof(Acc))) != KG_OK) return st;
  if (d_out_first && !t->first && (st = alloc(ctx, reinterpret_cast<void**>(&t->first), (nslots + 1) * 4)) != KG_OK)
    return st;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ctx, 2 * (nslots + 1), 4)), dim3(kJBlock), 0, ctx->stream,
                     reinterpret_cast<unsigned long long*>(t->acc), 0ull, 2 * (nslots + 1));
  KG_LAUNCH_CHECK(ctx);
  if (d_out_first) {
    hipLaunchKernelGGL(k_fill_i32j, dim3(grid_for(ctx, nslots + 1, 4)), dim3(kJBlock), 0, ctx->stream, t->first,
                       INT32_MAX, nslots + 1);
    KG_LAUNCH_CHECK(ctx);
  }
  if (nrows > 0) {
    ProbeOut o{};
    o.acc = t->acc;
    o.first = t->first;
    o.vals = d_vals;
    kg_time_begin(ctx);
    if (d_out_first)
      hipLaunchKernelGGL((k_ht_probe<kAgg, true>), dim3(grid_for(ctx, nrows, 64)), dim3(kJBlock), 0, ctx->stream,
                         t->view(), *key, nullptr, nrows, o);
    else
      hipLaunchKernelGGL((k_ht_probe<kAgg, false>), dim3(grid_for(ctx, nrows, 64)), dim3(kJBlock), 0, ctx->stream,
                         t->view(), *key, nullptr, nrows, o);
    kg_time_end(ctx);
    KG_LAUNCH_CHECK(ctx);
  }
  unsigned long long* cnt = t->counters + 3;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, cnt, 0ull, 1);
  KG_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_collect, dim3(grid_for(ctx, nslots + 1, 4)), dim3(kJBlock), 0, ctx->stream, t->keys, t->acc,
                     t->first, nslots, nslots, cnt, max_groups, d_out_key, d_out_sum, d_out_count, d_out_first);
  KG_LAUNCH_CHECK(ctx);
  unsigned long long h = 0;
  if ((st = read_counter(ctx, cnt, &h)) != KG_OK) return st;
  *out_ngroups = static_cast<int64_t>(h);
  if (*out_ngroups > max_groups)
    return kg_fail(ctx, KG_ERR_CAPACITY, "%lld groups exceed max_groups %lld", (long long)*out_ngroups,
                   (long long)max_groups);
  return KG_OK;
}

extern "C" int kg_ht_size(kg_ctx* ctx, const kg_hash_table* ht, int64_t* out_rows) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, ht && out_rows, "null pointer argument");
  unsigned long long h = 0;
  int st = read_counter(ctx, ht->counters, &h);
  if (st != KG_OK) return st;
  *out_rows = static_cast<int64_t>(h);
  return KG_OK;
}

extern "C" int kg_ht_free(kg_ctx* ctx, kg_hash_table* ht) {
  if (!ht) return KG_OK;
  if (ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  if (ht->keys) hipFree(ht->keys);
  if (ht->rows) hipFree(ht->rows);
  if (ht->acc) hipFree(ht->acc);
  if (ht->first) hipFree(ht->first);
  if (ht->side_rows) hipFree(ht->side_rows);
  if (ht->counters) hipFree(ht->counters);
  delete ht;
  return KG_OK;
}

extern "C" int kg_group_by_i64(kg_ctx* ctx, const int64_t* d_keys, const int64_t* d_vals, int64_t n,
                               int64_t max_groups, int64_t* d_out_key, int64_t* d_out_sum, int64_t* d_out_count,
                               int32_t* d_out_first, int64_t* out_ngroups) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count out of range");
  KG_REQUIRE(ctx, out_ngroups, "out_ngroups is NULL");
  KG_REQUIRE(ctx, n == 0 || (d_keys && d_vals), "input arrays are NULL");
  KG_REQUIRE(ctx, max_groups >= 0, "negative max_groups");
  KG_REQUIRE(ctx, max_groups == 0 || (d_out_key && d_out_sum && d_out_count), "output arrays are NULL");
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  *out_ngroups = 0;
  // Table sized for at most n groups at load factor 0.5.
  const uint64_t nbuckets = static_cast<uint64_t>(std::max<int64_t>(ceil_div(std::max<int64_t>(n, 1) * 2, kSlots), 1));
  const int64_t nslots = static_cast<int64_t>(nbuckets) * kSlots;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_keys = carve(nslots * 8), o_acc = carve((nslots + 1) * sizeof(Acc)),
               o_first = carve((nslots + 1) * 4), o_cnt = carve(64);
  int st = kg_work_reserve(ctx, off);
  if (st != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  int64_t* keys = reinterpret_cast<int64_t*>(w + o_keys);
  Acc* acc = reinterpret_cast<Acc*>(w + o_acc);
  int32_t* first = reinterpret_cast<int32_t*>(w + o_first);
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(w + o_cnt);
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(ctx, nslots, 4)), dim3(kJBlock), 0, ctx->stream, keys, kEmpty, nslots);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ctx, 2 * (nslots + 1), 4)), dim3(kJBlock), 0, ctx->stream,
                     reinterpret_cast<unsigned long long*>(acc), 0ull, 2 * (nslots + 1));
  hipLaunchKernelGGL(k_fill_i32j, dim3(grid_for(ctx, nslots + 1, 4)), dim3(kJBlock), 0, ctx->stream, first,
                     INT32_MAX, nslots + 1);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, cnt, 0ull, 8);
  KG_LAUNCH_CHECK(ctx);
  int* overflow = reinterpret_cast<int*>(cnt + 1);
  if (n > 0) {
    kg_time_begin(ctx);
    if (d_out_first)
      hipLaunchKernelGGL(k_group_by_i64<true>, dim3(grid_for(ctx, n, 4)), dim3(kJBlock), 0, ctx->stream, keys, acc,
                         first, nbuckets, d_keys, d_vals, n, overflow);
    else
      hipLaunchKernelGGL(k_group_by_i64<false>, dim3(grid_for(ctx, n, 4)), dim3(kJBlock), 0, ctx->stream, keys, acc,
                         first, nbuckets, d_keys, d_vals, n, overflow);
    kg_time_end(ctx);
    KG_LAUNCH_CHECK(ctx);
  }
  hipLaunchKernelGGL(k_collect, dim3(grid_for(ctx, nslots + 1, 4)), dim3(kJBlock), 0, ctx->stream, keys, acc, first,
                     nslots, nslots, cnt, max_groups, d_out_key, d_out_sum, d_out_count, d_out_first);
  KG_LAUNCH_CHECK(ctx);
  unsigned long long h = 0, of = 0;
  if ((st = read_counter(ctx, cnt, &h)) != KG_OK) return st;
  if ((st = read_counter(ctx, cnt + 1, &of)) != KG_OK) return st;
  if (of) return kg_fail(ctx, KG_ERR_CAPACITY, "group table overflow");
  *out_ngroups = static_cast<int64_t>(h);
  if (*out_ngroups > max_groups)
    return kg_fail(ctx, KG_ERR_CAPACITY, "%lld groups exceed max_groups %lld", (long long)*out_ngroups,
                   (long long)max_groups);
  return KG_OK;
}

extern "C" int kg_scan_join_group_agg(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols,
                                      const kg_pred* preds, int32_t npreds, const kg_hash_table* ht,
                                      int32_t probe_key_col, const kg_col* build_codes, int32_t group_card,
                                      const kg_agg* aggs, int32_t naggs, int64_t max_groups, int64_t* out_ngroups,
                                      int32_t* out_keys, int64_t* out_vals, int8_t* out_nulls) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, ht, "table is NULL");
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count out of range");
  KG_REQUIRE(ctx, out_ngroups && out_keys && (naggs == 0 || (aggs && out_vals && out_nulls)), "null output pointer");
  KG_REQUIRE(ctx, naggs >= 0 && naggs <= 16, "at most 16 aggregates");
  KG_REQUIRE(ctx, probe_key_col >= 0 && probe_key_col < ncols, "bad probe key column %d", probe_key_col);
  KG_REQUIRE(ctx, group_card > 0, "group cardinality must be positive");
  // No code column: the ungrouped AggregateOperator over the join output.
  KG_REQUIRE(ctx, build_codes || group_card == 1, "build group codes are NULL");
  if (build_codes) {
    KG_REQUIRE(ctx, build_codes->data, "build group codes have no data");
    if (build_codes->type != KG_INT && build_codes->type != KG_ENUM)
      return kg_fail(ctx, KG_ERR_UNSUPPORTED, "build group codes must be INT/ENUM codes");
    if (build_codes->nulls) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "nullable build group codes");
  }
  const int32_t* codes_data = build_codes ? static_cast<const int32_t*>(build_codes->data) : nullptr;
  int st = check_key(ctx, &cols[probe_key_col], n);
  if (st != KG_OK) return st;
  PredSet P;
  if ((st = build_preds(ctx, cols, ncols, preds, npreds, 1, n, P)) != KG_OK) return st;
  MeasureSet M;
  std::vector<AggPlan> plan;
  bool minmax = false;
  if ((st = build_measures(ctx, cols, ncols, aggs, naggs, n, M, plan, minmax)) != KG_OK) return st;
  const int G = group_card;
  const int nm = M.nm;
  const size_t lds_bytes = (static_cast<size_t>(G) * nm * 2 + G * 2 + (minmax ? 2 * G * nm : 0)) * 8;
  if (lds_bytes > 64 * 1024)
    return kg_fail(ctx, KG_ERR_UNSUPPORTED, "group table of %zu bytes exceeds the LDS budget", lds_bytes);
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  *out_ngroups = 0;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_fin = carve(static_cast<size_t>(G) * (nm > 0 ? nm : 1) * sizeof(AggFinal)), o_rows = carve(G * 8),
               o_first = carve(G * 8), o_fnull = carve(G + 16);
  if ((st = kg_work_reserve(ctx, off)) != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  AggFinal* dfin = reinterpret_cast<AggFinal*>(w + o_fin);
  unsigned long long* drows = reinterpret_cast<unsigned long long*>(w + o_rows);
  unsigned long long* dfirst = reinterpret_cast<unsigned long long*>(w + o_first);
  const int init_threads = std::max(G * nm, G);
  hipLaunchKernelGGL(k_init_join_final, dim3(ceil_div(init_threads, 256)), dim3(256), 0, ctx->stream, dfin, G, M,
                     drows, dfirst);
  KG_LAUNCH_CHECK(ctx);
  if (n > 0) {
    const int nb = grid_for(ctx, n, 16);
    kg_time_begin(ctx);
    if (minmax)
      hipLaunchKernelGGL(k_join_group_agg<true>, dim3(nb), dim3(kJBlock), lds_bytes, ctx->stream, P,
                         cols[probe_key_col], ht->view(), codes_data, G, M, n,
                         dfin, drows, dfirst);
    else
      hipLaunchKernelGGL(k_join_group_agg<false>, dim3(nb), dim3(kJBlock), lds_bytes, ctx->stream, P,
                         cols[probe_key_col], ht->view(), codes_data, G, M, n,
                         dfin, drows, dfirst);
    kg_time_end(ctx);
    KG_LAUNCH_CHECK(ctx);
  }
  std::vector<AggFinal> fin(static_cast<size_t>(G) * (nm > 0 ? nm : 1));
  std::vector<uint64_t> rowsv(G), firstu(G);
  std::vector<std::vector<int8_t>> first_null(naggs);
  KG_HIP_CHECK(ctx, hipMemcpyAsync(fin.data(), dfin, static_cast<size_t>(G) * nm * sizeof(AggFinal),
                                   hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipMemcpyAsync(rowsv.data(), drows, G * 8, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipMemcpyAsync(firstu.data(), dfirst, G * 8, hipMemcpyDeviceToHost, ctx->stream));
  int8_t* dfn = reinterpret_cast<int8_t*>(w + o_fnull);
  for (int a = 0; a < naggs; a++) {
    if (plan[a].kind != KG_AGG_COUNT || !plan[a].count_nulls) continue;
    first_null[a].resize(G);
    hipLaunchKernelGGL(k_first_null_join, dim3(ceil_div(G, 256)), dim3(256), 0, ctx->stream, dfirst, G,
                       plan[a].count_nulls, dfn);
    KG_LAUNCH_CHECK(ctx);
    KG_HIP_CHECK(ctx, hipMemcpyAsync(first_null[a].data(), dfn, G, hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  }
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<int64_t> first64(G);
  // (probe row << 32 | build row) < 2^63 for every present group.
  for (int g = 0; g < G; g++) first64[g] = static_cast<int64_t>(firstu[g]);
  return emit_groups(ctx, G, build_codes ? 1 : 0, 0, M, plan, fin, rowsv, first64, first_null, max_groups,
                     out_ngroups, out_keys, out_vals, out_nulls);
}
