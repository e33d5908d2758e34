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
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <new>
#include <type_traits>

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
    if (t.narrow) insert_narrow(t, k, static_cast<int32_t>(r), overflow);
    else insert_wide(t, k, static_cast<int32_t>(r), overflow);
  }
  // One atomic per wave for the row count.
  for (int s = kWave / 2; s >= 1; s >>= 1) local += __shfl_xor(local, s);
  if ((threadIdx.x & (kWave - 1)) == 0 && local) atomicAdd(inserted, local);
}

// Minimum and maximum non-NULL key of the rows to insert (integer keys).
__global__ __launch_bounds__(kJBlock) void k_key_range(kg_col key, const int32_t* sel, int64_t n,
                                                        unsigned long long* out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kJBlock;
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  constexpr int U = 8;  // loads in flight per thread
  for (int64_t i0 = static_cast<int64_t>(blockIdx.x) * kJBlock + threadIdx.x; i0 < n; i0 += stride * U) {
    int64_t kk[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + u * stride;
      ok[u] = i < n && load_key_nt(key, sel ? sel[i] : i, kk[u]);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (!ok[u]) continue;
      lo = kk[u] < lo ? kk[u] : lo;
      hi = kk[u] > hi ? kk[u] : hi;
    }
  }
  for (int s = kWave / 2; s >= 1; s >>= 1) {
    const int64_t a = static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(lo), s));
    const int64_t c = static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(hi), s));
    lo = a < lo ? a : lo;
    hi = c > hi ? c : hi;
  }
  // Across the workgroup's waves, then one atomic pair per workgroup (a pair
  // per wave on one address serializes thousands of atomics).
  __shared__ int64_t wlo[kJBlock / kWave], whi[kJBlock / kWave];
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    wlo[wave] = lo;
    whi[wave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kJBlock / kWave; w++) {
      lo = wlo[w] < lo ? wlo[w] : lo;
      hi = whi[w] > hi ? whi[w] : hi;
    }
    if (lo <= hi) {
      atomicMin(reinterpret_cast<long long*>(out), static_cast<long long>(lo));
      atomicMax(reinterpret_cast<long long*>(out + 1), static_cast<long long>(hi));
    }
  }
}

// Re-inserts the occupied slots of a narrow table (copied to keys/rows) into
// the wide table t.
__global__ __launch_bounds__(kJBlock) void k_widen(TableView t, const uint32_t* keys, const int32_t* rows,
                                                    int64_t nslots, int64_t base, int* overflow) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kJBlock;
  for (int64_t s = static_cast<int64_t>(blockIdx.x) * kJBlock + threadIdx.x; s < nslots; s += stride) {
    const uint32_t k32 = keys[s];
    if (k32 == kEmpty32) continue;
    insert_wide(t, static_cast<int64_t>(static_cast<uint64_t>(base) + k32), rows[s], overflow);
  }
}

enum ProbeMode { kCount = 0, kEmit = 1, kAgg = 2 };

struct ProbeOut {
  unsigned long long* count;   // kCount / kEmit: total pairs (kEmit: set to the workgroup's LDS cursor)
  int32_t* out_probe;          // kEmit
  int32_t* out_build;          // kEmit
  int64_t cap;                 // kEmit
  Acc* acc;                    // kAgg: per-slot accumulators (+1 side slot at the end)
  int32_t* first;              // kAgg: per-slot first probe row (optional)
  const int64_t* vals;         // kAgg: probe-side values
  unsigned long long* block_counts;      // kCount: per-workgroup totals instead of *count
  const unsigned long long* block_offs;  // kEmit: each workgroup's first output position
  // kAgg: the hot keys (device list) and, in the kernel, their LDS state.
  const int64_t* hot_keys;    // kAgg: device list of hot probe keys (k_hot_keys), or NULL
  const int32_t* hot_n;       // its length
  int32_t nhot;
  const int16_t* hidx;
  const int64_t* hkey;
  int64_t* hslot;
  unsigned long long* hsum;
  unsigned long long* hcnt;
  int32_t* hfirst;
};

// kAgg under skew: the hot probe keys (found by k_hot_keys from a sample,
// at most kHot) are accumulated per workgroup in LDS and added to the global
// accumulators once when the workgroup ends; every other key updates the
// global accumulators directly. Under Zipf theta = 1 about 6% of all probes
// carry one key, and their device atomics on one slot would serialize.
constexpr int kHot = 256;
constexpr int kHotHash = 512;
constexpr int kHotSample = 1 << 16;   // probe keys sampled by k_hot_keys
constexpr int kHotTable = 4096;       // distinct sampled keys counted
constexpr int kHotMinCount = 3;       // sampled occurrences that make a key hot
constexpr int kHotProbes = 16;        // table entries a sampled key may try

__device__ __forceinline__ uint32_t hot_pos(int64_t k) {
  return static_cast<uint32_t>(mix64(static_cast<uint64_t>(k)) >> 32) & (kHotHash - 1);
}

// The sample: kHotSample probe keys at an even stride (kEmpty where the row
// has no key), gathered by many workgroups so that the random reads overlap.
__global__ __launch_bounds__(256) void k_hot_sample(kg_col key, const int32_t* sel, int64_t n, int64_t* out) {
  const int64_t stride = n > kHotSample ? n / kHotSample : 1;
  const int64_t s = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (s >= kHotSample) return;
  int64_t k = kEmpty;
  if (s * stride < n) {
    int64_t v;
    if (load_key(key, sel ? sel[s * stride] : s * stride, v)) k = v;
  }
  out[s] = k;
}

// One workgroup: the sampled probe keys counted in an LDS
// open-addressing table (a key is claimed by a 64-bit CAS of the free marker;
// keys that find no entry within kHotProbes probes are not counted, nor is the
// free marker itself); keys seen at least kHotMinCount times become the hot
// list.
__global__ __launch_bounds__(1024) void k_hot_keys(const int64_t* sample, int64_t* hot, int32_t* nhot) {
  __shared__ unsigned long long tkey[kHotTable];
  __shared__ uint32_t tcnt[kHotTable];
  __shared__ int32_t out_n;
  for (int i = threadIdx.x; i < kHotTable; i += 1024) {
    tkey[i] = static_cast<unsigned long long>(kEmpty);
    tcnt[i] = 0;
  }
  if (threadIdx.x == 0) out_n = 0;
  __syncthreads();
  for (int s = threadIdx.x; s < kHotSample; s += 1024) {
    const int64_t k = sample[s];
    if (k == kEmpty) continue;
    uint32_t p = static_cast<uint32_t>(mix64(static_cast<uint64_t>(k))) & (kHotTable - 1);
    for (int probes = 0; probes < kHotProbes; probes++) {
      // A plain read first: once the table is full (more distinct keys than
      // entries, as under uniform keys) most probes end without a CAS.
      unsigned long long prev = tkey[p];
      if (prev == static_cast<unsigned long long>(kEmpty))
        prev = atomicCAS(&tkey[p], static_cast<unsigned long long>(kEmpty), static_cast<unsigned long long>(k));
      if (prev == static_cast<unsigned long long>(kEmpty) || prev == static_cast<unsigned long long>(k)) {
        atomicAdd(&tcnt[p], 1u);
        break;
      }
      p = (p + 1) & (kHotTable - 1);
    }
  }
  // The most frequent keys: a histogram of floor(log2(count)) picks the
  // smallest power-of-two threshold that admits at most kHot keys.
  __shared__ uint32_t lhist[32];
  __shared__ uint32_t thresh;
  if (threadIdx.x < 32) lhist[threadIdx.x] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < kHotTable; i += 1024)
    if (tcnt[i] >= kHotMinCount) atomicAdd(&lhist[31 - __clz(tcnt[i])], 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t above = 0, b = 32;
    while (b > 0 && above + lhist[b - 1] <= kHot) above += lhist[--b];
    thresh = b >= 32 ? 0xFFFFFFFFu : (1u << b);
    if (thresh < kHotMinCount) thresh = kHotMinCount;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHotTable; i += 1024) {
    if (tcnt[i] < thresh) continue;
    const int j = atomicAdd(&out_n, 1);
    if (j < kHot) hot[j] = static_cast<int64_t>(tkey[i]);
  }
  __syncthreads();
  if (threadIdx.x == 0) *nhot = out_n < kHot ? out_n : kHot;
}

// Device bytes of a table's hot-key buffer: the list (kHot keys), its length
// (in the next 64 bytes), then the sample.
constexpr size_t kHotBufBytes = kHot * 8 + 64 + static_cast<size_t>(kHotSample) * 8;

void launch_hot_keys(hipStream_t stream, const kg_col& key, const int32_t* sel, int64_t n, int64_t* hot) {
  int64_t* sample = hot + kHot + 8;
  hipLaunchKernelGGL(k_hot_sample, dim3(kHotSample / 256), dim3(256), 0, stream, key, sel, n, sample);
  hipLaunchKernelGGL(k_hot_keys, dim3(1), dim3(1024), 0, stream, sample, hot, reinterpret_cast<int32_t*>(hot + kHot));
}

template <bool FIRST>
__device__ __forceinline__ void agg_add(const ProbeOut& o, int64_t slot, int64_t k, unsigned long long sum,
                                        unsigned n, int32_t r) {
  if (o.nhot > 0) {
    for (uint32_t p = hot_pos(k);; p = (p + 1) & (kHotHash - 1)) {
      const int j = o.hidx[p];
      if (j < 0) break;
      if (o.hkey[j] == k) {
        o.hslot[j] = slot;
        atomicAdd(&o.hsum[j], sum);
        atomicAdd(&o.hcnt[j], static_cast<unsigned long long>(n));
        if (FIRST) atomicMin(&o.hfirst[j], r);
        return;
      }
    }
  }
  Acc* a = &o.acc[slot];
  atomicAdd(&a->sum, sum);
  atomicAdd(&a->count, static_cast<unsigned long long>(n));
  if (FIRST) atomicMin(&o.first[slot], r);
}

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
// several probes' first loads are in flight together. r is the probe row (for
// outputs), vi the index of its value (kAgg). Must be called by the whole wave
// (ballots).
template <int MODE, bool FIRST>
__device__ __forceinline__ void probe_one(const TableView& t, const ProbeOut& o, int64_t r, int64_t vi, int64_t k,
                                          bool active, uint64_t b, longlong2 kv0, int sub, int grp,
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
  if (MODE == kAgg && sub == 0 && nmatch)
    agg_add<FIRST>(o, first_slot, k, static_cast<unsigned long long>(o.vals[vi]) * nmatch, nmatch,
                   static_cast<int32_t>(r));
}

// The same for a narrow table: each lane of the group holds 4 of the bucket's
// 16 offsets (16 bytes); slot order within the bucket is lane-major.
// kv0_valid = false: kv0 is not used, bucket b is read from memory (a chain
// continued past its first bucket); max_buckets bounds the buckets visited.
template <int MODE, bool FIRST>
__device__ __forceinline__ void probe_one32(const TableView& t, const ProbeOut& o, int64_t r, int64_t vi, uint32_t k32,
                                            bool active, uint64_t b, uint4 kv0, int sub, int grp,
                                            unsigned long long& local, bool kv0_valid = true,
                                            uint64_t max_buckets = ~uint64_t{0}) {
  const uint32_t* keys = keys32(t);
  const uint64_t nb = nbuckets32(t);
  const uint64_t limit = max_buckets < nb ? max_buckets : nb;
  unsigned nmatch = 0;
  int64_t first_slot = -1;
  uint64_t scanned = 0;
  bool first_bucket = kv0_valid;
  if (limit == 0) active = false;
  while (__any(active)) {
    bool m0 = false, m1 = false, m2 = false, m3 = false, e = false;
    int64_t slot0 = 0;
    if (active) {
      slot0 = static_cast<int64_t>(b) * kSlots32 + sub * 4;
      const uint4 kv = first_bucket ? kv0 : *reinterpret_cast<const uint4*>(keys + slot0);
      m0 = kv.x == k32;
      m1 = kv.y == k32;
      m2 = kv.z == k32;
      m3 = kv.w == k32;
      e = kv.x == kEmpty32 || kv.y == kEmpty32 || kv.z == kEmpty32 || kv.w == kEmpty32;
    }
    first_bucket = false;
    const unsigned sh = grp * kProbeLanes;
    const unsigned g0 = (__ballot(m0) >> sh) & 0xF, g1 = (__ballot(m1) >> sh) & 0xF;
    const unsigned g2 = (__ballot(m2) >> sh) & 0xF, g3 = (__ballot(m3) >> sh) & 0xF;
    const unsigned ge = (__ballot(e) >> sh) & 0xF;
    if (MODE == kCount) {
      if (active && sub == 0) local += __popc(g0) + __popc(g1) + __popc(g2) + __popc(g3);
    } else if (MODE == kEmit) {
      const unsigned want = (m0 ? 1u : 0u) + (m1 ? 1u : 0u) + (m2 ? 1u : 0u) + (m3 ? 1u : 0u);
      const unsigned long long pos = wave_append(o.count, active ? want : 0u);
      unsigned j = 0;
      const bool mm[4] = {m0, m1, m2, m3};
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (!mm[q]) continue;
        if (pos + j < static_cast<unsigned long long>(o.cap)) {
          o.out_probe[pos + j] = static_cast<int32_t>(r);
          o.out_build[pos + j] = t.rows[slot0 + q];
        }
        j++;
      }
    } else {
      if (active) {
        nmatch += __popc(g0) + __popc(g1) + __popc(g2) + __popc(g3);
        if (first_slot < 0 && (g0 | g1 | g2 | g3)) {
          unsigned bits = 0;
          for (int q = 0; q < kProbeLanes; q++)
            bits |= (((g0 >> q) & 1u) << (4 * q)) | (((g1 >> q) & 1u) << (4 * q + 1)) |
                    (((g2 >> q) & 1u) << (4 * q + 2)) | (((g3 >> q) & 1u) << (4 * q + 3));
          first_slot = static_cast<int64_t>(b) * kSlots32 + (__ffs(bits) - 1);
        }
      }
    }
    if (active) {
      scanned++;
      if (ge || scanned >= limit) active = false;
      b = (b + 1 == nb) ? 0 : b + 1;
    }
  }
  if (MODE == kAgg && sub == 0 && nmatch)
    agg_add<FIRST>(o, first_slot, static_cast<int64_t>(static_cast<uint64_t>(t.base) + k32),
                   static_cast<unsigned long long>(o.vals[vi]) * nmatch, nmatch, static_cast<int32_t>(r));
}

// Emitting probes of one wave step into a narrow table: the first buckets of
// all U probes are already loaded; the row ids of every first-bucket match are
// loaded together (U * 4 independent reads in flight instead of one probe's at
// a time), one wave-aggregated append reserves the step's output, and only
// chains that run past a full first bucket continue bucket by bucket.
template <int U>
__device__ __forceinline__ void emit_first32(const TableView& t, const ProbeOut& o, const int64_t (&r)[U],
                                             const uint32_t (&k32)[U], const bool (&active)[U],
                                             const uint64_t (&b)[U], const uint4 (&kv)[U], int sub, int grp,
                                             unsigned long long& local) {
  const uint64_t nb = nbuckets32(t);
  const unsigned sh = grp * kProbeLanes;
  int32_t rows[U][4];
  unsigned mask[U];
  bool more[U];
  unsigned want = 0;
#pragma unroll
  for (int u = 0; u < U; u++) {
    unsigned mk = 0;
    bool e = false;
    if (active[u]) {
      mk = (kv[u].x == k32[u] ? 1u : 0u) | (kv[u].y == k32[u] ? 2u : 0u) | (kv[u].z == k32[u] ? 4u : 0u) |
           (kv[u].w == k32[u] ? 8u : 0u);
      e = kv[u].x == kEmpty32 || kv[u].y == kEmpty32 || kv[u].z == kEmpty32 || kv[u].w == kEmpty32;
    }
    const unsigned ge = (__ballot(e) >> sh) & 0xF;
    more[u] = active[u] && ge == 0;
    mask[u] = mk;
    want += __popc(mk);
    const int64_t slot0 = static_cast<int64_t>(b[u]) * kSlots32 + sub * 4;
#pragma unroll
    for (int q = 0; q < 4; q++) rows[u][q] = ((mk >> q) & 1u) ? t.rows[slot0 + q] : 0;
  }
  unsigned long long pos = wave_append(o.count, want);
  const unsigned long long cap = static_cast<unsigned long long>(o.cap);
#pragma unroll
  for (int u = 0; u < U; u++)
#pragma unroll
    for (int q = 0; q < 4; q++)
      if ((mask[u] >> q) & 1u) {
        if (pos < cap) {
          o.out_probe[pos] = static_cast<int32_t>(r[u]);
          o.out_build[pos] = rows[u][q];
        }
        pos++;
      }
  // Chains past a full first bucket: the following buckets, at most nb - 1 of
  // them (the first bucket is not visited again).
#pragma unroll
  for (int u = 0; u < U; u++)
    probe_one32<kEmit, false>(t, o, r[u], r[u], k32[u], more[u], b[u] + 1 == nb ? 0 : b[u] + 1, kv[u], sub, grp,
                              local, false, nb - 1);
}

// Probe kernel: every 4-lane group handles kUnroll probes per wave step; the
// first-bucket loads of all of them are issued before any is examined, so a
// wave keeps 16 * kUnroll random 64-byte reads in flight. Counting probes keep
// 8 per group in flight (256M probes into a 16M-key narrow table: 4.69 ms at 8,
// 4.87 ms at 4, 4.86 at 2, 5.4 at 12 where registers cut occupancy).
template <int MODE>
constexpr int probe_unroll() { return MODE == kCount ? 8 : 4; }

constexpr int kXcds = 8;

// Partitioned input (see k_part_scatter): probe keys grouped by table region,
// regions laid out XCD-major; workgroup b (on XCD b mod 8) probes only its XCD's
// regions, in order, so each region's slice of the table stays in that XCD's
// L2 while its probes run.
struct PartIn {
  const uint32_t* offs;   // scanned histogram [P][nblk] (+ total), null: not partitioned
  const int32_t* rowmap;  // probe row of each partitioned key (kEmit / kAgg first rows)
  int32_t nblk;
  int32_t P;
};

template <int MODE, bool FIRST, bool NARROW, int UNROLL = probe_unroll<MODE>()>
__global__ __launch_bounds__(kJBlock) void k_ht_probe(TableView t, kg_col key, const int32_t* sel, int64_t n,
                                                       ProbeOut o, PartIn part) {
  // (o is this workgroup's copy: the emit cursor and the cache live in LDS.)
  const int lane = threadIdx.x & (kWave - 1);
  const int sub = lane & (kProbeLanes - 1);
  const int grp = lane / kProbeLanes;
  constexpr int kGroups = kWave / kProbeLanes;
  constexpr int kUnroll = UNROLL;
  constexpr int kStep = kGroups * kUnroll;
  int64_t lo = 0, hi = n, gwave, nwaves;
  if (part.offs) {
    const int x = blockIdx.x % kXcds;
    const int per = part.P / kXcds;
    lo = part.offs[static_cast<int64_t>(x) * per * part.nblk];
    hi = part.offs[static_cast<int64_t>(x + 1) * per * part.nblk];
    gwave = (static_cast<int64_t>(blockIdx.x / kXcds) * kJBlock + threadIdx.x) / kWave;
    nwaves = static_cast<int64_t>(gridDim.x / kXcds) * kJBlock / kWave;
  } else {
    gwave = (static_cast<int64_t>(blockIdx.x) * kJBlock + threadIdx.x) / kWave;
    nwaves = static_cast<int64_t>(gridDim.x) * kJBlock / kWave;
  }
  unsigned long long local = 0;
  // Workgroup-local state: the emit cursor, the count partials, the
  // accumulator cache.
  __shared__ unsigned long long s_cur;
  __shared__ unsigned long long s_wsum[kJBlock / kWave];
  constexpr int kH = MODE == kAgg ? kHot : 1, kHH = MODE == kAgg ? kHotHash : 1;
  __shared__ int16_t s_hidx[kHH];
  __shared__ int64_t s_hkey[kH], s_hslot[kH];
  __shared__ unsigned long long s_hsum[kH], s_hcnt[kH];
  __shared__ int32_t s_hfirst[kH];
  if (MODE == kEmit && o.block_offs) {
    if (threadIdx.x == 0) s_cur = o.block_offs[blockIdx.x];
    o.count = &s_cur;
  }
  if (MODE == kAgg) {
    o.nhot = o.hot_keys ? *o.hot_n : 0;
    for (int e = threadIdx.x; e < kHH; e += kJBlock) s_hidx[e] = -1;
    for (int j = threadIdx.x; j < o.nhot; j += kJBlock) {
      s_hkey[j] = o.hot_keys[j];
      s_hslot[j] = -1;
      s_hsum[j] = 0;
      s_hcnt[j] = 0;
      s_hfirst[j] = INT32_MAX;
    }
    __syncthreads();
    if (threadIdx.x == 0)  // the hot keys are distinct: plain linear probing
      for (int j = 0; j < o.nhot; j++) {
        uint32_t p = hot_pos(s_hkey[j]);
        while (s_hidx[p] >= 0) p = (p + 1) & (kHotHash - 1);
        s_hidx[p] = static_cast<int16_t>(j);
      }
    o.hidx = s_hidx;
    o.hkey = s_hkey;
    o.hslot = s_hslot;
    o.hsum = s_hsum;
    o.hcnt = s_hcnt;
    o.hfirst = s_hfirst;
  }
  __syncthreads();
  // Keys of the next step are loaded while this step's buckets are in flight.
  int64_t nr[kUnroll], nv[kUnroll], nk[kUnroll];
  bool nact[kUnroll];
  auto load_step = [&](int64_t base) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const int64_t i = base + u * kGroups + grp;
      nr[u] = -1;
      nv[u] = -1;
      nk[u] = 0;
      nact[u] = false;
      if (i < hi) {
        if (part.offs) {
          nv[u] = i;
          nr[u] = part.rowmap ? part.rowmap[i] : i;
        } else {
          nr[u] = sel ? sel[i] : i;
          nv[u] = nr[u];
        }
        nact[u] = load_key_nt(key, part.offs ? i : nr[u], nk[u]);
      }
    }
  };
  int64_t base = lo + gwave * kStep;
  if (base < hi) load_step(base);
  for (; base < hi; base += nwaves * kStep) {
    int64_t r[kUnroll], v[kUnroll], k[kUnroll];
    bool active[kUnroll];
    uint64_t b[kUnroll];
    uint32_t k32[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      r[u] = nr[u];
      v[u] = nv[u];
      k[u] = nk[u];
      active[u] = nact[u];
      k32[u] = 0;
      // A key outside a narrow table's range cannot match: no memory access.
      if (NARROW) active[u] = active[u] && narrow_key(t, k[u], k32[u]);
      b[u] = active[u] ? bucket_of(k[u], NARROW ? nbuckets32(t) : t.nbuckets) : 0;
    }
    if (NARROW) {
      uint4 kv[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        kv[u] = make_uint4(kEmpty32, kEmpty32, kEmpty32, kEmpty32);
        if (active[u])
          kv[u] = *reinterpret_cast<const uint4*>(keys32(t) + static_cast<int64_t>(b[u]) * kSlots32 + sub * 4);
      }
      if (base + nwaves * kStep < hi) load_step(base + nwaves * kStep);
      if constexpr (MODE == kEmit) {
        emit_first32<kUnroll>(t, o, r, k32, active, b, kv, sub, grp, local);
      } else {
#pragma unroll
        for (int u = 0; u < kUnroll; u++)
          probe_one32<MODE, FIRST>(t, o, r[u], v[u], k32[u], active[u], b[u], kv[u], sub, grp, local);
      }
    } else {
      longlong2 kv[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        kv[u] = make_longlong2(kEmpty, kEmpty);
        if (active[u] && k[u] != kEmpty)
          kv[u] = *reinterpret_cast<const longlong2*>(t.keys + static_cast<int64_t>(b[u]) * kSlots + sub * 2);
      }
      if (base + nwaves * kStep < hi) load_step(base + nwaves * kStep);
#pragma unroll
      for (int u = 0; u < kUnroll; u++)
        probe_one<MODE, FIRST>(t, o, r[u], v[u], k[u], active[u], b[u], kv[u], sub, grp, local);
    }
  }
  if (MODE == kCount) {
    for (int s = kWave / 2; s >= 1; s >>= 1) local += __shfl_xor(local, s);
    if (o.block_counts) {
      if (lane == 0) s_wsum[threadIdx.x / kWave] = local;
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kJBlock / kWave; w++) t += s_wsum[w];
        o.block_counts[blockIdx.x] = t;
      }
    } else if (lane == 0 && local) {
      atomicAdd(o.count, local);
    }
  }
  if (MODE == kAgg) {
    __syncthreads();
    for (int j = threadIdx.x; j < o.nhot; j += kJBlock) {
      const int64_t slot = s_hslot[j];
      if (slot < 0 || s_hcnt[j] == 0) continue;
      Acc* a = &o.acc[slot];
      atomicAdd(&a->sum, s_hsum[j]);
      atomicAdd(&a->count, s_hcnt[j]);
      if (FIRST) atomicMin(&o.first[slot], s_hfirst[j]);
    }
  }
}

// ---- radix partitioning of probe keys by table region -------------------------
// Region of a key = floor(bucket * P / nbuckets): regions are contiguous bucket
// ranges, 1/P of the table each. Regions are stored XCD-major (region p at
// position (p mod 8) * P/8 + p / 8) so that XCD x's regions are contiguous.
// Keys that cannot match (NULL, NaN, outside a narrow table's range) are
// dropped here. Each workgroup owns a contiguous range of input rows; the
// histogram and the scatter walk it identically.
constexpr int kPartTile = 2048;   // keys staged in LDS per scatter step

__device__ __forceinline__ int key_region(const TableView& t, const kg_col& key, int64_t r, uint32_t P, int64_t& k) {
  if (!load_key(key, r, k)) return -1;
  uint64_t nb = t.nbuckets;
  if (t.narrow) {
    uint32_t k32;
    if (!narrow_key(t, k, k32)) return -1;
    nb = nbuckets32(t);
  }
  const uint32_t p = static_cast<uint32_t>((bucket_of(k, nb) * P) / nb);
  return static_cast<int>((p % kXcds) * (P / kXcds) + p / kXcds);
}

__device__ __forceinline__ void block_range(int64_t n, int64_t& lo, int64_t& hi) {
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  lo = static_cast<int64_t>(blockIdx.x) * chunk;
  hi = lo + chunk < n ? lo + chunk : n;
  if (lo > n) lo = n;
}

__global__ __launch_bounds__(kJBlock) void k_part_hist(TableView t, kg_col key, const int32_t* sel, int64_t n,
                                                        int P, uint32_t* hist) {
  extern __shared__ uint32_t cnt[];
  for (int q = threadIdx.x; q < P; q += kJBlock) cnt[q] = 0;
  __syncthreads();
  int64_t lo, hi;
  block_range(n, lo, hi);
  constexpr int U = 8;  // loads in flight per thread
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += static_cast<int64_t>(kJBlock) * U) {
    int q[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + static_cast<int64_t>(u) * kJBlock;
      int64_t k;
      q[u] = i < hi ? key_region(t, key, sel ? sel[i] : i, P, k) : -1;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (q[u] >= 0) atomicAdd(&cnt[q[u]], 1u);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < P; q += kJBlock) hist[static_cast<int64_t>(q) * gridDim.x + blockIdx.x] = cnt[q];
}

// Exclusive scan of m counts in one workgroup of 1024 threads; out[m] = total.
__global__ __launch_bounds__(1024) void k_scan_u32(const uint32_t* in, uint32_t* out, int64_t m) {
  __shared__ uint32_t part[1024];
  const int64_t per = (m + 1023) / 1024;
  const int64_t lo = threadIdx.x * per;
  const int64_t hi = lo + per < m ? lo + per : m;
  uint32_t sum = 0;
  for (int64_t i = lo; i < hi; i++) sum += in[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int s = 1; s < 1024; s <<= 1) {
    const uint32_t v = threadIdx.x >= s ? part[threadIdx.x - s] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - sum;
  for (int64_t i = lo; i < hi; i++) {
    const uint32_t c = in[i];
    out[i] = run;
    run += c;
  }
  if (threadIdx.x == 1023) out[m] = part[1023];
}

// Exclusive prefix sums of cnt[0..P) into off[0..P) by one kJBlock-thread
// workgroup: each thread sums a run of consecutive counters, the run totals
// are scanned across the wave (shuffles) and the waves (wtot in LDS), and each
// thread writes its run's offsets. Ends with the workgroup synchronized.
template <int NT = kJBlock>
__device__ __forceinline__ void block_excl_scan(const uint32_t* cnt, uint32_t* off, int P, uint32_t* wtot) {
  const int per = (P + NT - 1) / NT;
  const int lo = min(static_cast<int>(threadIdx.x) * per, P), hi = min(lo + per, P);
  uint32_t s = 0;
  for (int q = lo; q < hi; q++) s += cnt[q];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  uint32_t incl = s;
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d);
    if (lane >= d) incl += o;
  }
  if (lane == kWave - 1) wtot[wave] = incl;
  __syncthreads();
  uint32_t run = incl - s;
  for (int w = 0; w < wave; w++) run += wtot[w];
  for (int q = lo; q < hi; q++) {
    off[q] = run;
    run += cnt[q];
  }
  __syncthreads();
}

// Scatter: each tile of TILE rows is counting-sorted by region in LDS and
// written out as one contiguous run per region (the workgroup's running cursor
// per region starts at its scanned histogram offset).
template <bool VALS, bool ROWS, int TILE = kPartTile>
__global__ __launch_bounds__(kJBlock) void k_part_scatter(TableView t, kg_col key, const int32_t* sel,
                                                           const int64_t* vals, int64_t n, int P,
                                                           const uint32_t* offs, int64_t* okeys, int64_t* ovals,
                                                           int32_t* orows) {
  extern __shared__ unsigned char smem[];
  __shared__ uint32_t wtot[kJBlock / kWave];
  uint32_t* cursor = reinterpret_cast<uint32_t*>(smem);
  uint32_t* lcnt = cursor + P;
  uint32_t* loff = lcnt + P;
  int64_t* tkeys = reinterpret_cast<int64_t*>(loff + P + (P & 1));
  int64_t* tvals = tkeys + TILE;
  int32_t* trows = reinterpret_cast<int32_t*>(tvals + (VALS ? TILE : 0));
  uint16_t* treg = reinterpret_cast<uint16_t*>(trows + (ROWS ? TILE : 0));
  constexpr int kPer = TILE / kJBlock;
  for (int q = threadIdx.x; q < P; q += kJBlock) cursor[q] = offs[static_cast<int64_t>(q) * gridDim.x + blockIdx.x];
  int64_t lo, hi;
  block_range(n, lo, hi);
  for (int64_t t0 = lo; t0 < hi; t0 += TILE) {
    for (int q = threadIdx.x; q < P; q += kJBlock) lcnt[q] = 0;
    __syncthreads();
    int reg[kPer];
    uint32_t rank[kPer];
    int64_t kk[kPer], rr[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int64_t i = t0 + threadIdx.x + static_cast<int64_t>(j) * kJBlock;
      reg[j] = -1;
      rr[j] = 0;
      if (i < hi) {
        rr[j] = sel ? sel[i] : i;
        reg[j] = key_region(t, key, rr[j], P, kk[j]);
        if (reg[j] >= 0) rank[j] = atomicAdd(&lcnt[reg[j]], 1u);
      }
    }
    __syncthreads();
    block_excl_scan(lcnt, loff, P, wtot);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (reg[j] < 0) continue;
      const uint32_t slot = loff[reg[j]] + rank[j];
      tkeys[slot] = kk[j];
      if (VALS) tvals[slot] = vals[rr[j]];
      if (ROWS) trows[slot] = static_cast<int32_t>(rr[j]);
      treg[slot] = static_cast<uint16_t>(reg[j]);
    }
    __syncthreads();
    const uint32_t total = loff[P - 1] + lcnt[P - 1];
    for (uint32_t slot = threadIdx.x; slot < total; slot += kJBlock) {
      const int q = treg[slot];
      const uint32_t g = cursor[q] + (slot - loff[q]);
      okeys[g] = tkeys[slot];
      if (VALS) ovals[g] = tvals[slot];
      if (ROWS) orows[g] = trows[slot];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < P; q += kJBlock) cursor[q] += lcnt[q];
    __syncthreads();
  }
}

// ---- partitioned build ------------------------------------------------------------
// A large build into a table beyond the caches would cost every row a random
// CAS and a random row-id store on lines far apart in HBM. Instead the
// (key, row) pairs are grouped by table region in two counting-sort passes
// and each region's slice of the table is built in one workgroup's LDS and
// written back with coalesced stores:
//   A  k_bp_hist      per workgroup (a contiguous range of input rows), the
//                     number of pairs per coarse region, and the key range
//                     that decides the table form (one pass over the keys);
//   B  k_bp_scatter1  the same rows, tile by tile counting-sorted by coarse
//                     region in LDS and written as one run per region;
//   C  k_bp_hist2     per chunk of a coarse region, the pairs per fine region;
//   D  k_bp_scatter2  each chunk counting-sorted by fine region;
//   E  k_bp_region    one workgroup per fine region: its slice of the table in
//                     LDS (free, or read back when the table already holds
//                     rows), one LDS atomic per pair on its bucket's fill count
//                     (occupied slots are always a prefix of a bucket, so the
//                     count is the first free slot), the slice written back.
// Fine region f of a key = multiply-high of its hash by PF, coarse region =
// f / (PF / PC). The table's bucket count (of the form in use) is a multiple
// of PF, so region f holds exactly the keys whose home buckets lie in
// [f m, (f + 1) m), m = buckets / PF: slices are disjoint. A key that finds
// every bucket from its home to the slice's end full is spilled, and inserted
// after all regions by the global-atomic path, which continues into the next
// slice exactly as a direct insert would; so is a wide table's free-marker key
// (it goes to the side list). The slots a key takes in a bucket may differ
// from a direct build's; the set of (key, row) per bucket chain does not.
constexpr int kBPCoarse = 128;         // coarse regions (at most)
constexpr int kBPTile = 8192;          // pairs per scatter tile
constexpr int kBPBlock = 256;          // threads of a histogram / region workgroup
constexpr int kBPSBlock = 512;         // threads of a scatter workgroup
constexpr int kBPHBlock = 1024;        // threads of a pass-A workgroup
// Dynamic LDS of a scatter workgroup: the tile's keys, row ids and bins.
constexpr size_t kBPScatterLds = kBPTile * (8 + 4 + 1);
constexpr int64_t kBPSliceBudget = 48 * 1024;  // LDS bytes of a region's slice (3 or more workgroups per CU)

struct BPGeo {
  uint64_t nbf;   // buckets of the table's form (16-slot buckets for a narrow table)
  int64_t m;      // buckets per fine region
  int32_t PF;     // fine regions, a power of two
  int32_t PC;     // coarse regions, a power of two dividing PF
  int32_t F2;     // fine regions per coarse region
};

__device__ __forceinline__ uint32_t bp_fine(const BPGeo& g, int64_t k) {
  return static_cast<uint32_t>(__umul64hi(mix64(static_cast<uint64_t>(k)), static_cast<uint64_t>(g.PF)));
}

// A workgroup's contiguous range of input rows, of even length (pairs of
// consecutive rows are read together).
__device__ __forceinline__ void bp_block_range(int64_t n, int64_t& lo, int64_t& hi) {
  int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  chunk += chunk & 1;
  lo = static_cast<int64_t>(blockIdx.x) * chunk;
  hi = lo + chunk < n ? lo + chunk : n;
  if (lo > n) lo = n;
}

// Build keys and row ids of input rows i and i + 1 (below hi). FAST: a
// non-NULL BIGINT key column, 16-byte aligned, no row selection, so both keys
// come from one 16-byte load. ok*: the row exists and its key is not NULL.
template <bool FAST>
__device__ __forceinline__ void bp_load2(const kg_col& key, const int32_t* sel, int64_t i, int64_t hi, int64_t& k0,
                                         int64_t& k1, int32_t& r0, int32_t& r1, bool& ok0, bool& ok1) {
  k0 = k1 = 0;
  if (FAST) {
    const int64_t* kp = static_cast<const int64_t*>(key.data);
    r0 = static_cast<int32_t>(i);
    r1 = static_cast<int32_t>(i + 1);
    ok0 = i < hi;
    ok1 = i + 1 < hi;
    if (ok1) {
      typedef long long v2i64 __attribute__((ext_vector_type(2)));
      const v2i64 v = __builtin_nontemporal_load(reinterpret_cast<const v2i64*>(kp + i));
      k0 = v[0];
      k1 = v[1];
    } else if (ok0) {
      k0 = __builtin_nontemporal_load(kp + i);
    }
    return;
  }
  r0 = i < hi ? (sel ? sel[i] : static_cast<int32_t>(i)) : 0;
  r1 = i + 1 < hi ? (sel ? sel[i + 1] : static_cast<int32_t>(i + 1)) : 0;
  ok0 = i < hi && load_key_nt(key, r0, k0);
  ok1 = i + 1 < hi && load_key_nt(key, r1, k1);
}

// Pass A. hist[c * gridDim.x + block] (and hist[PC * gridDim.x] = 0);
// range[0] / range[1] receive the smallest / largest non-NULL key (signed, by
// atomicMin / atomicMax).
template <bool FAST>
__global__ __launch_bounds__(kBPHBlock) void k_bp_hist(kg_col key, const int32_t* sel, int64_t n, BPGeo g,
                                                        uint32_t* hist, unsigned long long* range) {
  constexpr int NW = kBPHBlock / kWave;
  __shared__ uint32_t cnt[NW][kBPCoarse];
  __shared__ int64_t wlo[NW], whi[NW];
  for (int q = threadIdx.x; q < NW * kBPCoarse; q += kBPHBlock) (&cnt[0][0])[q] = 0;
  __syncthreads();
  const int wave = threadIdx.x / kWave;
  const int shift = __builtin_ctz(static_cast<unsigned>(g.F2));
  int64_t lo, hi;
  bp_block_range(n, lo, hi);
  int64_t kmin = INT64_MAX, kmax = INT64_MIN;
  constexpr int U = 4;  // row pairs in flight per thread
  for (int64_t i0 = lo + 2 * threadIdx.x; i0 < hi; i0 += static_cast<int64_t>(kBPHBlock) * 2 * U) {
    int64_t kk[2 * U];
    bool ok[2 * U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      int32_t r0, r1;
      bp_load2<FAST>(key, sel, i0 + static_cast<int64_t>(u) * 2 * kBPHBlock, hi, kk[2 * u], kk[2 * u + 1], r0, r1,
                     ok[2 * u], ok[2 * u + 1]);
    }
#pragma unroll
    for (int u = 0; u < 2 * U; u++) {
      if (!ok[u]) continue;
      kmin = kk[u] < kmin ? kk[u] : kmin;
      kmax = kk[u] > kmax ? kk[u] : kmax;
      atomicAdd(&cnt[wave][bp_fine(g, kk[u]) >> shift], 1u);
    }
  }
  for (int s = kWave / 2; s >= 1; s >>= 1) {
    const int64_t a = static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(kmin), s));
    const int64_t b = static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(kmax), s));
    kmin = a < kmin ? a : kmin;
    kmax = b > kmax ? b : kmax;
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    wlo[wave] = kmin;
    whi[wave] = kmax;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < g.PC; q += kBPHBlock) {
    uint32_t c = 0;
    for (int w = 0; w < NW; w++) c += cnt[w][q];
    hist[static_cast<int64_t>(q) * gridDim.x + blockIdx.x] = c;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) hist[static_cast<int64_t>(g.PC) * gridDim.x] = 0;  // the scan's last input
  if (threadIdx.x == 0) {
    for (int w = 1; w < NW; w++) {
      kmin = wlo[w] < kmin ? wlo[w] : kmin;
      kmax = whi[w] > kmax ? whi[w] : kmax;
    }
    if (kmin <= kmax) {
      atomicMin(reinterpret_cast<long long*>(&range[0]), static_cast<long long>(kmin));
      atomicMax(reinterpret_cast<long long*>(&range[1]), static_cast<long long>(kmax));
    }
  }
}

// Counting sort of one tile of up to kBPTile pairs (kk / rr, reg = bin or -1
// per item, kPer items per thread) into LDS, then one run per bin written at
// the bin's cursor; cursors advance by the bin counts. Ends synchronized.
template <int kPer, typename BinT, int NT = kBPSBlock>
__device__ __forceinline__ void bp_tile_scatter(const int64_t (&kk)[kPer], const int32_t (&rr)[kPer],
                                                const int (&reg)[kPer], int nbin, uint32_t* cursor, uint32_t* lcnt,
                                                uint32_t* loff, uint32_t* wtot, int64_t* tkeys, int32_t* trows,
                                                BinT* tbin, int64_t* okeys, int32_t* orows) {
  uint32_t rank[kPer];
#pragma unroll
  for (int j = 0; j < kPer; j++)
    if (reg[j] >= 0) rank[j] = atomicAdd(&lcnt[reg[j]], 1u);
  __syncthreads();
  block_excl_scan<NT>(lcnt, loff, nbin, wtot);
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    if (reg[j] < 0) continue;
    const uint32_t slot = loff[reg[j]] + rank[j];
    tkeys[slot] = kk[j];
    trows[slot] = rr[j];
    tbin[slot] = static_cast<BinT>(reg[j]);
  }
  __syncthreads();
  const uint32_t total = loff[nbin - 1] + lcnt[nbin - 1];
  for (uint32_t slot = threadIdx.x; slot < total; slot += NT) {
    const int q = tbin[slot];
    const uint32_t o = cursor[q] + (slot - loff[q]);
    okeys[o] = tkeys[slot];
    orows[o] = trows[slot];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nbin; q += NT) {
    cursor[q] += lcnt[q];
    lcnt[q] = 0;
  }
  __syncthreads();
}

// Pass B: rows of the workgroup's range into coarse runs at the scanned pass-A
// offsets (off[c * gridDim.x + block]).
template <bool FAST>
__global__ __launch_bounds__(kBPSBlock) void k_bp_scatter1(kg_col key, const int32_t* sel, int64_t n, BPGeo g,
                                                           const uint32_t* off, int64_t* okeys, int32_t* orows) {
  __shared__ uint32_t cursor[kBPCoarse], lcnt[kBPCoarse], loff[kBPCoarse];
  __shared__ uint32_t wtot[kBPSBlock / kWave];
  extern __shared__ int64_t tkeys[];  // kBPTile keys, then kBPTile row ids and bins
  int32_t* trows = reinterpret_cast<int32_t*>(tkeys + kBPTile);
  uint8_t* tbin = reinterpret_cast<uint8_t*>(trows + kBPTile);
  constexpr int kPer = kBPTile / kBPSBlock;
  const int shift = __builtin_ctz(static_cast<unsigned>(g.F2));
  for (int q = threadIdx.x; q < g.PC; q += kBPSBlock) {
    cursor[q] = off[static_cast<int64_t>(q) * gridDim.x + blockIdx.x];
    lcnt[q] = 0;
  }
  __syncthreads();
  int64_t lo, hi;
  bp_block_range(n, lo, hi);
  for (int64_t t0 = lo; t0 < hi; t0 += kBPTile) {
    int64_t kk[kPer];
    int32_t rr[kPer];
    int reg[kPer];
#pragma unroll
    for (int j = 0; j < kPer / 2; j++) {
      bool ok0, ok1;
      bp_load2<FAST>(key, sel, t0 + 2 * threadIdx.x + static_cast<int64_t>(j) * 2 * kBPSBlock, hi, kk[2 * j],
                     kk[2 * j + 1], rr[2 * j], rr[2 * j + 1], ok0, ok1);
      reg[2 * j] = ok0 ? static_cast<int>(bp_fine(g, kk[2 * j]) >> shift) : -1;
      reg[2 * j + 1] = ok1 ? static_cast<int>(bp_fine(g, kk[2 * j + 1]) >> shift) : -1;
    }
    bp_tile_scatter<kPer>(kk, rr, reg, g.PC, cursor, lcnt, loff, wtot, tkeys, trows, tbin, okeys, orows);
  }
}

// Workgroup (c, q) of passes C and D: the q-th of Q chunks of coarse region
// c's pairs (off1: the scanned pass-A histogram over nblk1 workgroups).
__device__ __forceinline__ void bp_chunk(const uint32_t* off1, int nblk1, int Q, int& c, int& q, int64_t& lo,
                                         int64_t& hi) {
  c = blockIdx.x / Q;
  q = blockIdx.x % Q;
  const int64_t s = off1[static_cast<int64_t>(c) * nblk1], e = off1[static_cast<int64_t>(c + 1) * nblk1];
  lo = s + (e - s) * q / Q;
  hi = s + (e - s) * (q + 1) / Q;
}

constexpr int kBPMaxF2 = 256;  // fine regions per coarse region (at most)

// Pass C: hist2[(c * F2 + fl) * Q + q] = pairs of chunk (c, q) in fine region
// c * F2 + fl (and hist2[PF * Q] = 0); clears the spill count of pass E.
__global__ __launch_bounds__(kBPHBlock) void k_bp_hist2(BPGeo g, const int64_t* keys, int nblk1, int Q,
                                                        const uint32_t* off1, uint32_t* hist2,
                                                        unsigned long long* spill_count) {
  __shared__ uint32_t fcnt[kBPMaxF2];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hist2[static_cast<int64_t>(g.PF) * Q] = 0;  // the scan's last input
    *spill_count = 0;
  }
  for (int f = threadIdx.x; f < g.F2; f += kBPHBlock) fcnt[f] = 0;
  __syncthreads();
  int c, q;
  int64_t lo, hi;
  bp_chunk(off1, nblk1, Q, c, q, lo, hi);
  const uint32_t base = static_cast<uint32_t>(c) * static_cast<uint32_t>(g.F2);
  constexpr int U = 8;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += static_cast<int64_t>(kBPHBlock) * U) {
    int64_t kk[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + static_cast<int64_t>(u) * kBPHBlock;
      kk[u] = i < hi ? __builtin_nontemporal_load(keys + i) : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t fl = bp_fine(g, kk[u]) - base;
      if (i0 + static_cast<int64_t>(u) * kBPHBlock < hi && fl < static_cast<uint32_t>(g.F2)) atomicAdd(&fcnt[fl], 1u);
    }
  }
  __syncthreads();
  for (int f = threadIdx.x; f < g.F2; f += kBPHBlock) hist2[(static_cast<int64_t>(base) + f) * Q + q] = fcnt[f];
}

// Pass D: chunk (c, q) counting-sorted by fine region into okeys / orows at
// the scanned pass-C offsets.
__global__ __launch_bounds__(kBPSBlock) void k_bp_scatter2(BPGeo g, const int64_t* keys, const int32_t* rows,
                                                           int nblk1, int Q, const uint32_t* off1,
                                                           const uint32_t* off2, int64_t* okeys, int32_t* orows) {
  __shared__ uint32_t cursor[kBPMaxF2], lcnt[kBPMaxF2], loff[kBPMaxF2];
  __shared__ uint32_t wtot[kBPSBlock / kWave];
  extern __shared__ int64_t tkeys[];  // kBPTile keys, then kBPTile row ids and bins
  int32_t* trows = reinterpret_cast<int32_t*>(tkeys + kBPTile);
  uint8_t* tbin = reinterpret_cast<uint8_t*>(trows + kBPTile);
  constexpr int kPer = kBPTile / kBPSBlock;
  int c, q;
  int64_t lo, hi;
  bp_chunk(off1, nblk1, Q, c, q, lo, hi);
  const uint32_t base = static_cast<uint32_t>(c) * static_cast<uint32_t>(g.F2);
  for (int f = threadIdx.x; f < g.F2; f += kBPSBlock) {
    cursor[f] = off2[(static_cast<int64_t>(base) + f) * Q + q];
    lcnt[f] = 0;
  }
  __syncthreads();
  for (int64_t t0 = lo; t0 < hi; t0 += kBPTile) {
    int64_t kk[kPer];
    int32_t rr[kPer];
    int reg[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int64_t i = t0 + threadIdx.x + static_cast<int64_t>(j) * kBPSBlock;
      const bool in = i < hi;
      kk[j] = in ? __builtin_nontemporal_load(keys + i) : 0;
      rr[j] = in ? __builtin_nontemporal_load(rows + i) : 0;
      const uint32_t fl = bp_fine(g, kk[j]) - base;
      reg[j] = (in && fl < static_cast<uint32_t>(g.F2)) ? static_cast<int>(fl) : -1;
    }
    bp_tile_scatter<kPer>(kk, rr, reg, g.F2, cursor, lcnt, loff, wtot, tkeys, trows, tbin, okeys, orows);
  }
}

// Pass E: workgroup f builds buckets [f m, (f + 1) m) of the table from fine
// region f's pairs [off2[f * Q], off2[(f + 1) * Q]). fresh: the table holds no
// rows yet, so the slice starts free (and every slot of the table is written,
// free markers included); otherwise it is read back. Spilled pairs go to
// spill_keys / spill_rows.
template <bool NARROW>
__global__ __launch_bounds__(kBPBlock) void k_bp_region(TableView t, BPGeo g, int Q, const uint32_t* off2,
                                                         const int64_t* keys, const int32_t* rows, int fresh,
                                                         unsigned long long* inserted, int64_t* spill_keys,
                                                         int32_t* spill_rows, unsigned long long* spill_count) {
  using KeyT = typename std::conditional<NARROW, uint32_t, unsigned long long>::type;
  constexpr int S = NARROW ? kSlots32 : kSlots;
  const KeyT kFree = NARROW ? static_cast<KeyT>(kEmpty32) : static_cast<KeyT>(kEmpty);
  extern __shared__ unsigned long long lds_words[];
  const int64_t m = g.m;
  const int64_t f = blockIdx.x;
  KeyT* lkey = reinterpret_cast<KeyT*>(lds_words);
  int32_t* lrow = reinterpret_cast<int32_t*>(lkey + m * S);
  int32_t* lfill = lrow + m * S;
  KeyT* gkey = NARROW ? reinterpret_cast<KeyT*>(keys32(t)) : reinterpret_cast<KeyT*>(t.keys);
  const int64_t g0 = f * m * S;  // first slot of the slice
  const int64_t ns = m * S;
  {
    uint4* lk4 = reinterpret_cast<uint4*>(lkey);
    uint4* lr4 = reinterpret_cast<uint4*>(lrow);
    const uint4* gk4 = reinterpret_cast<const uint4*>(gkey + g0);
    const uint4* gr4 = reinterpret_cast<const uint4*>(t.rows + g0);
    const int64_t nk4 = ns * static_cast<int64_t>(sizeof(KeyT)) / 16, nr4 = ns / 4;
    const uint4 free4 = NARROW ? make_uint4(kEmpty32, kEmpty32, kEmpty32, kEmpty32)
                               : make_uint4(0u, 0x80000000u, 0u, 0x80000000u);
    if (fresh) {
      for (int64_t i = threadIdx.x; i < nk4; i += kBPBlock) lk4[i] = free4;
      for (int64_t i = threadIdx.x; i < nr4; i += kBPBlock) lr4[i] = make_uint4(0u, 0u, 0u, 0u);
      for (int64_t b = threadIdx.x; b < m; b += kBPBlock) lfill[b] = 0;
    } else {
      for (int64_t i = threadIdx.x; i < nk4; i += kBPBlock) lk4[i] = gk4[i];
      for (int64_t i = threadIdx.x; i < nr4; i += kBPBlock) lr4[i] = gr4[i];
    }
  }
  __syncthreads();
  if (!fresh) {
    // Fill count of each bucket: its first free slot.
    for (int64_t b = threadIdx.x; b < m; b += kBPBlock) {
      int c = 0;
      while (c < S && lkey[b * S + c] != kFree) c++;
      lfill[b] = c;
    }
    __syncthreads();
  }
  const int64_t a = off2[f * Q], e = off2[(f + 1) * Q];
  const uint64_t first = static_cast<uint64_t>(f * m);
  // One pair: the first bucket from its home with a free slot (its fill count
  // below S), else the spill list.
  auto insert_one = [&](int64_t k, int32_t r) {
    bool ok;
    KeyT kv;
    if (NARROW) {
      uint32_t k32;
      ok = narrow_key(t, k, k32);
      kv = static_cast<KeyT>(k32);
    } else {
      ok = k != kEmpty;
      kv = static_cast<KeyT>(k);
    }
    const int64_t home = static_cast<int64_t>(bucket_of(k, g.nbf)) - static_cast<int64_t>(first);
    if (ok && home >= 0) {
      for (int64_t bb = home; bb < m; bb++) {
        const int c = atomicAdd(&lfill[bb], 1);
        if (c < S) {
          lkey[bb * S + c] = kv;
          lrow[bb * S + c] = r;
          return;
        }
      }
    }
    const unsigned long long pos = atomicAdd(spill_count, 1ull);
    spill_keys[pos] = k;
    spill_rows[pos] = r;
  };
  constexpr int U = 8;  // pairs in flight per thread
  for (int64_t i0 = a + threadIdx.x; i0 < e; i0 += static_cast<int64_t>(kBPBlock) * U) {
    int64_t kk[U];
    int32_t rr[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + static_cast<int64_t>(u) * kBPBlock;
      kk[u] = i < e ? __builtin_nontemporal_load(keys + i) : 0;
      rr[u] = i < e ? __builtin_nontemporal_load(rows + i) : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (i0 + static_cast<int64_t>(u) * kBPBlock < e) insert_one(kk[u], rr[u]);
  }
  __syncthreads();
  {
    const uint4* lk4 = reinterpret_cast<const uint4*>(lkey);
    const uint4* lr4 = reinterpret_cast<const uint4*>(lrow);
    uint4* gk4 = reinterpret_cast<uint4*>(gkey + g0);
    uint4* gr4 = reinterpret_cast<uint4*>(t.rows + g0);
    const int64_t nk4 = ns * static_cast<int64_t>(sizeof(KeyT)) / 16, nr4 = ns / 4;
    for (int64_t i = threadIdx.x; i < nk4; i += kBPBlock) gk4[i] = lk4[i];
    for (int64_t i = threadIdx.x; i < nr4; i += kBPBlock) gr4[i] = lr4[i];
  }
  if (threadIdx.x == 0 && e > a) atomicAdd(inserted, static_cast<unsigned long long>(e - a));
}

// Inserts the spilled (key, row) pairs with the global-atomic path.
__global__ __launch_bounds__(kJBlock) void k_build_spill(TableView t, const int64_t* keys, const int32_t* rows,
                                                          const unsigned long long* count, int* overflow) {
  const int64_t n = static_cast<int64_t>(*count);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kJBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kJBlock + threadIdx.x; i < n; i += stride) {
    if (t.narrow) insert_narrow(t, keys[i], rows[i], overflow);
    else insert_wide(t, keys[i], rows[i], overflow);
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
__global__ __launch_bounds__(kJBlock) void k_collect(TableView t, const Acc* acc, const int32_t* first,
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
        out_key[p] = s == side_key_slot ? kEmpty
                     : (t.narrow ? static_cast<int64_t>(static_cast<uint64_t>(t.base) + keys32(t)[s]) : t.keys[s]);
        out_sum[p] = static_cast<int64_t>(acc[s].sum);
        out_cnt[p] = static_cast<int64_t>(acc[s].count);
        if (out_first) out_first[p] = first[s];
      }
    }
    pos += tot;
  }
}

// ---- probe + GROUP BY matched key, partitioned by table region -------------------
// The atomic form above adds every matched probe row into a per-slot
// accumulator array the size of the table: two device atomics per match on
// random lines of an array far beyond the caches. Here the probe rows are
// first grouped by table region (fine region f = the multiply-high of the
// key's hash by PF, which fixes a range of home buckets), in two counting-sort
// passes (coarse regions
// of kAggFine fine regions, then fine regions inside each coarse one); then
// one workgroup per fine region probes its rows against that slice of the
// table (read from L2), keeps SUM and COUNT per matched slot in LDS, and writes
// its groups straight to the output. All rows of a key share its home bucket,
// so a key's group is owned by exactly one region, also when the key was
// placed in a later bucket of its chain: the LDS window covers kAggMargin
// buckets past the region's last home bucket, and a match beyond that (or a
// wide table's free-marker key, whose rows live in the side list) sets a flag
// on which the host runs the atomic form instead.
// Frequent keys (the k_hot_keys sample) skip the partition: each scatter
// workgroup sums their rows in LDS and adds the totals to a small device
// array, which the region owning each such key folds in before it writes.
constexpr int kAggFine = 128;          // fine regions per coarse region
constexpr int kAggMargin = 8;          // chain buckets past a region covered by its window
constexpr int kAggTile = 2048;         // rows per scatter tile
constexpr int kAggBlock = 1024;        // threads of a region workgroup
constexpr int64_t kAggLdsBudget = 96 * 1024;

struct AggGeo {
  uint64_t nb;  // buckets (16-slot buckets for a narrow table)
  int32_t PF;   // fine regions, a multiple of kAggFine
  int32_t S;    // slots per bucket
};

// Fine region of a key: the multiply-high of the key's hash (the one that
// picks its home bucket) by PF, so region f holds exactly the keys whose home
// buckets lie in [floor(f * nb / PF), ceil((f + 1) * nb / PF)); neighbouring
// regions may share a boundary bucket (a key's group still has one owner).
__device__ __forceinline__ int agg_fine(const AggGeo& g, int64_t k) {
  return static_cast<int>(__umul64hi(mix64(static_cast<uint64_t>(k)), static_cast<uint64_t>(g.PF)));
}

// Probe key of row r; false when the row cannot match (NULL, NaN, outside a
// narrow table's range). A wide table's free-marker key sets *fallback.
__device__ __forceinline__ bool agg_key(const TableView& t, const kg_col& key, int64_t r, int64_t& k,
                                        int* fallback) {
  if (!load_key_nt(key, r, k)) return false;
  if (t.narrow) {
    uint32_t k32;
    return narrow_key(t, k, k32);
  }
  if (k == kEmpty) {
    atomicOr(fallback, 1);
    return false;
  }
  return true;
}

// The hot-key list in LDS: an open-addressing index (positions by hot_pos,
// claimed in parallel with LDS CAS; kHotHash > kHot leaves a free entry, so
// every search ends). Returns the list length. Ends synchronized.
__device__ __forceinline__ int hot_setup(const int64_t* hot, const int32_t* hot_n, int32_t* hidx, int64_t* hkey) {
  const int nhot = hot ? *hot_n : 0;
  for (int e = threadIdx.x; e < kHotHash; e += blockDim.x) hidx[e] = -1;
  for (int j = threadIdx.x; j < nhot; j += blockDim.x) hkey[j] = hot[j];
  __syncthreads();
  for (int j = threadIdx.x; j < nhot; j += blockDim.x) {
    uint32_t p = hot_pos(hkey[j]);
    for (int probes = 0; probes < kHotHash; probes++) {
      if (atomicCAS(&hidx[p], -1, j) == -1) break;
      p = (p + 1) & (kHotHash - 1);
    }
  }
  __syncthreads();
  return nhot;
}

__device__ __forceinline__ int hot_find(const int32_t* hidx, const int64_t* hkey, int nhot, int64_t k) {
  if (nhot == 0) return -1;
  uint32_t p = hot_pos(k);
  for (int probes = 0; probes < kHotHash; probes++) {
    const int j = hidx[p];
    if (j < 0) return -1;
    if (hkey[j] == k) return j;
    p = (p + 1) & (kHotHash - 1);
  }
  return -1;
}

// Pass 1: per workgroup (a contiguous range of probe rows), the number of
// non-hot matchable rows per coarse region; hist[c * gridDim.x + block].
__global__ __launch_bounds__(kJBlock) void k_agg_hist(TableView t, kg_col key, int64_t n, AggGeo g, int PC,
                                                       const int64_t* hot, const int32_t* hot_n, uint32_t* hist,
                                                       int* fallback) {
  extern __shared__ uint32_t ccnt[];  // [kJBlock / kWave][PC]: one histogram per wave
  __shared__ int32_t hidx[kHotHash];
  __shared__ int64_t hkey[kHot];
  constexpr int NW = kJBlock / kWave;
  for (int q = threadIdx.x; q < NW * PC; q += kJBlock) ccnt[q] = 0;
  const int nhot = hot_setup(hot, hot_n, hidx, hkey);
  uint32_t* wcnt = ccnt + (threadIdx.x / kWave) * PC;
  int64_t lo, hi;
  block_range(n, lo, hi);
  constexpr int U = 8;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += static_cast<int64_t>(kJBlock) * U) {
    int64_t kk[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + static_cast<int64_t>(u) * kJBlock;
      ok[u] = i < hi && agg_key(t, key, i, kk[u], fallback);
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (ok[u] && hot_find(hidx, hkey, nhot, kk[u]) < 0) atomicAdd(&wcnt[agg_fine(g, kk[u]) / kAggFine], 1u);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < PC; q += kJBlock) {
    uint32_t c = 0;
    for (int w = 0; w < NW; w++) c += ccnt[w * PC + q];
    hist[static_cast<int64_t>(q) * gridDim.x + blockIdx.x] = c;
  }
}

// Pass 2: the same rows, tile by tile counting-sorted by coarse region in LDS
// and written as one run per region from the workgroup's scanned offsets
// (okeys/ovals); hot rows are summed in LDS instead and added to hot_acc
// ({sum, count} per hot key) when the workgroup ends.
__global__ __launch_bounds__(kJBlock) void k_agg_scatter1(TableView t, kg_col key, const int64_t* vals, int64_t n,
                                                           AggGeo g, int PC, const int64_t* hot, const int32_t* hot_n,
                                                           const uint32_t* off, int64_t* okeys, int64_t* ovals,
                                                           unsigned long long* hot_acc, int* fallback) {
  extern __shared__ unsigned char smem[];
  __shared__ uint32_t wtot[kJBlock / kWave];
  __shared__ int32_t hidx[kHotHash];
  __shared__ int64_t hkey[kHot];
  __shared__ unsigned long long hsum[kHot], hcnt[kHot];
  uint32_t* cursor = reinterpret_cast<uint32_t*>(smem);
  uint32_t* lcnt = cursor + PC;
  uint32_t* loff = lcnt + PC;
  int64_t* tkeys = reinterpret_cast<int64_t*>(smem + ((static_cast<size_t>(3 * PC) * 4 + 15) / 16) * 16);
  int64_t* tvals = tkeys + kAggTile;
  uint16_t* treg = reinterpret_cast<uint16_t*>(tvals + kAggTile);
  constexpr int kPer = kAggTile / kJBlock;
  for (int q = threadIdx.x; q < PC; q += kJBlock) cursor[q] = off[static_cast<int64_t>(q) * gridDim.x + blockIdx.x];
  for (int j = threadIdx.x; j < kHot; j += kJBlock) {
    hsum[j] = 0;
    hcnt[j] = 0;
  }
  const int nhot = hot_setup(hot, hot_n, hidx, hkey);
  int64_t lo, hi;
  block_range(n, lo, hi);
  for (int64_t t0 = lo; t0 < hi; t0 += kAggTile) {
    for (int q = threadIdx.x; q < PC; q += kJBlock) lcnt[q] = 0;
    __syncthreads();
    int reg[kPer];
    uint32_t rank[kPer];
    int64_t kk[kPer], vv[kPer];
    bool ok[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int64_t i = t0 + threadIdx.x + static_cast<int64_t>(j) * kJBlock;
      ok[j] = i < hi && agg_key(t, key, i, kk[j], fallback);
      vv[j] = ok[j] ? __builtin_nontemporal_load(vals + i) : 0;
    }
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      reg[j] = -1;
      if (!ok[j]) continue;
      const int h = hot_find(hidx, hkey, nhot, kk[j]);
      if (h >= 0) {
        atomicAdd(&hsum[h], static_cast<unsigned long long>(vv[j]));
        atomicAdd(&hcnt[h], 1ull);
        continue;
      }
      reg[j] = agg_fine(g, kk[j]) / kAggFine;
      rank[j] = atomicAdd(&lcnt[reg[j]], 1u);
    }
    __syncthreads();
    block_excl_scan(lcnt, loff, PC, wtot);
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (reg[j] < 0) continue;
      const uint32_t slot = loff[reg[j]] + rank[j];
      tkeys[slot] = kk[j];
      tvals[slot] = vv[j];
      treg[slot] = static_cast<uint16_t>(reg[j]);
    }
    __syncthreads();
    const uint32_t total = loff[PC - 1] + lcnt[PC - 1];
    for (uint32_t slot = threadIdx.x; slot < total; slot += kJBlock) {
      const int q = treg[slot];
      const uint32_t o = cursor[q] + (slot - loff[q]);
      okeys[o] = tkeys[slot];
      ovals[o] = tvals[slot];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < PC; q += kJBlock) cursor[q] += lcnt[q];
    __syncthreads();
  }
  for (int j = threadIdx.x; j < nhot; j += kJBlock) {
    if (hcnt[j] == 0) continue;
    atomicAdd(&hot_acc[2 * j], hsum[j]);
    atomicAdd(&hot_acc[2 * j + 1], hcnt[j]);
  }
}

// Workgroup (c, q) of passes 3 and 4: the q-th of Q chunks of coarse region
// c's rows (off1: the scanned pass-1 histogram, nblk1 workgroups).
__device__ __forceinline__ void agg_chunk(const uint32_t* off1, int nblk1, int Q, int& c, int& q, int64_t& lo,
                                          int64_t& hi) {
  c = blockIdx.x / Q;
  q = blockIdx.x % Q;
  const int64_t s = off1[static_cast<int64_t>(c) * nblk1], e = off1[static_cast<int64_t>(c + 1) * nblk1];
  lo = s + (e - s) * q / Q;
  hi = s + (e - s) * (q + 1) / Q;
}

// Pass 3: per chunk, the number of rows per fine region of its coarse region;
// hist2[(c * kAggFine + fl) * Q + q], so that one exclusive scan gives every
// (fine region, chunk) its output position.
__global__ __launch_bounds__(kJBlock) void k_agg_hist2(AggGeo g, const int64_t* keys, int nblk1, int Q,
                                                        const uint32_t* off1, uint32_t* hist2) {
  __shared__ uint32_t fcnt[kAggFine];
  for (int f = threadIdx.x; f < kAggFine; f += kJBlock) fcnt[f] = 0;
  __syncthreads();
  int c, q;
  int64_t lo, hi;
  agg_chunk(off1, nblk1, Q, c, q, lo, hi);
  constexpr int U = 4;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += static_cast<int64_t>(kJBlock) * U) {
    int64_t kk[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + static_cast<int64_t>(u) * kJBlock;
      kk[u] = i < hi ? keys[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i0 + static_cast<int64_t>(u) * kJBlock >= hi) continue;
      const uint32_t fl = static_cast<uint32_t>(agg_fine(g, kk[u]) - c * kAggFine);
      if (fl < static_cast<uint32_t>(kAggFine)) atomicAdd(&fcnt[fl], 1u);
    }
  }
  __syncthreads();
  for (int f = threadIdx.x; f < kAggFine; f += kJBlock)
    hist2[(static_cast<int64_t>(c) * kAggFine + f) * Q + q] = fcnt[f];
}

// Pass 4: each chunk counting-sorted by fine region (LDS tiles) into
// keys2/vals2 at the positions of the scanned pass-3 histogram.
__global__ __launch_bounds__(kJBlock) void k_agg_scatter2(AggGeo g, const int64_t* keys, const int64_t* vals,
                                                           int nblk1, int Q, const uint32_t* off1,
                                                           const uint32_t* off2, int64_t* okeys, int64_t* ovals) {
  __shared__ uint32_t cursor[kAggFine], lcnt[kAggFine], loff[kAggFine];
  __shared__ uint32_t wtot[kJBlock / kWave];
  __shared__ int64_t tkeys[kAggTile], tvals[kAggTile];
  __shared__ uint8_t treg[kAggTile];
  constexpr int kPer = kAggTile / kJBlock;
  int c, q;
  int64_t lo, hi;
  agg_chunk(off1, nblk1, Q, c, q, lo, hi);
  for (int f = threadIdx.x; f < kAggFine; f += kJBlock)
    cursor[f] = off2[(static_cast<int64_t>(c) * kAggFine + f) * Q + q];
  for (int64_t t0 = lo; t0 < hi; t0 += kAggTile) {
    for (int f = threadIdx.x; f < kAggFine; f += kJBlock) lcnt[f] = 0;
    __syncthreads();
    int reg[kPer];
    uint32_t rank[kPer];
    int64_t kk[kPer], vv[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int64_t i = t0 + threadIdx.x + static_cast<int64_t>(j) * kJBlock;
      kk[j] = i < hi ? keys[i] : 0;
      vv[j] = i < hi ? vals[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      reg[j] = -1;
      if (t0 + threadIdx.x + static_cast<int64_t>(j) * kJBlock >= hi) continue;
      const uint32_t fl = static_cast<uint32_t>(agg_fine(g, kk[j]) - c * kAggFine);
      if (fl >= static_cast<uint32_t>(kAggFine)) continue;
      reg[j] = static_cast<int>(fl);
      rank[j] = atomicAdd(&lcnt[fl], 1u);
    }
    __syncthreads();
    block_excl_scan(lcnt, loff, kAggFine, wtot);
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (reg[j] < 0) continue;
      const uint32_t slot = loff[reg[j]] + rank[j];
      tkeys[slot] = kk[j];
      tvals[slot] = vv[j];
      treg[slot] = static_cast<uint8_t>(reg[j]);
    }
    __syncthreads();
    const uint32_t total = loff[kAggFine - 1] + lcnt[kAggFine - 1];
    for (uint32_t slot = threadIdx.x; slot < total; slot += kJBlock) {
      const int f = treg[slot];
      const uint32_t o = cursor[f] + (slot - loff[f]);
      okeys[o] = tkeys[slot];
      ovals[o] = tvals[slot];
    }
    __syncthreads();
    for (int f = threadIdx.x; f < kAggFine; f += kJBlock) cursor[f] += lcnt[f];
    __syncthreads();
  }
}

// Matches of k in its bucket chain from home bucket b (the probe rule: until a
// bucket with a free slot): their number, and the first matching slot.
template <bool NARROW>
__device__ __forceinline__ unsigned agg_find(const TableView& t, uint64_t nb, int64_t k, uint64_t b,
                                             int64_t& first) {
  unsigned nm = 0;
  first = -1;
  if (NARROW) {
    uint32_t k32;
    if (!narrow_key(t, k, k32)) return 0;
    const uint32_t* keys = keys32(t);
    for (uint64_t scanned = 0; scanned < nb; scanned++) {
      const uint4* bk = reinterpret_cast<const uint4*>(keys + b * kSlots32);
      bool empty = false;
#pragma unroll
      for (int q = 0; q < kSlots32 / 4; q++) {
        const uint4 x = bk[q];
        const uint32_t v[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
          if (v[e] == k32) {
            if (first < 0) first = static_cast<int64_t>(b) * kSlots32 + 4 * q + e;
            nm++;
          }
          empty |= v[e] == kEmpty32;
        }
      }
      if (empty) break;
      b = (b + 1 == nb) ? 0 : b + 1;
    }
  } else {
    for (uint64_t scanned = 0; scanned < nb; scanned++) {
      const longlong2* bk = reinterpret_cast<const longlong2*>(t.keys + b * kSlots);
      bool empty = false;
#pragma unroll
      for (int q = 0; q < kSlots / 2; q++) {
        const longlong2 x = bk[q];
        if (x.x == k) {
          if (first < 0) first = static_cast<int64_t>(b) * kSlots + 2 * q;
          nm++;
        }
        if (x.y == k) {
          if (first < 0) first = static_cast<int64_t>(b) * kSlots + 2 * q + 1;
          nm++;
        }
        empty |= x.x == kEmpty || x.y == kEmpty;
      }
      if (empty) break;
      b = (b + 1 == nb) ? 0 : b + 1;
    }
  }
  return nm;
}

// Pass 5: workgroup f aggregates fine region f. LDS holds {sum} and {count}
// (64-bit) for the slots of the region's home buckets and the kAggMargin
// buckets after them; a matched row adds val * matches and matches to its
// key's first matching slot (the group of the atomic form). Hot keys of this
// region add their sampled totals the same way. Then the live slots are
// written out (one output range per workgroup, slot order inside it).
template <bool NARROW>
__global__ __launch_bounds__(kAggBlock) void k_agg_region(TableView t, AggGeo g, int Q, const uint32_t* off2,
                                                           const int64_t* keys, const int64_t* vals,
                                                           const int64_t* hot, const int32_t* hot_n,
                                                           const unsigned long long* hot_acc,
                                                           unsigned long long* count, int64_t cap, int64_t* out_key,
                                                           int64_t* out_sum, int64_t* out_cnt, int* fallback) {
  extern __shared__ unsigned long long lacc[];
  constexpr int NW = kAggBlock / kWave;
  __shared__ unsigned int wsum[NW];
  __shared__ unsigned int wc[2][NW];
  __shared__ unsigned long long obase;
  const uint64_t nb = g.nb;
  const int S = g.S;
  const uint64_t f = blockIdx.x;
  const uint64_t blo = (f * nb) / g.PF, bhi = ((f + 1) * nb + g.PF - 1) / g.PF;
  // The window may run past the table's last bucket onto its first ones (a
  // chain wraps around there); slot positions are taken modulo the table.
  const uint64_t wb = min(bhi - blo + kAggMargin, nb);
  const int64_t tslots = static_cast<int64_t>(nb) * S;
  const int64_t nsl = static_cast<int64_t>(wb) * S;
  const int64_t slot0 = static_cast<int64_t>(blo) * S;
  unsigned long long* lsum = lacc;
  unsigned long long* lcnt = lacc + nsl;
  for (int64_t s = threadIdx.x; s < 2 * nsl; s += kAggBlock) lacc[s] = 0;
  __syncthreads();
  const int64_t lo = off2[f * Q], hi = off2[(f + 1) * Q];
  constexpr int U = 4;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += static_cast<int64_t>(kAggBlock) * U) {
    int64_t kk[U], vv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + static_cast<int64_t>(u) * kAggBlock;
      kk[u] = i < hi ? __builtin_nontemporal_load(keys + i) : 0;
      vv[u] = i < hi ? __builtin_nontemporal_load(vals + i) : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i0 + static_cast<int64_t>(u) * kAggBlock >= hi) continue;
      int64_t first;
      const unsigned nm = agg_find<NARROW>(t, nb, kk[u], bucket_of(kk[u], nb), first);
      if (nm == 0) continue;
      int64_t rel = first - slot0;
      if (rel < 0) rel += tslots;
      if (rel >= nsl) {
        atomicOr(fallback, 1);
        continue;
      }
      atomicAdd(&lsum[rel], static_cast<unsigned long long>(vv[u]) * nm);
      atomicAdd(&lcnt[rel], static_cast<unsigned long long>(nm));
    }
  }
  const int nhot = hot ? *hot_n : 0;
  for (int j = threadIdx.x; j < nhot; j += kAggBlock) {
    const int64_t k = hot[j];
    if (hot_acc[2 * j + 1] == 0 || agg_fine(g, k) != static_cast<int>(f)) continue;
    int64_t first;
    const unsigned nm = agg_find<NARROW>(t, nb, k, bucket_of(k, nb), first);
    if (nm == 0) continue;
    int64_t rel = first - slot0;
    if (rel < 0) rel += tslots;
    if (rel >= nsl) {
      atomicOr(fallback, 1);
      continue;
    }
    atomicAdd(&lsum[rel], hot_acc[2 * j] * nm);
    atomicAdd(&lcnt[rel], hot_acc[2 * j + 1] * nm);
  }
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  unsigned int mine = 0;
  for (int64_t s = threadIdx.x; s < nsl; s += kAggBlock) mine += lcnt[s] > 0 ? 1u : 0u;
  for (int st = kWave / 2; st >= 1; st >>= 1) mine += __shfl_xor(mine, st);
  if (lane == 0) wsum[wave] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < NW; w++) tot += wsum[w];
    obase = tot ? atomicAdd(count, tot) : 0;
  }
  __syncthreads();
  unsigned long long pos = obase;
  int it = 0;
  for (int64_t s0 = 0; s0 < nsl; s0 += kAggBlock, it++) {
    const int64_t s = s0 + threadIdx.x;
    const bool live = s < nsl && lcnt[s] > 0;
    const uint64_t m = __ballot(live);
    if (lane == 0) wc[it & 1][wave] = __popcll(m);
    __syncthreads();
    unsigned int before = 0, tot = 0;
    for (int w = 0; w < NW; w++) {
      const unsigned int cw = wc[it & 1][w];
      before += w < wave ? cw : 0;
      tot += cw;
    }
    if (live) {
      const unsigned long long p = pos + before + __popcll(m & ((1ull << lane) - 1ull));
      if (p < static_cast<unsigned long long>(cap)) {
        const int64_t gs = slot0 + s < tslots ? slot0 + s : slot0 + s - tslots;
        out_key[p] = NARROW ? static_cast<int64_t>(static_cast<uint64_t>(t.base) + keys32(t)[gs]) : t.keys[gs];
        out_sum[p] = static_cast<int64_t>(lsum[s]);
        out_cnt[p] = static_cast<int64_t>(lcnt[s]);
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
  // Each LDS word is set by exactly one of these loops (two loops writing
  // one word from different waves would race without a barrier).
  const int zero_words = G * nm * 2 + G;  // sums, non-NULL counts, rows
  for (int i = threadIdx.x; i < zero_words; i += kJBlock) lds[i] = 0;
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
    table_matches(t, k, consume);
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

// Deterministic mode: the join output rows of the star pipeline as items in
// probe-row order. Pass 1 counts each probe row's matches whose group code is
// a group; pass 2 writes (group, probe row) per match at the row's offset.
// Matches of one probe row carry the same measure values, so their order among
// themselves does not change any group's sum.
__global__ void k_star_det_count(PredSet P, kg_col key, TableView t, const int32_t* codes, int G, int64_t n,
                                 int64_t* cnt) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r > n) return;
  if (r == n) {  // the scan's extra element
    cnt[n] = 0;
    return;
  }
  int64_t c = 0;
  int64_t k;
  if (eval_preds(P, r) && load_key(key, r, k)) {
    table_matches(t, k, [&](int32_t brow) {
      const int g = codes ? codes[brow] : 0;
      if (g >= 0 && g < G) c++;
    });
  }
  cnt[r] = c;
}

__global__ void k_star_det_items(PredSet P, kg_col key, TableView t, const int32_t* codes, int G, int64_t n,
                                 const int64_t* off, int32_t* gid, int32_t* src) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t o = off[r];
  const int64_t end = off[r + 1];
  int64_t k;
  if (o == end || !eval_preds(P, r) || !load_key(key, r, k)) return;
  table_matches(t, k, [&](int32_t brow) {
    const int g = codes ? codes[brow] : 0;
    if (g >= 0 && g < G && o < end) {
      gid[o] = g;
      src[o] = static_cast<int32_t>(r);
      o++;
    }
  });
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
  int64_t* hot = nullptr;      // probe_agg: hot probe keys (kHot) + their count, lazily allocated
  int32_t* first = nullptr;    // lazily allocated for probe_agg with first rows
  int32_t* side_rows = nullptr;
  unsigned long long* counters = nullptr;  // [0] inserted rows, [1] side count, [2] overflow flag, [3] scratch,
                                           // [4..5] key range of a build
  int device = 0;
  int form = -1;          // -1 until the first build with keys, then 0 wide / 1 narrow
  bool narrow_unfilled = false;  // narrow form chosen, its free markers not yet written
  int64_t alloc_slots = 0;       // slots the key / row memory holds (kg::ht_reset reuses it)
  int64_t acc_slots = 0;         // entries of acc (and first, when allocated)
  int64_t base = 0;       // narrow: smallest key
  TableView view() const {
    TableView v{keys, rows, nbuckets, side_rows, counters + 1, form == 1 ? 1 : 0, 0, base};
    return v;
  }
};

namespace {

// Table memory comes from the context's block cache (kg_dev_alloc) and goes
// back to it (kg_dev_release).
int alloc(kg_ctx* ctx, void** p, size_t bytes) { return kg_dev_alloc(ctx, p, bytes); }

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

// Converts a narrow table to the wide form, re-inserting its entries.
int widen(kg_ctx* ctx, kg_hash_table* ht) {
  const int64_t nslots = static_cast<int64_t>(ht->nbuckets) * kSlots;
  uint32_t* tk = nullptr;
  int32_t* tr = nullptr;
  int st;
  if ((st = alloc(ctx, reinterpret_cast<void**>(&tk), nslots * 4)) != KG_OK) return st;
  if ((st = alloc(ctx, reinterpret_cast<void**>(&tr), nslots * 4)) != KG_OK) {
    kg_dev_release(ctx, tk);
    return st;
  }
  auto done = [&](int s2) {
    kg_dev_release(ctx, tk);
    kg_dev_release(ctx, tr);
    return s2;
  };
  if (hipMemcpyAsync(tk, ht->keys, nslots * 4, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(tr, ht->rows, nslots * 4, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
    return done(kg_fail(ctx, KG_ERR_HIP, "copy of the narrow table failed"));
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(ctx, nslots, 4)), dim3(kJBlock), 0, ctx->stream, ht->keys, kEmpty,
                     nslots);
  ht->form = 0;
  hipLaunchKernelGGL(k_widen, dim3(grid_for(ctx, nslots, 4)), dim3(kJBlock), 0, ctx->stream, ht->view(), tk, tr,
                     nslots, ht->base, reinterpret_cast<int*>(ht->counters + 2));
  if (hipGetLastError() != hipSuccess) return done(kg_fail(ctx, KG_ERR_HIP, "widen launch failed"));
  return done(KG_OK);
}

// Picks the table's form before inserting a build batch: the first batch with
// integer keys whose range spans at most 2^32 - 2 makes the table narrow; a
// later batch outside that range (or a REAL key) turns it wide.
int choose_form(kg_ctx* ctx, kg_hash_table* ht, const kg_col* key, const int32_t* d_rows, int64_t nrows) {
  if (ht->form == 0) return KG_OK;
  if (key->type == KG_REAL) return ht->form == 1 ? widen(ctx, ht) : (ht->form = 0, KG_OK);
  unsigned long long* rng = ht->counters + 4;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, rng, static_cast<unsigned long long>(INT64_MAX),
                     1);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, rng + 1,
                     static_cast<unsigned long long>(INT64_MIN), 1);
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_key_range, dim3(std::min(grid_for(ctx, nrows, 16), ctx->num_cus * 8)), dim3(kJBlock), 0,
                     ctx->stream, *key, d_rows, nrows, rng);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  int st = kg_pinned_reserve(ctx, 64);
  if (st != KG_OK) return st;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(ctx->pinned, rng, 16, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  const int64_t lo = static_cast<const int64_t*>(ctx->pinned)[0], hi = static_cast<const int64_t*>(ctx->pinned)[1];
  if (lo > hi) return KG_OK;  // no non-NULL keys in this batch
  const uint64_t limit = static_cast<uint64_t>(kEmpty32) - 1;
  if (ht->form == -1) {
    if (static_cast<uint64_t>(hi) - static_cast<uint64_t>(lo) <= limit) {
      // The narrow free marker is written by kg_ht_build before it inserts
      // (or not at all when a fresh region build writes every slot).
      ht->form = 1;
      ht->base = lo;
      ht->narrow_unfilled = true;
    } else {
      ht->form = 0;
    }
    return KG_OK;
  }
  if (lo < ht->base || static_cast<uint64_t>(hi) - static_cast<uint64_t>(ht->base) > limit) return widen(ctx, ht);
  return KG_OK;
}

// Partitioned build (passes A-E above). Used for builds of at least 2^20 rows
// into tables of at least 32 MB (kg_ctx_set_build_regions: 0 never, > 0 for
// every build, with at least that many fine regions where the table allows).
bool use_part_build(const kg_ctx* ctx, const kg_hash_table* ht, int64_t nrows) {
  if (ctx->build_regions == 0) return false;
  if (ctx->build_regions > 0) return true;
  return nrows >= (1 << 20) && ht->nbuckets * kSlots * 8 >= (32ull << 20);
}

// Fine regions for a table form of nbf buckets of bpb LDS bytes each: the
// smallest power of two dividing nbf whose slices fit kBPSliceBudget (or the
// requested count), at most 32768.
int bp_regions(const kg_ctx* ctx, uint64_t nbf, int64_t bpb) {
  const uint64_t pmax = std::min<uint64_t>(nbf & (~nbf + 1), 32768);
  uint64_t pf = 1;
  while (pf < pmax && static_cast<int64_t>(nbf / pf) * bpb > kBPSliceBudget) pf <<= 1;
  while (pf < pmax && ctx->build_regions > 0 && pf < static_cast<uint64_t>(ctx->build_regions)) pf <<= 1;
  return static_cast<int>(pf);
}

// LDS bytes per bucket of a region slice: keys, row ids, fill count.
int64_t bp_bucket_bytes(bool narrow) { return narrow ? kSlots32 * 8 + 4 : kSlots * 12 + 4; }

// Sets the table form from the key range [lo, hi] of this build (the rule of
// choose_form); widens a narrow table the keys leave.
int form_from_range(kg_ctx* ctx, kg_hash_table* ht, int32_t key_type, int64_t lo, int64_t hi) {
  if (ht->form == 0) return KG_OK;
  if (key_type == KG_REAL) return ht->form == 1 ? widen(ctx, ht) : (ht->form = 0, KG_OK);
  if (lo > hi) return KG_OK;  // no non-NULL keys
  const uint64_t limit = static_cast<uint64_t>(kEmpty32) - 1;
  if (ht->form == -1) {
    if (static_cast<uint64_t>(hi) - static_cast<uint64_t>(lo) <= limit) {
      ht->form = 1;
      ht->base = lo;
      ht->narrow_unfilled = true;
    } else {
      ht->form = 0;
    }
    return KG_OK;
  }
  if (lo < ht->base || static_cast<uint64_t>(hi) - static_cast<uint64_t>(ht->base) > limit) return widen(ctx, ht);
  return KG_OK;
}

// Inserts directly (one CAS per row): the form from a key-range pass, the
// narrow free markers when first needed, then k_ht_build.
int direct_build(kg_ctx* ctx, kg_hash_table* ht, const kg_col* key, const int32_t* d_rows, int64_t nrows) {
  int st;
  if ((st = choose_form(ctx, ht, key, d_rows, nrows)) != KG_OK) return st;
  if (ht->narrow_unfilled) {
    ht->narrow_unfilled = false;
    const int64_t words = static_cast<int64_t>(ht->nbuckets) * kSlots / 2;  // 32-bit slots / 2
    kg_time_begin(ctx);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ctx, words, 4)), dim3(kJBlock), 0, ctx->stream,
                       reinterpret_cast<unsigned long long*>(ht->keys), ~0ull, words);
    kg_time_end(ctx);
    KG_LAUNCH_CHECK(ctx);
  }
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_ht_build, dim3(grid_for(ctx, nrows, 4)), dim3(kJBlock), 0, ctx->stream, ht->view(), *key,
                     d_rows, nrows, ht->counters, reinterpret_cast<int*>(ht->counters + 2));
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  return KG_OK;
}

// The partitioned build; falls back to direct_build when the table's bucket
// count cannot be split into slices that fit the LDS.
int part_build(kg_ctx* ctx, kg_hash_table* ht, const kg_col* key, const int32_t* sel, int64_t n, bool fresh) {
  const uint64_t nb = ht->nbuckets;
  const int pf_wide = bp_regions(ctx, nb, bp_bucket_bytes(false));
  const int pf_narrow = bp_regions(ctx, nb / 2, bp_bucket_bytes(true));
  const int PC = std::min(kBPCoarse, std::min(pf_wide, pf_narrow));
  const int nblk = static_cast<int>(std::min<int64_t>(4096, std::max<int64_t>(8, ceil_div(n, 4 * kBPTile))));
  const int Q = std::max(1, nblk / PC);
  const int pf_cap = std::max(pf_wide, pf_narrow);
  const int64_t m1 = static_cast<int64_t>(PC) * nblk, m2 = static_cast<int64_t>(pf_cap) * Q;
  size_t tmp1 = 0, tmp2 = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp1, static_cast<uint32_t*>(nullptr),
                                         static_cast<uint32_t*>(nullptr), static_cast<int>(m1 + 1), ctx->stream);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, static_cast<uint32_t*>(nullptr),
                                         static_cast<uint32_t*>(nullptr), static_cast<int>(m2 + 1), ctx->stream);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o2 = off;
    off += (bytes + 255) / 256 * 256;
    return o2;
  };
  const size_t o_h1 = carve((m1 + 1) * 4), o_o1 = carve((m1 + 1) * 4), o_h2 = carve((m2 + 1) * 4),
               o_o2 = carve((m2 + 1) * 4), o_k1 = carve(n * 8), o_r1 = carve(n * 4), o_k2 = carve(n * 8),
               o_r2 = carve(n * 4), o_sc = carve(8), o_tmp = carve(std::max(tmp1, tmp2));
  int st = kg_work_reserve(ctx, off);
  if (st != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  uint32_t* h1 = reinterpret_cast<uint32_t*>(w + o_h1);
  uint32_t* o1 = reinterpret_cast<uint32_t*>(w + o_o1);
  uint32_t* h2 = reinterpret_cast<uint32_t*>(w + o_h2);
  uint32_t* o2 = reinterpret_cast<uint32_t*>(w + o_o2);
  int64_t* k1 = reinterpret_cast<int64_t*>(w + o_k1);
  int32_t* r1 = reinterpret_cast<int32_t*>(w + o_r1);
  int64_t* k2 = reinterpret_cast<int64_t*>(w + o_k2);
  int32_t* r2 = reinterpret_cast<int32_t*>(w + o_r2);
  unsigned long long* scnt = reinterpret_cast<unsigned long long*>(w + o_sc);
  unsigned long long* rng = ht->counters + 4;
  // Pass A (coarse regions only: PF = PC, F2 = 1) and the key range.
  BPGeo g{0, 0, PC, PC, 1};
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, rng, static_cast<unsigned long long>(INT64_MAX),
                     1);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, rng + 1,
                     static_cast<unsigned long long>(INT64_MIN), 1);
  kg_time_begin(ctx);
  const bool fast = key->type == KG_BIGINT && !key->nulls && !sel && (reinterpret_cast<uintptr_t>(key->data) & 15) == 0;
  if (fast)
    hipLaunchKernelGGL(k_bp_hist<true>, dim3(nblk), dim3(kBPHBlock), 0, ctx->stream, *key, sel, n, g, h1, rng);
  else
    hipLaunchKernelGGL(k_bp_hist<false>, dim3(nblk), dim3(kBPHBlock), 0, ctx->stream, *key, sel, n, g, h1, rng);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  if ((st = kg_pinned_reserve(ctx, 64)) != KG_OK) return st;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(ctx->pinned, rng, 16, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  const int64_t lo = static_cast<const int64_t*>(ctx->pinned)[0], hi = static_cast<const int64_t*>(ctx->pinned)[1];
  if ((st = form_from_range(ctx, ht, key->type, lo, hi)) != KG_OK) return st;
  if (ht->form < 0) return KG_OK;  // every key is NULL: nothing to insert
  const bool narrow = ht->form == 1;
  g.nbf = narrow ? nb / 2 : nb;
  g.PF = std::min(narrow ? pf_narrow : pf_wide, PC * kBPMaxF2);
  g.F2 = g.PF / PC;
  g.m = static_cast<int64_t>(g.nbf / g.PF);
  const size_t lds = static_cast<size_t>(g.m * bp_bucket_bytes(narrow)) + 64;
  if (g.nbf % g.PF != 0 || static_cast<int64_t>(lds) > ctx->max_lds) return direct_build(ctx, ht, key, sel, n);
  const void* region_kernel =
      narrow ? reinterpret_cast<const void*>(&k_bp_region<true>) : reinterpret_cast<const void*>(&k_bp_region<false>);
  if ((st = kg_allow_lds(ctx, region_kernel, lds)) != KG_OK ||
      (st = kg_allow_lds(ctx, reinterpret_cast<const void*>(fast ? &k_bp_scatter1<true> : &k_bp_scatter1<false>),
                         kBPScatterLds)) != KG_OK ||
      (st = kg_allow_lds(ctx, reinterpret_cast<const void*>(&k_bp_scatter2), kBPScatterLds)) != KG_OK)
    return st;
  if (ht->narrow_unfilled) {
    ht->narrow_unfilled = false;
    if (!fresh) {  // slices are read back: the free markers must be there first
      const int64_t words = static_cast<int64_t>(nb) * kSlots / 2;
      hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ctx, words, 4)), dim3(kJBlock), 0, ctx->stream,
                         reinterpret_cast<unsigned long long*>(ht->keys), ~0ull, words);
    }
  }
  const TableView v = ht->view();
  kg_time_begin(ctx);
  size_t tb = tmp1;
  if (hipcub::DeviceScan::ExclusiveSum(w + o_tmp, tb, h1, o1, static_cast<int>(m1 + 1), ctx->stream) != hipSuccess)
    return kg_fail(ctx, KG_ERR_HIP, "build histogram scan failed");
  if (fast)
    hipLaunchKernelGGL(k_bp_scatter1<true>, dim3(nblk), dim3(kBPSBlock), kBPScatterLds, ctx->stream, *key, sel, n, g,
                       o1, k1, r1);
  else
    hipLaunchKernelGGL(k_bp_scatter1<false>, dim3(nblk), dim3(kBPSBlock), kBPScatterLds, ctx->stream, *key, sel, n,
                       g, o1, k1, r1);
  hipLaunchKernelGGL(k_bp_hist2, dim3(PC * Q), dim3(kBPHBlock), 0, ctx->stream, g, k1, nblk, Q, o1, h2, scnt);
  tb = tmp2;
  const int64_t mf = static_cast<int64_t>(g.PF) * Q;
  if (hipcub::DeviceScan::ExclusiveSum(w + o_tmp, tb, h2, o2, static_cast<int>(mf + 1), ctx->stream) != hipSuccess)
    return kg_fail(ctx, KG_ERR_HIP, "build region histogram scan failed");
  hipLaunchKernelGGL(k_bp_scatter2, dim3(PC * Q), dim3(kBPSBlock), kBPScatterLds, ctx->stream, g, k1, r1, nblk, Q, o1,
                     o2, k2,
                     r2);
  // The coarse pairs are consumed: their buffers take the spilled pairs.
  if (narrow)
    hipLaunchKernelGGL(k_bp_region<true>, dim3(g.PF), dim3(kBPBlock), lds, ctx->stream, v, g, Q, o2, k2, r2,
                       fresh ? 1 : 0, ht->counters, k1, r1, scnt);
  else
    hipLaunchKernelGGL(k_bp_region<false>, dim3(g.PF), dim3(kBPBlock), lds, ctx->stream, v, g, Q, o2, k2, r2,
                       fresh ? 1 : 0, ht->counters, k1, r1, scnt);
  hipLaunchKernelGGL(k_build_spill, dim3(ctx->num_cus), dim3(kJBlock), 0, ctx->stream, v, k1, r1, scnt,
                     reinterpret_cast<int*>(ht->counters + 2));
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  return KG_OK;
}

}  // namespace

namespace kg {

// Buckets of a table for capacity_rows at load_factor (an even count: the
// narrow form has half as many buckets of 16 slots).
// Tables of 2^15 buckets or more take a multiple of a power of two g of at
// least nb / 256 buckets (at most 0.8% more), so that the partitioned build
// can split either form into equal slices.
uint64_t table_buckets_for(int64_t capacity_rows, double load_factor) {
  const double slots = std::max<double>(static_cast<double>(capacity_rows) / load_factor, kSlots);
  uint64_t nb = static_cast<uint64_t>(ceil_div(static_cast<int64_t>(slots), kSlots));
  if (nb >= (uint64_t{1} << 15)) {
    uint64_t g = 1;
    while (g * 256 < nb) g <<= 1;
    nb = (nb + g - 1) / g * g;
  }
  return nb + (nb & 1);
}

double tuned_load_factor(const kg_ctx* ctx, int64_t build_rows) {
  int lg = 0;
  while ((int64_t{1} << lg) < build_rows && lg < 63) lg++;
  double lf = 0.5;
  int best = 1 << 30;
  for (const auto& t : ctx->join_tunes) {
    if (static_cast<int>(t.cls >> 8) != lg) continue;
    const int rank = static_cast<int>((t.cls >> 4) & 0xF) * 4 + static_cast<int>(t.cls & 0xF);  // unique, uniform first
    if (rank < best) {
      best = rank;
      lf = t.load_factor;
    }
  }
  return lf;
}

bool ht_fits(const kg_hash_table* t, int64_t capacity_rows, double load_factor) {
  if (load_factor == 0) load_factor = 0.5;
  return t && load_factor > 0 && load_factor <= 0.9 && capacity_rows >= 0 && capacity_rows < INT32_MAX &&
         static_cast<int64_t>(table_buckets_for(capacity_rows, load_factor) * kSlots) <= t->alloc_slots;
}

int ht_reset(kg_ctx* ctx, kg_hash_table* t, int64_t capacity_rows, double load_factor) {
  KG_REQUIRE(ctx, ht_fits(t, capacity_rows, load_factor), "table memory does not fit %lld rows",
             (long long)capacity_rows);
  if (load_factor == 0) load_factor = 0.5;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  t->capacity = capacity_rows;
  t->nbuckets = table_buckets_for(capacity_rows, load_factor);
  t->form = -1;
  t->base = 0;
  t->narrow_unfilled = false;
  const int64_t nslots = static_cast<int64_t>(t->nbuckets) * kSlots;
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(ctx, nslots, 4)), dim3(kJBlock), 0, ctx->stream, t->keys, kEmpty,
                     nslots);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, t->counters, 0ull, 8);
  KG_LAUNCH_CHECK(ctx);
  return KG_OK;
}

}  // namespace kg

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
  t->nbuckets = kg::table_buckets_for(capacity_rows, load_factor);
  const size_t nslots = t->nbuckets * kSlots;
  t->alloc_slots = static_cast<int64_t>(nslots);
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
  st = use_part_build(ctx, ht, nrows) ? part_build(ctx, ht, key, d_rows, nrows, inserted == 0)
                                      : direct_build(ctx, ht, key, d_rows, nrows);
  if (st != KG_OK) return st;
  unsigned long long overflow = 0;
  if ((st = read_counter(ctx, ht->counters + 2, &overflow)) != KG_OK) return st;
  if (overflow) return kg_fail(ctx, KG_ERR_CAPACITY, "join table overflow (side list or buckets full)");
  return KG_OK;
}

namespace {

// Probe-key partitioning: the region count (0 = probe directly). Only an
// explicit setting (kg_ctx_set_probe_regions, e.g. from kg_join_tune)
// partitions: measured on MI355X, the histogram + scatter passes cost more than
// the L2 locality they buy for a 16M-key table probed 256M times (6.3 ms with
// 128 regions against 4.8 ms direct), so the default is direct.
int probe_regions(const kg_ctx* ctx) { return ctx->probe_regions > 0 ? ctx->probe_regions : 0; }

// Probe workgroups per CU (kg_join_tune), 32 by default.
int probe_grid(kg_ctx* ctx, int64_t n) {
  const int per_cu = ctx->probe_wg_per_cu > 0 ? ctx->probe_wg_per_cu : 32;
  const int64_t g = ceil_div(std::max<int64_t>(n, 1), static_cast<int64_t>(kJBlock) * 64);
  return static_cast<int>(std::min<int64_t>(g, static_cast<int64_t>(ctx->num_cus) * per_cu));
}

template <int MODE, bool FIRST, int UNROLL = probe_unroll<MODE>()>
void launch_probe_kernel(kg_ctx* ctx, const kg_hash_table* ht, const kg_col& key, const int32_t* sel, int64_t n,
                         const ProbeOut& o, const PartIn& part, int grid) {
  if (ht->form == 1)
    hipLaunchKernelGGL((k_ht_probe<MODE, FIRST, true, UNROLL>), dim3(grid), dim3(kJBlock), 0, ctx->stream, ht->view(),
                       key, sel, n, o, part);
  else
    hipLaunchKernelGGL((k_ht_probe<MODE, FIRST, false, UNROLL>), dim3(grid), dim3(kJBlock), 0, ctx->stream,
                       ht->view(), key, sel, n, o, part);
}

// One probe pass over the given (possibly partitioned) keys. An emitting
// probe runs twice over the same workgroup layout: counting first (one total
// per workgroup), then, after an exclusive scan of the totals, emitting each
// workgroup's pairs from its own offset through an LDS cursor, so no pair
// goes through a device-wide atomic counter; the grand total goes to *o.count.
__global__ void k_copy_u64(const unsigned long long* src, unsigned long long* dst) { *dst = *src; }

template <int MODE, bool FIRST>
int run_probe(kg_ctx* ctx, const kg_hash_table* ht, const kg_col& key, const int32_t* sel, int64_t n, ProbeOut o,
              const PartIn& part, int grid, char* scratch) {
  if (MODE != kEmit) {
    launch_probe_kernel<MODE, FIRST>(ctx, ht, key, sel, n, o, part, grid);
    return KG_OK;
  }
  unsigned long long* bc = reinterpret_cast<unsigned long long*>(scratch);
  unsigned long long* bo = bc + grid + 1;
  void* tmp = bo + grid + 1;
  size_t tb = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tb, bc, bo, grid + 1, ctx->stream);
  ProbeOut oc{};
  oc.block_counts = bc;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, bc + grid, 0ull, 1);
  // The same rows per workgroup as the emitting pass: the same unroll.
  launch_probe_kernel<kCount, false, probe_unroll<kEmit>()>(ctx, ht, key, sel, n, oc, part, grid);
  if (hipcub::DeviceScan::ExclusiveSum(tmp, tb, bc, bo, grid + 1, ctx->stream) != hipSuccess)
    return kg_fail(ctx, KG_ERR_HIP, "probe count scan failed");
  hipLaunchKernelGGL(k_copy_u64, dim3(1), dim3(1), 0, ctx->stream, bo + grid, o.count);
  o.block_offs = bo;
  launch_probe_kernel<kEmit, FIRST>(ctx, ht, key, sel, n, o, part, grid);
  return KG_OK;
}

// Device scratch bytes run_probe needs for `grid` workgroups.
size_t run_probe_scratch(kg_ctx* ctx, int grid) {
  size_t tb = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tb, static_cast<unsigned long long*>(nullptr),
                                         static_cast<unsigned long long*>(nullptr), grid + 1, ctx->stream);
  return (static_cast<size_t>(grid) + 1) * 16 + tb + 256;
}

// Probes n keys (rows sel[i] or i), directly or through a region partition of
// the keys (histogram, scan, scatter, then an XCD-ranged probe). The whole
// sequence is the timed region.
template <int MODE, bool FIRST>
int launch_probe(kg_ctx* ctx, const kg_hash_table* ht, const kg_col& key, const int32_t* sel, int64_t n,
                 ProbeOut o) {
  const int P = probe_regions(ctx);
  // kAgg: the hot-key sample runs first, inside the timed region.
  auto hot_pass = [&]() {
    if (MODE == kAgg && o.hot_keys)
      launch_hot_keys(ctx->stream, key, sel, n, const_cast<int64_t*>(o.hot_keys));
  };
  if (P == 0) {
    const int grid = probe_grid(ctx, n);
    char* scratch = nullptr;
    if (MODE == kEmit) {
      const int st = kg_work_reserve(ctx, run_probe_scratch(ctx, grid));
      if (st != KG_OK) return st;
      scratch = static_cast<char*>(ctx->work);
    }
    kg_time_begin(ctx);
    hot_pass();
    const int st = run_probe<MODE, FIRST>(ctx, ht, key, sel, n, o, PartIn{}, grid, scratch);
    kg_time_end(ctx);
    if (st != KG_OK) return st;
    KG_LAUNCH_CHECK(ctx);
    return KG_OK;
  }
  KG_REQUIRE(ctx, P % kXcds == 0 && P <= 65536, "region count %d must be a multiple of %d", P, kXcds);
  const bool vals = MODE == kAgg, rows = MODE == kEmit || (MODE == kAgg && FIRST);
  const int nblk = static_cast<int>(std::min<int64_t>(2048, std::max<int64_t>(kXcds, ceil_div(n, 1 << 16))));
  const int64_t m = static_cast<int64_t>(P) * nblk;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o2 = off;
    off += (bytes + 255) / 256 * 256;
    return o2;
  };
  const int grid = (probe_grid(ctx, n) + kXcds - 1) / kXcds * kXcds;
  const size_t o_hist = carve(m * 4), o_offs = carve((m + 1) * 4), o_keys = carve(n * 8),
               o_vals = carve(vals ? n * 8 : 8), o_rows = carve(rows ? n * 4 : 8),
               o_run = carve(MODE == kEmit ? run_probe_scratch(ctx, grid) : 8);
  int st = kg_work_reserve(ctx, off);
  if (st != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  uint32_t* hist = reinterpret_cast<uint32_t*>(w + o_hist);
  uint32_t* offs = reinterpret_cast<uint32_t*>(w + o_offs);
  int64_t* pkeys = reinterpret_cast<int64_t*>(w + o_keys);
  int64_t* pvals = reinterpret_cast<int64_t*>(w + o_vals);
  int32_t* prows = reinterpret_cast<int32_t*>(w + o_rows);
  const size_t scatter_lds = static_cast<size_t>(3 * P + 1) * 4 + 8 + kPartTile * 8 + (vals ? kPartTile * 8 : 0) +
                             (rows ? kPartTile * 4 : 0) + kPartTile * 2;
  KG_REQUIRE(ctx, scatter_lds <= 64 * 1024, "region count %d too large", P);
  kg_time_begin(ctx);
  hot_pass();
  hipLaunchKernelGGL(k_part_hist, dim3(nblk), dim3(kJBlock), P * 4, ctx->stream, ht->view(), key, sel, n, P, hist);
  hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, ctx->stream, hist, offs, m);
  if (vals && rows)
    hipLaunchKernelGGL((k_part_scatter<true, true>), dim3(nblk), dim3(kJBlock), scatter_lds, ctx->stream, ht->view(),
                       key, sel, o.vals, n, P, offs, pkeys, pvals, prows);
  else if (vals)
    hipLaunchKernelGGL((k_part_scatter<true, false>), dim3(nblk), dim3(kJBlock), scatter_lds, ctx->stream,
                       ht->view(), key, sel, o.vals, n, P, offs, pkeys, pvals, prows);
  else if (rows)
    hipLaunchKernelGGL((k_part_scatter<false, true>), dim3(nblk), dim3(kJBlock), scatter_lds, ctx->stream,
                       ht->view(), key, sel, o.vals, n, P, offs, pkeys, pvals, prows);
  else
    hipLaunchKernelGGL((k_part_scatter<false, false>), dim3(nblk), dim3(kJBlock), scatter_lds, ctx->stream,
                       ht->view(), key, sel, o.vals, n, P, offs, pkeys, pvals, prows);
  kg_col pk{KG_BIGINT, 0, pkeys, nullptr};
  ProbeOut po = o;
  if (vals) po.vals = pvals;
  PartIn part{offs, rows ? prows : nullptr, nblk, P};
  st = run_probe<MODE, FIRST>(ctx, ht, pk, nullptr, n, po, part, grid, w + o_run);
  kg_time_end(ctx);
  if (st != KG_OK) return st;
  KG_LAUNCH_CHECK(ctx);
  return KG_OK;
}

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
  if ((st = launch_probe<kCount, false>(ctx, ht, *key, d_rows, nrows, o)) != KG_OK) return st;
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
  if ((st = launch_probe<kEmit, false>(ctx, ht, *key, d_rows, nrows, o)) != KG_OK) return st;
  unsigned long long h = 0;
  if ((st = read_counter(ctx, cnt, &h)) != KG_OK) return st;
  *out_count = static_cast<int64_t>(h);
  if (*out_count > out_cap)
    return kg_fail(ctx, KG_ERR_CAPACITY, "%lld join pairs exceed out_cap %lld", (long long)*out_count,
                   (long long)out_cap);
  return KG_OK;
}

namespace {

// Layout of a partitioned probe + GROUP BY (see k_agg_region); PF = 0 means
// the atomic form. The LDS window of a region (its home buckets plus
// kAggMargin, 16 bytes per slot) fits kAggLdsBudget.
struct RegionAggPlan {
  int PC = 0, PF = 0, Q = 0, nblk1 = 0;
  int64_t span = 0;  // window buckets of the largest region
};

RegionAggPlan region_agg_plan(kg_ctx* ctx, const kg_hash_table* ht, int64_t n) {
  RegionAggPlan p;
  if (ctx->agg_regions == 0 || ht->form < 0) return p;
  const bool narrow = ht->form == 1;
  const int S = narrow ? kSlots32 : kSlots;
  const int64_t nb = static_cast<int64_t>(narrow ? ht->nbuckets / 2 : ht->nbuckets);
  if (ctx->agg_regions < 0 && (n < (1 << 22) || nb * S * 16 < (64ll << 20))) return p;
  const int64_t budget = std::min<int64_t>(kAggLdsBudget, ctx->max_lds - 4096);
  const int64_t win = budget / (16 * S);  // window buckets that fit
  if (win <= kAggMargin + 2) return p;
  int64_t PF = ceil_div(nb, win - kAggMargin - 1);
  if (ctx->agg_regions > 0) PF = std::max<int64_t>(PF, ctx->agg_regions);
  const int64_t PC = ceil_div(PF, kAggFine);
  if (PC > 2048) return p;
  PF = PC * kAggFine;
  p.PC = static_cast<int>(PC);
  p.PF = static_cast<int>(PF);
  p.nblk1 = static_cast<int>(std::min<int64_t>(2048, std::max<int64_t>(kXcds, ceil_div(n, 4 * kAggTile))));
  p.Q = static_cast<int>(std::max<int64_t>(1, ceil_div(4 * ctx->num_cus, PC)));
  p.span = std::min<int64_t>(ceil_div(nb, PF) + 1 + kAggMargin, nb);
  return p;
}

// Runs the partitioned form; *fallback is set when a key's group lies outside
// its region's window (the caller then runs the atomic form). The whole
// sequence, from the hot-key sample to the written groups, is the timed region.
int run_agg_regions(kg_ctx* ctx, kg_hash_table* t, const kg_col& key, const int64_t* vals, int64_t n,
                    const RegionAggPlan& p, int64_t cap, int64_t* out_key, int64_t* out_sum, int64_t* out_cnt,
                    unsigned long long* count, bool* fallback) {
  const bool narrow = t->form == 1;
  const AggGeo g{narrow ? t->nbuckets / 2 : t->nbuckets, p.PF, narrow ? kSlots32 : kSlots};
  const int64_t m1 = static_cast<int64_t>(p.PC) * p.nblk1, m2 = static_cast<int64_t>(p.PF) * p.Q;
  size_t cub1 = 0, cub2 = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, cub1, static_cast<uint32_t*>(nullptr),
                                         static_cast<uint32_t*>(nullptr), static_cast<int>(m1 + 1), ctx->stream);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, cub2, static_cast<uint32_t*>(nullptr),
                                         static_cast<uint32_t*>(nullptr), static_cast<int>(m2 + 1), ctx->stream);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o2 = off;
    off += (bytes + 255) / 256 * 256;
    return o2;
  };
  const size_t o_h1 = carve((m1 + 1) * 4), o_o1 = carve((m1 + 1) * 4), o_h2 = carve((m2 + 1) * 4),
               o_o2 = carve((m2 + 1) * 4), o_k1 = carve(n * 8), o_v1 = carve(n * 8), o_k2 = carve(n * 8),
               o_v2 = carve(n * 8), o_hot = carve(kHot * 16), o_flag = carve(8),
               o_cub = carve(std::max(cub1, cub2));
  int st = kg_work_reserve(ctx, off);
  if (st != KG_OK) return st;
  if (!t->hot && (st = alloc(ctx, reinterpret_cast<void**>(&t->hot), kHotBufBytes)) != KG_OK) return st;
  char* w = static_cast<char*>(ctx->work);
  uint32_t* h1 = reinterpret_cast<uint32_t*>(w + o_h1);
  uint32_t* o1 = reinterpret_cast<uint32_t*>(w + o_o1);
  uint32_t* h2 = reinterpret_cast<uint32_t*>(w + o_h2);
  uint32_t* o2 = reinterpret_cast<uint32_t*>(w + o_o2);
  int64_t* k1 = reinterpret_cast<int64_t*>(w + o_k1);
  int64_t* v1 = reinterpret_cast<int64_t*>(w + o_v1);
  int64_t* k2 = reinterpret_cast<int64_t*>(w + o_k2);
  int64_t* v2 = reinterpret_cast<int64_t*>(w + o_v2);
  unsigned long long* hot_acc = reinterpret_cast<unsigned long long*>(w + o_hot);
  int* flag = reinterpret_cast<int*>(w + o_flag);
  const int64_t* hot = t->hot;
  const int32_t* hot_n = reinterpret_cast<const int32_t*>(t->hot + kHot);
  const size_t lds1 = static_cast<size_t>(p.PC) * 4 * (kJBlock / kWave);
  const size_t lds2 = (static_cast<size_t>(3 * p.PC) * 4 + 15) / 16 * 16 + static_cast<size_t>(kAggTile) * 18;
  const size_t lds5 = static_cast<size_t>(p.span) * g.S * 16;
  if ((st = kg_allow_lds(ctx, reinterpret_cast<const void*>(&k_agg_scatter1), lds2)) != KG_OK) return st;
  const void* region_kernel = narrow ? reinterpret_cast<const void*>(&k_agg_region<true>)
                                     : reinterpret_cast<const void*>(&k_agg_region<false>);
  if ((st = kg_allow_lds(ctx, region_kernel, lds5)) != KG_OK) return st;
  const TableView v = t->view();
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_fill_i32j, dim3(1), dim3(64), 0, ctx->stream, reinterpret_cast<int32_t*>(h1 + m1), 0, 1);
  hipLaunchKernelGGL(k_fill_i32j, dim3(1), dim3(64), 0, ctx->stream, reinterpret_cast<int32_t*>(h2 + m2), 0, 1);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kJBlock), 0, ctx->stream, hot_acc, 0ull, 2 * kHot);
  hipLaunchKernelGGL(k_fill_i32j, dim3(1), dim3(64), 0, ctx->stream, flag, 0, 1);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, count, 0ull, 1);
  launch_hot_keys(ctx->stream, key, nullptr, n, t->hot);
  hipLaunchKernelGGL(k_agg_hist, dim3(p.nblk1), dim3(kJBlock), lds1, ctx->stream, v, key, n, g, p.PC, hot, hot_n, h1,
                     flag);
  size_t tb = std::max(cub1, cub2);
  if (hipcub::DeviceScan::ExclusiveSum(w + o_cub, tb, h1, o1, static_cast<int>(m1 + 1), ctx->stream) != hipSuccess)
    return kg_fail(ctx, KG_ERR_HIP, "aggregation histogram scan failed");
  hipLaunchKernelGGL(k_agg_scatter1, dim3(p.nblk1), dim3(kJBlock), lds2, ctx->stream, v, key, vals, n, g, p.PC, hot,
                     hot_n, o1, k1, v1, hot_acc, flag);
  const int grid2 = p.PC * p.Q;
  hipLaunchKernelGGL(k_agg_hist2, dim3(grid2), dim3(kJBlock), 0, ctx->stream, g, k1, p.nblk1, p.Q, o1, h2);
  tb = std::max(cub1, cub2);
  if (hipcub::DeviceScan::ExclusiveSum(w + o_cub, tb, h2, o2, static_cast<int>(m2 + 1), ctx->stream) != hipSuccess)
    return kg_fail(ctx, KG_ERR_HIP, "aggregation histogram scan failed");
  hipLaunchKernelGGL(k_agg_scatter2, dim3(grid2), dim3(kJBlock), 0, ctx->stream, g, k1, v1, p.nblk1, p.Q, o1, o2, k2,
                     v2);
  if (narrow)
    hipLaunchKernelGGL(k_agg_region<true>, dim3(p.PF), dim3(kAggBlock), lds5, ctx->stream, v, g, p.Q, o2, k2, v2,
                       hot, hot_n, hot_acc, count, cap, out_key, out_sum, out_cnt, flag);
  else
    hipLaunchKernelGGL(k_agg_region<false>, dim3(p.PF), dim3(kAggBlock), lds5, ctx->stream, v, g, p.Q, o2, k2, v2,
                       hot, hot_n, hot_acc, count, cap, out_key, out_sum, out_cnt, flag);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  if ((st = kg_pinned_reserve(ctx, 64)) != KG_OK) return st;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(ctx->pinned, flag, 4, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *fallback = *static_cast<const int32_t*>(ctx->pinned) != 0;
  return KG_OK;
}

}  // namespace

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
  unsigned long long* cnt = t->counters + 3;
  const RegionAggPlan plan = d_out_first || nrows == 0 ? RegionAggPlan{} : region_agg_plan(ctx, t, nrows);
  bool fallback = true;
  if (plan.PF > 0) {
    st = run_agg_regions(ctx, t, *key, d_vals, nrows, plan, max_groups, d_out_key, d_out_sum, d_out_count, cnt,
                         &fallback);
    if (st != KG_OK) return st;
    ctx->last_path = fallback ? "hashed" : "regions";
  }
  if (!fallback) {
    unsigned long long h = 0;
    if ((st = read_counter(ctx, cnt, &h)) != KG_OK) return st;
    *out_ngroups = static_cast<int64_t>(h);
    if (*out_ngroups > max_groups)
      return kg_fail(ctx, KG_ERR_CAPACITY, "%lld groups exceed max_groups %lld", (long long)*out_ngroups,
                     (long long)max_groups);
    return KG_OK;
  }
  if (plan.PF == 0) ctx->last_path = "hashed";
  const int64_t nslots = static_cast<int64_t>(t->nbuckets) * kSlots;
  if (t->acc_slots < nslots + 1) {  // a reset table may have grown past its accumulators
    kg_dev_release(ctx, t->acc);
    kg_dev_release(ctx, t->first);
    t->acc = nullptr;
    t->first = nullptr;
    t->acc_slots = 0;
  }
  if (!t->acc) {
    if ((st = alloc(ctx, reinterpret_cast<void**>(&t->acc), (t->alloc_slots + 1) * sizeof(Acc))) != KG_OK) return st;
    t->acc_slots = t->alloc_slots + 1;
  }
  if (d_out_first && !t->first &&
      (st = alloc(ctx, reinterpret_cast<void**>(&t->first), t->acc_slots * 4)) != KG_OK)
    return st;
  // The atomic form: accumulator clear, probe, collect; all timed.
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ctx, 2 * (nslots + 1), 4)), dim3(kJBlock), 0, ctx->stream,
                     reinterpret_cast<unsigned long long*>(t->acc), 0ull, 2 * (nslots + 1));
  KG_LAUNCH_CHECK(ctx);
  if (d_out_first) {
    hipLaunchKernelGGL(k_fill_i32j, dim3(grid_for(ctx, nslots + 1, 4)), dim3(kJBlock), 0, ctx->stream, t->first,
                       INT32_MAX, nslots + 1);
    KG_LAUNCH_CHECK(ctx);
  }
  kg_time_end(ctx);
  if (nrows > 0) {
    if (!t->hot && (st = alloc(ctx, reinterpret_cast<void**>(&t->hot), kHotBufBytes)) != KG_OK) return st;
    ProbeOut o{};
    o.acc = t->acc;
    o.first = t->first;
    o.vals = d_vals;
    o.hot_keys = t->hot;
    o.hot_n = reinterpret_cast<const int32_t*>(t->hot + kHot);
    st = d_out_first ? launch_probe<kAgg, true>(ctx, t, *key, nullptr, nrows, o)
                     : launch_probe<kAgg, false>(ctx, t, *key, nullptr, nrows, o);
    if (st != KG_OK) return st;
  }
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, cnt, 0ull, 1);
  KG_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_collect, dim3(grid_for(ctx, nslots + 1, 4)), dim3(kJBlock), 0, ctx->stream, t->view(), t->acc,
                     t->first, nslots, nslots, cnt, max_groups, d_out_key, d_out_sum, d_out_count, d_out_first);
  kg_time_end(ctx);
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
  if (ctx) (void)hipSetDevice(ctx->device);
  // Back to the context's block cache (freed when the table is not the
  // context's or there is no context).
  for (void* p : {static_cast<void*>(ht->keys), static_cast<void*>(ht->rows), static_cast<void*>(ht->acc),
                  static_cast<void*>(ht->hot), static_cast<void*>(ht->first), static_cast<void*>(ht->side_rows),
                  static_cast<void*>(ht->counters)})
    kg_dev_release(ctx && ctx->device == ht->device ? ctx : nullptr, p);
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
  TableView gv{};
  gv.keys = keys;
  hipLaunchKernelGGL(k_collect, dim3(grid_for(ctx, nslots + 1, 4)), dim3(kJBlock), 0, ctx->stream, gv, acc, first,
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

namespace {

// The star pipeline's REAL sums in the deterministic order (det_fp_sums over
// its join output rows), written into dfin.
int star_det_sums(kg_ctx* ctx, int64_t n, const PredSet& P, const kg_col& key, const TableView& t,
                  const int32_t* codes, int G, const MeasureSet& M, AggFinal* dfin) {
  hipStream_t s = ctx->stream;
  size_t scan_tmp = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp, static_cast<int64_t*>(nullptr),
                                         static_cast<int64_t*>(nullptr), static_cast<int>(n + 1), s);
  char* buf = nullptr;
  const size_t cnt_b = (n + 1) * 8, off_b = (n + 1) * 8;
  if (kg_dev_alloc(ctx, reinterpret_cast<void**>(&buf), cnt_b + off_b + scan_tmp + 256) != KG_OK)
    return kg_fail(ctx, KG_ERR_OOM, "deterministic star: device allocation failed");
  int64_t* cnt = reinterpret_cast<int64_t*>(buf);
  int64_t* off = reinterpret_cast<int64_t*>(buf + cnt_b);
  void* tmp = buf + cnt_b + off_b;
  int64_t N = 0;
  int st = KG_OK;
  hipLaunchKernelGGL(k_star_det_count, dim3(ceil_div(n + 1, 256)), dim3(256), 0, s, P, key, t, codes, G, n, cnt);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    size_t tb = scan_tmp;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, off, static_cast<int>(n + 1), s);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&N, off + n, 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  char* items = nullptr;
  if (e == hipSuccess && N > 0) {
    if (kg_dev_alloc(ctx, reinterpret_cast<void**>(&items), N * 8) != KG_OK) {
      st = kg_fail(ctx, KG_ERR_OOM, "deterministic star: %lld join rows", (long long)N);
    } else {
      int32_t* gid = reinterpret_cast<int32_t*>(items);
      int32_t* src = gid + N;
      hipLaunchKernelGGL(k_star_det_items, dim3(ceil_div(n, 256)), dim3(256), 0, s, P, key, t, codes, G, n, off,
                         gid, src);
      e = hipGetLastError();
      if (e == hipSuccess) st = det_fp_sums(ctx, N, gid, src, G, M, reinterpret_cast<uint64_t*>(dfin),
                                            sizeof(AggFinal) / 8);
    }
  }
  (void)hipStreamSynchronize(s);
  kg_dev_release(ctx, items);
  kg_dev_release(ctx, buf);
  if (st != KG_OK) return st;
  if (e != hipSuccess) return kg_fail(ctx, KG_ERR_HIP, "deterministic star failed: %s", hipGetErrorString(e));
  return KG_OK;
}

}  // namespace

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
  if (static_cast<int64_t>(lds_bytes) > ctx->max_lds)
    return kg_fail(ctx, KG_ERR_UNSUPPORTED, "group table of %zu bytes exceeds the LDS budget", lds_bytes);
  {
    const void* kfn = minmax ? reinterpret_cast<const void*>(&k_join_group_agg<true>)
                             : reinterpret_cast<const void*>(&k_join_group_agg<false>);
    if ((st = kg_allow_lds(ctx, kfn, lds_bytes)) != KG_OK) return st;
  }
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
    bool any_real = false;
    for (int m = 0; m < nm; m++) any_real |= M.is_real[m] != 0;
    if (ctx->deterministic && any_real) {
      // Deterministic mode: the REAL sums again, in the fixed order over the
      // join output rows.
      if ((st = star_det_sums(ctx, n, P, cols[probe_key_col], ht->view(), codes_data, G, M, dfin)) != KG_OK)
        return st;
    }
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

// ---- join autotuner -------------------------------------------------------------

namespace {

// Copies up to `cap` keys (rows 0..) of a column to the host as int64 (NULL /
// NaN keys dropped) for the class statistics.
int sample_keys(kg_ctx* ctx, const kg_col* key, int64_t n, int64_t cap, std::vector<int64_t>& out) {
  out.clear();
  const int64_t m = std::min(n, cap);
  if (m <= 0) return KG_OK;
  const int w = type_width(key->type);
  std::vector<unsigned char> raw(static_cast<size_t>(m) * w);
  std::vector<int8_t> nul;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(raw.data(), key->data, raw.size(), hipMemcpyDeviceToHost, ctx->stream));
  if (key->nulls) {
    nul.resize(m);
    KG_HIP_CHECK(ctx, hipMemcpyAsync(nul.data(), key->nulls, m, hipMemcpyDeviceToHost, ctx->stream));
  }
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  out.reserve(m);
  for (int64_t i = 0; i < m; i++) {
    if (!nul.empty() && nul[i]) continue;
    const unsigned char* p = raw.data() + i * w;
    int64_t v = 0;
    switch (key->type) {
      case KG_SMALLINT: v = *reinterpret_cast<const int16_t*>(p); break;
      case KG_BOOLEAN: v = *reinterpret_cast<const int8_t*>(p); break;
      case KG_BIGINT: v = *reinterpret_cast<const int64_t*>(p); break;
      case KG_REAL: {
        const double d = *reinterpret_cast<const double*>(p);
        if (d != d) continue;
        std::memcpy(&v, &d, 8);
        break;
      }
      default: v = *reinterpret_cast<const int32_t*>(p); break;
    }
    out.push_back(v);
  }
  return KG_OK;
}

// The tuning class: log2 build rows (6 bits), distinct fraction class of the
// build sample (0 unique, 1 few duplicates, 2 many), top-key share class of the
// probe sample (0 < 0.1%, 1 < 5%, 2 heavier).
uint64_t tuning_class(int64_t nbuild, std::vector<int64_t> b, std::vector<int64_t> p) {
  int lg = 0;
  while ((int64_t{1} << lg) < nbuild && lg < 63) lg++;
  int card = 0;
  if (!b.empty()) {
    std::sort(b.begin(), b.end());
    const double d = static_cast<double>(std::unique(b.begin(), b.end()) - b.begin()) / b.size();
    card = d >= 0.99 ? 0 : (d >= 0.5 ? 1 : 2);
  }
  int skew = 0;
  if (!p.empty()) {
    std::sort(p.begin(), p.end());
    int64_t best = 0, run = 0;
    for (size_t i = 0; i < p.size(); i++) {
      run = (i > 0 && p[i] == p[i - 1]) ? run + 1 : 1;
      best = std::max(best, run);
    }
    const double top = static_cast<double>(best) / p.size();
    skew = top < 0.001 ? 0 : (top < 0.05 ? 1 : 2);
  }
  return (static_cast<uint64_t>(lg) << 8) | (static_cast<uint64_t>(card) << 4) | static_cast<uint64_t>(skew);
}

// Mean time of `reps` probe_count calls (after one warm call) in ms.
int time_probe(kg_ctx* ctx, const kg_hash_table* ht, const kg_col* key, int64_t n, int reps, double* ms) {
  // Kernel time (the sum of one call's timed spans, not the host round trips
  // around it), the fastest of `reps` calls after one untimed call.
  int64_t cnt = 0;
  int st = kg_ht_probe_count(ctx, ht, key, nullptr, n, &cnt);
  if (st != KG_OK) return st;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> saved;
  saved.swap(ctx->ev_pending);
  const int timing = ctx->timing;
  ctx->timing = 1;
  double best = 1e30;
  for (int i = 0; i < reps && st == KG_OK; i++) {
    st = kg_ht_probe_count(ctx, ht, key, nullptr, n, &cnt);
    double total = 0;
    int64_t spans = 0;
    if (st == KG_OK) st = kg_ctx_kernel_stats(ctx, &total, &spans, 1);
    if (st == KG_OK) best = std::min(best, total);
  }
  for (auto& pr : ctx->ev_pending) {  // left over only when a call failed
    ctx->ev_pool.push_back(pr.first);
    ctx->ev_pool.push_back(pr.second);
  }
  ctx->ev_pending.swap(saved);
  ctx->timing = timing;
  *ms = best;
  return st;
}

// Kernel time of one build of n keys into ht: the sum of the build's timed
// spans (the context's own pending timings are set aside meanwhile).
int time_build(kg_ctx* ctx, kg_hash_table* ht, const kg_col* key, int64_t n, double* ms) {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> saved;
  saved.swap(ctx->ev_pending);
  const int timing = ctx->timing;
  ctx->timing = 1;
  int st = kg_ht_build(ctx, ht, key, nullptr, n);
  double total = 0;
  int64_t spans = 0;
  const int st2 = kg_ctx_kernel_stats(ctx, &total, &spans, 1);
  for (auto& pr : ctx->ev_pending) {  // left over only when the stats failed
    ctx->ev_pool.push_back(pr.first);
    ctx->ev_pool.push_back(pr.second);
  }
  ctx->ev_pending.swap(saved);
  ctx->timing = timing;
  *ms = total;
  return st != KG_OK ? st : st2;
}

}  // namespace

extern "C" int kg_join_tune(kg_ctx* ctx, const kg_col* build_key, int64_t nbuild, const kg_col* probe_key,
                            int64_t nprobe, kg_join_tuning* out) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, out && build_key && probe_key, "null argument");
  KG_REQUIRE(ctx, nbuild >= 0 && nbuild < INT32_MAX && nprobe >= 0 && nprobe < INT32_MAX, "row count out of range");
  int st;
  if ((st = check_key(ctx, build_key, nbuild)) != KG_OK || (st = check_key(ctx, probe_key, nprobe)) != KG_OK)
    return st;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  std::memset(out, 0, sizeof(*out));
  std::vector<int64_t> bs, ps;
  if ((st = sample_keys(ctx, build_key, nbuild, 1 << 16, bs)) != KG_OK) return st;
  if ((st = sample_keys(ctx, probe_key, nprobe, 1 << 16, ps)) != KG_OK) return st;
  const uint64_t cls = tuning_class(nbuild, bs, ps);
  out->tuning_class = cls;
  for (const auto& t : ctx->join_tunes)
    if (t.cls == cls) {
      out->load_factor = t.load_factor;
      out->probe_wg_per_cu = t.wg_per_cu;
      out->probe_regions = t.regions;
      out->probe_ms = t.probe_ms;
      out->from_cache = 1;
      ctx->probe_wg_per_cu = t.wg_per_cu;
      ctx->probe_regions = t.regions;
      return KG_OK;
    }
  const int64_t nb = std::min<int64_t>(nbuild, int64_t{1} << 26);
  const int64_t np = std::min<int64_t>(nprobe, int64_t{1} << 27);
  double best_lf = 0.5, best_ms = 1e30;
  int best_wg = 32, best_regions = 0, trials = 0;
  kg_hash_table* keep = nullptr;
  const int saved_regions = ctx->probe_regions, saved_wg = ctx->probe_wg_per_cu;
  auto restore = [&](int s2) {
    if (keep) kg_ht_free(ctx, keep);
    ctx->probe_regions = saved_regions;
    ctx->probe_wg_per_cu = saved_wg;
    return s2;
  };
  // A setting replaces the incumbent (the defaults first) only when it is at
  // least 3% faster: equal settings time within about that of each other, and
  // a switch on noise (16 instead of 32 probe workgroups per CU) cost 3% at
  // full size.
  constexpr double kMargin = 0.97;
  // 1. load factor (direct probes, 32 workgroups per CU), by the build time
  // plus the probe time scaled to all nprobe rows: a load factor whose probes
  // are no faster but whose build is (fewer table lines to write) wins.
  ctx->probe_regions = 0;
  ctx->probe_wg_per_cu = 32;
  double best_score = 1e30;
  for (double lf : {0.5, 0.35, 0.65}) {
    kg_hash_table* ht = nullptr;
    if ((st = kg_ht_create(ctx, std::max<int64_t>(nb, 1), lf, &ht)) != KG_OK) return restore(st);
    double bms = 0, ms = 0;
    st = time_build(ctx, ht, build_key, nb, &bms);
    if (st == KG_OK) st = time_probe(ctx, ht, probe_key, np, 3, &ms);
    if (st != KG_OK) {
      kg_ht_free(ctx, ht);
      return restore(st);
    }
    trials++;
    const double score = bms * static_cast<double>(nbuild) / static_cast<double>(std::max<int64_t>(nb, 1)) +
                         ms * static_cast<double>(nprobe) / static_cast<double>(std::max<int64_t>(np, 1));
    if (score < best_score * kMargin) {
      best_score = score;
      best_ms = ms;
      best_lf = lf;
      if (keep) kg_ht_free(ctx, keep);
      keep = ht;
    } else {
      kg_ht_free(ctx, ht);
    }
  }
  // 2. workgroups per CU, then 3. regions, on the best table
  for (int wg : {16, 8}) {
    ctx->probe_wg_per_cu = wg;
    double ms = 0;
    if ((st = time_probe(ctx, keep, probe_key, np, 3, &ms)) != KG_OK) return restore(st);
    trials++;
    if (ms < best_ms * kMargin) {
      best_ms = ms;
      best_wg = wg;
    }
  }
  ctx->probe_wg_per_cu = best_wg;
  for (int reg : {64, 128}) {
    ctx->probe_regions = reg;
    double ms = 0;
    if ((st = time_probe(ctx, keep, probe_key, np, 3, &ms)) != KG_OK) return restore(st);
    trials++;
    if (ms < best_ms * kMargin) {
      best_ms = ms;
      best_regions = reg;
    }
  }
  restore(KG_OK);
  ctx->probe_wg_per_cu = best_wg;
  ctx->probe_regions = best_regions;
  ctx->join_tunes.push_back({cls, best_lf, best_wg, best_regions, best_ms});
  out->load_factor = best_lf;
  out->probe_wg_per_cu = best_wg;
  out->probe_regions = best_regions;
  out->probe_ms = best_ms;
  out->trials = trials;
  return KG_OK;
}
