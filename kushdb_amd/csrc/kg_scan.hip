// Scan-select and scan-select + group-by aggregation over dense code keys.
//
// Reference behaviour: ScanSelectTranslator (compile/translators/scan_select_translator.cc:92-156),
// SimdScanSelectTranslator (simd_scan_select_translator.cc:94-532),
// GroupByAggregateTranslator (group_by_aggregate_translator.cc:36-111),
// AggregateTranslator (aggregate_translator.cc:34-126), aggregators
// (compile/proxy/aggregator.cc:14-298).
//
// Layout and kernels:
//  * Columns are resident in HBM as plain typed arrays (one per column) with an
//    optional int8 null array, the .kdb / _null.kdb layout.
//  * Selection: a tile of 4096 rows per workgroup; each wave evaluates 64 rows,
//    ballots the result and compacts with a popcount prefix (the wave-wide form
//    of the reference's 8-wide mask -> permutation-table compaction).
//  * Low-cardinality aggregation (<= 8 dense groups): every lane keeps private
//    per-group partials in registers, folds its rows in a fixed order, and the
//    workgroup combines them with a fixed tree (lanes, then waves); a second
//    kernel adds the per-workgroup partials in workgroup order. The order
//    depends only on n, so FP64 results are reproducible and the CPU oracle
//    restates it exactly.
//  * Up to kMaxLdsGroups groups: a workgroup-private table in LDS updated with
//    LDS atomics, merged into a global table with global atomics.
#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <cstring>
#include <vector>

#include "kg_ops.h"

namespace kg {


// ---------------------------------------------------------------------------
// Selection

__global__ __launch_bounds__(kBlock) void k_select_count(PredSet p, int64_t n, int32_t* tile_counts) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kSelTile;
  int cnt = 0;
#pragma unroll 4
  for (int k = 0; k < kSelTile / kBlock; k++) {
    const int64_t r = base + k * kBlock + threadIdx.x;
    if (r < n && eval_preds(p, r)) cnt++;
  }
  // Wave reduction, then across the workgroup's waves.
  for (int s = kWave / 2; s >= 1; s >>= 1) cnt += __shfl_xor(cnt, s);
  __shared__ int wsum[kBlock / kWave];
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kBlock / kWave; w++) t += wsum[w];
    tile_counts[blockIdx.x] = t;
  }
}

// Exclusive scan of ntiles counts by one workgroup; total written to *total.
__global__ __launch_bounds__(1024) void k_exclusive_scan(const int32_t* in, int64_t* out, int64_t m,
                                                         int64_t* total) {
  __shared__ int64_t partial[1024];
  const int64_t per = (m + 1023) / 1024;
  const int64_t lo = threadIdx.x * per;
  const int64_t hi = lo + per < m ? lo + per : m;
  int64_t s = 0;
  for (int64_t i = lo; i < hi; i++) s += in[i];
  partial[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int64_t v = threadIdx.x >= off ? partial[threadIdx.x - off] : 0;
    __syncthreads();
    partial[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t run = threadIdx.x == 0 ? 0 : partial[threadIdx.x - 1];
  for (int64_t i = lo; i < hi; i++) {
    out[i] = run;
    run += in[i];
  }
  if (threadIdx.x == 1023) *total = partial[1023];
}

__global__ __launch_bounds__(kBlock) void k_select_write(PredSet p, int64_t n, const int64_t* tile_offsets,
                                                         int32_t* out_rows) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kSelTile;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  __shared__ int wcnt[2][kBlock / kWave];
  int64_t pos = tile_offsets[blockIdx.x];
  for (int k = 0; k < kSelTile / kBlock; k++) {
    const int64_t r = base + k * kBlock + threadIdx.x;
    const bool pass = r < n && eval_preds(p, r);
    const uint64_t mask = __ballot(pass);
    const int below = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wcnt[k & 1][wave] = __popcll(mask);
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < kBlock / kWave; w++) {
      const int c = wcnt[k & 1][w];
      before += w < wave ? c : 0;
      total += c;
    }
    if (pass) out_rows[pos + before + below] = static_cast<int32_t>(r);
    pos += total;
  }
}

// ---------------------------------------------------------------------------
// Aggregation over dense group codes


// NN: some measure is nullable, so non-NULL counts are kept per measure
// (otherwise every measure's count is the group's row count).
// G is the register budget (compile time); Grt <= G is the number of dense
// groups, which sizes the per-workgroup output records.
template <int G, bool MINMAX, bool NN>
__global__ __launch_bounds__(kBlock) void k_group_agg_reg(PredSet P, KeySet K, MeasureSet M, int64_t n, int Grt,
                                                           AggOut out) {
  uint64_t sum[G][kMaxMeasures];
  uint64_t mn[MINMAX ? G : 1][MINMAX ? kMaxMeasures : 1];
  uint64_t mx[MINMAX ? G : 1][MINMAX ? kMaxMeasures : 1];
  uint32_t nn[NN ? G : 1][NN ? kMaxMeasures : 1];
  uint32_t rows[G];
  int32_t first[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    rows[g] = 0;
    first[g] = INT32_MAX;
#pragma unroll
    for (int m = 0; m < kMaxMeasures; m++) {
      sum[g][m] = 0;  // +0.0 and int64 0 share the bit pattern
      if constexpr (NN) nn[g][m] = 0;
      if constexpr (MINMAX) {
        mn[g][m] = 0;
        mx[g][m] = 0;
      }
    }
  }
  const int64_t T = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; r < n; r += T) {
    if (!eval_preds(P, r)) continue;
    const int grp = group_of(K, r);
    if (grp < 0 || grp >= Grt) continue;  // codes outside [0, card) are not groups
    uint64_t v[kMaxMeasures];
    bool vnull[kMaxMeasures];
#pragma unroll
    for (int m = 0; m < kMaxMeasures; m++)
      if (m < M.nm) v[m] = load_measure(M, m, r, vnull[m]);
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (grp != g) continue;
      const bool first_row = rows[g] == 0;
      rows[g]++;
      first[g] = min(first[g], static_cast<int32_t>(r));
#pragma unroll
      for (int m = 0; m < kMaxMeasures; m++) {
        if (m >= M.nm || vnull[m]) continue;
        bool fresh = first_row;
        if constexpr (NN) {
          fresh = nn[g][m] == 0;
          nn[g][m]++;
        }
        sum[g][m] = M.is_real[m] ? fadd_bits(sum[g][m], v[m]) : sum[g][m] + v[m];
        if constexpr (MINMAX) {
          // Written as selects: the equivalent conditional stores
          // `if (fresh || less) mn = v` were compiled so that rows after a
          // thread's first grid-stride step never updated MIN/MAX.
          const uint64_t cur_mn = mn[g][m], cur_mx = mx[g][m];
          const bool lt = val_less(M.is_real[m], v[m], cur_mn);
          const bool gt = val_less(M.is_real[m], cur_mx, v[m]);
          mn[g][m] = (fresh || lt) ? v[m] : cur_mn;
          mx[g][m] = (fresh || gt) ? v[m] : cur_mx;
        }
      }
    }
  }

  // Fixed-order workgroup reduction: lanes by p[l] += p[l ^ s], s = 32..1
  // (every lane ends with the same tree result), then waves by the same tree.
  __shared__ uint64_t wred[kBlock / kWave][G * kMaxMeasures * (MINMAX ? 3 : 1)];
  __shared__ uint32_t wcnt[kBlock / kWave][G * (kMaxMeasures + 1)];
  __shared__ int32_t wfirst[kBlock / kWave][G];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int g = 0; g < G; g++) {
#pragma unroll
    for (int m = 0; m < kMaxMeasures; m++) {
      if (m >= M.nm) continue;
      uint64_t s = sum[g][m];
      uint32_t c = NN ? nn[g][m < kMaxMeasures ? m : 0] : rows[g];
      uint64_t lo = MINMAX ? mn[g][m] : 0, hi = MINMAX ? mx[g][m] : 0;
      for (int st = kWave / 2; st >= 1; st >>= 1) {
        const uint64_t os = shfl_xor_u64(s, st);
        const uint32_t oc = __shfl_xor(c, st);
        // Tree order: the lower lane of each pair is the left operand.
        if (M.is_real[m]) {
          s = (lane & st) ? fadd_bits(os, s) : fadd_bits(s, os);
        } else {
          s = s + os;
        }
        if constexpr (MINMAX) {
          const uint64_t olo = shfl_xor_u64(lo, st), ohi = shfl_xor_u64(hi, st);
          const bool take_lo = oc > 0 && (c == 0 || val_less(M.is_real[m], olo, lo));
          const bool take_hi = oc > 0 && (c == 0 || val_less(M.is_real[m], hi, ohi));
          lo = take_lo ? olo : lo;
          hi = take_hi ? ohi : hi;
        }
        c += oc;
      }
      if (lane == 0) {
        wred[wave][(g * kMaxMeasures + m) * (MINMAX ? 3 : 1)] = s;
        if constexpr (MINMAX) {
          wred[wave][(g * kMaxMeasures + m) * 3 + 1] = lo;
          wred[wave][(g * kMaxMeasures + m) * 3 + 2] = hi;
        }
        wcnt[wave][g * (kMaxMeasures + 1) + m] = c;
      }
    }
    uint32_t rc = rows[g];
    int32_t fr = first[g];
    for (int st = kWave / 2; st >= 1; st >>= 1) {
      rc += __shfl_xor(rc, st);
      fr = min(fr, __shfl_xor(fr, st));
    }
    if (lane == 0) {
      wcnt[wave][g * (kMaxMeasures + 1) + kMaxMeasures] = rc;
      wfirst[wave][g] = fr;
    }
  }
  __syncthreads();
  constexpr int NW = kBlock / kWave;
  for (int idx = threadIdx.x; idx < G * kMaxMeasures; idx += kBlock) {
    const int g = idx / kMaxMeasures, m = idx % kMaxMeasures;
    if (m >= M.nm || g >= Grt) continue;
    const int stride = MINMAX ? 3 : 1;
    uint64_t p[NW], plo[NW], phi[NW];
    uint32_t pc[NW];
    for (int w = 0; w < NW; w++) {
      p[w] = wred[w][idx * stride];
      pc[w] = wcnt[w][g * (kMaxMeasures + 1) + m];
      if constexpr (MINMAX) {
        plo[w] = wred[w][idx * 3 + 1];
        phi[w] = wred[w][idx * 3 + 2];
      }
    }
    for (int st = NW / 2; st >= 1; st >>= 1) {
      for (int w = 0; w < st; w++) {
        p[w] = M.is_real[m] ? fadd_bits(p[w], p[w + st]) : p[w] + p[w + st];
        if constexpr (MINMAX) {
          const bool take_lo = pc[w + st] > 0 && (pc[w] == 0 || val_less(M.is_real[m], plo[w + st], plo[w]));
          const bool take_hi = pc[w + st] > 0 && (pc[w] == 0 || val_less(M.is_real[m], phi[w], phi[w + st]));
          plo[w] = take_lo ? plo[w + st] : plo[w];
          phi[w] = take_hi ? phi[w + st] : phi[w];
        }
        pc[w] += pc[w + st];
      }
    }
    const int64_t o = (static_cast<int64_t>(g) * M.nm + m) * gridDim.x + blockIdx.x;
    out.sum[o] = p[0];
    out.nn[o] = pc[0];
    if constexpr (MINMAX) {
      out.mn[o] = plo[0];
      out.mx[o] = phi[0];
    }
  }
  for (int g = threadIdx.x; g < Grt; g += kBlock) {
    uint32_t rc = 0;
    int32_t fr = INT32_MAX;
    for (int w = 0; w < NW; w++) {
      rc += wcnt[w][g * (kMaxMeasures + 1) + kMaxMeasures];
      fr = min(fr, wfirst[w][g]);
    }
    out.rows[static_cast<int64_t>(blockIdx.x) * Grt + g] = rc;
    if (fr != INT32_MAX) atomicMin(&out.first[g], fr);
  }
}

// Adds the per-workgroup partials in workgroup order. One thread per
// (group, measure); thread G*nm + g handles the row count of group g.

// Adds the per-workgroup partials in workgroup order. One 64-lane workgroup
// per output: (group, measure) outputs first, then the row count of each
// group. The lanes stage all partials of that output in LDS in parallel, then
// lane 0 folds them in workgroup order (the order that defines the FP64
// result); the integer row counts are summed in any order.
__global__ __launch_bounds__(64) void k_group_agg_final(const AggOut in, int nblocks, int G, MeasureSet M,
                                                        bool minmax, AggFinal* fin, uint64_t* rows_out) {
  extern __shared__ uint64_t stage[];
  const int idx = blockIdx.x;
  const int nm = M.nm;
  const int lane = threadIdx.x;
  if (idx < G * nm) {
    const int g = idx / nm, m = idx % nm;
    uint64_t* ssum = stage;
    uint64_t* snn = stage + nblocks;
    uint64_t* smn = stage + 2 * nblocks;
    uint64_t* smx = stage + 3 * nblocks;
    for (int b = lane; b < nblocks; b += 64) {
      const int64_t o = (static_cast<int64_t>(g) * nm + m) * nblocks + b;
      ssum[b] = in.sum[o];
      snn[b] = in.nn[o];
      if (minmax) {
        smn[b] = in.mn[o];
        smx[b] = in.mx[o];
      }
    }
    __syncthreads();
    if (lane == 0) {
      // A workgroup without rows contributes +0.0 (its partial is never
      // -0.0: every lane starts from +0.0), so folding every workgroup's
      // partial unconditionally gives the same bits and keeps the LDS reads
      // independent of the branch.
      const bool real = M.is_real[m];
      uint64_t s = 0, lo = 0, hi = 0, c = 0;
      if (real) {
        double d = 0.0;
        for (int b = 0; b < nblocks; b++) d += __longlong_as_double(static_cast<long long>(ssum[b]));
        s = static_cast<uint64_t>(__double_as_longlong(d));
      } else {
        for (int b = 0; b < nblocks; b++) s += ssum[b];
      }
      for (int b = 0; b < nblocks; b++) {
        const uint64_t bc = snn[b];
        if (minmax && bc) {
          const bool take_lo = c == 0 || val_less(real, smn[b], lo);
          const bool take_hi = c == 0 || val_less(real, hi, smx[b]);
          lo = take_lo ? smn[b] : lo;
          hi = take_hi ? smx[b] : hi;
        }
        c += bc;
      }
      fin[idx] = AggFinal{s, lo, hi, c};
    }
  } else if (idx < G * nm + G) {
    const int g = idx - G * nm;
    uint64_t rc = 0;
    for (int b = lane; b < nblocks; b += 64) rc += in.rows[static_cast<int64_t>(b) * G + g];
    for (int st = 32; st >= 1; st >>= 1) rc += __shfl_xor(rc, st);
    if (lane == 0) rows_out[g] = rc;
  }
}

// Fast path of the dense group-by: every predicate and key column is a
// non-NULL 32-bit integer column, every measure a non-NULL REAL or INT
// (32-bit) column, all 16-byte aligned; conjunctive predicates; no MIN/MAX.
// Each lane takes kVec consecutive rows per step with 16-byte loads; rows of a
// chunk are folded in ascending order, chunks c = t + k*T in ascending k (the
// last, partial chunk belongs to lane nfull % T after its full chunks).
// REAL and INT measures are accumulated in separate register arrays whose
// sizes (NF, NI) are compile-time.
constexpr int kVec = 4;
constexpr int kFastPreds = 2;

struct FastSet {
  int32_t npred;
  int32_t op[kFastPreds];
  int64_t lit[kFastPreds];
  const int32_t* pcol[kFastPreds];
  int32_t nk;
  const int32_t* key[2];
  int32_t card1;
  int32_t nf, ni;              // REAL / INT measures in use (<= NF / NI)
  const double* fcol[4];
  const int32_t* icol[4];
  int32_t fidx[4], iidx[4];    // their positions in the MeasureSet order
  int32_t nm;
};

template <int G, int NF, int NI>
struct FastAcc {
  double f[G][NF > 0 ? NF : 1];
  int64_t i[G][NI > 0 ? NI : 1];
  uint32_t rows[G];
  int32_t first[G];
};

template <int G, int NF, int NI, bool FULL>
__device__ __forceinline__ void fast_chunk(const FastSet& S, int64_t r0, int64_t n, int Grt,
                                           FastAcc<G, NF, NI>& A) {
  int32_t pv[kFastPreds][kVec], kv[2][kVec], iv[NI > 0 ? NI : 1][kVec];
  double fv[NF > 0 ? NF : 1][kVec];
  auto ld32 = [&](const int32_t* p, int32_t (&o)[kVec]) {
    if (FULL) {
      const int4 x = *reinterpret_cast<const int4*>(p + r0);
      o[0] = x.x;
      o[1] = x.y;
      o[2] = x.z;
      o[3] = x.w;
    } else {
#pragma unroll
      for (int v = 0; v < kVec; v++) o[v] = r0 + v < n ? p[r0 + v] : 0;
    }
  };
#pragma unroll
  for (int q = 0; q < kFastPreds; q++)
    if (q < S.npred) ld32(S.pcol[q], pv[q]);
#pragma unroll
  for (int q = 0; q < 2; q++)
    if (q < S.nk) ld32(S.key[q], kv[q]);
#pragma unroll
  for (int m = 0; m < NI; m++)
    if (m < S.ni) ld32(S.icol[m], iv[m]);
#pragma unroll
  for (int m = 0; m < NF; m++) {
    if (m >= S.nf) continue;
    const double* p = S.fcol[m];
    if (FULL) {
      const double2 x = *reinterpret_cast<const double2*>(p + r0);
      const double2 y = *reinterpret_cast<const double2*>(p + r0 + 2);
      fv[m][0] = x.x;
      fv[m][1] = x.y;
      fv[m][2] = y.x;
      fv[m][3] = y.y;
    } else {
#pragma unroll
      for (int v = 0; v < kVec; v++) fv[m][v] = r0 + v < n ? p[r0 + v] : 0.0;
    }
  }
#pragma unroll
  for (int v = 0; v < kVec; v++) {
    bool pass = FULL || r0 + v < n;
#pragma unroll
    for (int q = 0; q < kFastPreds; q++)
      if (q < S.npred) pass = pass && cmp_i64(S.op[q], pv[q][v], S.lit[q]);
    const int grp = S.nk == 0 ? 0 : (S.nk == 1 ? kv[0][v] : kv[0][v] * S.card1 + kv[1][v]);
    pass = pass && grp >= 0 && grp < Grt;
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (pass && grp == g) {
        A.rows[g]++;
        A.first[g] = min(A.first[g], static_cast<int32_t>(r0 + v));
#pragma unroll
        for (int m = 0; m < NF; m++) A.f[g][m] += fv[m][v];
#pragma unroll
        for (int m = 0; m < NI; m++) A.i[g][m] += iv[m][v];
      }
    }
  }
}

template <int G, int NF, int NI>
__global__ __launch_bounds__(kBlock) void k_group_agg_v4(FastSet S, int64_t n, int Grt, AggOut out) {
  FastAcc<G, NF, NI> A;
#pragma unroll
  for (int g = 0; g < G; g++) {
    A.rows[g] = 0;
    A.first[g] = INT32_MAX;
#pragma unroll
    for (int m = 0; m < NF; m++) A.f[g][m] = 0.0;
#pragma unroll
    for (int m = 0; m < NI; m++) A.i[g][m] = 0;
  }
  const int64_t T = static_cast<int64_t>(gridDim.x) * kBlock;
  const int64_t nfull = n / kVec;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
#pragma unroll 2
  for (int64_t c = t; c < nfull; c += T) fast_chunk<G, NF, NI, true>(S, c * kVec, n, Grt, A);
  if (nfull * kVec < n && t == nfull % T) fast_chunk<G, NF, NI, false>(S, nfull * kVec, n, Grt, A);

  // Same fixed-order workgroup reduction as k_group_agg_reg.
  constexpr int NFI = (NF > 0 ? NF : 1) + (NI > 0 ? NI : 1);
  __shared__ uint64_t wred[kBlock / kWave][G * NFI];
  __shared__ uint32_t wcnt[kBlock / kWave][G];
  __shared__ int32_t wfirst[kBlock / kWave][G];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int g = 0; g < G; g++) {
#pragma unroll
    for (int m = 0; m < NF; m++) {
      double sv = A.f[g][m];
      for (int st = kWave / 2; st >= 1; st >>= 1) {
        const double os = __longlong_as_double(
            static_cast<long long>(shfl_xor_u64(static_cast<uint64_t>(__double_as_longlong(sv)), st)));
        sv = (lane & st) ? os + sv : sv + os;
      }
      if (lane == 0) wred[wave][g * NFI + m] = static_cast<uint64_t>(__double_as_longlong(sv));
    }
#pragma unroll
    for (int m = 0; m < NI; m++) {
      int64_t sv = A.i[g][m];
      for (int st = kWave / 2; st >= 1; st >>= 1) sv += static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(sv), st));
      if (lane == 0) wred[wave][g * NFI + (NF > 0 ? NF : 1) + m] = static_cast<uint64_t>(sv);
    }
    uint32_t rc = A.rows[g];
    int32_t fr = A.first[g];
    for (int st = kWave / 2; st >= 1; st >>= 1) {
      rc += __shfl_xor(rc, st);
      fr = min(fr, __shfl_xor(fr, st));
    }
    if (lane == 0) {
      wcnt[wave][g] = rc;
      wfirst[wave][g] = fr;
    }
  }
  __syncthreads();
  constexpr int NW = kBlock / kWave;
  for (int idx = threadIdx.x; idx < G * NFI; idx += kBlock) {
    const int g = idx / NFI, j = idx % NFI;
    const bool is_f = j < (NF > 0 ? NF : 1);
    const int m = is_f ? j : j - (NF > 0 ? NF : 1);
    if (g >= Grt || (is_f ? m >= S.nf : m >= S.ni)) continue;
    uint64_t p[NW];
    for (int w = 0; w < NW; w++) p[w] = wred[w][idx];
    for (int st = NW / 2; st >= 1; st >>= 1)
      for (int w = 0; w < st; w++) p[w] = is_f ? fadd_bits(p[w], p[w + st]) : p[w] + p[w + st];
    uint32_t cnt = 0;
    for (int w = 0; w < NW; w++) cnt += wcnt[w][g];
    const int mi = is_f ? S.fidx[m] : S.iidx[m];
    const int64_t o = (static_cast<int64_t>(g) * S.nm + mi) * gridDim.x + blockIdx.x;
    out.sum[o] = p[0];
    out.nn[o] = cnt;
  }
  for (int g = threadIdx.x; g < Grt; g += kBlock) {
    uint32_t rc = 0;
    int32_t fr = INT32_MAX;
    for (int w = 0; w < NW; w++) {
      rc += wcnt[w][g];
      fr = min(fr, wfirst[w][g]);
    }
    out.rows[static_cast<int64_t>(blockIdx.x) * Grt + g] = rc;
    if (fr != INT32_MAX) atomicMin(&out.first[g], fr);
  }
}

// Workgroup-private LDS table for up to kMaxLdsGroups groups (any order of
// FP additions: not reproducible bit for bit, within FP64 rounding of it).
template <bool MINMAX>
__global__ __launch_bounds__(kBlock) void k_group_agg_lds(PredSet P, KeySet K, MeasureSet M, int64_t n,
                                                           int G, AggFinal* gfin, unsigned long long* grows,
                                                           int32_t* gfirst) {
  extern __shared__ unsigned long long lds[];
  const int nm = M.nm;
  unsigned long long* lsum = lds;                        // [G][nm]
  unsigned long long* lnn = lsum + G * nm;               // [G][nm]
  unsigned long long* lrows = lnn + G * nm;              // [G]
  unsigned long long* lfirst = lrows + G;                // [G]
  unsigned long long* lmn = lfirst + G;                  // [G][nm] (MINMAX)
  unsigned long long* lmx = lmn + (MINMAX ? G * nm : 0); // [G][nm] (MINMAX)
  // Each LDS word is set by exactly one of these loops (two loops writing
  // one word from different waves would race without a barrier).
  const int zero_words = G * nm * 2 + G;  // sums, non-NULL counts, rows
  for (int i = threadIdx.x; i < zero_words; i += kBlock) lds[i] = 0;
  for (int i = threadIdx.x; i < G; i += kBlock) lfirst[i] = INT32_MAX;
  if (MINMAX)
    for (int i = threadIdx.x; i < G * nm; i += kBlock) {
      lmn[i] = M.is_real[i % nm] ? static_cast<unsigned long long>(__double_as_longlong(__builtin_inf()))
                                 : static_cast<unsigned long long>(INT64_MAX);
      lmx[i] = M.is_real[i % nm] ? static_cast<unsigned long long>(__double_as_longlong(-__builtin_inf()))
                                 : static_cast<unsigned long long>(INT64_MIN);
    }
  __syncthreads();
  const int64_t T = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; r < n; r += T) {
    if (!eval_preds(P, r)) continue;
    const int g = group_of(K, r);
    if (g < 0 || g >= G) continue;  // codes outside [0, card) are not groups
    atomicAdd(&lrows[g], 1ull);
    atomicMin(&lfirst[g], static_cast<unsigned long long>(r));
    for (int m = 0; m < nm; m++) {
      bool vn;
      const uint64_t v = load_measure(M, m, r, vn);
      if (vn) continue;
      atomicAdd(&lnn[g * nm + m], 1ull);
      if (M.is_real[m]) {
        atomicAdd(reinterpret_cast<double*>(&lsum[g * nm + m]), __longlong_as_double(v));
        if (MINMAX) {
          atomicMin(reinterpret_cast<double*>(&lmn[g * nm + m]), __longlong_as_double(v));
          atomicMax(reinterpret_cast<double*>(&lmx[g * nm + m]), __longlong_as_double(v));
        }
      } else {
        atomicAdd(&lsum[g * nm + m], static_cast<unsigned long long>(v));
        if (MINMAX) {
          atomicMin(reinterpret_cast<long long*>(&lmn[g * nm + m]), static_cast<long long>(v));
          atomicMax(reinterpret_cast<long long*>(&lmx[g * nm + m]), static_cast<long long>(v));
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < G; i += kBlock) {
    if (lrows[i] == 0) continue;
    atomicAdd(&grows[i], lrows[i]);
    atomicMin(&gfirst[i], static_cast<int32_t>(lfirst[i]));
  }
  for (int i = threadIdx.x; i < G * nm; i += kBlock) {
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

// Deterministic mode: the group of every row (-1 when the row does not
// qualify or its code is not a group), the items of det_fp_sums.
__global__ void k_det_gid_dense(PredSet P, KeySet K, int64_t n, int G, int32_t* gid) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int g = -1;
  if (eval_preds(P, r)) {
    g = group_of(K, r);
    if (g < 0 || g >= G) g = -1;
  }
  gid[r] = g;
}

__global__ void k_init_lds_final(AggFinal* fin, int G, MeasureSet M, unsigned long long* rows, int32_t* first) {
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
    first[i] = INT32_MAX;
  }
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// For COUNT(col): whether each group's first row has a NULL in `nulls`.
__global__ void k_first_null(const int32_t* first, int G, const int8_t* nulls, int8_t* out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G) out[g] = (first[g] != INT32_MAX && nulls) ? nulls[first[g]] : 0;
}

}  // namespace kg

using namespace kg;

namespace kg {

int build_preds(kg_ctx* ctx, const kg_col* cols, int32_t ncols, const kg_pred* preds, int32_t npreds,
                int32_t conj, int64_t n, PredSet& p) {
  KG_REQUIRE(ctx, npreds >= 0 && npreds <= kMaxPreds, "at most %d predicates (got %d)", kMaxPreds, npreds);
  KG_REQUIRE(ctx, npreds == 0 || preds, "preds is NULL");
  std::memset(&p, 0, sizeof(p));
  p.n = npreds;
  p.conj = conj ? 1 : 0;
  for (int i = 0; i < npreds; i++) {
    const kg_pred& q = preds[i];
    KG_REQUIRE(ctx, q.col >= 0 && q.col < ncols, "predicate %d references column %d of %d", i, q.col, ncols);
    const kg_col& c = cols[q.col];
    KG_REQUIRE(ctx, n == 0 || c.data, "predicate column %d has no data", q.col);
    if (c.type == KG_TEXT) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "TEXT predicates are not supported here");
    KG_REQUIRE(ctx, type_width(c.type) > 0, "predicate column %d has bad type %d", q.col, c.type);
    KG_REQUIRE(ctx, q.op >= KG_EQ && q.op <= KG_GEQ, "bad comparison op %d", q.op);
    p.type[i] = c.type;
    p.op[i] = q.op;
    p.data[i] = c.data;
    p.nulls[i] = c.nulls;
    p.ilit[i] = q.ilit;
    p.flit[i] = q.flit;
  }
  return KG_OK;
}

}  // namespace kg

extern "C" int kg_scan_select(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols,
                              const kg_pred* preds, int32_t npreds, int32_t conjunction,
                              int32_t* d_out_rows, int64_t* out_count) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count %lld out of range", (long long)n);
  KG_REQUIRE(ctx, out_count, "out_count is NULL");
  KG_REQUIRE(ctx, n == 0 || d_out_rows, "d_out_rows is NULL");
  PredSet p;
  int st = build_preds(ctx, cols, ncols, preds, npreds, conjunction, n, p);
  if (st != KG_OK) return st;
  *out_count = 0;
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const int64_t ntiles = ceil_div(n, kSelTile);
  const size_t need = ntiles * 4 + ntiles * 8 + 64;
  if ((st = kg_work_reserve(ctx, need)) != KG_OK) return st;
  if ((st = kg_pinned_reserve(ctx, 64)) != KG_OK) return st;
  int32_t* counts = static_cast<int32_t*>(ctx->work);
  int64_t* offsets = reinterpret_cast<int64_t*>(static_cast<char*>(ctx->work) + ((ntiles * 4 + 7) / 8) * 8);
  int64_t* total = offsets + ntiles;
  hipLaunchKernelGGL(k_select_count, dim3(ntiles), dim3(kBlock), 0, ctx->stream, p, n, counts);
  KG_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(1024), 0, ctx->stream, counts, offsets, ntiles, total);
  KG_LAUNCH_CHECK(ctx);
  kg_time_begin(ctx);
  hipLaunchKernelGGL(k_select_write, dim3(ntiles), dim3(kBlock), 0, ctx->stream, p, n, offsets, d_out_rows);
  kg_time_end(ctx);
  KG_LAUNCH_CHECK(ctx);
  KG_HIP_CHECK(ctx, hipMemcpyAsync(ctx->pinned, total, 8, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *out_count = *static_cast<int64_t*>(ctx->pinned);
  return KG_OK;
}

// kg_scan_select with its tile counts, offsets and total in the device's
// static scratch array (ctx->scratch) instead of the context workspace, and
// the total read back into a stack variable: no device or pinned allocation,
// no event. The same kernels; every scratch word it reads was written earlier
// in the same call (counts by k_select_count, offsets and total by the scan).
extern "C" int kg_scan_select_noalloc(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols,
                                      const kg_pred* preds, int32_t npreds, int32_t conjunction,
                                      int32_t* d_out_rows, int64_t* out_count) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count %lld out of range", (long long)n);
  KG_REQUIRE(ctx, out_count, "out_count is NULL");
  KG_REQUIRE(ctx, n == 0 || d_out_rows, "d_out_rows is NULL");
  const int64_t ntiles = ceil_div(n, kSelTile);
  if (ntiles > kNoAllocMaxTiles)
    return kg_fail(ctx, KG_ERR_CAPACITY,
                   "kg_scan_select_noalloc: %lld rows exceed its scratch limit of %lld rows (use kg_scan_select)",
                   (long long)n, (long long)(kNoAllocMaxTiles * kSelTile));
  KG_REQUIRE(ctx, ctx->scratch, "no device scratch for device %d", ctx->device);
  PredSet p;
  int st = build_preds(ctx, cols, ncols, preds, npreds, conjunction, n, p);
  if (st != KG_OK) return st;
  *out_count = 0;
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  char* s = static_cast<char*>(ctx->scratch);
  int32_t* counts = reinterpret_cast<int32_t*>(s);
  int64_t* offsets = reinterpret_cast<int64_t*>(s + kNoAllocMaxTiles * 4);
  int64_t* total = reinterpret_cast<int64_t*>(s + kNoAllocMaxTiles * 12);
  hipLaunchKernelGGL(k_select_count, dim3(ntiles), dim3(kBlock), 0, ctx->stream, p, n, counts);
  KG_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(1024), 0, ctx->stream, counts, offsets, ntiles, total);
  KG_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_select_write, dim3(ntiles), dim3(kBlock), 0, ctx->stream, p, n, offsets, d_out_rows);
  KG_LAUNCH_CHECK(ctx);
  int64_t count = 0;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(&count, total, 8, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *out_count = count;
  return KG_OK;
}

namespace {

template <int G, bool MM>
void launch_reg(hipStream_t s, int nb, const PredSet& P, const KeySet& K, const MeasureSet& M, int64_t n,
                int Grt, const AggOut& o) {
  bool nullable = false;
  for (int m = 0; m < M.nm; m++) nullable |= M.nulls[m] != nullptr || M.nulls2[m] != nullptr;
  if (nullable)
    hipLaunchKernelGGL((k_group_agg_reg<G, MM, true>), dim3(nb), dim3(kBlock), 0, s, P, K, M, n, Grt, o);
  else
    hipLaunchKernelGGL((k_group_agg_reg<G, MM, false>), dim3(nb), dim3(kBlock), 0, s, P, K, M, n, Grt, o);
}

bool launch_reg_any(int G, bool mm, hipStream_t s, int nb, const PredSet& P, const KeySet& K,
                    const MeasureSet& M, int64_t n, const AggOut& o) {
  if (mm) {
    if (G <= 1) launch_reg<1, true>(s, nb, P, K, M, n, G, o);
    else if (G <= 2) launch_reg<2, true>(s, nb, P, K, M, n, G, o);
    else if (G <= 4) launch_reg<4, true>(s, nb, P, K, M, n, G, o);
    else return false;
  } else {
    if (G <= 1) launch_reg<1, false>(s, nb, P, K, M, n, G, o);
    else if (G <= 2) launch_reg<2, false>(s, nb, P, K, M, n, G, o);
    else if (G <= 4) launch_reg<4, false>(s, nb, P, K, M, n, G, o);
    else if (G <= 6) launch_reg<6, false>(s, nb, P, K, M, n, G, o);
    else if (G <= 8) launch_reg<8, false>(s, nb, P, K, M, n, G, o);
    else return false;
  }
  return true;
}

template <int G, int NF>
void launch_v4_gf(hipStream_t s, int nb, const FastSet& S, int64_t n, int Grt, const AggOut& o) {
  if (S.ni == 0)
    hipLaunchKernelGGL((k_group_agg_v4<G, NF, 0>), dim3(nb), dim3(kBlock), 0, s, S, n, Grt, o);
  else if (S.ni <= 1)
    hipLaunchKernelGGL((k_group_agg_v4<G, NF, 1>), dim3(nb), dim3(kBlock), 0, s, S, n, Grt, o);
  else if (S.ni <= 3)
    hipLaunchKernelGGL((k_group_agg_v4<G, NF, 3>), dim3(nb), dim3(kBlock), 0, s, S, n, Grt, o);
  else
    hipLaunchKernelGGL((k_group_agg_v4<G, NF, 4>), dim3(nb), dim3(kBlock), 0, s, S, n, Grt, o);
}

template <int G>
void launch_v4_g(hipStream_t s, int nb, const FastSet& S, int64_t n, int Grt, const AggOut& o) {
  if (S.nf == 0) launch_v4_gf<G, 0>(s, nb, S, n, Grt, o);
  else if (S.nf <= 1) launch_v4_gf<G, 1>(s, nb, S, n, Grt, o);
  else if (S.nf <= 2) launch_v4_gf<G, 2>(s, nb, S, n, Grt, o);
  else launch_v4_gf<G, 4>(s, nb, S, n, Grt, o);
}

bool launch_v4(int G, hipStream_t s, int nb, const FastSet& S, int64_t n, const AggOut& o) {
  if (G <= 1) launch_v4_g<1>(s, nb, S, n, G, o);
  else if (G <= 2) launch_v4_g<2>(s, nb, S, n, G, o);
  else if (G <= 4) launch_v4_g<4>(s, nb, S, n, G, o);
  else if (G <= 6) launch_v4_g<6>(s, nb, S, n, G, o);
  else if (G <= 8) launch_v4_g<8>(s, nb, S, n, G, o);
  else return false;
  return true;
}

// The k_group_agg_v4 instance launch_v4 runs for a group count and measure set.
template <int G, int NF>
const void* v4_kernel_gf(const FastSet& S) {
  if (S.ni == 0) return reinterpret_cast<const void*>(&k_group_agg_v4<G, NF, 0>);
  if (S.ni <= 1) return reinterpret_cast<const void*>(&k_group_agg_v4<G, NF, 1>);
  if (S.ni <= 3) return reinterpret_cast<const void*>(&k_group_agg_v4<G, NF, 3>);
  return reinterpret_cast<const void*>(&k_group_agg_v4<G, NF, 4>);
}

template <int G>
const void* v4_kernel_g(const FastSet& S) {
  if (S.nf == 0) return v4_kernel_gf<G, 0>(S);
  if (S.nf <= 1) return v4_kernel_gf<G, 1>(S);
  if (S.nf <= 2) return v4_kernel_gf<G, 2>(S);
  return v4_kernel_gf<G, 4>(S);
}

const void* v4_kernel(int G, const FastSet& S) {
  if (G <= 1) return v4_kernel_g<1>(S);
  if (G <= 2) return v4_kernel_g<2>(S);
  if (G <= 4) return v4_kernel_g<4>(S);
  if (G <= 6) return v4_kernel_g<6>(S);
  return v4_kernel_g<8>(S);
}

// Workgroups of a kernel that fit on one CU at once (its registers and LDS),
// cached per kernel; 0 when the runtime cannot tell.
int resident_per_cu(const void* kernel) {
  static std::mutex mu;
  static std::unordered_map<const void*, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(kernel);
  if (it != cache.end()) return it->second;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess) {
    (void)hipGetLastError();
    per_cu = 0;
  }
  cache[kernel] = per_cu;
  return per_cu;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// The fast path applies to this predicate / key / measure set.
bool fast_set(const PredSet& P, const KeySet& K, const MeasureSet& M, bool minmax, FastSet& S) {
  if (minmax || !P.conj || P.n > kFastPreds) return false;
  std::memset(&S, 0, sizeof(S));
  auto is_i32 = [](int32_t t) { return t == KG_INT || t == KG_DATE || t == KG_ENUM; };
  S.npred = P.n;
  for (int i = 0; i < P.n; i++) {
    if (!is_i32(P.type[i]) || P.nulls[i] || !aligned16(P.data[i])) return false;
    S.op[i] = P.op[i];
    S.lit[i] = P.ilit[i];
    S.pcol[i] = static_cast<const int32_t*>(P.data[i]);
  }
  S.nk = K.nk;
  for (int k = 0; k < K.nk; k++) {
    if (!aligned16(K.data[k])) return false;
    S.key[k] = K.data[k];
  }
  S.card1 = K.card1;
  S.nm = M.nm;
  for (int m = 0; m < M.nm; m++) {
    if (M.nulls[m] || M.data2[m] || !aligned16(M.data[m])) return false;
    if (M.type[m] == KG_REAL) {
      if (S.nf == 4) return false;
      S.fidx[S.nf] = m;
      S.fcol[S.nf++] = static_cast<const double*>(M.data[m]);
    } else if (M.type[m] == KG_INT) {
      if (S.ni == 4) return false;
      S.iidx[S.ni] = m;
      S.icol[S.ni++] = static_cast<const int32_t*>(M.data[m]);
    } else {
      return false;
    }
  }
  return true;
}

}  // namespace

// Number of workgroups of the fixed-order aggregation for n rows; the CPU
// oracle's deterministic restatement takes the same value.
extern "C" int kg_agg_blocks_for(int64_t n) {
  int64_t nb = ceil_div(n, static_cast<int64_t>(kBlock) * kAggRowsPerThreadMin * 4);
  if (nb < 1) nb = 1;
  if (nb > kAggMaxBlocks) nb = kAggMaxBlocks;
  return static_cast<int>(nb);
}

extern "C" int kg_ctx_agg_order(const kg_ctx* ctx, int64_t n, int32_t* nblocks, int32_t* vec) {
  if (!ctx || !nblocks || !vec) return KG_ERR_INVALID;
  *nblocks = kg_agg_blocks_for(n);
  *vec = ctx->last_agg_vec;
  return KG_OK;
}

namespace {

// kg_scan_group_agg for a dense key space too large for an on-chip table:
// kg_hash_group_agg over the same key columns, then each group's key codes
// read at its first row.
int dense_as_hashed(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols, const kg_pred* preds, int32_t npreds,
                    const int32_t* key_cols, int32_t nkeys, const kg_agg* aggs, int32_t naggs, int64_t max_groups,
                    int64_t* out_ngroups, int32_t* out_keys, int64_t* out_vals, int8_t* out_nulls) {
  std::vector<int32_t> first(static_cast<size_t>(std::max<int64_t>(max_groups, 1)));
  int st = kg_hash_group_agg(ctx, n, cols, ncols, preds, npreds, key_cols, nkeys, aggs, naggs, max_groups,
                             out_ngroups, first.data(), out_vals, out_nulls);
  if (st != KG_OK || *out_ngroups == 0) return st;
  const int64_t G = *out_ngroups;
  int32_t* d = nullptr;
  if (hipMalloc(&d, static_cast<size_t>(G) * 8) != hipSuccess) {
    (void)hipGetLastError();
    return kg_fail(ctx, KG_ERR_OOM, "device allocation of %lld bytes failed", (long long)G * 8);
  }
  std::vector<int32_t> codes(G);
  hipError_t e = hipMemcpyAsync(d, first.data(), G * 4, hipMemcpyHostToDevice, ctx->stream);
  for (int k = 0; k < nkeys && e == hipSuccess; k++) {
    st = kg_gather(ctx, &cols[key_cols[k]], d, G, d + G, nullptr);
    if (st != KG_OK) break;
    e = hipMemcpyAsync(codes.data(), d + G, G * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    for (int64_t g = 0; g < G && e == hipSuccess; g++) out_keys[g * nkeys + k] = codes[g];
  }
  (void)hipFree(d);
  if (st != KG_OK) return st;
  if (e != hipSuccess) return kg_fail(ctx, KG_ERR_HIP, "group key read-back failed: %s", hipGetErrorString(e));
  return KG_OK;
}

}  // namespace

extern "C" int kg_scan_group_agg(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols,
                                 const kg_pred* preds, int32_t npreds, const int32_t* key_cols,
                                 const int32_t* key_card, int32_t nkeys, const kg_agg* aggs, int32_t naggs,
                                 int64_t max_groups, int64_t* out_ngroups, int32_t* out_keys,
                                 int64_t* out_vals, int8_t* out_nulls) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count %lld out of range", (long long)n);
  KG_REQUIRE(ctx, out_ngroups && (naggs == 0 || (aggs && out_vals && out_nulls)), "null output pointer");
  KG_REQUIRE(ctx, nkeys >= 0 && nkeys <= 2, "1 or 2 dense key columns (got %d)", nkeys);
  KG_REQUIRE(ctx, naggs >= 0 && naggs <= 16, "at most 16 aggregates (got %d)", naggs);
  KG_REQUIRE(ctx, nkeys == 0 || (key_cols && key_card && out_keys), "key arrays are NULL");
  PredSet P;
  int st = build_preds(ctx, cols, ncols, preds, npreds, 1, n, P);
  if (st != KG_OK) return st;

  KeySet K;
  std::memset(&K, 0, sizeof(K));
  K.nk = nkeys;
  int64_t G = 1;
  for (int k = 0; k < nkeys; k++) {
    KG_REQUIRE(ctx, key_cols[k] >= 0 && key_cols[k] < ncols, "bad key column %d", key_cols[k]);
    const kg_col& c = cols[key_cols[k]];
    if (c.type != KG_INT && c.type != KG_ENUM && c.type != KG_DATE)
      return kg_fail(ctx, KG_ERR_UNSUPPORTED, "dense group keys must be 32-bit codes (type %d)", c.type);
    if (c.nulls) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "nullable dense group keys are not supported");
    KG_REQUIRE(ctx, key_card[k] > 0, "key cardinality must be positive");
    K.data[k] = static_cast<const int32_t*>(c.data);
    G *= key_card[k];
  }
  if (nkeys == 2) K.card1 = key_card[1];
  // A dense table that cannot stay on chip (more than kMaxLdsGroups groups, or
  // more LDS than one workgroup gets: 160 KB on gfx950) runs as the hashed
  // GROUP BY over the same key columns: the same groups, aggregates and
  // insertion order.
  bool hashed = G > kMaxLdsGroups;

  MeasureSet M;
  std::vector<AggPlan> plan;
  bool minmax = false;
  if ((st = build_measures(ctx, cols, ncols, aggs, naggs, n, M, plan, minmax)) != KG_OK) return st;

  *out_ngroups = 0;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (!hashed && G > (minmax ? 4 : 8)) {
    const size_t lds_bytes = (G * M.nm * 2 + G * 2 + (minmax ? 2 * G * M.nm : 0)) * 8;
    hashed = static_cast<int64_t>(lds_bytes) > ctx->max_lds;
  }
  if (hashed) {
    st = dense_as_hashed(ctx, n, cols, ncols, preds, npreds, key_cols, nkeys, aggs, naggs, max_groups, out_ngroups,
                         out_keys, out_vals, out_nulls);
    ctx->last_path = "hashed";
    return st;
  }
  const int nm = M.nm;
  std::vector<AggFinal> fin(G * (nm > 0 ? nm : 1));
  std::vector<uint64_t> rowsv(G);
  std::vector<int32_t> firstv(G);
  std::vector<int64_t> first64(G);
  std::vector<std::vector<int8_t>> first_null(naggs);

  const bool use_reg = G <= (minmax ? 4 : 8);
  ctx->last_path = use_reg ? "register" : "lds";
  if (use_reg) {
    const int nb = kg_agg_blocks_for(n);
    const size_t per_block = static_cast<size_t>(G) * (nm > 0 ? nm : 1);
    size_t off = 0;
    auto carve = [&](size_t bytes) {
      size_t o = off;
      off += (bytes + 255) / 256 * 256;
      return o;
    };
    const size_t o_sum = carve(nb * per_block * 8), o_mn = carve(minmax ? nb * per_block * 8 : 8),
                 o_mx = carve(minmax ? nb * per_block * 8 : 8), o_nn = carve(nb * per_block * 4),
                 o_rows = carve(static_cast<size_t>(nb) * G * 4), o_first = carve(G * 4),
                 o_fin = carve(G * (nm > 0 ? nm : 1) * sizeof(AggFinal)), o_rfin = carve(G * 8),
                 o_fnull = carve(G * 16 + 16);
    if ((st = kg_work_reserve(ctx, off)) != KG_OK) return st;
    char* w = static_cast<char*>(ctx->work);
    AggOut o;
    o.sum = reinterpret_cast<uint64_t*>(w + o_sum);
    o.mn = reinterpret_cast<uint64_t*>(w + o_mn);
    o.mx = reinterpret_cast<uint64_t*>(w + o_mx);
    o.nn = reinterpret_cast<uint32_t*>(w + o_nn);
    o.rows = reinterpret_cast<uint32_t*>(w + o_rows);
    o.first = reinterpret_cast<int32_t*>(w + o_first);
    hipLaunchKernelGGL(k_fill_i32, dim3(ceil_div(G, 256)), dim3(256), 0, ctx->stream, o.first, INT32_MAX, G);
    KG_LAUNCH_CHECK(ctx);
    FastSet S;
    const bool fast = fast_set(P, K, M, minmax, S);
    // Outside the deterministic mode the vector kernel takes one wave of
    // resident workgroups (at most nb): a second round of workgroups per CU
    // leaves a tail (60M rows, 6 groups: 1024 workgroups 0.617 ms, 512 at two
    // resident per CU 0.599 ms). The deterministic mode keeps nb, which with
    // the vector width defines its FP64 addition order.
    int nb_run = nb;
    if (fast && !ctx->deterministic) {
      const int per_cu = resident_per_cu(v4_kernel(static_cast<int>(G), S));
      if (per_cu > 0) nb_run = static_cast<int>(std::min<int64_t>(nb, static_cast<int64_t>(per_cu) * ctx->num_cus));
    }
    kg_time_begin(ctx);
    const bool launched = fast ? launch_v4(static_cast<int>(G), ctx->stream, nb_run, S, n, o)
                               : launch_reg_any(static_cast<int>(G), minmax, ctx->stream, nb, P, K, M, n, o);
    kg_time_end(ctx);
    ctx->last_agg_vec = fast ? kVec : 1;
    if (!launched)
      return kg_fail(ctx, KG_ERR_UNSUPPORTED, "internal: no register kernel for %lld groups", (long long)G);
    KG_LAUNCH_CHECK(ctx);
    AggFinal* dfin = reinterpret_cast<AggFinal*>(w + o_fin);
    uint64_t* drows = reinterpret_cast<uint64_t*>(w + o_rfin);
    const int outputs = static_cast<int>(G * nm + G);
    const size_t stage_bytes = static_cast<size_t>(nb_run) * 8 * (minmax ? 4 : 2);
    hipLaunchKernelGGL(k_group_agg_final, dim3(outputs), dim3(64), stage_bytes, ctx->stream, o, nb_run,
                       static_cast<int>(G), M, minmax, dfin, drows);
    KG_LAUNCH_CHECK(ctx);
    KG_HIP_CHECK(ctx, hipMemcpyAsync(fin.data(), dfin, G * nm * sizeof(AggFinal), hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP_CHECK(ctx, hipMemcpyAsync(rowsv.data(), drows, G * 8, hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP_CHECK(ctx, hipMemcpyAsync(firstv.data(), o.first, G * 4, hipMemcpyDeviceToHost, ctx->stream));
    int8_t* dfn = reinterpret_cast<int8_t*>(w + o_fnull);
    for (int a = 0; a < naggs; a++) {
      if (plan[a].kind != KG_AGG_COUNT || !plan[a].count_nulls) continue;
      first_null[a].resize(G);
      hipLaunchKernelGGL(k_first_null, dim3(ceil_div(G, 256)), dim3(256), 0, ctx->stream, o.first,
                         static_cast<int>(G), plan[a].count_nulls, dfn);
      KG_LAUNCH_CHECK(ctx);
      KG_HIP_CHECK(ctx, hipMemcpyAsync(first_null[a].data(), dfn, G, hipMemcpyDeviceToHost, ctx->stream));
      KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  } else {
    size_t off = 0;
    auto carve = [&](size_t bytes) {
      size_t o = off;
      off += (bytes + 255) / 256 * 256;
      return o;
    };
    bool any_real = false;
    for (int m = 0; m < nm; m++) any_real |= M.is_real[m] != 0;
    const bool det = ctx->deterministic && any_real && n > 0;
    const size_t o_fin = carve(G * (nm > 0 ? nm : 1) * sizeof(AggFinal)), o_rows = carve(G * 8),
                 o_first = carve(G * 4), o_fnull = carve(G + 16), o_gid = carve(det ? n * 4 : 0);
    if ((st = kg_work_reserve(ctx, off)) != KG_OK) return st;
    char* w = static_cast<char*>(ctx->work);
    AggFinal* dfin = reinterpret_cast<AggFinal*>(w + o_fin);
    unsigned long long* drows = reinterpret_cast<unsigned long long*>(w + o_rows);
    int32_t* dfirst = reinterpret_cast<int32_t*>(w + o_first);
    const int init_threads = static_cast<int>(std::max<int64_t>(G * nm, G));
    hipLaunchKernelGGL(k_init_lds_final, dim3(ceil_div(init_threads, 256)), dim3(256), 0, ctx->stream, dfin,
                       static_cast<int>(G), M, drows, dfirst);
    KG_LAUNCH_CHECK(ctx);
    const size_t lds_bytes = (G * nm * 2 + G * 2 + (minmax ? 2 * G * nm : 0)) * 8;
    // Up to the whole LDS of a CU (160 KB on gfx950): one workgroup per CU
    // when the table is large, more when it is small.
    const void* kfn = minmax ? reinterpret_cast<const void*>(&k_group_agg_lds<true>)
                             : reinterpret_cast<const void*>(&k_group_agg_lds<false>);
    if ((st = kg_allow_lds(ctx, kfn, lds_bytes)) != KG_OK) return st;
    int nb = static_cast<int>(std::min<int64_t>(std::max<int64_t>(ceil_div(n, kBlock * 16), 1), ctx->num_cus * 4));
    kg_time_begin(ctx);
    if (minmax)
      hipLaunchKernelGGL(k_group_agg_lds<true>, dim3(nb), dim3(kBlock), lds_bytes, ctx->stream, P, K, M, n,
                         static_cast<int>(G), dfin, drows, dfirst);
    else
      hipLaunchKernelGGL(k_group_agg_lds<false>, dim3(nb), dim3(kBlock), lds_bytes, ctx->stream, P, K, M, n,
                         static_cast<int>(G), dfin, drows, dfirst);
    kg_time_end(ctx);
    KG_LAUNCH_CHECK(ctx);
    if (det) {
      // Deterministic mode: the REAL sums again, in the fixed order.
      int32_t* gid = reinterpret_cast<int32_t*>(w + o_gid);
      hipLaunchKernelGGL(k_det_gid_dense, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, P, K, n,
                         static_cast<int>(G), gid);
      KG_LAUNCH_CHECK(ctx);
      if ((st = det_fp_sums(ctx, n, gid, nullptr, G, M, reinterpret_cast<uint64_t*>(dfin),
                            sizeof(AggFinal) / 8)) != KG_OK)
        return st;
    }
    KG_HIP_CHECK(ctx, hipMemcpyAsync(fin.data(), dfin, G * nm * sizeof(AggFinal), hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP_CHECK(ctx, hipMemcpyAsync(rowsv.data(), drows, G * 8, hipMemcpyDeviceToHost, ctx->stream));
    KG_HIP_CHECK(ctx, hipMemcpyAsync(firstv.data(), dfirst, G * 4, hipMemcpyDeviceToHost, ctx->stream));
    int8_t* dfn = reinterpret_cast<int8_t*>(w + o_fnull);
    for (int a = 0; a < naggs; a++) {
      if (plan[a].kind != KG_AGG_COUNT || !plan[a].count_nulls) continue;
      first_null[a].resize(G);
      hipLaunchKernelGGL(k_first_null, dim3(ceil_div(G, 256)), dim3(256), 0, ctx->stream, dfirst,
                         static_cast<int>(G), plan[a].count_nulls, dfn);
      KG_LAUNCH_CHECK(ctx);
      KG_HIP_CHECK(ctx, hipMemcpyAsync(first_null[a].data(), dfn, G, hipMemcpyDeviceToHost, ctx->stream));
      KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  }

  for (int64_t g = 0; g < G; g++) first64[g] = firstv[g];
  return emit_groups(ctx, G, nkeys, K.card1, M, plan, fin, rowsv, first64, first_null, max_groups, out_ngroups,
                     out_keys, out_vals, out_nulls);
}

namespace kg {

int build_measures(kg_ctx* ctx, const kg_col* cols, int32_t ncols, const kg_agg* aggs, int32_t naggs, int64_t n,
                   MeasureSet& M, std::vector<AggPlan>& plan, bool& minmax) {
  std::memset(&M, 0, sizeof(M));
  plan.assign(naggs, AggPlan{});
  minmax = false;
  auto check_col = [&](int32_t ci, const char* what) -> int {
    KG_REQUIRE(ctx, ci >= 0 && ci < ncols, "%s references column %d of %d", what, ci, ncols);
    const kg_col& c = cols[ci];
    KG_REQUIRE(ctx, n == 0 || c.data, "%s column %d has no data", what, ci);
    if (c.type == KG_TEXT || c.type == KG_BOOLEAN)
      return kg_fail(ctx, KG_ERR_UNSUPPORTED, "aggregate over type %d is not supported here", c.type);
    KG_REQUIRE(ctx, type_width(c.type) > 0, "column %d has bad type %d", ci, c.type);
    return KG_OK;
  };
  for (int a = 0; a < naggs; a++) {
    const kg_agg& q = aggs[a];
    AggPlan& pl = plan[a];
    pl.kind = q.kind;
    pl.measure = -1;
    pl.count_nulls = nullptr;
    KG_REQUIRE(ctx, q.kind >= KG_AGG_SUM && q.kind <= KG_AGG_COUNT, "bad aggregate kind %d", q.kind);
    if (q.kind == KG_AGG_COUNT && q.col < 0) {
      pl.out_type = KG_BIGINT;
      continue;
    }
    int st = check_col(q.col, "aggregate");
    if (st != KG_OK) return st;
    const kg_col& c = cols[q.col];
    const kg_col* c2 = nullptr;
    if (q.col2 >= 0) {
      if ((st = check_col(q.col2, "aggregate factor")) != KG_OK) return st;
      c2 = &cols[q.col2];
      if (c.type == KG_DATE || c.type == KG_ENUM || c2->type == KG_DATE || c2->type == KG_ENUM)
        return kg_fail(ctx, KG_ERR_INVALID, "products of non-numeric columns");
    }
    // The input's SQL type: a product is REAL (price * Float(qty)).
    const int32_t in_type = c2 ? KG_REAL : c.type;
    if (q.kind == KG_AGG_COUNT) {
      pl.out_type = KG_BIGINT;
      pl.count_nulls = c2 ? nullptr : c.nulls;
      if (c2 && (c.nulls || c2->nulls))
        return kg_fail(ctx, KG_ERR_UNSUPPORTED, "COUNT over a product of nullable columns");
    } else if (q.kind == KG_AGG_AVG) {
      pl.out_type = KG_REAL;
    } else {
      if (q.kind == KG_AGG_SUM && (in_type == KG_DATE || in_type == KG_ENUM))
        return kg_fail(ctx, KG_ERR_INVALID, "cannot compute sum of non-numeric col");
      if ((q.kind == KG_AGG_MIN || q.kind == KG_AGG_MAX) && in_type == KG_ENUM)
        return kg_fail(ctx, KG_ERR_UNSUPPORTED, "MIN/MAX over ENUM compares decoded strings");
      pl.out_type = in_type;
    }
    if (q.kind == KG_AGG_MIN || q.kind == KG_AGG_MAX) minmax = true;
    int found = -1;
    for (int m = 0; m < M.nm; m++)
      if (M.data[m] == c.data && M.data2[m] == (c2 ? c2->data : nullptr)) found = m;
    if (found < 0) {
      KG_REQUIRE(ctx, M.nm < kMaxMeasures, "at most %d distinct aggregate inputs", kMaxMeasures);
      found = M.nm++;
      M.is_real[found] = in_type == KG_REAL;
      M.type[found] = c.type;
      M.data[found] = c.data;
      M.nulls[found] = c.nulls;
      if (c2) {
        M.type2[found] = c2->type;
        M.data2[found] = c2->data;
        M.nulls2[found] = c2->nulls;
      }
    }
    pl.measure = found;
  }
  return KG_OK;
}

int emit_groups(kg_ctx* ctx, int64_t G, int32_t nkeys, int32_t card1, const MeasureSet& M,
                const std::vector<AggPlan>& plan, const std::vector<AggFinal>& fin,
                const std::vector<uint64_t>& rowsv, const std::vector<int64_t>& firstv,
                const std::vector<std::vector<int8_t>>& first_null, int64_t max_groups, int64_t* out_ngroups,
                int32_t* out_keys, int64_t* out_vals, int8_t* out_nulls) {
  const int naggs = static_cast<int>(plan.size());
  const int nm = M.nm;
  std::vector<int32_t> present;
  for (int64_t g = 0; g < G; g++)
    if (rowsv[g] > 0) present.push_back(static_cast<int32_t>(g));
  std::sort(present.begin(), present.end(), [&](int32_t a, int32_t b) { return firstv[a] < firstv[b]; });
  const int64_t ng = static_cast<int64_t>(present.size());
  *out_ngroups = ng;
  if (ng > max_groups)
    return kg_fail(ctx, KG_ERR_CAPACITY, "%lld groups exceed max_groups %lld", (long long)ng, (long long)max_groups);
  for (int64_t i = 0; i < ng; i++) {
    const int32_t g = present[i];
    if (nkeys == 1) out_keys[i] = g;
    if (nkeys == 2) {
      out_keys[i * 2] = g / card1;
      out_keys[i * 2 + 1] = g % card1;
    }
    for (int a = 0; a < naggs; a++) {
      const AggPlan& pl = plan[a];
      int64_t& v = out_vals[i * naggs + a];
      int8_t& vn = out_nulls[i * naggs + a];
      vn = 0;
      if (pl.kind == KG_AGG_COUNT) {
        if (pl.measure < 0) {
          v = static_cast<int64_t>(rowsv[g]);
        } else {
          // CountAggregator::Initialize stores 1 even for a NULL first value
          // (compile/proxy/aggregator.cc:285-287).
          const AggFinal& f = fin[g * nm + pl.measure];
          v = static_cast<int64_t>(f.nn) + (first_null[a].empty() ? 0 : first_null[a][g]);
        }
        continue;
      }
      const AggFinal& f = fin[g * nm + pl.measure];
      if (f.nn == 0) {
        vn = 1;
        v = 0;
        continue;
      }
      const bool real = M.is_real[pl.measure];
      switch (pl.kind) {
        case KG_AGG_SUM:
          // SUM keeps the child type and wraps in it (aggregator.cc:44-71).
          if (real) {
            v = static_cast<int64_t>(f.sum);
          } else if (pl.out_type == KG_INT) {
            v = static_cast<int32_t>(static_cast<uint32_t>(f.sum));
          } else if (pl.out_type == KG_SMALLINT) {
            v = static_cast<int16_t>(static_cast<uint16_t>(f.sum));
          } else {
            v = static_cast<int64_t>(f.sum);
          }
          break;
        case KG_AGG_MIN: v = static_cast<int64_t>(f.mn); break;
        case KG_AGG_MAX: v = static_cast<int64_t>(f.mx); break;
        case KG_AGG_AVG: {
          double s;
          if (real) {
            std::memcpy(&s, &f.sum, 8);
          } else {
            s = static_cast<double>(static_cast<int64_t>(f.sum));
          }
          const double avg = s / static_cast<double>(f.nn);
          std::memcpy(&v, &avg, 8);
          break;
        }
      }
    }
  }
  return KG_OK;
}

}  // namespace kg
