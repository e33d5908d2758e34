// Shared device-side pieces of the scan / aggregate / join kernels:
// predicate sets, aggregate inputs ("measures"), and the host-side plan that
// turns kg_agg requests into measures and per-group results back into SQL
// values (compile/proxy/aggregator.cc semantics).
#pragma once

#include <vector>

#include "kg_internal.h"

namespace kg {

constexpr int kMaxPreds = 8;
constexpr int kMaxMeasures = 8;
constexpr int kBlock = 256;
constexpr int kSelTile = 4096;          // rows per workgroup in the selection kernels
constexpr int kAggMaxBlocks = 1024;     // workgroups of the fixed-order aggregation
constexpr int kAggRowsPerThreadMin = 8; // n / (kBlock * this) workgroups, capped above
constexpr int kMaxLdsGroups = 1024;

struct PredSet {
  int32_t n;
  int32_t conj;
  int32_t type[kMaxPreds];
  int32_t op[kMaxPreds];
  const void* data[kMaxPreds];
  const int8_t* nulls[kMaxPreds];
  int64_t ilit[kMaxPreds];
  double flit[kMaxPreds];
};

__device__ __forceinline__ bool eval_preds(const PredSet& p, int64_t r) {
  bool pass = p.conj ? true : (p.n == 0);
  for (int i = 0; i < p.n; i++) {
    bool v;
    if (p.type[i] == KG_REAL) {
      v = cmp_f64(p.op[i], static_cast<const double*>(p.data[i])[r], p.flit[i]);
    } else {
      v = cmp_i64(p.op[i], load_int(p.type[i], p.data[i], r), p.ilit[i]);
    }
    if (p.nulls[i] && p.nulls[i][r]) v = false;
    pass = p.conj ? (pass && v) : (pass || v);
  }
  return pass;
}

// Aggregate inputs.
// One measure column: the aggregate input cols[col] (times Float(cols[col2])
// when col2 is set).
struct MeasureSet {
  int32_t nm;
  int32_t is_real[kMaxMeasures];  // accumulate as double (else exact int64)
  int32_t type[kMaxMeasures];
  int32_t type2[kMaxMeasures];
  const void* data[kMaxMeasures];
  const void* data2[kMaxMeasures];  // optional second factor
  const int8_t* nulls[kMaxMeasures];
  const int8_t* nulls2[kMaxMeasures];
};

struct KeySet {
  int32_t nk;
  const int32_t* data[2];
  const int8_t* nulls[2];
  int32_t card1;  // group = code0 * card1 + code1
};

__device__ __forceinline__ int group_of(const KeySet& k, int64_t r) {
  if (k.nk == 0) return 0;
  if (k.nk == 1) return k.data[0][r];
  return k.data[0][r] * k.card1 + k.data[1][r];
}

// Loads measure m of row r: bits of a double (is_real) or an int64, and its NULL flag.
__device__ __forceinline__ uint64_t load_measure(const MeasureSet& M, int m, int64_t r, bool& null) {
  null = (M.nulls[m] && M.nulls[m][r]) || (M.data2[m] && M.nulls2[m] && M.nulls2[m][r]);
  if (M.is_real[m]) {
    double x = M.type[m] == KG_REAL ? static_cast<const double*>(M.data[m])[r]
                                    : static_cast<double>(load_int(M.type[m], M.data[m], r));
    if (M.data2[m]) {
      double y = M.type2[m] == KG_REAL ? static_cast<const double*>(M.data2[m])[r]
                                       : static_cast<double>(load_int(M.type2[m], M.data2[m], r));
      x = x * y;
    }
    return static_cast<uint64_t>(__double_as_longlong(x));
  }
  return static_cast<uint64_t>(load_int(M.type[m], M.data[m], r));
}

// Per-workgroup output record for group g: counts[g], first row, and per
// measure: sum (double bits or int64), non-NULL count, min, max.
struct AggOut {
  uint64_t* sum;     // [G][nm][nblocks]
  uint64_t* mn;      // [G][nm][nblocks] (minmax only)
  uint64_t* mx;      // [G][nm][nblocks] (minmax only)
  uint32_t* nn;      // [G][nm][nblocks] non-NULL counts
  uint32_t* rows;    // [nblocks][G]
  int32_t* first;    // [G] atomicMin across workgroups (init INT32_MAX)
};

__device__ __forceinline__ uint64_t fadd_bits(uint64_t a, uint64_t b) {
  return static_cast<uint64_t>(__double_as_longlong(__longlong_as_double(a) + __longlong_as_double(b)));
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int s) {
  const uint32_t lo = __shfl_xor(static_cast<uint32_t>(v), s);
  const uint32_t hi = __shfl_xor(static_cast<uint32_t>(v >> 32), s);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Order-preserving min/max on the 8-byte value (double or int64).
__device__ __forceinline__ bool val_less(bool is_real, uint64_t a, uint64_t b) {
  return is_real ? __longlong_as_double(a) < __longlong_as_double(b)
                 : static_cast<int64_t>(a) < static_cast<int64_t>(b);
}

// Final per-group, per-measure state: sum (double bits or wrapping int64),
// min, max and the non-NULL count.
struct AggFinal {
  uint64_t sum, mn, mx;
  uint64_t nn;
};

// Host view of one requested aggregate.
struct AggPlan {
  int32_t kind;
  int32_t measure;            // index into MeasureSet, -1 for COUNT(*)
  int32_t out_type;           // SQL type of the result
  const int8_t* count_nulls;  // COUNT(col): the column's NULL array
};

// Validates predicates against cols and fills a PredSet.
// (Column data may be NULL only when n == 0.)
int build_preds(kg_ctx* ctx, const kg_col* cols, int32_t ncols, const kg_pred* preds, int32_t npreds,
                int32_t conj, int64_t n, PredSet& p);

// Validates aggregates, dedups their inputs into measures, and returns the
// plan; minmax tells whether any MIN/MAX is requested.
int build_measures(kg_ctx* ctx, const kg_col* cols, int32_t ncols, const kg_agg* aggs, int32_t naggs, int64_t n,
                   MeasureSet& M, std::vector<AggPlan>& plan, bool& minmax);

// Emits the groups with rows[g] > 0 in ascending first[g] order (insertion
// order) as SQL values: out_vals[i*naggs+a] / out_nulls, and the dense codes
// of group g split into nkeys codes (code0 * card1 + code1).
int emit_groups(kg_ctx* ctx, int64_t G, int32_t nkeys, int32_t card1, const MeasureSet& M,
                const std::vector<AggPlan>& plan, const std::vector<AggFinal>& fin,
                const std::vector<uint64_t>& rows, const std::vector<int64_t>& first,
                const std::vector<std::vector<int8_t>>& first_null, int64_t max_groups, int64_t* out_ngroups,
                int32_t* out_keys, int64_t* out_vals, int8_t* out_nulls);

// Deterministic mode (kg_detsum.hip): for every REAL measure m of M and every
// group g in [0, G) with items, writes the fixed-order FP64 sum of m over the
// group's items to dst[(g * M.nm + m) * stride]. Item k (0..N-1) belongs to
// group gid[k] (outside [0, G): none) and reads its values at row src[k]
// (src NULL: row k); items are taken in ascending k. Synchronizes the stream.
int det_fp_sums(kg_ctx* ctx, int64_t N, const int32_t* gid, const int32_t* src, int64_t G, const MeasureSet& M,
                uint64_t* dst, int32_t stride);


}  // namespace kg
