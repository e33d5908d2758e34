// GROUP BY over hashed keys for any number of groups: the device form of
// runtime::AggregateHashTable + proxy::AggregateHashTable::UpdateOrInsert
// (runtime/aggregate_hash_table.cc:9-135, compile/proxy/aggregate_hash_table.cc:
// 187-343) for keys that are not small dense codes.
//
// One pass over the qualifying rows: each row hashes its key tuple (1..4
// fixed-width columns, NULL flags included) into an open-addressing table of
// group slots. A slot holds the row id of a representative row, claimed with
// a CAS; a row matches a slot when its keys equal the representative's, read
// back from the (immutable) key columns, so no partially written key is ever
// compared. The row then adds into its group's accumulators with device
// atomics (64-bit integer adds that wrap, FP64 adds, 64-bit min/max, CAS loops
// for FP64 min/max) and lowers the group's first row. Groups are compacted on
// the device and emitted in first-row order (ForEach's insertion order,
// aggregate_hash_table.cc:245-291) through the same aggregate plan as the dense
// path (SUM wraps in the child type, AVG = sum / count, COUNT(col) counts the
// first row whatever its NULL flag).
//
// Key equality follows the generated comparison: integers by value, REAL by
// `==` (so -0.0 and 0.0 are one group and a NaN key matches nothing), a NULL
// key component equal only to NULL.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "kg_ops.h"

namespace kg {
namespace {

constexpr int kGBlock = 256;
constexpr int kMaxHashKeys = 4;

struct HKeys {
  int32_t nk;
  int32_t type[kMaxHashKeys];
  const void* data[kMaxHashKeys];
  const int8_t* nulls[kMaxHashKeys];
};

// A key component as 8 comparable bytes: integers sign-extended, REAL with
// -0.0 folded into +0.0 (equal under ==).
__device__ __forceinline__ uint64_t key_bits(const HKeys& K, int k, int64_t r) {
  if (K.type[k] == KG_REAL) {
    const double x = static_cast<const double*>(K.data[k])[r];
    return static_cast<uint64_t>(__double_as_longlong(x == 0.0 ? 0.0 : x));
  }
  return static_cast<uint64_t>(load_int(K.type[k], K.data[k], r));
}

__device__ __forceinline__ bool key_null(const HKeys& K, int k, int64_t r) {
  return K.nulls[k] && K.nulls[k][r];
}

__device__ __forceinline__ uint64_t key_hash(const HKeys& K, int64_t r) {
  uint64_t h = 0x2545F4914F6CDD1Dull;
  for (int k = 0; k < K.nk; k++) {
    const bool nl = key_null(K, k, r);
    h = mix64(h ^ (nl ? 0x9E3779B97F4A7C15ull : key_bits(K, k, r)) ^ (static_cast<uint64_t>(k) << 56));
  }
  return h;
}

__device__ __forceinline__ bool keys_equal(const HKeys& K, int64_t a, int64_t b) {
  for (int k = 0; k < K.nk; k++) {
    const bool na = key_null(K, k, a), nb = key_null(K, k, b);
    if (na != nb) return false;
    if (na) continue;
    if (K.type[k] == KG_REAL) {
      if (!(static_cast<const double*>(K.data[k])[a] == static_cast<const double*>(K.data[k])[b])) return false;
    } else if (load_int(K.type[k], K.data[k], a) != load_int(K.type[k], K.data[k], b)) {
      return false;
    }
  }
  return true;
}

struct GTable {
  int32_t* owner;               // [nslots] representative row, -1 free
  unsigned long long* rows;     // [nslots]
  int32_t* first;               // [nslots] first qualifying row, INT32_MAX if none
  unsigned long long* sum;      // [nslots * nm]
  unsigned long long* nn;       // [nslots * nm]
  unsigned long long* mn;       // [nslots * nm] (minmax)
  unsigned long long* mx;       // [nslots * nm] (minmax)
  uint64_t mask;
};

__device__ __forceinline__ void atomic_min_f64(unsigned long long* p, double x) {
  unsigned long long old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (x < __longlong_as_double(static_cast<long long>(old))) {
    const unsigned long long prev = atomicCAS(p, old, static_cast<unsigned long long>(__double_as_longlong(x)));
    if (prev == old) break;
    old = prev;
  }
}

__device__ __forceinline__ void atomic_max_f64(unsigned long long* p, double x) {
  unsigned long long old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (x > __longlong_as_double(static_cast<long long>(old))) {
    const unsigned long long prev = atomicCAS(p, old, static_cast<unsigned long long>(__double_as_longlong(x)));
    if (prev == old) break;
    old = prev;
  }
}

// Every claimed slot is one group; the claim that takes the count past
// max_groups raises *overflow, and every thread stops at its next row once it
// sees the flag, so an input with far more distinct keys than max_groups ends
// after about one table's worth of claims instead of probing a full table
// for each remaining row.
__global__ __launch_bounds__(kGBlock) void k_hash_group(PredSet P, HKeys K, MeasureSet M, int32_t minmax, int64_t n,
                                                         GTable T, int* overflow, unsigned long long* claimed,
                                                         int64_t max_groups, int32_t* row_slot) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kGBlock;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kGBlock + threadIdx.x; r < n; r += stride) {
    if (__hip_atomic_load(overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    if (row_slot) row_slot[r] = -1;
    if (!eval_preds(P, r)) continue;
    uint64_t s = key_hash(K, r) & T.mask;
    int64_t slot = -1;
    for (uint64_t probes = 0; probes <= T.mask; probes++) {
      int32_t o = __hip_atomic_load(&T.owner[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (o < 0) {
        const int32_t prev = atomicCAS(&T.owner[s], -1, static_cast<int32_t>(r));
        if (prev < 0) {
          slot = static_cast<int64_t>(s);
          if (static_cast<int64_t>(atomicAdd(claimed, 1ull)) >= max_groups) atomicOr(overflow, 1);
          break;
        }
        o = prev;
      }
      if (keys_equal(K, r, o)) {
        slot = static_cast<int64_t>(s);
        break;
      }
      s = (s + 1) & T.mask;
    }
    if (slot < 0) {
      atomicOr(overflow, 1);
      continue;
    }
    if (row_slot) row_slot[r] = static_cast<int32_t>(slot);
    atomicAdd(&T.rows[slot], 1ull);
    atomicMin(&T.first[slot], static_cast<int32_t>(r));
    for (int m = 0; m < M.nm; m++) {
      bool null;
      const uint64_t v = load_measure(M, m, r, null);
      if (null) continue;
      const int64_t i = slot * M.nm + m;
      atomicAdd(&T.nn[i], 1ull);
      if (M.is_real[m]) {
        const double x = __longlong_as_double(static_cast<long long>(v));
        atomicAdd(reinterpret_cast<double*>(&T.sum[i]), x);
        if (minmax) {
          atomic_min_f64(&T.mn[i], x);
          atomic_max_f64(&T.mx[i], x);
        }
      } else {
        atomicAdd(&T.sum[i], static_cast<unsigned long long>(v));
        if (minmax) {
          atomicMin(reinterpret_cast<long long*>(&T.mn[i]), static_cast<long long>(v));
          atomicMax(reinterpret_cast<long long*>(&T.mx[i]), static_cast<long long>(v));
        }
      }
    }
  }
}

__global__ void k_gtable_init(GTable T, int64_t nslots, int32_t nm, int32_t minmax, MeasureSet M) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < nslots; s += stride) {
    T.owner[s] = -1;
    T.rows[s] = 0;
    T.first[s] = INT32_MAX;
    for (int m = 0; m < nm; m++) {
      T.sum[s * nm + m] = M.is_real[m] ? static_cast<unsigned long long>(__double_as_longlong(0.0)) : 0ull;
      T.nn[s * nm + m] = 0;
      if (minmax) {
        T.mn[s * nm + m] = M.is_real[m] ? static_cast<unsigned long long>(__double_as_longlong(__builtin_inf()))
                                        : static_cast<unsigned long long>(INT64_MAX);
        T.mx[s * nm + m] = M.is_real[m] ? static_cast<unsigned long long>(__double_as_longlong(-__builtin_inf()))
                                        : static_cast<unsigned long long>(INT64_MIN);
      }
    }
  }
}

__global__ void k_gtable_used(GTable T, int64_t nslots, uint8_t* used, int32_t* ids) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < nslots; s += stride) {
    used[s] = T.rows[s] > 0 ? 1 : 0;
    ids[s] = static_cast<int32_t>(s);
  }
}

// Compact record per used group: rows, first row, and per measure the final
// state; NULL flags of the first row for the COUNT(col) inputs.
struct CountNulls {
  int32_t n;
  int32_t agg[16];
  const int8_t* nulls[16];
};

__global__ void k_gtable_pack(GTable T, const int32_t* slots, int64_t G, int32_t nm, int32_t minmax, uint64_t* rows,
                              int32_t* first, AggFinal* fin, CountNulls cn, int8_t* fnull) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += stride) {
    const int64_t s = slots[g];
    rows[g] = T.rows[s];
    first[g] = T.first[s];
    for (int m = 0; m < nm; m++) {
      AggFinal f;
      f.sum = T.sum[s * nm + m];
      f.nn = T.nn[s * nm + m];
      f.mn = minmax ? T.mn[s * nm + m] : 0;
      f.mx = minmax ? T.mx[s * nm + m] : 0;
      fin[g * nm + m] = f;
    }
    for (int c = 0; c < cn.n; c++) fnull[c * G + g] = cn.nulls[c][T.first[s]] ? 1 : 0;
  }
}

int galloc(kg_ctx* ctx, void** p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return kg_fail(ctx, KG_ERR_OOM, "device allocation of %zu bytes failed", bytes);
  }
  return KG_OK;
}

}  // namespace
}  // namespace kg

using namespace kg;

extern "C" int kg_hash_group_agg(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols, const kg_pred* preds,
                                 int32_t npreds, const int32_t* key_cols, int32_t nkeys, const kg_agg* aggs,
                                 int32_t naggs, int64_t max_groups, int64_t* out_ngroups, int32_t* out_first_rows,
                                 int64_t* out_vals, int8_t* out_nulls) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count %lld out of range", (long long)n);
  KG_REQUIRE(ctx, out_ngroups && (naggs == 0 || (aggs && out_vals && out_nulls)), "null output pointer");
  KG_REQUIRE(ctx, nkeys >= 1 && nkeys <= kMaxHashKeys && key_cols, "1..%d key columns (got %d)", kMaxHashKeys, nkeys);
  KG_REQUIRE(ctx, naggs >= 0 && naggs <= 16, "at most 16 aggregates (got %d)", naggs);
  KG_REQUIRE(ctx, max_groups >= 0 && max_groups < (1ll << 30), "max_groups out of range");
  KG_REQUIRE(ctx, max_groups == 0 || out_first_rows, "out_first_rows is NULL");
  *out_ngroups = 0;
  PredSet P;
  int st = build_preds(ctx, cols, ncols, preds, npreds, 1, n, P);
  if (st != KG_OK) return st;
  HKeys K;
  std::memset(&K, 0, sizeof(K));
  K.nk = nkeys;
  for (int k = 0; k < nkeys; k++) {
    KG_REQUIRE(ctx, key_cols[k] >= 0 && key_cols[k] < ncols, "bad key column %d", key_cols[k]);
    const kg_col& c = cols[key_cols[k]];
    if (type_width(c.type) == 0)
      return kg_fail(ctx, KG_ERR_UNSUPPORTED, "group keys must be fixed-width (type %d; TEXT keys group by "
                     "kg_text_hash codes)", c.type);
    KG_REQUIRE(ctx, n == 0 || c.data, "key column %d has no data", key_cols[k]);
    K.type[k] = c.type;
    K.data[k] = c.data;
    K.nulls[k] = c.nulls;
  }
  MeasureSet M;
  std::vector<AggPlan> plan;
  bool minmax = false;
  if ((st = build_measures(ctx, cols, ncols, aggs, naggs, n, M, plan, minmax)) != KG_OK) return st;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  ctx->last_path = "hashed";
  const int nm = M.nm;
  const int64_t cap = std::max<int64_t>(std::min<int64_t>(n, max_groups + 1), 1);
  uint64_t nslots = 64;
  while (nslots < static_cast<uint64_t>(cap) * 2) nslots <<= 1;
  GTable T{};
  T.mask = nslots - 1;
  const size_t mslots = nslots * static_cast<size_t>(nm > 0 ? nm : 1);
  bool any_real = false;
  for (int m = 0; m < nm; m++) any_real |= M.is_real[m] != 0;
  // Deterministic mode: each row's slot, for the fixed-order REAL sums.
  const bool det = ctx->deterministic && any_real && n > 0;
  void* base[9] = {};
  auto release = [&]() {
    for (void* p : base)
      if (p) (void)hipFree(p);
  };
  if ((st = galloc(ctx, &base[0], nslots * 4)) != KG_OK || (st = galloc(ctx, &base[1], nslots * 8)) != KG_OK ||
      (st = galloc(ctx, &base[2], nslots * 4)) != KG_OK || (st = galloc(ctx, &base[3], mslots * 8)) != KG_OK ||
      (st = galloc(ctx, &base[4], mslots * 8)) != KG_OK ||
      (st = galloc(ctx, &base[5], minmax ? mslots * 8 : 8)) != KG_OK ||
      (st = galloc(ctx, &base[6], minmax ? mslots * 8 : 8)) != KG_OK || (st = galloc(ctx, &base[7], 16)) != KG_OK ||
      (st = galloc(ctx, &base[8], det ? n * 4 : 8)) != KG_OK) {
    release();
    return st;
  }
  T.owner = static_cast<int32_t*>(base[0]);
  T.rows = static_cast<unsigned long long*>(base[1]);
  T.first = static_cast<int32_t*>(base[2]);
  T.sum = static_cast<unsigned long long*>(base[3]);
  T.nn = static_cast<unsigned long long*>(base[4]);
  T.mn = static_cast<unsigned long long*>(base[5]);
  T.mx = static_cast<unsigned long long*>(base[6]);
  int* overflow = static_cast<int*>(base[7]);
  auto grid = [&](int64_t work) {
    return static_cast<int>(std::min<int64_t>(std::max<int64_t>(ceil_div(work, kGBlock * 4), 1),
                                              static_cast<int64_t>(ctx->num_cus) * 16));
  };
  auto fail = [&](int s2) {
    release();
    return s2;
  };
  if (hipMemsetAsync(overflow, 0, 16, ctx->stream) != hipSuccess) return fail(kg_fail(ctx, KG_ERR_HIP, "memset"));
  hipLaunchKernelGGL(k_gtable_init, dim3(grid(nslots)), dim3(kGBlock), 0, ctx->stream, T, static_cast<int64_t>(nslots),
                     nm, minmax ? 1 : 0, M);
  if (hipGetLastError() != hipSuccess) return fail(kg_fail(ctx, KG_ERR_HIP, "init launch failed"));
  kg_time_begin(ctx);
  if (n > 0)
    hipLaunchKernelGGL(k_hash_group, dim3(grid(n)), dim3(kGBlock), 0, ctx->stream, P, K, M, minmax ? 1 : 0, n, T,
                       overflow, reinterpret_cast<unsigned long long*>(static_cast<char*>(base[7]) + 8), max_groups,
                       det ? static_cast<int32_t*>(base[8]) : nullptr);
  kg_time_end(ctx);
  if (hipGetLastError() != hipSuccess) return fail(kg_fail(ctx, KG_ERR_HIP, "group launch failed"));
  int hov = 0;
  if (hipMemcpyAsync(&hov, overflow, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return fail(kg_fail(ctx, KG_ERR_HIP, "overflow read-back failed"));
  if (hov) return fail(kg_fail(ctx, KG_ERR_CAPACITY, "more than %lld groups", (long long)max_groups));
  if (det && (st = det_fp_sums(ctx, n, static_cast<int32_t*>(base[8]), nullptr, static_cast<int64_t>(nslots), M,
                               reinterpret_cast<uint64_t*>(T.sum), 1)) != KG_OK)
    return fail(st);

  // Used slots, compacted.
  size_t cub = 0;
  hipcub::DeviceSelect::Flagged(nullptr, cub, static_cast<int32_t*>(nullptr), static_cast<uint8_t*>(nullptr),
                                static_cast<int32_t*>(nullptr), static_cast<int64_t*>(nullptr),
                                static_cast<int>(nslots), ctx->stream);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_used = carve(nslots), o_ids = carve(nslots * 4), o_slots = carve(nslots * 4), o_cnt = carve(16),
               o_cub = carve(cub);
  if ((st = kg_work_reserve(ctx, off)) != KG_OK) return fail(st);
  char* w = static_cast<char*>(ctx->work);
  uint8_t* used = reinterpret_cast<uint8_t*>(w + o_used);
  int32_t* ids = reinterpret_cast<int32_t*>(w + o_ids);
  int32_t* slots = reinterpret_cast<int32_t*>(w + o_slots);
  int64_t* cnt = reinterpret_cast<int64_t*>(w + o_cnt);
  hipLaunchKernelGGL(k_gtable_used, dim3(grid(nslots)), dim3(kGBlock), 0, ctx->stream, T,
                     static_cast<int64_t>(nslots), used, ids);
  size_t cb = cub;
  if (hipcub::DeviceSelect::Flagged(w + o_cub, cb, ids, used, slots, cnt, static_cast<int>(nslots), ctx->stream) !=
      hipSuccess)
    return fail(kg_fail(ctx, KG_ERR_HIP, "group compaction failed"));
  int64_t G = 0;
  if (hipMemcpyAsync(&G, cnt, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return fail(kg_fail(ctx, KG_ERR_HIP, "group count read-back failed"));
  *out_ngroups = G;
  if (G > max_groups)
    return fail(kg_fail(ctx, KG_ERR_CAPACITY, "%lld groups exceed max_groups %lld", (long long)G,
                        (long long)max_groups));
  if (G == 0) {
    release();
    return KG_OK;
  }
  CountNulls cn{};
  for (int a = 0; a < naggs; a++)
    if (plan[a].kind == KG_AGG_COUNT && plan[a].count_nulls) {
      cn.agg[cn.n] = a;
      cn.nulls[cn.n++] = plan[a].count_nulls;
    }
  void* pk[4] = {};
  auto release2 = [&]() {
    for (void* p : pk)
      if (p) (void)hipFree(p);
  };
  const size_t gm = static_cast<size_t>(G) * (nm > 0 ? nm : 1);
  if ((st = galloc(ctx, &pk[0], G * 8)) != KG_OK || (st = galloc(ctx, &pk[1], G * 4)) != KG_OK ||
      (st = galloc(ctx, &pk[2], gm * sizeof(AggFinal))) != KG_OK ||
      (st = galloc(ctx, &pk[3], static_cast<size_t>(G) * (cn.n > 0 ? cn.n : 1))) != KG_OK) {
    release2();
    return fail(st);
  }
  hipLaunchKernelGGL(k_gtable_pack, dim3(grid(G)), dim3(kGBlock), 0, ctx->stream, T, slots, G, nm, minmax ? 1 : 0,
                     static_cast<uint64_t*>(pk[0]), static_cast<int32_t*>(pk[1]), static_cast<AggFinal*>(pk[2]), cn,
                     static_cast<int8_t*>(pk[3]));
  std::vector<uint64_t> rowsv(G);
  std::vector<int32_t> first32(G);
  std::vector<AggFinal> fin(gm);
  std::vector<int8_t> fnull(static_cast<size_t>(G) * cn.n);
  hipError_t e = hipMemcpyAsync(rowsv.data(), pk[0], G * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(first32.data(), pk[1], G * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && nm > 0)
    e = hipMemcpyAsync(fin.data(), pk[2], gm * sizeof(AggFinal), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && cn.n > 0) e = hipMemcpyAsync(fnull.data(), pk[3], fnull.size(), hipMemcpyDeviceToHost,
                                                      ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  release2();
  release();
  if (e != hipSuccess) return kg_fail(ctx, KG_ERR_HIP, "group read-back failed: %s", hipGetErrorString(e));
  // Insertion order: ascending first row.
  std::vector<int64_t> order(G);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return first32[a] < first32[b]; });
  std::vector<uint64_t> rows_o(G);
  std::vector<int64_t> first_o(G);
  std::vector<AggFinal> fin_o(gm);
  std::vector<std::vector<int8_t>> fn_o(naggs);
  for (int c = 0; c < cn.n; c++) fn_o[cn.agg[c]].resize(G);
  for (int64_t i = 0; i < G; i++) {
    const int64_t g = order[i];
    rows_o[i] = rowsv[g];
    first_o[i] = first32[g];
    for (int m = 0; m < nm; m++) fin_o[i * nm + m] = fin[g * nm + m];
    for (int c = 0; c < cn.n; c++) fn_o[cn.agg[c]][i] = fnull[c * G + g];
    out_first_rows[i] = first32[g];
  }
  return emit_groups(ctx, G, 0, 1, M, plan, fin_o, rows_o, first_o, fn_o, max_groups, out_ngroups, nullptr, out_vals,
                     out_nulls);
}
