// Fixed-order FP64 sums for the deterministic mode (kg_ctx_set_deterministic).
//
// The reference folds a group's FP64 SUM sequentially in arrival order
// (compile/proxy/aggregator.cc:67-71; AVG is a running mean, :239-268). The
// device paths that add through LDS or global atomics (dense tables beyond
// registers, the star pipeline, hashed GROUP BY) add in a timing-dependent
// order. In the deterministic mode those paths take their REAL sums from here
// instead, in an order that depends only on the input:
//
//   Items (qualifying rows, or join output rows) keep their input order. Each
//   group's items are split into consecutive chunks of 256. A chunk is summed
//   by 64 lanes: lane l folds items 4l..4l+3 left to right, then the lanes
//   combine pairwise at distances 32, 16, 8, 4, 2, 1 (lane l adds lane l+d,
//   lower lane first). When a group has more than 256 items, its chunk sums
//   are summed the same way, until one value is left. Missing items and NULL
//   values count as -0.0, the exact additive identity of IEEE addition, so
//   they change no bit of the result.
//
// Device plan: a stable radix sort of (group, item) pairs groups the items
// and keeps their order; segment bounds come from key changes; each level is
// one wave per chunk (group found by a binary search over the chunk offsets).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "kg_ops.h"

namespace kg {
namespace {

constexpr int kDetChunk = 256;
constexpr int kDetBlock = 256;

__global__ void k_det_keys(const int32_t* gid, int64_t N, uint32_t G, uint32_t* keys, int32_t* idx) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int32_t g = gid[k];
  keys[k] = (g >= 0 && static_cast<uint32_t>(g) < G) ? static_cast<uint32_t>(g) : G;
  idx[k] = static_cast<int32_t>(k);
}

__global__ void k_det_zero(int64_t* a, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// start[g] = first sorted position of g, end[g] = one past its last.
__global__ void k_det_bounds(const uint32_t* skeys, int64_t N, uint32_t G, int64_t* start, int64_t* end) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const uint32_t g = skeys[k];
  if (g >= G) return;
  if (k == 0 || skeys[k - 1] != g) start[g] = k;
  if (k == N - 1 || skeys[k + 1] != g) end[g] = k + 1;
}

// cnt[g] = end[g] - start[g]; nch[g] = chunks of the next level; nch[G] = 0.
__global__ void k_det_counts(const int64_t* start, const int64_t* end, int64_t G, int64_t* cnt, int64_t* nch) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g > G) return;
  if (g == G) {
    nch[G] = 0;
    return;
  }
  const int64_t c = end[g] - start[g];
  cnt[g] = c;
  nch[g] = (c + kDetChunk - 1) / kDetChunk;
}

// Next level's segments: the chunk sums of group g sit at [choff[g], +nch[g]).
__global__ void k_det_next(const int64_t* choff, const int64_t* nch_in, int64_t G, int64_t* start, int64_t* cnt,
                           int64_t* nch_out) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g > G) return;
  if (g == G) {
    nch_out[G] = 0;
    return;
  }
  start[g] = choff[g];
  cnt[g] = nch_in[g];
  nch_out[g] = (nch_in[g] + kDetChunk - 1) / kDetChunk;
}

struct RealSet {
  int32_t nr;
  int32_t m[kMaxMeasures];  // measure index of each REAL measure
};

__device__ __forceinline__ double det_value(const MeasureSet& M, int m, int64_t row) {
  bool null;
  const uint64_t v = load_measure(M, m, row, null);
  return null ? -0.0 : __longlong_as_double(static_cast<long long>(v));
}

// One wave per chunk w < W. GATHER: the chunk's items are sorted item ids
// (sidx), their rows src[item] (or the item id), the values those rows'
// measures; otherwise the values are in[pos * nr + j].
template <bool GATHER>
__global__ __launch_bounds__(kDetBlock) void k_det_level(int64_t W, const int64_t* choff, int64_t G,
                                                         const int64_t* seg_start, const int64_t* seg_cnt,
                                                         const int32_t* sidx, const int32_t* src, MeasureSet M,
                                                         RealSet R, const uint64_t* in, uint64_t* out) {
  const int64_t w = (static_cast<int64_t>(blockIdx.x) * kDetBlock + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (w >= W) return;
  int64_t lo = 0, hi = G;  // last g with choff[g] <= w (choff[G] = W > w)
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (choff[mid] <= w) lo = mid;
    else hi = mid;
  }
  const int64_t c = w - choff[lo];
  const int64_t base = seg_start[lo] + c * kDetChunk;
  const int64_t count = min(static_cast<int64_t>(kDetChunk), seg_cnt[lo] - c * kDetChunk);
  int64_t row[4];
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int64_t pos = lane * 4 + e;
    row[e] = -1;
    if (pos < count) {
      if (GATHER) {
        const int32_t item = sidx[base + pos];
        row[e] = src ? src[item] : item;
      } else {
        row[e] = base + pos;
      }
    }
  }
  for (int j = 0; j < R.nr; j++) {
    double s = -0.0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      double x = -0.0;
      if (row[e] >= 0) x = GATHER ? det_value(M, R.m[j], row[e]) : __longlong_as_double(static_cast<long long>(in[row[e] * R.nr + j]));
      s = e == 0 ? x : s + x;
    }
    for (int d = kWave / 2; d >= 1; d >>= 1) {
      const double o = __longlong_as_double(
          static_cast<long long>(shfl_xor_u64(static_cast<uint64_t>(__double_as_longlong(s)), d)));
      s = (lane & d) ? o + s : s + o;
    }
    if (lane == 0) out[w * R.nr + j] = static_cast<uint64_t>(__double_as_longlong(s));
  }
}

// dst[(g * nm + m) * stride] = the group's final sum, for groups with items.
__global__ void k_det_store(const uint64_t* fin, const int64_t* choff, const int64_t* cnt, int64_t G, RealSet R,
                            int32_t nm, uint64_t* dst, int32_t stride) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= G || cnt[g] == 0) return;
  for (int j = 0; j < R.nr; j++) dst[(g * nm + R.m[j]) * stride] = fin[choff[g] * R.nr + j];
}

}  // namespace

int det_fp_sums(kg_ctx* ctx, int64_t N, const int32_t* gid, const int32_t* src, int64_t G, const MeasureSet& M,
                uint64_t* dst, int32_t stride) {
  RealSet R{};
  for (int m = 0; m < M.nm; m++)
    if (M.is_real[m]) R.m[R.nr++] = m;
  if (R.nr == 0 || N == 0 || G == 0) return KG_OK;
  KG_REQUIRE(ctx, G < (1ll << 31) && N < (1ll << 31), "deterministic sums: %lld items, %lld groups out of range",
             (long long)N, (long long)G);
  hipStream_t s = ctx->stream;
  int end_bit = 1;
  while ((1ll << end_bit) <= G) end_bit++;
  size_t sort_tmp = 0, scan_tmp = 0, max_tmp = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, sort_tmp, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                     static_cast<int32_t*>(nullptr), static_cast<int32_t*>(nullptr),
                                     static_cast<int>(N), 0, end_bit, s);
  hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp, static_cast<int64_t*>(nullptr), static_cast<int64_t*>(nullptr),
                                   static_cast<int>(G + 1), s);
  hipcub::DeviceReduce::Max(nullptr, max_tmp, static_cast<int64_t*>(nullptr), static_cast<int64_t*>(nullptr),
                            static_cast<int>(G), s);
  const size_t tmp = std::max(sort_tmp, std::max(scan_tmp, max_tmp));
  // Level buffers: chunk sums of level L are at most N / 256^(L+1) + G each.
  const int64_t w0 = N / kDetChunk + G + 1;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_k0 = carve(N * 4), o_k1 = carve(N * 4), o_i0 = carve(N * 4), o_i1 = carve(N * 4),
               o_tmp = carve(tmp), o_start = carve((G + 1) * 8), o_end = carve((G + 1) * 8),
               o_cnt = carve((G + 1) * 8), o_nch = carve((G + 1) * 8), o_nch2 = carve((G + 1) * 8),
               o_choff = carve((G + 1) * 8), o_max = carve(8), o_pa = carve(w0 * R.nr * 8),
               o_pb = carve(w0 * R.nr * 8);
  char* base = nullptr;
  if (hipMalloc(&base, off) != hipSuccess) {
    (void)hipGetLastError();
    return kg_fail(ctx, KG_ERR_OOM, "deterministic sums: device allocation of %zu bytes failed", off);
  }
  auto P = [&](size_t o) { return static_cast<void*>(base + o); };
  uint32_t* k0 = static_cast<uint32_t*>(P(o_k0));
  uint32_t* k1 = static_cast<uint32_t*>(P(o_k1));
  int32_t* i0 = static_cast<int32_t*>(P(o_i0));
  int32_t* i1 = static_cast<int32_t*>(P(o_i1));
  int64_t* start = static_cast<int64_t*>(P(o_start));
  int64_t* endp = static_cast<int64_t*>(P(o_end));
  int64_t* cnt = static_cast<int64_t*>(P(o_cnt));
  int64_t* nch = static_cast<int64_t*>(P(o_nch));
  int64_t* nch2 = static_cast<int64_t*>(P(o_nch2));
  int64_t* choff = static_cast<int64_t*>(P(o_choff));
  int64_t* dmax = static_cast<int64_t*>(P(o_max));
  uint64_t* pa = static_cast<uint64_t*>(P(o_pa));
  uint64_t* pb = static_cast<uint64_t*>(P(o_pb));
  auto blocks = [](int64_t work) { return dim3(static_cast<unsigned>(std::max<int64_t>(ceil_div(work, 256), 1))); };
  hipError_t e = hipSuccess;
  int st = KG_OK;
  do {
    hipLaunchKernelGGL(k_det_keys, blocks(N), dim3(256), 0, s, gid, N, static_cast<uint32_t>(G), k0, i0);
    size_t t = tmp;
    if ((e = hipcub::DeviceRadixSort::SortPairs(P(o_tmp), t, k0, k1, i0, i1, static_cast<int>(N), 0, end_bit, s)) !=
        hipSuccess)
      break;
    hipLaunchKernelGGL(k_det_zero, blocks(G + 1), dim3(256), 0, s, start, G + 1);
    hipLaunchKernelGGL(k_det_zero, blocks(G + 1), dim3(256), 0, s, endp, G + 1);
    hipLaunchKernelGGL(k_det_bounds, blocks(N), dim3(256), 0, s, k1, N, static_cast<uint32_t>(G), start, endp);
    hipLaunchKernelGGL(k_det_counts, blocks(G + 1), dim3(256), 0, s, start, endp, G, cnt, nch);
    t = tmp;
    if ((e = hipcub::DeviceReduce::Max(P(o_tmp), t, cnt, dmax, static_cast<int>(G), s)) != hipSuccess) break;
    int64_t maxcnt = 0;
    if ((e = hipMemcpyAsync(&maxcnt, dmax, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
    // Level loop: seg (start, cnt) -> chunk sums at choff.
    const uint64_t* in = nullptr;
    uint64_t* out = pa;
    bool gather = true;
    for (;;) {
      t = tmp;
      if ((e = hipcub::DeviceScan::ExclusiveSum(P(o_tmp), t, nch, choff, static_cast<int>(G + 1), s)) != hipSuccess)
        break;
      int64_t W = 0;
      if ((e = hipMemcpyAsync(&W, choff + G, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
      if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
      if (W > w0) {
        st = kg_fail(ctx, KG_ERR_HIP, "internal: deterministic level of %lld chunks", (long long)W);
        break;
      }
      const dim3 grid(static_cast<unsigned>(std::max<int64_t>(ceil_div(W * kWave, kDetBlock), 1)));
      if (gather)
        hipLaunchKernelGGL(k_det_level<true>, grid, dim3(kDetBlock), 0, s, W, choff, G, start, cnt, i1, src, M, R,
                           nullptr, out);
      else
        hipLaunchKernelGGL(k_det_level<false>, grid, dim3(kDetBlock), 0, s, W, choff, G, start, cnt, i1, src, M, R,
                           in, out);
      if ((e = hipGetLastError()) != hipSuccess) break;
      if (maxcnt <= kDetChunk) break;
      maxcnt = (maxcnt + kDetChunk - 1) / kDetChunk;
      hipLaunchKernelGGL(k_det_next, blocks(G + 1), dim3(256), 0, s, choff, nch, G, start, cnt, nch2);
      std::swap(nch, nch2);
      in = out;
      out = out == pa ? pb : pa;
      gather = false;
    }
    if (e != hipSuccess || st != KG_OK) break;
    hipLaunchKernelGGL(k_det_store, blocks(G), dim3(256), 0, s, out, choff, cnt, G, R, M.nm, dst, stride);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  } while (false);
  (void)hipStreamSynchronize(s);  // nothing may still use the buffers
  (void)hipFree(base);
  if (st != KG_OK) return st;
  if (e != hipSuccess) return kg_fail(ctx, KG_ERR_HIP, "deterministic sums failed: %s", hipGetErrorString(e));
  return KG_OK;
}

}  // namespace kg
