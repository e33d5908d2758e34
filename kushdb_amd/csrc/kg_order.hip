// ORDER BY and result rows.
//
// OrderByTranslator buffers its input and sorts it with runtime::Vector::Sort
// (compile/translators/order_by_translator.cc:35-110), a top-down merge sort
// whose merge takes the right run's element whenever comp(left, right) is
// false (runtime/vector.cc:45-110). comp tests the keys in order with
// LessThan (compile/proxy/evaluate.cc:339-400): a NULL is never less than
// anything and a non-NULL is less than a NULL; ENUM compares the decoded
// strings; BOOLEAN cannot be ordered. Rows with equal keys therefore end up in
// REVERSE input order (every merge puts the right run's equal elements first).
//
// Here: a least-significant-key-first sequence of stable radix sorts
// (hipcub::DeviceRadixSort::SortPairs) over order-preserving key encodings,
// starting from the reversed input positions, so equal keys keep reverse input
// order. Encodings: integers with the sign bit flipped; REAL with -0.0 folded
// into +0.0 (they are equal under <) and the usual sign-magnitude flip; ENUM as
// the rank of its string (given by the caller); descending keys complemented;
// NULL as a key bit above the value (set for ascending: NULLs last; clear for
// descending: NULLs first). A TEXT key (runtime::String::LessThan: unsigned
// bytes, then length) is first turned into ranks by a stable device merge
// sort of the row positions under that comparison; equal strings share a rank.
//
// kg_format_rows prints rows as runtime/printer.cc:13-35 and
// compile/translators/output_translator.cc:25-31 do: every value followed by
// '|', NULL as the empty string, REAL with std::fixed (six decimals), DATE as
// YYYY-MM-DD, BOOLEAN as t / f, a newline after each row.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kg_internal.h"

using namespace kg;

namespace {

__global__ void k_order_iota_rev(int32_t* perm, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) perm[i] = static_cast<int32_t>(n - 1 - i);
}

__device__ __forceinline__ uint64_t enc_value(int32_t type, const void* col, const int32_t* rank, int32_t r) {
  switch (type) {
    case KG_SMALLINT:
      return static_cast<uint64_t>(static_cast<uint32_t>(static_cast<int32_t>(static_cast<const int16_t*>(col)[r])) ^
                                   0x80000000u);
    case KG_INT:
    case KG_DATE:
      return static_cast<uint64_t>(static_cast<uint32_t>(static_cast<const int32_t*>(col)[r]) ^ 0x80000000u);
    case KG_ENUM:
      return static_cast<uint64_t>(static_cast<uint32_t>(rank[static_cast<const int32_t*>(col)[r]]));
    case KG_BIGINT:
      return static_cast<uint64_t>(static_cast<const int64_t*>(col)[r]) ^ 0x8000000000000000ull;
    default: {  // KG_REAL
      uint64_t b = static_cast<uint64_t>(__double_as_longlong(static_cast<const double*>(col)[r]));
      if (b == 0x8000000000000000ull) b = 0;  // -0.0 == +0.0
      return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
    }
  }
}

// Key of position j of the current order for a 32-bit value type: NULL bit at
// bit 32 (ascending: NULLs last; descending: NULLs first), the value below.
__global__ void k_order_key32(int32_t type, const void* col, const int8_t* nulls, const int32_t* rank,
                              const int32_t* perm, int64_t n, int32_t asc, uint64_t* key) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int32_t r = perm[j];
  const bool null = nulls && nulls[r];
  uint64_t v = null ? 0 : enc_value(type, col, rank, r);
  if (!asc) v = ~v & 0xFFFFFFFFull;
  const uint64_t nb = (null == static_cast<bool>(asc)) ? 1ull : 0ull;
  key[j] = (nb << 32) | v;
}

// 64-bit value types: the value key, then (pass == 1) the NULL key alone.
__global__ void k_order_key64(int32_t type, const void* col, const int8_t* nulls, const int32_t* perm, int64_t n,
                              int32_t asc, int32_t pass, uint64_t* key) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int32_t r = perm[j];
  const bool null = nulls && nulls[r];
  if (pass == 1) {
    key[j] = (null == static_cast<bool>(asc)) ? 1ull : 0ull;
    return;
  }
  const uint64_t v = null ? 0 : enc_value(type, col, nullptr, r);
  key[j] = asc ? v : ~v;
}

// String order of TEXT rows a, b (unsigned bytes, then the shorter first).
struct TextLess {
  const int64_t* off;
  const uint8_t* bytes;
  __device__ bool operator()(int32_t a, int32_t b) const {
    const int64_t la = off[a + 1] - off[a], lb = off[b + 1] - off[b];
    const uint8_t* pa = bytes + off[a];
    const uint8_t* pb = bytes + off[b];
    const int64_t m = la < lb ? la : lb;
    for (int64_t i = 0; i < m; i++)
      if (pa[i] != pb[i]) return pa[i] < pb[i];
    return la < lb;
  }
};

__global__ void k_iota(int32_t* v, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = static_cast<int32_t>(i);
}

// flag[j] = 1 when sorted row j's string differs from row j - 1's (0 for j = 0).
__global__ void k_text_newrun(const int32_t* sorted, int64_t n, TextLess lt, int32_t* flag) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  flag[j] = (j > 0 && lt(sorted[j - 1], sorted[j])) ? 1 : 0;
}

__global__ void k_text_scatter_rank(const int32_t* sorted, const int32_t* run, int64_t n, int32_t* rank) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < n) rank[sorted[j]] = run[j];
}

// Julian day number to proleptic Gregorian year, month, day (the inverse of
// the day count from 4713 BC of Date::BuildDate, runtime/date.cc:10-25).
void split_julian(int32_t jd, int32_t* y, int32_t* m, int32_t* d) {
  // Days from 1 March of a year 4800 BC (Gregorian), counted in 400-year
  // cycles, centuries, 4-year cycles and years with March as the first month.
  const int64_t a = static_cast<int64_t>(jd) + 32044;
  const int64_t b = (4 * a + 3) / 146097;
  const int64_t c = a - 146097 * b / 4;
  const int64_t e = (4 * c + 3) / 1461;
  const int64_t f = c - 1461 * e / 4;
  const int64_t g = (5 * f + 2) / 153;
  *d = static_cast<int32_t>(f - (153 * g + 2) / 5 + 1);
  *m = static_cast<int32_t>(g + 3 - 12 * (g / 10));
  *y = static_cast<int32_t>(100 * b + e - 4800 + g / 10);
}

}  // namespace

namespace {

// rank[r] = the rank of TEXT row r's string among rows 0..n-1 (equal strings
// share a rank).
int text_ranks(kg_ctx* ctx, const kg_text_col& t, int64_t n, int32_t* rank) {
  const int ni = static_cast<int>(n);
  const int grid = static_cast<int>((n + 255) / 256);
  TextLess lt{t.offsets, t.bytes};
  size_t sort_bytes = 0, scan_bytes = 0;
  KG_HIP_CHECK(ctx, hipcub::DeviceMergeSort::StableSortKeys(nullptr, sort_bytes, static_cast<int32_t*>(nullptr), ni,
                                                            lt, ctx->stream));
  KG_HIP_CHECK(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, static_cast<int32_t*>(nullptr),
                                                     static_cast<int32_t*>(nullptr), ni, ctx->stream));
  const size_t ib = (static_cast<size_t>(n) * 4 + 255) / 256 * 256;
  char* w = nullptr;
  if (hipMalloc(&w, 3 * ib + sort_bytes + scan_bytes) != hipSuccess) {
    (void)hipGetLastError();
    return kg_fail(ctx, KG_ERR_OOM, "TEXT ranking scratch");
  }
  int32_t* sorted = reinterpret_cast<int32_t*>(w);
  int32_t* flag = reinterpret_cast<int32_t*>(w + ib);
  int32_t* run = reinterpret_cast<int32_t*>(w + 2 * ib);
  void* cub = w + 3 * ib;
  bool ok = true;
  hipLaunchKernelGGL(k_iota, dim3(grid), dim3(256), 0, ctx->stream, sorted, n);
  size_t sb = sort_bytes;
  ok = ok && hipcub::DeviceMergeSort::StableSortKeys(cub, sb, sorted, ni, lt, ctx->stream) == hipSuccess;
  hipLaunchKernelGGL(k_text_newrun, dim3(grid), dim3(256), 0, ctx->stream, sorted, n, lt, flag);
  size_t cb = scan_bytes;
  ok = ok && hipcub::DeviceScan::InclusiveSum(cub, cb, flag, run, ni, ctx->stream) == hipSuccess;
  hipLaunchKernelGGL(k_text_scatter_rank, dim3(grid), dim3(256), 0, ctx->stream, sorted, run, n, rank);
  ok = ok && hipStreamSynchronize(ctx->stream) == hipSuccess;
  (void)hipFree(w);
  if (!ok) {
    (void)hipGetLastError();
    return kg_fail(ctx, KG_ERR_HIP, "TEXT ranking failed");
  }
  return KG_OK;
}

}  // namespace

extern "C" int kg_enum_ranks(const char* strings, const int64_t* offsets, int32_t count, int32_t* out_rank) {
  if (count < 0 || (count > 0 && (!strings || !offsets || !out_rank))) return KG_ERR_INVALID;
  std::vector<int32_t> idx(static_cast<size_t>(count));
  for (int32_t i = 0; i < count; i++) idx[i] = i;
  auto view = [&](int32_t c) {
    return std::string(strings + offsets[c], static_cast<size_t>(offsets[c + 1] - offsets[c]));
  };
  // std::string compares as unsigned bytes (char_traits<char>::lt).
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return view(a) < view(b); });
  for (int32_t j = 0; j < count; j++) out_rank[idx[j]] = j;
  return KG_OK;
}

extern "C" int kg_order_by(kg_ctx* ctx, int64_t n, const kg_col* keys, int32_t nkeys, const int32_t* ascending,
                           const int32_t* const* enum_rank, const kg_text_col* text_keys, int32_t* d_perm) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX, "row count %lld out of range", (long long)n);
  KG_REQUIRE(ctx, nkeys >= 1 && keys && ascending, "at least one key with its direction");
  KG_REQUIRE(ctx, n == 0 || d_perm, "d_perm is NULL");
  for (int k = 0; k < nkeys; k++) {
    const int32_t t = keys[k].type;
    if (t == KG_BOOLEAN) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "BOOLEAN keys cannot be ordered (key %d)", k);
    if (t == KG_TEXT) {
      KG_REQUIRE(ctx, text_keys && (n == 0 || text_keys[k].offsets), "TEXT key %d needs its text column", k);
      continue;
    }
    KG_REQUIRE(ctx, n == 0 || keys[k].data, "key %d has no data", k);
    KG_REQUIRE(ctx, t == KG_SMALLINT || t == KG_INT || t == KG_DATE || t == KG_ENUM || t == KG_BIGINT || t == KG_REAL,
               "bad key type %d", t);
    if (t == KG_ENUM) KG_REQUIRE(ctx, enum_rank && enum_rank[k], "ENUM key %d needs its string ranks", k);
  }
  if (n == 0) return KG_OK;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const int ni = static_cast<int>(n);
  size_t cub_bytes = 0;
  KG_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, cub_bytes, static_cast<uint64_t*>(nullptr),
                                                       static_cast<uint64_t*>(nullptr),
                                                       static_cast<int32_t*>(nullptr),
                                                       static_cast<int32_t*>(nullptr), ni, 0, 64, ctx->stream));
  const size_t kb = (static_cast<size_t>(n) * 8 + 255) / 256 * 256;
  const size_t pb = (static_cast<size_t>(n) * 4 + 255) / 256 * 256;
  char* w = nullptr;
  if (hipMalloc(&w, 2 * kb + 2 * pb + cub_bytes) != hipSuccess) {
    (void)hipGetLastError();
    return kg_fail(ctx, KG_ERR_OOM, "ORDER BY scratch of %zu bytes", 2 * kb + 2 * pb + cub_bytes);
  }
  uint64_t* k0 = reinterpret_cast<uint64_t*>(w);
  uint64_t* k1 = reinterpret_cast<uint64_t*>(w + kb);
  int32_t* p1 = reinterpret_cast<int32_t*>(w + 2 * kb);
  int32_t* trank = reinterpret_cast<int32_t*>(w + 2 * kb + pb);  // ranks of a TEXT key
  void* cub = w + 2 * kb + 2 * pb;
  int32_t* perm = d_perm;  // the current order lives in d_perm between passes
  int st = KG_OK;
  const int grid = static_cast<int>((n + 255) / 256);
  auto sort_pass = [&](int end_bit) -> bool {
    size_t cb = cub_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(cub, cb, k0, k1, perm, p1, ni, 0, end_bit, ctx->stream) != hipSuccess)
      return false;
    return hipMemcpyAsync(perm, p1, static_cast<size_t>(n) * 4, hipMemcpyDeviceToDevice, ctx->stream) == hipSuccess;
  };
  hipLaunchKernelGGL(k_order_iota_rev, dim3(grid), dim3(256), 0, ctx->stream, perm, n);
  for (int k = nkeys - 1; k >= 0 && st == KG_OK; k--) {
    kg_col c = keys[k];
    const int32_t asc = ascending[k] ? 1 : 0;
    if (c.type == KG_TEXT) {  // order by the rows' string ranks
      st = text_ranks(ctx, text_keys[k], n, trank);
      if (st != KG_OK) break;
      c.type = KG_INT;
      c.data = trank;
      c.nulls = text_keys[k].nulls;
    }
    if (c.type == KG_BIGINT || c.type == KG_REAL) {
      hipLaunchKernelGGL(k_order_key64, dim3(grid), dim3(256), 0, ctx->stream, c.type, c.data, c.nulls, perm, n,
                         asc, 0, k0);
      if (!sort_pass(64)) st = KG_ERR_HIP;
      if (st == KG_OK && c.nulls) {
        hipLaunchKernelGGL(k_order_key64, dim3(grid), dim3(256), 0, ctx->stream, c.type, c.data, c.nulls, perm, n,
                           asc, 1, k0);
        if (!sort_pass(1)) st = KG_ERR_HIP;
      }
    } else {
      hipLaunchKernelGGL(k_order_key32, dim3(grid), dim3(256), 0, ctx->stream, c.type, c.data, c.nulls,
                         enum_rank ? enum_rank[k] : nullptr, perm, n, asc, k0);
      if (!sort_pass(33)) st = KG_ERR_HIP;
    }
  }
  if (st == KG_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) st = KG_ERR_HIP;
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(w);
  if (st == KG_ERR_HIP) {
    (void)hipGetLastError();
    return kg_fail(ctx, st, "ORDER BY sort failed");
  }
  return st;
}

extern "C" int kg_format_rows(const kg_fmt_col* cols, int32_t ncols, int64_t n, const int32_t* order, char* buf,
                              int64_t cap, int64_t* out_len) {
  if (!out_len || (ncols > 0 && !cols) || n < 0) return KG_ERR_INVALID;
  std::string s;
  char tmp[64];
  for (int64_t i = 0; i < n; i++) {
    const int64_t r = order ? order[i] : i;
    if (r < 0 || r >= n) return KG_ERR_INVALID;
    for (int32_t c = 0; c < ncols; c++) {
      const kg_fmt_col& f = cols[c];
      if (f.nulls && f.nulls[r]) {
        s += '|';
        continue;
      }
      switch (f.type) {
        case KG_SMALLINT: s += std::to_string(static_cast<const int16_t*>(f.data)[r]); break;
        case KG_INT: s += std::to_string(static_cast<const int32_t*>(f.data)[r]); break;
        case KG_BIGINT: s += std::to_string(static_cast<const int64_t*>(f.data)[r]); break;
        case KG_BOOLEAN: s += static_cast<const int8_t*>(f.data)[r] ? 't' : 'f'; break;
        case KG_REAL:
          std::snprintf(tmp, sizeof(tmp), "%.6f", static_cast<const double*>(f.data)[r]);
          s += tmp;
          break;
        case KG_DATE: {
          int32_t y, m, d;
          split_julian(static_cast<const int32_t*>(f.data)[r], &y, &m, &d);
          std::snprintf(tmp, sizeof(tmp), "%d-%02d-%02d", y, m, d);
          s += tmp;
          break;
        }
        case KG_ENUM: {
          const int32_t code = static_cast<const int32_t*>(f.data)[r];
          if (!f.dict || !f.dict_off || code < 0 || code >= f.dict_n) return KG_ERR_INVALID;
          s.append(f.dict + f.dict_off[code], static_cast<size_t>(f.dict_off[code + 1] - f.dict_off[code]));
          break;
        }
        case KG_TEXT: {
          if (!f.dict || !f.dict_off) return KG_ERR_INVALID;
          s.append(f.dict + f.dict_off[r], static_cast<size_t>(f.dict_off[r + 1] - f.dict_off[r]));
          break;
        }
        default: return KG_ERR_INVALID;
      }
      s += '|';
    }
    s += '\n';
  }
  *out_len = static_cast<int64_t>(s.size());
  if (!buf) return KG_OK;
  if (static_cast<int64_t>(s.size()) > cap) return KG_ERR_CAPACITY;
  std::memcpy(buf, s.data(), s.size());
  return KG_OK;
}
