// Multi-way chain query (KG_Q_CHAIN) as one native plan over the library's
// own entry points: every child filtered to a selection (ScanSelect, then a
// TEXT predicate), a column index on the join column of each next table over
// its selection (MemoryColumnIndex, compile/proxy/memory_column_index.cc:
// 282-388), index-nested-loop expansion one table at a time
// (permutable_skinner_join_translator.cc:359-433) or the Skinner driver
// (kg_skinner_join), the result tuples deduplicated in a tuple-id set (the
// valid-tuple handler, :842-912), then the ungrouped MIN aggregates
// (aggregate_translator.cc:34-126; MinMaxAggregator, aggregator.cc:155-171).
// A table with a stored index on its fk column (over all rows) is joined
// through that index instead, and its own predicates are evaluated on the
// expanded tuples afterwards (the reference's table functions likewise test
// flagged predicates after the index probe, :663-832); its tuple ids are then
// row ids. Results are the same either way. Device memory comes from the
// context's block cache, so a served stream of these queries allocates
// nothing once warm.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "kg_internal.h"
#include "kg_ops.h"
#include "kg_strops.h"

namespace {

constexpr int kCBlock = 256;
constexpr int kChainPattern = 240;

// A table's TEXT predicate for the device (offsets NULL: none).
struct TextPred {
  kg_text_col col;
  int32_t op;
  int32_t len;
  uint8_t p[kChainPattern];
};

// keep[i] = the table's predicates hold at row tuples[i * w + pos]; idx[i] = i.
__global__ __launch_bounds__(kCBlock) void k_tuple_keep(const int32_t* tuples, int32_t w, int32_t pos, int64_t m,
                                                         kg::PredSet P, TextPred F, uint8_t* keep, int32_t* idx) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kCBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kCBlock + threadIdx.x; i < m; i += stride) {
    const int64_t r = tuples[i * w + pos];
    bool ok = kg::eval_preds(P, r);
    if (ok && F.col.offsets) {
      if (F.col.nulls && F.col.nulls[r]) {
        ok = false;
      } else {
        const int64_t a = F.col.offsets[r], b = F.col.offsets[r + 1];
        ok = F.op == KG_TEXT_LIKE ? kg::like_match(F.col.bytes + a, b - a, F.p, F.len)
                                  : kg::text_match(F.op, F.col.bytes + a, b - a, F.p, F.len);
      }
    }
    keep[i] = ok ? 1 : 0;
    idx[i] = static_cast<int32_t>(i);
  }
}

// out[j] = tuples[idx[j]] (w ids each), j < *count.
__global__ __launch_bounds__(kCBlock) void k_tuple_take(const int32_t* tuples, int32_t w, const int32_t* idx,
                                                         const int64_t* count, int32_t* out) {
  const int64_t m = *count * w;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kCBlock;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * kCBlock + threadIdx.x; e < m; e += stride) {
    const int64_t j = e / w;
    out[e] = tuples[static_cast<int64_t>(idx[j]) * w + (e - j * w)];
  }
}

int cgrid(kg_ctx* ctx, int64_t work) {
  return static_cast<int>(std::min<int64_t>(kg::ceil_div(std::max<int64_t>(work, 1), kCBlock),
                                            static_cast<int64_t>(ctx->num_cus) * 16));
}

// Device blocks of one query, returned to the context's cache at the end.
struct Blocks {
  kg_ctx* ctx;
  std::vector<void*> held;
  explicit Blocks(kg_ctx* c) : ctx(c) {}
  ~Blocks() {
    for (void* p : held) kg_dev_release(ctx, p);
  }
  template <class T>
  int get(int64_t n, T** p) {
    void* v = nullptr;
    const int st = kg_dev_alloc(ctx, &v, static_cast<size_t>(std::max<int64_t>(n, 1)) * sizeof(T));
    if (st == KG_OK) held.push_back(v);
    *p = static_cast<T*>(v);
    return st;
  }
};

struct IndexGuard {
  kg_ctx* ctx;
  kg_index* idx = nullptr;
  ~IndexGuard() {
    if (idx) kg_index_free(ctx, idx);
  }
};

struct SetGuard {
  kg_ctx* ctx;
  kg_tuple_set* s = nullptr;
  ~SetGuard() {
    if (s) kg_tuple_set_free(ctx, s);
  }
};

bool valid_col(const kg_chain_table& t, int32_t c) { return c >= 0 && c < t.ncols && t.cols[c].data; }

int check_chain(kg_ctx* ctx, const kg_chain* q) {
  KG_REQUIRE(ctx, q && q->tables, "chain query is NULL");
  const int k = q->ntables;
  KG_REQUIRE(ctx, k >= 2 && k <= 8, "2..8 chain tables (got %d)", k);
  for (int i = 0; i < k; i++) {
    const kg_chain_table& t = q->tables[i];
    KG_REQUIRE(ctx, t.n >= 0 && t.n < INT32_MAX, "table %d: row count out of range", i);
    KG_REQUIRE(ctx, t.ncols > 0 && t.cols, "table %d has no columns", i);
    KG_REQUIRE(ctx, t.npreds == 0 || t.preds, "table %d: predicates are NULL", i);
    KG_REQUIRE(ctx, !t.text || (t.text_pattern || t.text_plen == 0), "table %d: TEXT pattern is NULL", i);
    if (i + 1 < k) KG_REQUIRE(ctx, valid_col(t, t.id_col), "table %d: bad id column %d", i, t.id_col);
    if (i > 0) KG_REQUIRE(ctx, valid_col(t, t.fk_col), "table %d: bad fk column %d", i, t.fk_col);
  }
  KG_REQUIRE(ctx, q->min_text_table < k && (q->min_text_table < 0 || q->min_text), "bad MIN(TEXT) table");
  KG_REQUIRE(ctx, q->min_int_table < k, "bad MIN table");
  if (q->min_int_table >= 0) {
    const kg_chain_table& t = q->tables[q->min_int_table];
    KG_REQUIRE(ctx, valid_col(t, q->min_int_col), "bad MIN column %d", q->min_int_col);
    if (kg::type_width(t.cols[q->min_int_col].type) == 0)
      return kg_fail(ctx, KG_ERR_UNSUPPORTED, "MIN column of type %d", t.cols[q->min_int_col].type);
  }
  KG_REQUIRE(ctx, q->skinner_budget >= 0, "negative Skinner budget");
  return KG_OK;
}

// Selection of table t: ascending row ids passing its predicates, or NULL
// (every row) when it has none.
int select_table(kg_ctx* ctx, Blocks& B, const kg_chain_table& t, const int32_t** sel, int64_t* cnt) {
  *sel = nullptr;
  *cnt = t.n;
  int st;
  if (t.npreds > 0) {
    int32_t* rows = nullptr;
    if ((st = B.get(t.n, &rows)) != KG_OK) return st;
    if ((st = kg_scan_select(ctx, t.n, t.cols, t.ncols, t.preds, t.npreds, 1, rows, cnt)) != KG_OK) return st;
    *sel = rows;
  }
  if (t.text) {
    int32_t* rows = nullptr;
    if ((st = B.get(*cnt, &rows)) != KG_OK) return st;
    int64_t m = 0;
    if ((st = kg_text_select(ctx, t.text, t.text_op, t.text_pattern, t.text_plen, *sel, *cnt, rows, &m)) != KG_OK)
      return st;
    *sel = rows;
    *cnt = m;
  }
  return KG_OK;
}

// Keeps the tuples (width w) whose row at position pos satisfies table t's
// predicates; *cur / *ncur are replaced by the kept tuples.
int filter_tuples(kg_ctx* ctx, Blocks& B, const kg_chain_table& t, int32_t w, int32_t pos, const int32_t** cur,
                  int64_t* ncur) {
  kg::PredSet P{};
  int st = kg::build_preds(ctx, t.cols, t.ncols, t.preds, t.npreds, 1, t.n, P);
  if (st != KG_OK) return st;
  TextPred F{};
  if (t.text) {
    KG_REQUIRE(ctx, t.text->offsets && t.text->bytes, "TEXT column is NULL");
    KG_REQUIRE(ctx, t.text_op >= KG_TEXT_EQ && t.text_op <= KG_TEXT_LIKE, "bad text operator %d", t.text_op);
    KG_REQUIRE(ctx, t.text_plen >= 0 && t.text_plen <= kChainPattern, "pattern length %d out of 0..%d", t.text_plen,
               kChainPattern);
    const uint8_t* pat = reinterpret_cast<const uint8_t*>(t.text_pattern);
    if (t.text_op == KG_TEXT_LIKE && !kg::like_pattern_ok(pat, t.text_plen)) {  // re2 rejects it: no match
      *ncur = 0;
      return KG_OK;
    }
    F.col = *t.text;
    F.op = t.text_op;
    F.len = t.text_plen;
    if (t.text_plen) std::memcpy(F.p, pat, t.text_plen);
  }
  const int64_t m = *ncur;
  uint8_t* keep = nullptr;
  int32_t *idx = nullptr, *kept = nullptr, *out = nullptr;
  int64_t* count = nullptr;
  size_t cub_bytes = 0;
  hipcub::DeviceSelect::Flagged(nullptr, cub_bytes, idx, keep, kept, count, static_cast<int>(m), ctx->stream);
  uint8_t* cub = nullptr;
  if ((st = B.get(m, &keep)) != KG_OK || (st = B.get(m, &idx)) != KG_OK || (st = B.get(m, &kept)) != KG_OK ||
      (st = B.get(1, &count)) != KG_OK || (st = B.get(static_cast<int64_t>(cub_bytes), &cub)) != KG_OK ||
      (st = B.get(m * w, &out)) != KG_OK)
    return st;
  hipLaunchKernelGGL(k_tuple_keep, dim3(cgrid(ctx, m)), dim3(kCBlock), 0, ctx->stream, *cur, w, pos, m, P, F, keep,
                     idx);
  KG_LAUNCH_CHECK(ctx);
  KG_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(cub, cub_bytes, idx, keep, kept, count, static_cast<int>(m),
                                                  ctx->stream));
  hipLaunchKernelGGL(k_tuple_take, dim3(cgrid(ctx, m * w)), dim3(kCBlock), 0, ctx->stream, *cur, w, kept, count,
                     out);
  KG_LAUNCH_CHECK(ctx);
  int64_t h = 0;
  KG_HIP_CHECK(ctx, hipMemcpyAsync(&h, count, 8, hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *cur = out;
  *ncur = h;
  return KG_OK;
}

}  // namespace

extern "C" int kg_chain_run(kg_ctx* ctx, const kg_chain* q, int64_t* out_count, int64_t* out_min_int,
                            int8_t* min_int_null, int64_t* out_min_text_row) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, out_count && out_min_int && min_int_null && out_min_text_row, "null output pointer");
  *out_count = 0;
  *out_min_int = 0;
  *min_int_null = 1;
  *out_min_text_row = -1;
  int st = check_chain(ctx, q);
  if (st != KG_OK) return st;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const int k = q->ntables;
  const kg_chain_table* T = q->tables;
  Blocks B(ctx);
  // Tables joined through a stored fk index keep no selection: their
  // predicates are applied to the expanded tuples.
  auto through_stored = [&](int i) { return q->skinner_budget == 0 && i > 0 && T[i].fk_index != nullptr; };
  std::vector<const int32_t*> sel(k, nullptr);
  std::vector<int64_t> cnt(k);
  for (int i = 0; i < k; i++) {
    cnt[i] = T[i].n;
    if (!through_stored(i) && (st = select_table(ctx, B, T[i], &sel[i], &cnt[i])) != KG_OK) return st;
  }

  // Result tuples: int32[k] tuple ids (positions in each table's selection),
  // table order, each once.
  SetGuard set{ctx};
  int64_t ntup = 0;
  if (q->skinner_budget > 0) {
    std::vector<kg_skinner_table> st_tabs(k);
    std::vector<kg_skinner_cond> conds(k - 1);
    for (int i = 0; i < k; i++) st_tabs[i] = kg_skinner_table{cnt[i], sel[i]};
    for (int i = 1; i < k; i++)
      conds[i - 1] = kg_skinner_cond{i - 1, i, T[i - 1].cols[T[i - 1].id_col], T[i].cols[T[i].fk_col]};
    kg_skinner_opts o{};
    o.budget_per_episode = q->skinner_budget;
    o.seed = q->seed > 0 ? q->seed : 1;
    kg_skinner_stats ss{};
    if ((st = kg_skinner_join(ctx, k, st_tabs.data(), k - 1, conds.data(), &o, &set.s, &ss)) != KG_OK) return st;
    ntup = ss.distinct_tuples;
  } else {
    // The order t0, t1, ...: t_i's index on its fk over its selection, probed
    // with t_(i-1).id of every partial tuple.
    const int32_t* cur = nullptr;  // NULL: the identity tuples of t0's selection
    int64_t ncur = cnt[0];
    for (int i = 1; i < k && ncur > 0; i++) {
      IndexGuard built{ctx};
      const kg_index* idx = through_stored(i) ? T[i].fk_index : nullptr;
      if (!idx) {
        if ((st = kg_index_build(ctx, &T[i].cols[T[i].fk_col], sel[i], cnt[i], &built.idx)) != KG_OK) return st;
        idx = built.idx;
      }
      // Over a selection with a unique id column each new row joins at most
      // one partial tuple, so the selection bounds the output; through a
      // stored index, the average fan-out (rows per referenced row) twice
      // over. Other cases retry at the total.
      int64_t cap = std::max<int64_t>(cnt[i], 1), total = 0;
      if (through_stored(i))
        cap = std::min<int64_t>(cap, ncur * (2 * kg::ceil_div(T[i].n, std::max<int64_t>(T[i - 1].n, 1)) + 2));
      int32_t* out = nullptr;
      if ((st = B.get(cap * (i + 1), &out)) != KG_OK) return st;
      const kg_col* src_key = &T[i - 1].cols[T[i - 1].id_col];
      st = kg_index_join_expand(ctx, ncur, i, cur, i - 1, src_key, sel[i - 1], idx, cap, out, &total);
      if (st == KG_ERR_CAPACITY) {
        cap = total;
        if ((st = B.get(cap * (i + 1), &out)) != KG_OK) return st;
        st = kg_index_join_expand(ctx, ncur, i, cur, i - 1, src_key, sel[i - 1], idx, cap, out, &total);
      }
      if (st != KG_OK) return st;
      cur = out;
      ncur = total;
      if (through_stored(i) && (T[i].npreds > 0 || T[i].text) && ncur > 0)
        if ((st = filter_tuples(ctx, B, T[i], i + 1, i, &cur, &ncur)) != KG_OK) return st;
    }
    if ((st = kg_tuple_set_create(ctx, k, std::max<int64_t>(ncur, 1), &set.s)) != KG_OK) return st;
    if (ncur > 0 && (st = kg_tuple_set_insert(ctx, set.s, cur, ncur, nullptr, &ntup)) != KG_OK) return st;
  }
  *out_count = ntup;
  if (ntup == 0) return KG_OK;
  int32_t* tuples = nullptr;
  if ((st = B.get(ntup * k, &tuples)) != KG_OK) return st;
  int64_t got = 0;
  if ((st = kg_tuple_set_extract(ctx, set.s, tuples, ntup, &got)) != KG_OK) return st;
  int32_t* rows = nullptr;
  if ((st = B.get(ntup, &rows)) != KG_OK) return st;
  if (q->min_text_table >= 0) {
    const int t = q->min_text_table;
    if ((st = kg_tuple_rows(ctx, tuples, k, t, sel[t], ntup, rows)) != KG_OK) return st;
    int64_t pos = -1, len = 0;
    if ((st = kg_text_minmax(ctx, q->min_text, rows, ntup, 0, &pos, nullptr, 0, &len)) != KG_OK) return st;
    if (pos >= 0) {
      int32_t r = 0;
      KG_HIP_CHECK(ctx, hipMemcpyAsync(&r, rows + pos, 4, hipMemcpyDeviceToHost, ctx->stream));
      KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      *out_min_text_row = r;
    }
  }
  if (q->min_int_table >= 0) {
    const int t = q->min_int_table;
    const kg_col& c = T[t].cols[q->min_int_col];
    if ((st = kg_tuple_rows(ctx, tuples, k, t, sel[t], ntup, rows)) != KG_OK) return st;
    const int w = kg::type_width(c.type);
    int64_t* vals = nullptr;  // ntup values of width w (8-byte blocks hold them)
    if ((st = B.get((ntup * w + 7) / 8, &vals)) != KG_OK) return st;
    int8_t* nulls = nullptr;
    if (c.nulls && (st = B.get(ntup, &nulls)) != KG_OK) return st;
    if ((st = kg_gather(ctx, &c, rows, ntup, vals, nulls)) != KG_OK) return st;
    const kg_col g{c.type, 0, vals, nulls};
    const kg_agg a{KG_AGG_MIN, 0, -1, 0};
    int64_t ng = 0, v = 0;
    int8_t nl = 1;
    if ((st = kg_scan_group_agg(ctx, ntup, &g, 1, nullptr, 0, nullptr, nullptr, 0, &a, 1, 1, &ng, nullptr, &v,
                                &nl)) != KG_OK)
      return st;
    if (ng == 1 && !nl) {
      *out_min_int = v;
      *min_int_null = 0;
    }
  }
  return KG_OK;
}
