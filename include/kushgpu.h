/*
 * kushgpu — MI355X (gfx950) execution of KushDB's pipeline operators behind a
 * C ABI.
 *
 * KushDB's generated pipelines call runtime C++ functions by absolute address
 * (khir ProgramBuilder::DeclareExternalFunction, khir/program_builder.h:125;
 * bound in compile/forward_declare.cc:18-60). Those calls are per tuple
 * (runtime/hash_table.h:22-38, runtime/aggregate_hash_table.h:28-49) and
 * several hot loops are inline IR rather than calls at all (the aggregate probe
 * loop, compile/proxy/aggregate_hash_table.cc:187-243; the join bucket loop,
 * compile/translators/hash_join_translator.cc:105-139). Per-tuple calls into a
 * GPU are pointless, so this ABI sits one level up, at the pipeline seam of
 * execution/pipeline.cc:23-35: one call per pipeline body over a whole column
 * range, replacing the body_N(start, end) morsel loop of
 * execution/executable_query.cc:118-144 for the operators below.
 *
 * Conventions
 *  - Every function returns a kg_status; on failure kg_last_error(ctx) holds a
 *    message. Nothing throws across this boundary (the reference throws C++
 *    exceptions, e.g. runtime/file_manager.cc:25-53).
 *  - Pointers prefixed d_ (and kg_col.data / kg_col.nulls) are device memory;
 *    everything else is host memory. Input buffers are read-only during a call.
 *  - Work is ordered on the context's HIP stream. Calls that return host
 *    values (counts, small result sets) synchronize that stream; calls that only
 *    write device buffers do not.
 *  - One context is externally synchronized: do not call it from two threads
 *    at once.
 */
#ifndef KUSHGPU_H_
#define KUSHGPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
enum kg_status {
  KG_OK = 0,
  KG_ERR_INVALID = 1,     /* bad argument */
  KG_ERR_HIP = 2,         /* HIP runtime error */
  KG_ERR_OOM = 3,         /* device allocation failed */
  KG_ERR_CAPACITY = 4,    /* an output or table capacity was exceeded */
  KG_ERR_UNSUPPORTED = 5, /* a type / operator combination this path does not run */
};

/* SQL type ids; same order as catalog::TypeId (catalog/sql_type.h:7-16). */
enum kg_type {
  KG_SMALLINT = 0,
  KG_INT = 1,
  KG_BIGINT = 2,
  KG_REAL = 3,
  KG_DATE = 4,
  KG_TEXT = 5,
  KG_BOOLEAN = 6,
  KG_ENUM = 7,
};

/* Comparison operators (plan::BinaryArithmeticExpressionType EQ..GEQ as used
 * by ScanSelect / SimdScanSelect, simd_scan_select_translator.cc:402-442). */
enum kg_cmp { KG_EQ = 0, KG_NEQ = 1, KG_LT = 2, KG_LEQ = 3, KG_GT = 4, KG_GEQ = 5 };

/* Aggregate kinds (plan::AggregateType; compile/proxy/aggregator.cc). */
enum kg_agg_kind { KG_AGG_SUM = 0, KG_AGG_MIN = 1, KG_AGG_MAX = 2, KG_AGG_AVG = 3, KG_AGG_COUNT = 4 };

/* A resident column: `data` holds n values of the C type of `type` (int16,
 * int32, int64, double, int8); `nulls` is an int8 array with 1 = NULL, or NULL
 * for a non-nullable column. Mirrors runtime::ColumnData::*ColumnData plus the
 * `_null.kdb` int8 column (runtime/column_data.h:20-58). */
typedef struct kg_col {
  int32_t type;
  int32_t reserved;
  const void* data;
  const int8_t* nulls;
} kg_col;

/* One comparison `cols[col] op literal`; ilit for integer-like columns, flit
 * for REAL. A NULL value fails the comparison (scan_select_translator.cc:129-133). */
typedef struct kg_pred {
  int32_t col;
  int32_t op;
  int64_t ilit;
  double flit;
} kg_pred;

/* One aggregate over cols[col]; col = -1 with KG_AGG_COUNT is COUNT(*).
 * col2 >= 0 makes the input the REAL product cols[col] * Float(cols[col2])
 * (e.g. SUM(price * Float(qty))); -1 otherwise. */
typedef struct kg_agg {
  int32_t kind;
  int32_t col;
  int32_t col2;
  int32_t reserved;
} kg_agg;

/* A resident TEXT column: value i is bytes[offsets[i] .. offsets[i+1]); the
 * slot table of the .kdb TEXT format (runtime/column_data.cc:175-229)
 * flattened to prefix offsets (n+1 of them). nulls as in kg_col. */
typedef struct kg_text_col {
  const int64_t* offsets;
  const uint8_t* bytes;
  const int8_t* nulls;
} kg_text_col;

/* TEXT operators: comparisons with a literal, then the string predicates of
 * runtime/string.cc:14-82 (StartsWith, EndsWith, Contains, Like). */
enum kg_text_op {
  KG_TEXT_EQ = 0, KG_TEXT_NEQ = 1, KG_TEXT_LT = 2, KG_TEXT_LEQ = 3, KG_TEXT_GT = 4, KG_TEXT_GEQ = 5,
  KG_TEXT_STARTS_WITH = 6, KG_TEXT_ENDS_WITH = 7, KG_TEXT_CONTAINS = 8, KG_TEXT_LIKE = 9,
};

typedef struct kg_ctx kg_ctx;
typedef struct kg_hash_table kg_hash_table;

/* ---- context -------------------------------------------------------------
 * Replaces the per-query state (execution/query_state.h:26-41) and the
 * process-wide FileManager column cache (runtime/file_manager.cc:22-64): a
 * context owns a device, a stream and a grow-only workspace. */
int kg_ctx_create(int device, kg_ctx** out);
int kg_ctx_destroy(kg_ctx* ctx);
const char* kg_last_error(const kg_ctx* ctx);
/* Order the context's work on the caller's hipStream_t (NULL = the null
 * stream). A context starts with a stream of its own. */
int kg_ctx_set_stream(kg_ctx* ctx, void* hip_stream);
int kg_ctx_synchronize(kg_ctx* ctx);
/* FP64 SUM/AVG addition order (compile/proxy/aggregator.cc:67-71,239-268
 * fold sequentially in arrival order; no switch there).
 * 0 (default): dense GROUP BYs with at most 8 groups (4 with MIN/MAX) add in a
 * fixed tree order that depends only on n; the LDS table (more groups), the
 * star pipeline (kg_scan_join_group_agg), hashed GROUP BY (kg_hash_group_agg)
 * and batches add in a timing-dependent order (within FP64 rounding of the
 * fixed order).
 * 1: those paths take their REAL sums in a fixed order that depends only on
 * the input: each group's rows (join output rows for the star pipeline) in
 * input order, summed in chunks of 256 by a fixed lane / pairwise tree,
 * chunk sums summed the same way (reproduced bit for bit by the CPU oracle,
 * kor_det_sum). Batches do not take queries with REAL sums in this mode
 * (KG_ERR_UNSUPPORTED per query; kg_serve runs them on the plain path). */
int kg_ctx_set_deterministic(kg_ctx* ctx, int on);
/* Join probes can first partition the probe keys by table region (histogram,
 * scan, LDS-staged scatter) and then probe one region at a time per XCD, so
 * each region's slice of the table is read from HBM once. regions: -1 or 0
 * (default) probes directly, otherwise the region count (a multiple of 8, at
 * most 1024), as kg_join_tune may choose. No counterpart in the reference (its
 * probes are per tuple). */
int kg_ctx_set_probe_regions(kg_ctx* ctx, int32_t regions);
/* Join builds (kg_ht_build) of many rows group their (key, row) pairs by
 * table region first (a histogram pass that also finds the key range, then two
 * LDS counting-sort passes: coarse regions, then fine regions), then build each
 * fine region's slice of the table in one workgroup's LDS (one LDS atomic per
 * row on its bucket's fill count) and write it back with coalesced stores,
 * instead of one random CAS insert per row; the table holds the same (key,
 * row) entries per bucket chain, and keys that overflow a slice's last bucket
 * are inserted directly afterwards. regions: -1 (default) builds by regions
 * when at least 2^20 rows go into a table of at least 32 MB; 0 always inserts
 * directly; > 0 partitions every build, with at least that many fine regions
 * where the table allows (a power of two dividing its bucket count, slices
 * within the LDS). No counterpart in the reference (its inserts are per
 * tuple, runtime/hash_table.cc:18-27). */
int kg_ctx_set_build_regions(kg_ctx* ctx, int32_t regions);
/* Probe + GROUP BY matched build key (kg_ht_probe_agg without first rows) of
 * many rows: the probe rows are grouped by the table region of their key's home
 * bucket (two counting-sort passes), then one workgroup per region probes its
 * rows against that slice of the table and keeps SUM/COUNT per matched slot in
 * LDS, writing its groups straight to the output, instead of two device atomics
 * per matched row on a table-sized accumulator array. Frequent keys (sampled)
 * are summed per workgroup during the partition and folded in by their region.
 * Results are the same (integer sums and counts are exact). regions: -1
 * (default) partitions when at least 2^22 rows probe a table of at least 64 MB
 * of accumulators; 0 never; > 0 partitions into at least that many regions (a
 * multiple of 128) for any probe. A key whose slot lies more than 4 buckets past
 * its home region's end, or a wide table's free-marker key, makes the call run
 * again on the atomic path. No counterpart in the reference (per-tuple
 * UpdateOrInsert, compile/proxy/aggregate_hash_table.cc:187-243). */
int kg_ctx_set_agg_regions(kg_ctx* ctx, int32_t regions);
/* The FP64 addition order of the last kg_scan_group_agg call that used the
 * fixed-order kernels: T = *nblocks * 256 lanes; lane t folds, from +0.0, the
 * qualifying rows of chunks c = t + k*T (k ascending) of *vec consecutive rows
 * (ascending); lanes combine by the tree p[l] += p[l+s], s = 32..1, within a
 * 64-lane wave, then the 4 waves by the same tree; workgroup partials are
 * added in workgroup order. No counterpart in the reference (whose order is
 * the arrival order); it lets a checker restate the GPU order exactly. */
int kg_ctx_agg_order(const kg_ctx* ctx, int64_t n, int32_t* nblocks, int32_t* vec);
/* Kernel timing (no counterpart in the reference, whose timing is wall clock
 * around whole queries, util/time_execute.h:22-57): when on, every entry point
 * brackets its dominant kernel with hipEvents on the context's stream.
 * kg_ctx_kernel_stats synchronizes, returns the summed kernel time in ms and
 * the number of timed launches since the last reset, and resets if asked. */
/* The GROUP BY path the last aggregation call on this context took:
 * "register" (at most 8 dense groups, 4 with MIN/MAX, per-lane registers),
 * "lds" (dense groups in a workgroup-private LDS table of up to the device's
 * LDS per workgroup, 160 KB on gfx950), "hashed" (kg_hash_group_agg), or ""
 * before any. No counterpart in the reference. */
const char* kg_ctx_last_path(const kg_ctx* ctx);
int kg_ctx_set_timing(kg_ctx* ctx, int on);
int kg_ctx_kernel_stats(kg_ctx* ctx, double* total_ms, int64_t* launches, int reset);

/* ---- scan-select -----------------------------------------------------------
 * ScanSelect (compile/translators/scan_select_translator.cc:92-156): rows where
 * every predicate holds (conjunction = 1), or SimdScanSelect's disjunctive form
 * (simd_scan_select_translator.cc:444-454, conjunction = 0). Writes the
 * qualifying row ids to d_out_rows in ascending order — the order the 8-wide
 * permutation compaction (util/permute.cc:10-25) produces — and their number to
 * *out_count. d_out_rows must hold n entries. n < 2^31. */
int kg_scan_select(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols,
                   const kg_pred* preds, int32_t npreds, int32_t conjunction,
                   int32_t* d_out_rows, int64_t* out_count);
/* kg_scan_select that allocates nothing. No counterpart in the reference; it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and errors as kg_scan_select, except: n is limited to
 * 536,870,912 rows (131072 tiles of 4096; beyond that KG_ERR_CAPACITY), and
 * the tile counts / offsets live in a 1.5 MiB static device array per GPU,
 * held until the process ends and shared by every call on that GPU — calls
 * that overlap on one GPU (other streams, other threads, other contexts) race
 * and must be serialized by the caller. Synchronizes the context's stream. */
int kg_scan_select_noalloc(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols,
                           const kg_pred* preds, int32_t npreds, int32_t conjunction,
                           int32_t* d_out_rows, int64_t* out_count);

/* ---- scan-select + group-by aggregate (dense code keys) -----------------------
 * ScanSelect -> GroupByAggregate (compile/translators/group_by_aggregate_translator.cc:36-111)
 * where every group key is a dense int32 code column (ENUM / small INT domain):
 * key_cols[k] indexes cols, key_card[k] bounds its codes to [0, key_card[k]).
 * The groups are emitted in insertion order (first qualifying row), which is
 * the order AggregateHashTable::ForEach walks its payloads
 * (compile/proxy/aggregate_hash_table.cc:245-291). For each group:
 * out_keys[g*nkeys+k] are the codes, out_vals[g*naggs+a] the aggregate as an
 * 8-byte value (integers sign-extended in the SUM/MIN/MAX child type, REAL and
 * AVG as IEEE double bits, COUNT as int64) and out_nulls[g*naggs+a] its NULL flag.
 * Semantics follow compile/proxy/aggregator.cc: SUM wraps in the child type
 * (plan/expression/aggregate_expression.cc:21-27), AVG is over non-NULL values,
 * COUNT(col) counts non-NULL values plus one when the group's first row is NULL
 * (CountAggregator::Initialize stores 1 unconditionally, aggregator.cc:285-287).
 * With nkeys = 0 this is the ungrouped AggregateOperator (aggregate_translator.cc:34-126):
 * one group, or zero groups on empty input. */
int kg_scan_group_agg(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols,
                      const kg_pred* preds, int32_t npreds,
                      const int32_t* key_cols, const int32_t* key_card, int32_t nkeys,
                      const kg_agg* aggs, int32_t naggs, int64_t max_groups,
                      int64_t* out_ngroups, int32_t* out_keys, int64_t* out_vals,
                      int8_t* out_nulls);

/* ---- hash join -----------------------------------------------------------------
 * HBM-resident open-addressing replacement for runtime::HashTable
 * (runtime/hash_table.h:12-38, .cc:13-67) as used by HashJoinTranslator
 * (compile/translators/hash_join_translator.cc:34-139). Keys are one column of
 * type SMALLINT/INT/BIGINT/DATE/ENUM/BOOLEAN/REAL (REAL keys compare as doubles:
 * -0.0 equals +0.0, NaN matches nothing). */

/* Create an empty table able to hold `capacity_rows` build rows
 * (runtime::HashTable::Create, hash_table.cc:13-16). Keys live in 64-byte
 * buckets of eight 8-byte slots; load_factor (0 < lf <= 0.9, 0 = default 0.5)
 * sets slots = capacity_rows / lf. */
int kg_ht_create(kg_ctx* ctx, int64_t capacity_rows, double load_factor, kg_hash_table** out);
/* Insert build rows (hash_join_translator.cc:85-103; runtime::HashTable::Insert,
 * hash_table.cc:18-27). d_rows (may be NULL) lists the build row ids to insert,
 * e.g. a kg_scan_select output; otherwise rows 0..nrows-1. Rows with a NULL key
 * are skipped. May be called several times before probing. */
int kg_ht_build(kg_ctx* ctx, kg_hash_table* ht, const kg_col* key, const int32_t* d_rows,
                int64_t nrows);
/* Probe (hash_join_translator.cc:105-139; runtime::HashTable::GetBucket,
 * hash_table.cc:29-36 plus CheckEquality :57-77): every (probe row, build row)
 * pair with equal keys. Pairs are written in no particular order; the set equals
 * the reference's after sorting. *out_count receives the total number of pairs
 * even when it exceeds out_cap (then KG_ERR_CAPACITY is returned and only
 * out_cap pairs are written). */
int kg_ht_probe(kg_ctx* ctx, const kg_hash_table* ht, const kg_col* key, const int32_t* d_rows,
                int64_t nrows, int32_t* d_out_probe, int32_t* d_out_build, int64_t out_cap,
                int64_t* out_count);
/* Join cardinality only (the number of pairs kg_ht_probe would emit). */
int kg_ht_probe_count(kg_ctx* ctx, const kg_hash_table* ht, const kg_col* key,
                      const int32_t* d_rows, int64_t nrows, int64_t* out_count);
/* Join fused with GroupByAggregate on the matched build key:
 * GROUP BY build.key with SUM(d_vals) (BIGINT, wrapping) and COUNT(*), over all
 * join pairs. Writes at most max_groups groups (key, sum, count) to the device
 * arrays; *out_ngroups receives the number of groups. Group order is unspecified
 * (the reference's insertion order is recovered by the caller, if needed, from
 * the first probe row per group written to d_out_first when it is non-NULL). */
int kg_ht_probe_agg(kg_ctx* ctx, const kg_hash_table* ht, const kg_col* key, const int64_t* d_vals,
                    int64_t nrows, int64_t max_groups, int64_t* d_out_key, int64_t* d_out_sum,
                    int64_t* d_out_count, int32_t* d_out_first, int64_t* out_ngroups);
/* ---- join autotuner ------------------------------------------------------------
 * Picks the join table load factor, the probe workgroups per CU and the
 * probe-key partition fan-out (kg_ctx_set_probe_regions) for a build/probe
 * pair, per class of (build rows, build key cardinality, probe key skew): the
 * classes are log2 of the build rows, the distinct fraction of a build-key
 * sample (unique / few duplicates / many) and the share of the most frequent
 * key in a probe-key sample (uniform / moderate / heavy). The first call for a
 * class builds a table over the build keys (at most 64M rows) and times probes
 * of a probe sample (at most 128M rows) under each candidate, by coordinate
 * descent over load factor {0.35, 0.5, 0.65} (by build time plus the probe
 * time scaled to all probe rows), workgroups per CU {8, 16, 32} and regions
 * {0, 64, 128} (by probe time); later calls for the class reuse the choice. The
 * probe settings are applied to the context; the load factor is for
 * kg_ht_create. No counterpart in the reference (fixed std::unordered_map). */
typedef struct kg_join_tuning {
  double load_factor;
  int32_t probe_wg_per_cu;
  int32_t probe_regions;
  double probe_ms;              /* probe time of the choice on the sample */
  uint64_t tuning_class;        /* the (rows, cardinality, skew) class */
  int32_t from_cache;           /* 1: an earlier call's choice */
  int32_t trials;               /* configurations timed by this call */
} kg_join_tuning;
int kg_join_tune(kg_ctx* ctx, const kg_col* build_key, int64_t nbuild, const kg_col* probe_key, int64_t nprobe,
                 kg_join_tuning* out);

/* Number of build rows inserted so far. */
int kg_ht_size(kg_ctx* ctx, const kg_hash_table* ht, int64_t* out_rows);
/* Release the table (runtime::HashTable::Free, hash_table.cc:53-59). Its
 * device memory goes back to ctx's block cache (at most 2 GiB per context,
 * released by kg_ctx_destroy), where the next kg_ht_create on ctx takes it, in
 * the order of ctx's stream: a plan that creates and frees a table per query
 * allocates nothing after its first query. */
int kg_ht_free(kg_ctx* ctx, kg_hash_table* ht);

/* ---- star pipeline: scan-select -> join probe -> group-by ----------------------
 * ScanSelect over the probe-side columns (cols/preds), HashJoin probe of
 * cols[probe_key_col] into `ht`, then GroupByAggregate keyed by the matched
 * build row's dense code build_codes[build_row] in [0, group_card), with
 * aggregates over probe-side columns (kg_agg as in kg_scan_group_agg, products
 * allowed). One kernel; the group table lives in LDS. Reference chain:
 * ScanSelectTranslator (scan_select_translator.cc:92-156) ->
 * HashJoinTranslator::Consume (hash_join_translator.cc:105-139) ->
 * GroupByAggregateTranslator::Consume (group_by_aggregate_translator.cc:102-111).
 * Groups come out in the reference's insertion order (first join output row);
 * out_keys[g] is the group code. FP64 sums are added in timing-dependent order
 * (within FP64 rounding of the sequential fold). With build_codes NULL and
 * group_card 1 this is the ungrouped AggregateOperator over the join output
 * (aggregate_translator.cc:34-126): one group, or none when nothing joins, and
 * out_keys is not written. */
int kg_scan_join_group_agg(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols,
                           const kg_pred* preds, int32_t npreds, const kg_hash_table* ht,
                           int32_t probe_key_col, const kg_col* build_codes, int32_t group_card,
                           const kg_agg* aggs, int32_t naggs, int64_t max_groups, int64_t* out_ngroups,
                           int32_t* out_keys, int64_t* out_vals, int8_t* out_nulls);

/* ---- high-cardinality group-by ---------------------------------------------------
 * GroupByAggregate on one BIGINT key with SUM(BIGINT) and COUNT(*), for group
 * counts too large for an on-chip table: replaces the salted linear-probe
 * runtime::AggregateHashTable (runtime/aggregate_hash_table.cc:9-135) and its
 * generated UpdateOrInsert loop (compile/proxy/aggregate_hash_table.cc:187-243)
 * with an HBM open-addressing table updated by 64-bit atomics. Writes at most
 * max_groups groups to the device arrays in unspecified order, and the first
 * input row of each group to d_out_first (if non-NULL) so the caller can restore
 * insertion order. */
int kg_group_by_i64(kg_ctx* ctx, const int64_t* d_keys, const int64_t* d_vals, int64_t n,
                    int64_t max_groups, int64_t* d_out_key, int64_t* d_out_sum,
                    int64_t* d_out_count, int32_t* d_out_first, int64_t* out_ngroups);

/* GROUP BY over hashed keys for any number of groups: runtime::AggregateHashTable
 * and the generated UpdateOrInsert loop (runtime/aggregate_hash_table.cc:9-135,
 * compile/proxy/aggregate_hash_table.cc:187-343) for keys that are not small
 * dense codes. Rows 0..n-1 passing preds (as kg_scan_group_agg) are grouped by
 * the key columns key_cols[0..nkeys) (1..4 fixed-width columns; NULL is a key
 * value of its own; REAL keys compare with ==, so -0.0 and 0.0 are one group).
 * Groups come out in insertion order (their first qualifying row ascending,
 * ForEach's order): out_first_rows[g] is that row (the group's key values are
 * the key columns there); out_vals / out_nulls [g * naggs + a] as
 * kg_scan_group_agg (SUM wraps in the child type, AVG = sum / count, COUNT(col)
 * counts the first row whatever its NULL flag). FP64 SUM/AVG add in a
 * timing-dependent order (within 1e-12 relative of the sequential fold).
 * max_groups sizes the table: more groups return KG_ERR_CAPACITY. Outputs are
 * host arrays. TEXT keys group by their kg_text_hash codes. */
int kg_hash_group_agg(kg_ctx* ctx, int64_t n, const kg_col* cols, int32_t ncols, const kg_pred* preds,
                      int32_t npreds, const int32_t* key_cols, int32_t nkeys, const kg_agg* aggs, int32_t naggs,
                      int64_t max_groups, int64_t* out_ngroups, int32_t* out_first_rows, int64_t* out_vals,
                      int8_t* out_nulls);

/* ---- column indexes (Skinner joins) --------------------------------------------
 * Device replacement for runtime::MemoryColumnIndex
 * (runtime/memory_column_index.h:18-56, compile/proxy/memory_column_index.cc:282-428)
 * and, for unfiltered children, the disk index runtime::ColumnIndex
 * (runtime/column_index.cc:73-95): key -> bucket of tuple ids in ascending
 * order. Keys are one fixed-width column (BOOLEAN, SMALLINT, INT, DATE, ENUM,
 * BIGINT, REAL; ColumnIndex::GetInt8..GetFloat64, runtime/column_index.h:50-54;
 * TEXT keys index their kg_text_hash codes). NULL keys are not indexed
 * (permutable_skinner_join_translator.cc:972-994); REAL keys compare as
 * doubles: -0.0 and 0.0 are one key and NaN is neither indexed nor found. The
 * tuple id of a row is its position in d_rows (its position in the
 * materialized child), or its row id when d_rows is NULL. */
typedef struct kg_index kg_index;
int kg_index_build(kg_ctx* ctx, const kg_col* key, const int32_t* d_rows, int64_t nrows, kg_index** out);
/* Device tuple-id array (grouped by key, ascending per key), its length and
 * the number of distinct keys. */
int kg_index_info(kg_ctx* ctx, const kg_index* idx, const int32_t** d_tids, int64_t* ntids, int64_t* nkeys);
/* Batched GetBucket (compile/proxy/memory_column_index.cc:282-330; EMPTY for a
 * miss, runtime/column_index_bucket.cc:13-14): for each probe key (rows of
 * `key`, or d_rows[i]) the bucket as a device pointer into the tid array and a
 * size (0 for a missing or NULL key). Writes device arrays only. */
int kg_index_lookup(kg_ctx* ctx, const kg_index* idx, const kg_col* key, const int32_t* d_rows, int64_t n,
                    const int32_t** d_out_ptr, int32_t* d_out_count);
int kg_index_free(kg_ctx* ctx, kg_index* idx);

/* Batched sorted intersection (runtime/column_index_bucket.cc:59-182,
 * BucketListSortedIntersectionInit + PopulateResult drained to the end): task
 * t intersects the k ascending lists d_lists[t*k .. t*k+k-1] (sizes d_sizes)
 * from d_next_tuple[t] on (NULL = from the start). Results are written
 * ascending per task at d_out + d_out_offsets[t]; d_out_offsets has ntasks+1
 * entries. The reference's chunked form can repeat the last id of a 64-id
 * chunk when three or more lists are intersected; this returns the set, which
 * is what the downstream tuple dedup keeps. */
int kg_intersect_batch(kg_ctx* ctx, int64_t ntasks, int32_t k, const int32_t* const* d_lists,
                       const int32_t* d_sizes, const int32_t* d_next_tuple, int64_t* d_out_offsets,
                       int32_t* d_out, int64_t out_cap, int64_t* out_total);

/* ---- tuple-id set ----------------------------------------------------------------
 * Device replacement for runtime::TupleIdxTable (runtime/tuple_idx_table/
 * tuple_idx_table.h:18-51), the result dedup of the Skinner valid-tuple handler
 * (permutable_skinner_join_translator.cc:842-912): a set of int32[len] tuples
 * (len <= 8). */
typedef struct kg_tuple_set kg_tuple_set;
int kg_tuple_set_create(kg_ctx* ctx, int32_t len, int64_t capacity, kg_tuple_set** out);
/* Insert n tuples (d_tuples[n*len]). d_is_new[i] (optional) = Insert()'s
 * result when the batch is inserted in order: 1 for the first occurrence of a
 * tuple not in the set before the call. *out_new = number of tuples added. */
int kg_tuple_set_insert(kg_ctx* ctx, kg_tuple_set* s, const int32_t* d_tuples, int64_t n, uint8_t* d_is_new,
                        int64_t* out_new);
/* Copy the distinct tuples to d_out (unordered; the reference's byte-wise
 * iteration order is a sort of this output). */
int kg_tuple_set_extract(kg_ctx* ctx, const kg_tuple_set* s, int32_t* d_out, int64_t cap, int64_t* out_n);
int kg_tuple_set_free(kg_ctx* ctx, kg_tuple_set* s);

/* ---- Skinner adaptive join -------------------------------------------------------
 * Replaces runtime::ExecutePermutableSkinnerJoin
 * (runtime/skinner_join_executor.h:16-25, skinner_join_executor.cc:840-885)
 * with its table functions and valid-tuple handler
 * (compile/translators/permutable_skinner_join_translator.cc:82-994): the
 * multi-way equi-join of ntables materialized children under the conditions
 * conds, with the join order chosen per episode by a UCT search over join
 * orders (the role of UctJoinAgent, skinner_join_executor.cc:559-838). An
 * episode joins the next slice of its leftmost table's remaining tuples
 * through the other tables on the device: a table linked by one condition
 * through that condition's column index, a table linked by several through
 * the sorted intersection of all their index buckets
 * (permutable_skinner_join_translator.cc:359-662), a table linked by none as a
 * cross product; the result tuples go into a tuple-id set, and the join ends
 * when a leftmost table is complete.
 * *out_set receives a new set (caller frees with kg_tuple_set_free) holding
 * each result tuple once: int32[ntables] tuple ids, tuple id i of table t being
 * position i of its child (row d_rows[i] of the key columns, or row i when
 * d_rows is NULL). The set is the same for every join order and seed. NULL
 * keys join nothing. Keys: SMALLINT, INT, BIGINT, REAL, DATE, BOOLEAN or ENUM
 * (TEXT: KG_ERR_UNSUPPORTED). */
typedef struct kg_skinner_table {
  int64_t card;            /* tuples of the materialized child (< 2^31) */
  const int32_t* d_rows;   /* child tuple -> row of its key columns; NULL: identity */
} kg_skinner_table;
typedef struct kg_skinner_cond {
  int32_t left_table;      /* conds: left.col = right.col (Eq(ColRef, ColRef)) */
  int32_t right_table;
  kg_col left;             /* key columns, indexed by row */
  kg_col right;
} kg_skinner_cond;
typedef struct kg_skinner_opts {
  int64_t budget_per_episode;  /* partial tuples per episode (--budget_per_episode); 0: 2^20 */
  int64_t seed;                /* > 0: fixed join seed (--join_seed); otherwise the clock */
  int64_t max_episodes;        /* 0: no limit; otherwise KG_ERR_CAPACITY when not done by then */
  int32_t nprefix;             /* tables fixed at the front of every order (prefix_order) */
  int32_t prefix[8];
} kg_skinner_opts;
typedef struct kg_skinner_stats {
  int64_t episodes;
  int64_t distinct_orders;     /* join orders tried */
  int64_t result_tuples;       /* valid tuples found, repeats included (num_result_tuples) */
  int64_t distinct_tuples;     /* tuples in *out_set */
  int64_t partial_tuples;      /* partial tuples produced by all join steps */
  int32_t last_order[8];       /* order of the last episode */
} kg_skinner_stats;
int kg_skinner_join(kg_ctx* ctx, int32_t ntables, const kg_skinner_table* tables, int32_t nconds,
                    const kg_skinner_cond* conds, const kg_skinner_opts* opts, kg_tuple_set** out_set,
                    kg_skinner_stats* stats);

/* The UCT join-order agent alone (the role of UctJoinAgent,
 * skinner_join_executor.cc:779-838, with its own search: a flat node array,
 * per-table statistics, UCB1 over the connected candidates, one new node per
 * episode, random connected-first completion), host only, for callers that run
 * their own episodes. kg_uct_next writes the next order (ntables
 * entries); kg_uct_reward feeds that episode's reward back (>= 0) before the
 * next call. adjacency[i*ntables+j] != 0 when a condition joins tables i and j
 * (either triangle suffices). Returns KG_ERR_INVALID on misuse. */
typedef struct kg_uct kg_uct;
int kg_uct_create(int32_t ntables, const uint8_t* adjacency, int64_t seed, const int32_t* prefix, int32_t nprefix,
                  kg_uct** out);
int kg_uct_next(kg_uct* u, int32_t* order);
int kg_uct_reward(kg_uct* u, double reward);
void kg_uct_free(kg_uct* u);

/* ---- TEXT ------------------------------------------------------------------------
 * ScanSelect with one TEXT predicate `col op pattern` over rows 0..n-1 or over
 * the selection d_rows_in[0..n): qualifying row ids in input order to
 * d_rows_out. Semantics: byte-wise string_view comparisons and
 * StartsWith/EndsWith/Contains (runtime/string.cc:14-82); KG_TEXT_LIKE is the
 * anchored match of the planner's regex (plan/planner.cc:284-323: '%' any
 * sequence, '_' one character, everything else literal). The planner's
 * rewrites of LIKE '%x' to STARTS_WITH(x) and 'x%' to ENDS_WITH(x)
 * (planner.cc:249-269) happen before this call. Pattern <= 240 bytes. */
int kg_text_select(kg_ctx* ctx, const kg_text_col* col, int32_t op, const char* pattern, int32_t plen,
                   const int32_t* d_rows_in, int64_t n, int32_t* d_rows_out, int64_t* out_count);
/* MIN (is_max = 0) or MAX over the non-NULL values of col at rows d_rows[0..n)
 * (or 0..n-1): *out_pos = the input position of the first minimal value (-1
 * when all are NULL / n = 0), its bytes copied to out_buf (up to buf_cap) and
 * its length to *out_len (MinMaxAggregator over TEXT, aggregator.cc:155-160). */
int kg_text_minmax(kg_ctx* ctx, const kg_text_col* col, const int32_t* d_rows, int64_t n, int32_t is_max,
                   int64_t* out_pos, char* out_buf, int64_t buf_cap, int64_t* out_len);

/* TEXT equi-join keys. A TEXT key has no fixed width, so a TEXT join runs as a
 * BIGINT join on 64-bit hashes of the strings followed by an exact pair check:
 * kg_text_hash writes the hash of every value (rows d_rows[i], or i when
 * d_rows is NULL) to d_out and its NULL flag to d_out_nulls (optional; a NULL
 * key joins nothing, hash_join_translator.cc:86-101). It replaces the
 * std::hash<string_view> key hash of TEXT values (compile/proxy/value/
 * string.cc:104-107 -> runtime/string.cc:74-77); equal strings hash equal.
 * kg_text_pairs_equal then keeps, of n candidate pairs (a[d_rows_a[i]],
 * b[d_rows_b[i]]), the positions i whose strings are byte-equal and not NULL
 * (CheckEquality over TEXT keys, hash_join_translator.cc:57-77, with
 * String::Equals, runtime/string.cc:42-46), ascending, in d_out_idx. */
int kg_text_hash(kg_ctx* ctx, const kg_text_col* col, const int32_t* d_rows, int64_t n, int64_t* d_out,
                 int8_t* d_out_nulls);
int kg_text_pairs_equal(kg_ctx* ctx, const kg_text_col* a, const int32_t* d_rows_a, const kg_text_col* b,
                        const int32_t* d_rows_b, int64_t n, int32_t* d_out_idx, int64_t* out_count);
/* Exactness of grouping (or joining) TEXT values by their kg_text_hash codes:
 * of the n values (rows d_rows[i], or i when d_rows is NULL) with codes
 * d_hash[i], *out_mismatches receives the number whose string differs from
 * the string of the smallest row with the same code (NULL values are
 * skipped). 0 means equal codes imply equal strings on these rows, so a GROUP
 * BY or join on the codes equals the one on the strings (String::Equals,
 * runtime/string.cc:42-46); otherwise two strings collided. One pass builds a
 * device table of codes (2n slots), a second compares each row with its
 * code's row: O(n). No counterpart in the reference, which compares strings
 * inside its hash tables (compile/proxy/aggregate_hash_table.cc:187-243). */
int kg_text_hash_verify(kg_ctx* ctx, const kg_text_col* col, const int32_t* d_rows, const int64_t* d_hash,
                        int64_t n, int64_t* out_mismatches);

/* ---- index-nested-loop join step (Skinner table functions) ----------------------
 * Extends each partial join tuple d_tuples[i*width .. +width) (tuple ids per
 * table) by every tuple id of idx's bucket for the key src_key[row], where
 * row = d_src_rows[tuple[src_pos]] (or tuple[src_pos] when d_src_rows is
 * NULL): the GetBucket + handler chain of
 * permutable_skinner_join_translator.cc:359-433. Output tuples have width+1
 * ids, grouped by input tuple in input order, bucket order (ascending ids)
 * inside a group. d_tuples NULL (width 1) stands for the n one-table tuples
 * (0), (1), ..., (n-1): the first table of a join order over its whole
 * selection. *out_n receives the total even when it exceeds out_cap
 * (KG_ERR_CAPACITY). */
int kg_index_join_expand(kg_ctx* ctx, int64_t n, int32_t width, const int32_t* d_tuples, int32_t src_pos,
                         const kg_col* src_key, const int32_t* d_src_rows, const kg_index* idx,
                         int64_t out_cap, int32_t* d_out, int64_t* out_n);

/* Row ids of one table in a tuple batch: d_out[i] = d_sel[tuple_i[pos]] (or
 * tuple_i[pos] when d_sel is NULL) — the tuple id -> materialized row step the
 * Skinner valid-tuple handler does before emitting
 * (permutable_skinner_join_translator.cc:842-912). Device output only. */
int kg_tuple_rows(kg_ctx* ctx, const int32_t* d_tuples, int32_t width, int32_t pos, const int32_t* d_sel,
                  int64_t n, int32_t* d_out);
/* Materialize src[d_rows[i]] (and its NULL flag when d_out_nulls is set) for a
 * fixed-width column (runtime::ColumnData::Get* per row,
 * runtime/column_data.cc:96-108, done once per batch). Device output only. */
int kg_gather(kg_ctx* ctx, const kg_col* src, const int32_t* d_rows, int64_t n, void* d_out, int8_t* d_out_nulls);

/* ---- batched execution -----------------------------------------------------------
 * Many queries' pipelines in one persistent kernel fed by device work queues,
 * the four kernels of a batch (fills, build phase, scan/probe phase, finalize)
 * captured once as a HIP graph per kg_batch. No counterpart in the reference,
 * which runs one query at a time (execution/executable_query.cc:199-255); each
 * batched query is the pipeline a host plan would otherwise run through the
 * plain entry points:
 *   KG_Q_SCAN_AGG  kg_scan_group_agg(n, cols, preds, key_cols, key_card, aggs);
 *   KG_Q_JOIN_AGG  kg_scan_select(build_cols, build_preds) -> kg_ht_build(build
 *                  key) -> kg_scan_join_group_agg(n, cols, preds, probe key,
 *                  build code column or none = one group, aggs).
 * Results use the plain entry points' formats and semantics (SUM wrapping, AVG,
 * COUNT(col), insertion order). Queries with equal table_id and n read the same
 * rows and share tile reads. FP64 sums are added in a timing-dependent order
 * (within FP64 rounding of the plain path); integers, counts and group order
 * are exact. MIN/MAX and more than 16 aggregates are not batched
 * (KG_ERR_UNSUPPORTED in that query's result; run it on the plain path). */
typedef struct kg_batch kg_batch;
/* KG_Q_CHAIN: the multi-way join shape of the reference's Skinner queries
 * (PermutableSkinnerJoinTranslator over filtered children,
 * compile/translators/permutable_skinner_join_translator.cc:82-994, then the
 * ungrouped MIN aggregates of aggregate_translator.cc:34-126) as one served
 * query: tables t0..t(k-1) joined along a chain t_i.fk = t_(i-1).id, each
 * child filtered by conjunctive predicates over its columns and at most one
 * TEXT predicate (kg_text_select), the result tuples deduplicated, then
 * MIN(TEXT) over one table's string column and MIN over one table's integer
 * column. Run by kg_chain_run; batches report it KG_ERR_UNSUPPORTED and the
 * serving loop runs it on the plain path. With skinner_budget > 0 the join
 * builds its own indexes (fk_index is not used). */
typedef struct kg_chain_table {
  int64_t n;                    /* rows */
  const kg_col* cols;
  int32_t ncols;
  int32_t npreds;
  const kg_pred* preds;         /* conjunctive, over cols */
  const kg_text_col* text;      /* TEXT predicate column, or NULL for none */
  const char* text_pattern;
  int32_t text_op;              /* kg_text_op */
  int32_t text_plen;
  int32_t id_col;               /* key the next table references (cols index; -1 for the last table) */
  int32_t fk_col;               /* key that references the previous table's id_col (-1 for the first) */
  const kg_index* fk_index;     /* optional stored index: kg_index_build of fk_col over all rows (the
                                   column index of permutable_skinner_join_translator.cc:116-132). The
                                   fixed order then joins through it and tests this table's predicates on
                                   the joined rows; NULL: filter first, index the selection per query */
} kg_chain_table;
typedef struct kg_chain {
  int32_t ntables;              /* 2..8 */
  int32_t min_text_table;       /* MIN over min_text at that table's rows, or -1 */
  const kg_chain_table* tables;
  const kg_text_col* min_text;
  int32_t min_int_table;        /* MIN over tables[min_int_table].cols[min_int_col] (an integer type), or -1 */
  int32_t min_int_col;
  int64_t skinner_budget;       /* 0: the order t0, t1, ... by index-nested-loop steps; > 0: kg_skinner_join
                                   with this budget per episode (UCT-chosen orders) */
  int64_t seed;                 /* kg_skinner_join seed (> 0) */
} kg_chain;
/* One KG_Q_CHAIN query: *out_count = distinct result tuples; *out_min_int =
 * the MIN (null flag *min_int_null = 1 when there is no tuple or no MIN
 * column); *out_min_text_row = the row of min_text holding the first minimal
 * string (-1 when there is none). Synchronizes the context's stream. */
int kg_chain_run(kg_ctx* ctx, const kg_chain* q, int64_t* out_count, int64_t* out_min_int, int8_t* min_int_null,
                 int64_t* out_min_text_row);
enum kg_query_kind { KG_Q_SCAN_AGG = 0, KG_Q_JOIN_AGG = 1, KG_Q_CHAIN = 2 };
typedef struct kg_query {
  int32_t kind;                 /* kg_query_kind */
  int32_t table_id;             /* scan/probe table identity for tile sharing */
  int64_t n;                    /* scan/probe rows */
  const kg_col* cols;           /* scan/probe columns */
  int32_t ncols;
  int32_t npreds;
  const kg_pred* preds;         /* conjunctive */
  const int32_t* key_cols;      /* SCAN_AGG: dense code keys, as kg_scan_group_agg */
  const int32_t* key_card;
  int32_t nkeys;
  int32_t naggs;
  const kg_agg* aggs;           /* over cols */
  int32_t probe_key_col;        /* JOIN_AGG: probe key in cols */
  int32_t build_key_col;        /* JOIN_AGG: build key in build_cols */
  int64_t build_n;
  const kg_col* build_cols;
  int32_t build_ncols;
  int32_t build_npreds;
  const kg_pred* build_preds;
  int32_t build_code_col;       /* group code column in build_cols, or -1: one group */
  int32_t group_card;           /* codes are in [0, group_card) */
  double load_factor;           /* join table load factor; 0 = what kg_join_tune chose on
                                   this context for builds of the same power-of-two size
                                   (unique keys preferred), else 0.5 */
  const kg_chain* chain;        /* KG_Q_CHAIN: the query (the fields above are unused); its
                                   result is one group of 3 values: COUNT of distinct tuples,
                                   MIN int (null when none), row of the MIN string (null when none) */
} kg_query;
typedef struct kg_query_result {
  int32_t status;               /* KG_OK or this query's error */
  int32_t reserved;
  int64_t max_groups;           /* capacity of keys/vals/nulls, in groups (input) */
  int64_t ngroups;              /* groups produced */
  int32_t* keys;                /* [max_groups * nkeys] (JOIN_AGG with codes: 1 key) */
  int64_t* vals;                /* [max_groups * naggs] */
  int8_t* nulls;                /* [max_groups * naggs] */
} kg_query_result;
typedef struct kg_batch_opts {
  int64_t arena_bytes;          /* device bytes for one batch's tables, accumulators, results; 0 = 1 GiB */
  int64_t reserved;
} kg_batch_opts;
/* A batch slot with its own stream, arena and captured graph. */
int kg_batch_create(kg_ctx* ctx, const kg_batch_opts* opts, kg_batch** out);
int kg_batch_free(kg_batch* b);
/* Projected device bytes of one query in a batch arena (join table included),
 * the admission measure of the serving loop. */
int kg_batch_footprint(const kg_query* q, int64_t* bytes);
int64_t kg_batch_arena_bytes(const kg_batch* b);
/* Launch up to 256 queries on the batch's stream and return. Per-query errors
 * (validation, arena capacity) are reported in kg_batch_collect's results. */
int kg_batch_submit(kg_batch* b, const kg_query* qs, int32_t nq);
/* *done = 1 when the submitted batch has finished on the device. */
int kg_batch_poll(kg_batch* b, int32_t* done);
/* Queries complete individually while the batch runs: the device publishes
 * query i's result as soon as its last work item is done (tables are scanned
 * smallest first). *ready = 1 once query i (submission order) can be
 * collected with kg_batch_collect_query. */
int kg_batch_query_ready(kg_batch* b, int32_t i, int32_t* ready);
int kg_batch_collect_query(kg_batch* b, int32_t i, kg_query_result* res);
/* Wait for the submitted batch and write each query's result (res[nq]). */
int kg_batch_collect(kg_batch* b, kg_query_result* res);
/* Wait for the submitted batch without writing results (after collecting
 * every query with kg_batch_collect_query); the slot can take a new batch. */
int kg_batch_release(kg_batch* b);
/* kg_batch_submit + kg_batch_collect. */
int kg_batch_run(kg_batch* b, const kg_query* qs, int32_t nq, kg_query_result* res);

/* ---- serving loop ----------------------------------------------------------------
 * No counterpart in the reference (one query per process run,
 * util/time_execute.h:22-57). Replays queries arriving at arrival_s[i] seconds
 * (ascending) after the call starts, on columns already resident in device
 * memory, and records each query's latency (completion - arrival) and result.
 * KG_SERVE_PLAIN runs queries one at a time, in arrival order, through the
 * plain entry points (join queries reuse one pooled join table, reset per
 * query, as the batched path reuses its arenas). KG_SERVE_BATCHED keeps up to `slots` batches in flight
 * on their own streams (continuous batching): whenever a slot is free, the
 * queries that have arrived are admitted in arrival order while their projected
 * device footprint (kg_batch_footprint) fits the slot's arena, the batch holds
 * at most max_batch queries and its rows (scanned + build) stay within
 * max_batch_rows (the first query is always admitted), which bounds how long a
 * batch keeps its slot; chain queries (KG_Q_CHAIN), queries larger than an
 * arena and queries a batch reports unsupported run on the plain path in a
 * lane: a host thread with a context and stream of its own (kept with ctx for
 * later calls), one query at a time, alongside the batches; with
 * lane_scan_rows, so do large scans that arrive when nothing is pending. Each
 * query completes as soon as its own work is done (kg_batch_query_ready), not
 * at the end of its batch. */
enum kg_serve_mode { KG_SERVE_PLAIN = 0, KG_SERVE_BATCHED = 1 };
typedef struct kg_serve_opts {
  int32_t mode;                 /* kg_serve_mode */
  int32_t slots;                /* batches in flight; 0 = 3 */
  int32_t max_batch;            /* queries per batch; 0 = 128 */
  int32_t lanes;                /* batched mode: plain lanes (host threads with their own contexts); 0 = 1 */
  int64_t arena_bytes;          /* per slot; 0 = 1 GiB */
  int64_t max_batch_rows;       /* projected work cap of a batch: scanned + build rows; 0 = none */
  int64_t lane_scan_rows;       /* batched mode: a KG_Q_SCAN_AGG query of at least this many rows that arrives
                                   with no backlog (nothing pending) runs on an idle lane; 0 = never */
} kg_serve_opts;
typedef struct kg_serve_stats {
  double wall_s;                /* first arrival to last completion */
  double qps;
  double p50_s, p99_s, max_s;   /* latency percentiles */
  double mean_batch;            /* queries per batch (batched mode) */
  int64_t batches;
  int64_t fallbacks;            /* batched mode: queries run on a plain lane because batches do not take them */
  int64_t lane_scans;           /* batched mode: large scans sent to an idle lane (lane_scan_rows) */
} kg_serve_stats;
int kg_serve(kg_ctx* ctx, const kg_query* qs, int32_t nq, const double* arrival_s, const kg_serve_opts* opts,
             kg_query_result* res, double* latency_s, kg_serve_stats* stats);

/* ---- on-disk columns (FileManager / ColumnData / Enum / ColumnIndex) ---------
 * The reference's files, loaded into HBM-resident columns of the context. */

/* Replaces FileManager::Open + ColumnData::Open{Int8,Int16,Int32,Int64,Float64}
 * (runtime/file_manager.cc:22-60, runtime/column_data.cc:24-56): the typed
 * .kdb array at `path` (2, 4, 8 or 1 bytes per value for SMALLINT,
 * INT/DATE/ENUM, BIGINT/REAL, BOOLEAN) and, when "<stem>_null.kdb" exists
 * beside it, its int8 NULL flags (util/load.h:70-90). The column stays
 * resident in device memory, cached by path, until kg_file_release or the
 * context is destroyed; later calls for the same path return it again. */
int kg_file_column(kg_ctx* ctx, const char* path, int32_t type, kg_col* out, int64_t* out_n);
/* ColumnData::OpenText + GetText (runtime/column_data.cc:56,110-114): a TEXT
 * column file (int32 cardinality, 16-byte {length, offset} slots, NUL-ended
 * strings; :175-229) as device offsets [n + 1] + bytes (+ NULL flags). */
int kg_file_text(kg_ctx* ctx, const char* path, kg_text_col* out, int64_t* out_n);
/* Frees the resident column loaded from `path` (NULL: every loaded column). */
int kg_file_release(kg_ctx* ctx, const char* path);
/* Enum dictionary (runtime/enum.cc:26-131, .kdbenum): *out_count = the number
 * of codes; with buf non-NULL the strings of codes 0..count-1 are copied back
 * to back into buf (host memory) with offsets[code .. code + 1] delimiting
 * each (offsets holds max_codes + 1 >= count + 1 entries). */
int kg_file_enum(kg_ctx* ctx, const char* path, char* buf, int64_t buf_cap, int64_t* offsets, int32_t max_codes,
                 int32_t* out_count);
/* ColumnIndex::Open + Get* (runtime/column_index.cc:150-240, .kdbindex with
 * SMALLINT, INT, DATE, ENUM, BIGINT or REAL keys; other key types return
 * KG_ERR_UNSUPPORTED): the file's key -> tuple-id lists as a device column
 * index, equal to kg_index_build over the listed tuples. */
int kg_file_index(kg_ctx* ctx, const char* path, int32_t key_type, kg_index** out);

/* ---- ORDER BY and result rows -------------------------------------------------- */

/* Replaces OrderByTranslator's buffer sort (order_by_translator.cc:35-110 over
 * runtime/vector.cc:45-110 with LessThan, compile/proxy/evaluate.cc:339-400):
 * d_perm[0..n) = the input positions in ORDER BY keys[0], ..., keys[nkeys-1]
 * order, ascending[k] != 0 for ASC. NULLs sort after every value ascending and
 * before descending; ENUM keys order by their decoded strings (enum_rank[k]: a
 * device array of each code's string rank, kg_enum_ranks); TEXT keys (type
 * KG_TEXT, the column in text_keys[k]) by unsigned bytes; BOOLEAN keys are
 * rejected as in the reference. Rows with equal keys come out in reverse input
 * order, which is what the reference's merge (the right run's element first
 * unless comp(left, right)) produces. enum_rank / text_keys may be NULL when
 * no key needs them. */
int kg_order_by(kg_ctx* ctx, int64_t n, const kg_col* keys, int32_t nkeys, const int32_t* ascending,
                const int32_t* const* enum_rank, const kg_text_col* text_keys, int32_t* d_perm);
/* Rank of each dictionary string (codes 0..count-1, strings back to back in
 * `strings` delimited by offsets[code .. code + 1]) in std::string order:
 * the order Enum values compare in (compile/proxy/evaluate.cc:390-393). Host
 * memory in and out. */
int kg_enum_ranks(const char* strings, const int64_t* offsets, int32_t count, int32_t* out_rank);

/* One column of result rows in host memory. ENUM: dict / dict_off[code ..
 * code + 1] delimit each code's string (dict_n codes); TEXT: dict / dict_off[row
 * .. row + 1] delimit each row's string. */
typedef struct kg_fmt_col {
  int32_t type;
  int32_t dict_n;
  const void* data;
  const int8_t* nulls;
  const char* dict;
  const int64_t* dict_off;
} kg_fmt_col;
/* The OutputTranslator's printing (output_translator.cc:25-31,
 * runtime/printer.cc:13-35) of n rows, in `order` (host positions) or input
 * order: each value followed by '|', NULL as the empty string, REAL with six
 * decimals (std::fixed), DATE as YYYY-MM-DD, BOOLEAN as t / f, '\n' after each
 * row. *out_len = the text's length; with buf NULL only the length is
 * computed (KG_ERR_CAPACITY when cap is smaller). Host memory only. */
int kg_format_rows(const kg_fmt_col* cols, int32_t ncols, int64_t n, const int32_t* order, char* buf, int64_t cap,
                   int64_t* out_len);

/* ---- general expressions ---------------------------------------------------
 * The ExpressionTranslator (compile/translators/expression_translator.cc:43-360)
 * and the evaluation rules it calls (compile/proxy/evaluate.cc:11-338) over a
 * column set, evaluated on the device. An expression is an array of nodes in
 * which every child precedes its parent; the last node is the root. Types
 * follow the plan's rules (plan/expression/arithmetic_expression.cc:11-78):
 * both operands of a binary node have the same type, except ENUM against a
 * TEXT literal; comparisons, AND/OR and the string predicates are BOOLEAN;
 * ADD/SUB/MUL keep the operand type (integer types wrap); DIV takes REAL.
 * NULL rules as the reference's: arithmetic, comparisons and the string
 * predicates are NULL when an operand is; AND is NULL when its left operand
 * is NULL, else FALSE when it is FALSE, else the right operand; OR is the
 * right operand when the left is NULL or FALSE, else TRUE; NOT keeps NULL;
 * CASE takes the ELSE branch for a NULL condition; IS_NULL is never NULL; an
 * integer-to-REAL conversion turns NULL into a non-NULL 0.0; LIKE and IN on
 * a NULL value are a non-NULL FALSE; EXTRACT(YEAR) keeps NULL. */
enum kg_expr_op {
  KG_E_COL = 0,          /* column a of the column set */
  KG_E_LIT = 1,          /* literal: ilit (integer types, DATE as a Julian day,
                            BOOLEAN, ENUM code), flit (REAL), str/str_len (TEXT);
                            flags & KG_E_NULL: the NULL literal */
  KG_E_ADD = 2, KG_E_SUB = 3, KG_E_MUL = 4, KG_E_DIV = 5,
  KG_E_EQ = 6, KG_E_NEQ = 7, KG_E_LT = 8, KG_E_LEQ = 9, KG_E_GT = 10, KG_E_GEQ = 11,
  KG_E_AND = 12, KG_E_OR = 13,
  KG_E_STARTS_WITH = 14, KG_E_ENDS_WITH = 15, KG_E_CONTAINS = 16,
  KG_E_NOT = 17, KG_E_IS_NULL = 18,
  KG_E_CASE = 19,        /* CASE WHEN a THEN b ELSE c */
  KG_E_INT_TO_REAL = 20,
  KG_E_EXTRACT_YEAR = 21,
  KG_E_LIKE = 22,        /* a LIKE str (the SQL pattern: '%', '_') */
  KG_E_IN = 23           /* ENUM a IN (values[0 .. nvalues)) */
};
#define KG_E_NULL 1

typedef struct kg_expr_node {
  int32_t op;            /* kg_expr_op */
  int32_t type;          /* result type (kg_type) */
  int32_t a, b, c;       /* children (node indices), or the column (KG_E_COL) */
  int32_t flags;
  int64_t ilit;
  double flit;
  const char* str;       /* host bytes */
  int32_t str_len;
  int32_t nvalues;
  const int32_t* values; /* host ENUM codes of KG_E_IN */
} kg_expr_node;

/* ENUM dictionary of a column (host strings by code), for the operators that
 * compare decoded strings (evaluate.cc:222-300, expression_translator.cc:275-296). */
typedef struct kg_enum_dict {
  const char* bytes;
  const int64_t* offsets;  /* count + 1 */
  int32_t count;
} kg_enum_dict;

/* Evaluates the expression for rows d_rows[0 .. n) (or 0 .. n) of the column
 * set into out (device data of the root type's width and int8 NULL flags,
 * both n long). cols[i].type KG_TEXT takes text_cols[i]; dicts may be NULL
 * when no ENUM string operator is used. At most 64 nodes. */
int kg_expr_eval(kg_ctx* ctx, int64_t n, const kg_col* cols, const kg_text_col* text_cols, const kg_enum_dict* dicts,
                 int32_t ncols, const kg_expr_node* nodes, int32_t nnodes, const int32_t* d_rows, kg_col* out);
/* SelectTranslator (compile/translators/select_translator.cc:23-37): the rows
 * (ascending) whose BOOLEAN expression is TRUE and not NULL. */
int kg_expr_select(kg_ctx* ctx, int64_t n, const kg_col* cols, const kg_text_col* text_cols,
                   const kg_enum_dict* dicts, int32_t ncols, const kg_expr_node* nodes, int32_t nnodes,
                   const int32_t* d_rows, int32_t* d_out_rows, int64_t* out_count);

#ifdef __cplusplus
}
#endif

#endif /* KUSHGPU_H_ */
