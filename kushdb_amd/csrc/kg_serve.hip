// Serving loop: replays a stream of query arrivals against resident columns,
// either one query at a time through the plain entry points, or with
// continuous batching over several kg_batch slots, and records per-query
// latency. No counterpart in the reference, which runs one query per process
// (util/time_execute.h:22-57).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "kg_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Grow-only device buffer for build-side selections of the plain path.
struct Scratch {
  int32_t* sel = nullptr;
  int64_t cap = 0;
  ~Scratch() {
    if (sel) (void)hipFree(sel);
  }
  int reserve(kg_ctx* ctx, int64_t n) {
    if (n <= cap) return KG_OK;
    if (sel) (void)hipFree(sel);
    sel = nullptr;
    cap = 0;
    if (hipMalloc(reinterpret_cast<void**>(&sel), static_cast<size_t>(std::max<int64_t>(n, 1)) * 4) != hipSuccess) {
      (void)hipGetLastError();
      return kg_fail(ctx, KG_ERR_OOM, "selection buffer of %lld rows", (long long)n);
    }
    cap = n;
    return KG_OK;
  }
};

// The plain path's join table: one table whose device memory is reused by
// every join query (reset to an empty table of the query's size), replaced
// by a larger one only when a query needs more slots, as the batched path
// reuses its arena.
struct TablePool {
  kg_hash_table* t = nullptr;
  kg_ctx* ctx = nullptr;
  ~TablePool() {
    if (t) kg_ht_free(ctx, t);
  }
  int get(kg_ctx* c, int64_t rows, double lf, kg_hash_table** out) {
    ctx = c;
    if (t && kg::ht_fits(t, rows, lf)) {
      *out = t;
      return kg::ht_reset(c, t, rows, lf);
    }
    if (t) kg_ht_free(c, t);
    t = nullptr;
    const int st = kg_ht_create(c, rows, lf, &t);
    *out = t;
    return st;
  }
};

// One query through the plain entry points (what a host plan would call).
int run_plain(kg_ctx* ctx, const kg_query& q, kg_query_result& R, Scratch& S, TablePool& T) {
  R.ngroups = 0;
  if (q.kind == KG_Q_CHAIN) {
    KG_REQUIRE(ctx, R.max_groups >= 1 && R.vals && R.nulls, "a chain query's result takes one group of 3 values");
    int64_t count = 0, min_int = 0, text_row = -1;
    int8_t int_null = 1;
    const int st = kg_chain_run(ctx, q.chain, &count, &min_int, &int_null, &text_row);
    if (st != KG_OK) return st;
    R.ngroups = 1;
    R.vals[0] = count;
    R.nulls[0] = 0;
    R.vals[1] = min_int;
    R.nulls[1] = int_null;
    R.vals[2] = text_row;
    R.nulls[2] = text_row < 0 ? 1 : 0;
    return KG_OK;
  }
  if (q.kind == KG_Q_SCAN_AGG)
    return kg_scan_group_agg(ctx, q.n, q.cols, q.ncols, q.preds, q.npreds, q.key_cols, q.key_card, q.nkeys, q.aggs,
                             q.naggs, R.max_groups, &R.ngroups, R.keys, R.vals, R.nulls);
  if (q.kind != KG_Q_JOIN_AGG) return kg_fail(ctx, KG_ERR_INVALID, "bad query kind %d", q.kind);
  KG_REQUIRE(ctx, q.build_cols && q.build_key_col >= 0 && q.build_key_col < q.build_ncols, "bad build key column");
  const int32_t* sel = nullptr;
  int64_t nb = q.build_n;
  int st;
  if (q.build_npreds > 0) {
    if ((st = S.reserve(ctx, q.build_n)) != KG_OK) return st;
    if ((st = kg_scan_select(ctx, q.build_n, q.build_cols, q.build_ncols, q.build_preds, q.build_npreds, 1, S.sel,
                             &nb)) != KG_OK)
      return st;
    sel = S.sel;
  }
  kg_hash_table* ht = nullptr;
  if ((st = T.get(ctx, nb, q.load_factor, &ht)) != KG_OK) return st;
  st = kg_ht_build(ctx, ht, &q.build_cols[q.build_key_col], sel, nb);
  if (st == KG_OK) {
    const kg_col* codes = q.build_code_col >= 0 ? &q.build_cols[q.build_code_col] : nullptr;
    st = kg_scan_join_group_agg(ctx, q.n, q.cols, q.ncols, q.preds, q.npreds, ht, q.probe_key_col, codes,
                                codes ? q.group_card : 1, q.aggs, q.naggs, R.max_groups, &R.ngroups, R.keys, R.vals,
                                R.nulls);
  }
  return st;
}

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  // Nearest rank.
  size_t k = static_cast<size_t>(std::ceil(p * static_cast<double>(v.size())));
  k = std::min(std::max<size_t>(k, 1), v.size());
  return v[k - 1];
}

// The batched mode's plain lanes: host threads that run the queries batches
// do not take (chain queries, unsupported or oversized ones, and large scans
// that arrive with no backlog) one at a time, each on its own context and
// stream, while the serving loop keeps admitting and collecting batches.
struct Lane {
  std::mutex m;
  std::condition_variable cv;
  std::deque<int32_t> q;
  bool stop = false;
  std::thread th;
  std::atomic<int32_t> load{0};  // queued + running
  kg_ctx* ctx = nullptr;
  Scratch S;
  TablePool T;
  void push(int32_t i) {
    load.fetch_add(1);
    {
      std::lock_guard<std::mutex> lk(m);
      q.push_back(i);
    }
    cv.notify_one();
  }
  void finish() {
    if (!th.joinable()) return;
    {
      std::lock_guard<std::mutex> lk(m);
      stop = true;
    }
    cv.notify_one();
    th.join();
  }
};

// Lane contexts: created once per serving context (kept for later calls),
// with the same execution settings (deterministic mode, partitioning choices,
// join tunes).
int lane_contexts(kg_ctx* ctx, int n) {
  while (static_cast<int>(ctx->serve_lanes.size()) < n) {
    kg_ctx* l = nullptr;
    const int st = kg_ctx_create(ctx->device, &l);
    if (st != KG_OK) return kg_fail(ctx, st, "serving lane context: status %d", st);
    ctx->serve_lanes.push_back(l);
  }
  for (kg_ctx* l : ctx->serve_lanes) {
    l->deterministic = ctx->deterministic;
    l->probe_regions = ctx->probe_regions;
    l->probe_wg_per_cu = ctx->probe_wg_per_cu;
    l->build_regions = ctx->build_regions;
    l->agg_regions = ctx->agg_regions;
    l->join_tunes = ctx->join_tunes;
  }
  return KG_OK;
}

struct Slot {
  kg_batch* b = nullptr;
  bool busy = false;
  std::vector<int32_t> ids;
  std::vector<kg_query> qbuf;
  std::vector<char> finished;
  int left = 0;
};

}  // namespace

extern "C" int kg_serve(kg_ctx* ctx, const kg_query* qs, int32_t nq, const double* arrival_s,
                        const kg_serve_opts* opts, kg_query_result* res, double* latency_s, kg_serve_stats* stats) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, nq >= 0 && (nq == 0 || (qs && arrival_s && res && latency_s)), "null argument");
  KG_REQUIRE(ctx, opts, "options are NULL");
  for (int i = 1; i < nq; i++) KG_REQUIRE(ctx, arrival_s[i] >= arrival_s[i - 1], "arrivals must be ascending");
  const bool batched = opts->mode == KG_SERVE_BATCHED;
  KG_REQUIRE(ctx, batched || opts->mode == KG_SERVE_PLAIN, "bad serving mode %d", opts->mode);
  const int nslots = opts->slots > 0 ? opts->slots : 3;
  const int max_batch = opts->max_batch > 0 ? std::min(opts->max_batch, 256) : 128;
  KG_REQUIRE(ctx, nslots <= 8, "at most 8 batch slots");
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // Join queries without a load factor take the context's tuned choice (the
  // batched path resolves them the same way, so footprints agree).
  std::vector<kg_query> resolved(qs, qs + nq);
  for (kg_query& q : resolved)
    if (q.kind == KG_Q_JOIN_AGG && !(q.load_factor > 0)) q.load_factor = kg::tuned_load_factor(ctx, q.build_n);
  qs = resolved.data();
  Scratch S;
  TablePool T;
  std::vector<Slot> slots;
  kg_serve_stats st_{};
  const int nlanes = batched ? (opts->lanes > 0 ? opts->lanes : 1) : 0;
  KG_REQUIRE(ctx, nlanes <= 8, "at most 8 plain lanes");
  std::vector<std::unique_ptr<Lane>> lanes;
  auto cleanup = [&]() {
    for (auto& l : lanes) l->finish();
    for (Slot& s : slots)
      if (s.b) kg_batch_free(s.b);
  };
  if (batched) {
    const int st = lane_contexts(ctx, nlanes);
    if (st != KG_OK) return st;
    for (int k = 0; k < nlanes; k++) {
      lanes.emplace_back(new Lane());
      lanes.back()->ctx = ctx->serve_lanes[k];
    }
    kg_batch_opts bo{};
    bo.arena_bytes = opts->arena_bytes;
    slots.resize(nslots);
    for (Slot& s : slots) {
      const int st = kg_batch_create(ctx, &bo, &s.b);
      if (st != KG_OK) {
        cleanup();
        return st;
      }
    }
  }
  // Pre-fault the plain path's code and the batch graphs before the clock
  // starts. After this synchronization every column is complete, so batches
  // need not wait for the context stream (which may share a hardware queue
  // with a slot or a lane).
  if (const hipError_t e = hipDeviceSynchronize(); e != hipSuccess) {
    cleanup();
    return kg_fail(ctx, KG_ERR_HIP, "hipDeviceSynchronize failed: %s", hipGetErrorString(e));
  }
  for (Slot& sl : slots) kg::batch_order_after_ctx(sl.b, false);
  std::vector<double> done_at(nq, -1.0);
  std::deque<int32_t> pending;
  int32_t next = 0;
  std::atomic<int32_t> ndone{0};
  const Clock::time_point t0 = Clock::now();
  auto finish = [&](int32_t i) {
    done_at[i] = seconds_since(t0);
    latency_s[i] = done_at[i] - arrival_s[i];
    ndone.fetch_add(1, std::memory_order_release);
  };
  // Plain mode: in the loop. Batched mode: handed to the least loaded lane,
  // whose thread writes res[i], done_at[i] and latency_s[i] of its queries
  // only.
  auto least_loaded = [&]() {
    Lane* best = lanes[0].get();
    for (auto& l : lanes)
      if (l->load.load() < best->load.load()) best = l.get();
    return best;
  };
  auto plain = [&](int32_t i) {
    if (batched) {
      least_loaded()->push(i);
      return;
    }
    res[i].status = run_plain(ctx, qs[i], res[i], S, T);
    finish(i);
  };
  for (auto& lp : lanes) {
    Lane* L = lp.get();
    L->th = std::thread([&, L]() {
      (void)hipSetDevice(L->ctx->device);
      for (;;) {
        int32_t i;
        {
          std::unique_lock<std::mutex> lk(L->m);
          L->cv.wait(lk, [&]() { return L->stop || !L->q.empty(); });
          if (L->q.empty()) return;
          i = L->q.front();
          L->q.pop_front();
        }
        res[i].status = run_plain(L->ctx, qs[i], res[i], L->S, L->T);
        finish(i);
        L->load.fetch_sub(1);
      }
    });
  }
  // Large scans that arrive with no backlog run plain on an idle lane: a lone
  // scan gains nothing from a batch (no other query shares its tiles), and the
  // plain scan kernel streams its table faster than a batch's shared scan.
  auto to_lane = [&](int32_t i) {
    if (!batched || opts->lane_scan_rows <= 0 || qs[i].kind != KG_Q_SCAN_AGG || qs[i].n < opts->lane_scan_rows ||
        !pending.empty())
      return false;
    return least_loaded()->load.load() == 0;
  };
  while (ndone.load(std::memory_order_acquire) < nq) {
    const double now = seconds_since(t0);
    while (next < nq && arrival_s[next] <= now) {
      const int32_t i = next++;
      if (to_lane(i)) {
        st_.lane_scans++;
        plain(i);
      } else {
        pending.push_back(i);
      }
    }
    bool progressed = false;
    if (!batched) {
      if (!pending.empty()) {
        const int32_t i = pending.front();
        pending.pop_front();
        plain(i);
        progressed = true;
      }
    } else {
      // Completions: each query as soon as the device has published it; the
      // slot when its batch has finished.
      for (Slot& s : slots) {
        if (!s.busy) continue;
        for (size_t k = 0; k < s.ids.size(); k++) {
          if (s.finished[k]) continue;
          int32_t ready = 0;
          int st = kg_batch_query_ready(s.b, static_cast<int32_t>(k), &ready);
          if (st != KG_OK) {
            cleanup();
            return st;
          }
          if (!ready) continue;
          const int32_t i = s.ids[k];
          if ((st = kg_batch_collect_query(s.b, static_cast<int32_t>(k), &res[i])) != KG_OK) {
            cleanup();
            return st;
          }
          s.finished[k] = 1;
          s.left--;
          progressed = true;
          if (res[i].status == KG_ERR_UNSUPPORTED) {
            st_.fallbacks++;
            plain(i);
          } else {
            finish(i);
          }
        }
        if (s.left == 0) {
          int32_t done = 0;
          int st = kg_batch_poll(s.b, &done);
          if (st == KG_OK && done) st = kg_batch_release(s.b);
          if (st != KG_OK) {
            cleanup();
            return st;
          }
          if (done) {
            s.busy = false;
            progressed = true;
          }
        }
      }
      // Admission into free slots.
      for (Slot& s : slots) {
        if (s.busy || pending.empty()) continue;
        s.ids.clear();
        s.qbuf.clear();
        const int64_t arena = kg_batch_arena_bytes(s.b);
        int64_t used = 0, rows = 0;
        std::vector<kg::BatchSource> sources;  // build sources already admitted to this batch
        while (!pending.empty() && static_cast<int>(s.ids.size()) < max_batch) {
          const int32_t i = pending.front();
          int64_t fp = 0;
          kg_batch_footprint(&qs[i], &fp);
          if (fp > arena || qs[i].kind == KG_Q_CHAIN) {
            // Never fits a slot, or a kind batches do not run: the plain path.
            pending.pop_front();
            st_.fallbacks++;
            plain(i);
            continue;
          }
          // A build source already in the batch is not counted again.
          kg::BatchSource src{};
          const bool has_src = kg::batch_source_id(qs[i], &src);
          if (has_src && std::find(sources.begin(), sources.end(), src) != sources.end())
            fp = static_cast<int64_t>(kg::batch_own_bytes(qs[i]));
          if (used + fp > arena) break;
          if (has_src) sources.push_back(src);
          const int64_t qrows = qs[i].n + (qs[i].kind == KG_Q_JOIN_AGG ? qs[i].build_n : 0);
          if (opts->max_batch_rows > 0 && !s.ids.empty() && rows + qrows > opts->max_batch_rows) break;
          rows += qrows;
          used += fp;
          pending.pop_front();
          s.ids.push_back(i);
          s.qbuf.push_back(qs[i]);
        }
        if (s.ids.empty()) continue;
        s.finished.assign(s.ids.size(), 0);
        s.left = static_cast<int>(s.ids.size());
        const int st = kg_batch_submit(s.b, s.qbuf.data(), static_cast<int32_t>(s.ids.size()));
        if (st != KG_OK) {
          cleanup();
          return st;
        }
        s.busy = true;
        st_.batches++;
        st_.mean_batch += static_cast<double>(s.ids.size());
        progressed = true;
      }
    }
    if (!progressed) {
      bool busy = false;
      for (const Slot& s : slots) busy |= s.busy;
      if (!busy && pending.empty() && next < nq) {
        // Idle until the next arrival (sleep most of the gap, spin the rest).
        const double gap = arrival_s[next] - seconds_since(t0);
        if (gap > 200e-6) std::this_thread::sleep_for(std::chrono::duration<double>(gap - 100e-6));
      } else {
        std::this_thread::yield();
      }
    }
  }
  cleanup();  // joins the lanes: their writes are visible from here on
  std::vector<double> lat(latency_s, latency_s + nq);
  double last = 0;
  for (double d : done_at) last = std::max(last, d);
  st_.wall_s = nq > 0 ? last - arrival_s[0] : 0;
  st_.qps = st_.wall_s > 0 ? nq / st_.wall_s : 0;
  st_.p50_s = percentile(lat, 0.50);
  st_.p99_s = percentile(lat, 0.99);
  st_.max_s = percentile(lat, 1.0);
  if (st_.batches > 0) st_.mean_batch /= static_cast<double>(st_.batches);
  if (stats) *stats = st_;
  return KG_OK;
}
