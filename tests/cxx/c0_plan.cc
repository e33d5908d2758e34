// The small star query (1M-row fact filtered to a 30% day range, joined with a
// 10k-row dimension on fk = pk, GROUP BY the dimension's 25 categories with
// SUM(price * Float(qty)) and COUNT(*), ORDER BY category) driven through the
// C ABI the way generated code would drive it: each pipeline's state is
// created in its init hook, used in its body and freed in its reset hook, and
// a predecessor is reset once its last successor has run
// (execution/executable_query.cc:70-95, InitializeOutput / CleanUpPredecessors).
//
//   pipeline 1 (build):  Scan(dim) -> HashJoin build on pk
//   pipeline 2 (probe):  Scan(fact) -> Select(day range) -> HashJoin probe
//                        -> GroupByAggregate(cat)      [successor of 1]
//   pipeline 3 (output): ORDER BY cat over pipeline 2's groups
//
// The plan runs three times with different day ranges on one context; each
// result is checked against the CPU oracle's restatement of the same query
// (kor_q_star): categories and counts exact, sums within 1e-12 relative.
// Exit status 0 and a last line "c0 plan ok" when every run matches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kush_oracle.h"
#include "kushgpu.h"

namespace {

#define CHECK_KG(ctx, call)                                                              \
  do {                                                                                   \
    int st_ = (call);                                                                    \
    if (st_ != KG_OK) {                                                                  \
      std::fprintf(stderr, "%s:%d %s -> %d: %s\n", __FILE__, __LINE__, #call, st_,       \
                   kg_last_error(ctx));                                                  \
      std::exit(1);                                                                      \
    }                                                                                    \
  } while (0)

#define CHECK_HIP(call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      std::fprintf(stderr, "%s:%d %s -> %s\n", __FILE__, __LINE__, #call,               \
                   hipGetErrorString(e_));                                               \
      std::exit(1);                                                                      \
    }                                                                                    \
  } while (0)

constexpr int64_t kFact = 1000000;
constexpr int64_t kDim = 10000;
constexpr int32_t kCats = 25;
constexpr int32_t kDays = 2557;

uint64_t next(uint64_t& s) {  // SplitMix64
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <class T>
T* upload(const std::vector<T>& h) {
  T* d = nullptr;
  CHECK_HIP(hipMalloc(&d, h.size() * sizeof(T)));
  CHECK_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// Resident base columns: loaded into HBM once, like kg_file_column would.
struct Tables {
  std::vector<int32_t> fk, qty, day, pk, cat;
  std::vector<double> price;
  int32_t *d_fk, *d_qty, *d_day, *d_pk, *d_cat;
  double* d_price;
};

Tables make_tables() {
  Tables t;
  uint64_t s = 42;
  t.fk.resize(kFact);
  t.qty.resize(kFact);
  t.day.resize(kFact);
  t.price.resize(kFact);
  for (int64_t i = 0; i < kFact; i++) {
    t.fk[i] = static_cast<int32_t>(next(s) % kDim);
    t.qty[i] = 1 + static_cast<int32_t>(next(s) % 50);
    t.price[i] = 1.0 + static_cast<double>(next(s) >> 11) * 0x1.0p-53 * (1e5 - 1.0);
    t.day[i] = static_cast<int32_t>(next(s) % kDays);
  }
  t.pk.resize(kDim);
  t.cat.resize(kDim);
  for (int64_t i = 0; i < kDim; i++) {
    t.pk[i] = static_cast<int32_t>(i);
    t.cat[i] = static_cast<int32_t>(next(s) % kCats);
  }
  t.d_fk = upload(t.fk);
  t.d_qty = upload(t.qty);
  t.d_day = upload(t.day);
  t.d_price = upload(t.price);
  t.d_pk = upload(t.pk);
  t.d_cat = upload(t.cat);
  return t;
}

// Pipeline state, one struct per pipeline as the generated code keeps globals.
struct BuildPipeline {
  kg_hash_table* ht = nullptr;
};
struct ProbePipeline {
  std::vector<int32_t> cats;
  std::vector<int64_t> vals;  // [group][2]: SUM bits, COUNT
  std::vector<int8_t> nulls;
  int64_t ngroups = 0;
};

void init_build(kg_ctx* ctx, BuildPipeline& p) { CHECK_KG(ctx, kg_ht_create(ctx, kDim, 0.5, &p.ht)); }

void body_build(kg_ctx* ctx, BuildPipeline& p, const Tables& t) {
  const kg_col pk{KG_INT, 0, t.d_pk, nullptr};
  CHECK_KG(ctx, kg_ht_build(ctx, p.ht, &pk, nullptr, kDim));
}

void reset_build(kg_ctx* ctx, BuildPipeline& p) {
  CHECK_KG(ctx, kg_ht_free(ctx, p.ht));
  p.ht = nullptr;
}

void init_probe(ProbePipeline& p) {
  p.cats.assign(64, 0);
  p.vals.assign(64 * 2, 0);
  p.nulls.assign(64 * 2, 0);
  p.ngroups = 0;
}

void body_probe(kg_ctx* ctx, ProbePipeline& p, const BuildPipeline& b, const Tables& t, int32_t lo, int32_t hi) {
  const kg_col cols[4] = {{KG_INT, 0, t.d_fk, nullptr},
                          {KG_INT, 0, t.d_qty, nullptr},
                          {KG_REAL, 0, t.d_price, nullptr},
                          {KG_DATE, 0, t.d_day, nullptr}};
  const kg_pred preds[2] = {{3, KG_GEQ, lo, 0.0}, {3, KG_LT, hi, 0.0}};
  const kg_agg aggs[2] = {{KG_AGG_SUM, 2, 1, 0}, {KG_AGG_COUNT, -1, -1, 0}};
  const kg_col codes{KG_ENUM, 0, t.d_cat, nullptr};
  CHECK_KG(ctx, kg_scan_join_group_agg(ctx, kFact, cols, 4, preds, 2, b.ht, 0, &codes, kCats, aggs, 2, 64,
                                       &p.ngroups, p.cats.data(), p.vals.data(), p.nulls.data()));
}

void reset_probe(ProbePipeline& p) {
  p.cats.clear();
  p.vals.clear();
  p.nulls.clear();
}

// Pipeline 3: ORDER BY cat over the groups (insertion order -> category order).
std::vector<int64_t> body_output(const ProbePipeline& p) {
  std::vector<int64_t> order(p.ngroups);
  for (int64_t i = 0; i < p.ngroups; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return p.cats[a] < p.cats[b]; });
  return order;
}

bool check(const ProbePipeline& p, const std::vector<int64_t>& order, const Tables& t, int32_t lo, int32_t hi) {
  std::vector<int32_t> ocat(kCats);
  std::vector<double> osum(kCats);
  std::vector<int64_t> ocnt(kCats);
  const int64_t ng = kor_q_star(kFact, t.fk.data(), t.qty.data(), t.price.data(), t.day.data(), lo, hi, kDim,
                                t.pk.data(), t.cat.data(), kCats, ocat.data(), osum.data(), ocnt.data());
  if (ng != p.ngroups) {
    std::fprintf(stderr, "days [%d,%d): %lld groups, oracle %lld\n", lo, hi, (long long)p.ngroups, (long long)ng);
    return false;
  }
  std::vector<int64_t> oorder(ng);
  for (int64_t i = 0; i < ng; i++) oorder[i] = i;
  std::sort(oorder.begin(), oorder.end(), [&](int64_t a, int64_t b) { return ocat[a] < ocat[b]; });
  for (int64_t i = 0; i < ng; i++) {
    const int64_t g = order[i], o = oorder[i];
    double sum;
    std::memcpy(&sum, &p.vals[g * 2], 8);
    const int64_t cnt = p.vals[g * 2 + 1];
    const double rel = std::fabs(sum - osum[o]) / std::max(std::fabs(osum[o]), 1e-300);
    if (p.cats[g] != ocat[o] || cnt != ocnt[o] || p.nulls[g * 2] || rel > 1e-12) {
      std::fprintf(stderr, "days [%d,%d) row %lld: cat %d/%d count %lld/%lld sum %.17g/%.17g\n", lo, hi,
                   (long long)i, p.cats[g], ocat[o], (long long)cnt, (long long)ocnt[o], sum, osum[o]);
      return false;
    }
  }
  return true;
}

}  // namespace

int main() {
  kg_ctx* ctx = nullptr;
  if (kg_ctx_create(0, &ctx) != KG_OK) {
    std::fprintf(stderr, "kg_ctx_create failed\n");
    return 1;
  }
  const Tables t = make_tables();
  const int32_t ranges[3][2] = {{0, 767}, {900, 1667}, {1790, 2557}};
  for (const auto& r : ranges) {
    std::printf("c0 plan: days [%d, %d)\n", r[0], r[1]);
    std::fflush(stdout);
    BuildPipeline build;
    ProbePipeline probe;
    // pipeline 1
    init_build(ctx, build);
    body_build(ctx, build, t);
    // pipeline 2; its predecessor (pipeline 1) has no other successor, so it
    // is reset right after.
    init_probe(probe);
    body_probe(ctx, probe, build, t, r[0], r[1]);
    reset_build(ctx, build);
    // pipeline 3
    const std::vector<int64_t> order = body_output(probe);
    const bool ok = check(probe, order, t, r[0], r[1]);
    reset_probe(probe);
    if (!ok) return 1;
    std::printf("c0 plan: %zu groups equal to the oracle\n", order.size());
  }
  CHECK_KG(ctx, kg_ctx_destroy(ctx));
  for (const void* p : {(const void*)t.d_fk, (const void*)t.d_qty, (const void*)t.d_day, (const void*)t.d_price,
                        (const void*)t.d_pk, (const void*)t.d_cat})
    CHECK_HIP(hipFree(const_cast<void*>(p)));
  std::printf("c0 plan ok\n");
  return 0;
}
