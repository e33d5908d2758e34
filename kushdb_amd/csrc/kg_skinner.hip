// Skinner adaptive multi-way join: UCT over join orders on the host, every
// episode's join work on the device.
//
// The reference (runtime/skinner_join_executor.cc) runs episodes of a
// depth-first nested loop with a budget of 10000 visited tuples
// (:20, budget_per_episode), keeps the interrupted loop's state in a progress
// tree per order prefix (:34-198), and lets a UCT agent pick the next order from
// the reward of the last one (:252-274, :559-838). A depth-first loop of 10000
// tuples is nothing for a GPU, so an episode here is breadth-first: the agent's
// order takes the next slice of its leftmost table's remaining tuples, and the
// slice is joined through every other table with batched index-nested-loop
// expansions (kg_index_join_expand), equality filters for the remaining
// conditions, and an insert of the finished tuples into the device tuple-id
// set (the TupleIdxTable dedup of the valid-tuple handler,
// permutable_skinner_join_translator.cc:842-912). A slice is always finished,
// so the progress tree's resume state for an interrupted inner loop never
// arises: what remains of the reference's JoinState is offset_ (:112-136), the
// last leftmost tuple of each table whose join is complete. A tuple of table t
// at or below offset[t] is in every result tuple already found that contains
// it, so later episodes skip it in every position, as SetResumeProgress's
// progress = max(offset, last completed) does (:384-393). The join is done when
// an episode finishes the last tuple of its leftmost table (IsJoinCompleted,
// :43-51, for a finished loop).
//
// The budget is in partial tuples per episode; the slice of an order is
// sized from the partial tuples per leftmost tuple its last episode produced.
// The reward is the reference's (:252-274) over the episode's start and end
// states. The order comes from a UCT search over join orders (OrderSearch
// below) behind an ask/tell interface, so that episodes run in the host loop.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <set>
#include <utility>
#include <vector>

#include "kg_index.h"

namespace kg {
namespace {

constexpr int kSBlock = 256;
constexpr int kMaxTables = 8;   // tuple-id set width limit
constexpr int kMaxChecks = 16;  // conditions checked after one expansion

// ---------------------------------------------------------------- UCT agent
// Join-order search by UCT (upper confidence bounds applied to trees) as the
// SkinnerDB paper describes it (Trummer et al., SIGMOD 2019, section 4): a
// search tree over order prefixes, UCB1 selection inside the tree, one new
// node per episode, random completion of the order below it, and the
// episode's reward credited along the selected path. The reference runs the
// same method (runtime/skinner_join_executor.cc:559-838); this agent is laid
// out for the host loop around device episodes instead:
//   * nodes live in one flat array; a node holds the bit set of the tables
//     its prefix has joined and its children's indices by table id;
//   * statistics are kept per table id (tries, reward sum), not per action;
//   * selection is a loop from the root, not a recursion; candidates are the
//     unjoined tables a join condition connects to the prefix (all unjoined
//     tables when none is connected); an untried candidate is taken first
//     (uniformly at random among them), then the largest
//     mean + kExploration * sqrt(ln visits / tries), ties broken uniformly at
//     random;
//   * the order is completed by drawing, at each position, uniformly among the
//     connected unjoined tables (all unjoined when none is);
//   * every random choice comes from one SplitMix64 stream per agent.

// Agent seed: a positive seed as given, otherwise the clock (the reference's
// --join_seed default, skinner_join_executor.cc:26-32).
uint64_t join_seed(int64_t seed) {
  if (seed > 0) return static_cast<uint64_t>(seed);
  return static_cast<uint64_t>(std::chrono::system_clock::now().time_since_epoch().count());
}

// SplitMix64 (Steele, Lea, Flood 2014).
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  int below(int n) { return static_cast<int>(next() % static_cast<uint64_t>(n)); }
};

struct JoinGraph {
  int n = 0;
  std::vector<uint64_t> nbr;  // nbr[t]: bit set of the tables a condition joins to t
};

// Tables a step may add after the prefix `joined`: those connected to it, or
// every unjoined table when none is.
inline uint64_t step_candidates(const JoinGraph& g, uint64_t joined) {
  const uint64_t all = (g.n == 64 ? ~0ull : ((1ull << g.n) - 1)) & ~joined;
  uint64_t conn = 0;
  for (uint64_t m = joined; m; m &= m - 1) conn |= g.nbr[__builtin_ctzll(m)];
  conn &= all;
  return conn ? conn : all;
}

inline int pick_bit(uint64_t set, SplitMix& rng) {
  int k = rng.below(__builtin_popcountll(set));
  for (uint64_t m = set;; m &= m - 1)
    if (k-- == 0) return __builtin_ctzll(m);
}

class OrderSearch {
 public:
  static constexpr double kExploration = 1e-5;  // rewards lie in [0, 1]

  OrderSearch(const JoinGraph& g, const std::vector<int>& prefix, uint64_t seed) : g_(g), prefix_(prefix) {
    rng_.s = seed;
    uint64_t joined = 0;
    for (int t : prefix) joined |= 1ull << t;
    add_node(joined);
  }

  // Next order (all tables) and the path of (node, table) steps it took.
  void next(std::vector<int>& order) {
    path_.clear();
    order.assign(prefix_.begin(), prefix_.end());
    int node = 0;
    bool expanded = false;
    uint64_t joined = nodes_[0].joined;
    while (static_cast<int>(order.size()) < g_.n) {
      const int t = choose(node, joined);
      path_.push_back({node, t});
      order.push_back(t);
      joined |= 1ull << t;
      Node& cur = nodes_[node];
      int child = cur.child[t];
      if (child < 0 && !expanded) {  // grow the tree by one node
        child = add_node(joined);
        nodes_[node].child[t] = child;
        expanded = true;
      } else if (child < 0) {
        break;
      }
      node = child;
    }
    // Below the tree: complete the order at random, connected tables first.
    while (static_cast<int>(order.size()) < g_.n) {
      const int t = pick_bit(step_candidates(g_, joined), rng_);
      order.push_back(t);
      joined |= 1ull << t;
    }
  }

  void reward(double r) {
    for (const Step& s : path_) {
      Node& nd = nodes_[s.node];
      nd.visits++;
      nd.tries[s.table]++;
      nd.sum[s.table] += r;
    }
    path_.clear();
  }

 private:
  struct Node {
    uint64_t joined;
    int64_t visits = 0;
    std::vector<int64_t> tries;
    std::vector<double> sum;
    std::vector<int> child;  // by table id, -1 = not in the tree
  };
  struct Step {
    int node, table;
  };

  int add_node(uint64_t joined) {
    Node nd;
    nd.joined = joined;
    nd.tries.assign(g_.n, 0);
    nd.sum.assign(g_.n, 0.0);
    nd.child.assign(g_.n, -1);
    nodes_.push_back(std::move(nd));
    return static_cast<int>(nodes_.size()) - 1;
  }

  int choose(int node, uint64_t joined) {
    const Node& nd = nodes_[node];
    const uint64_t cand = step_candidates(g_, joined);
    uint64_t untried = 0;
    for (uint64_t m = cand; m; m &= m - 1)
      if (nd.tries[__builtin_ctzll(m)] == 0) untried |= m & (~m + 1);
    if (untried) return pick_bit(untried, rng_);
    const double lnv = std::log(static_cast<double>(nd.visits));
    double best = -1.0;
    uint64_t best_set = 0;
    for (uint64_t m = cand; m; m &= m - 1) {
      const int t = __builtin_ctzll(m);
      const double q = nd.sum[t] / nd.tries[t] + kExploration * std::sqrt(lnv / nd.tries[t]);
      if (q > best) {
        best = q;
        best_set = 1ull << t;
      } else if (q == best) {
        best_set |= 1ull << t;
      }
    }
    return pick_bit(best_set, rng_);
  }

  const JoinGraph& g_;
  std::vector<int> prefix_;
  SplitMix rng_;
  std::vector<Node> nodes_;
  std::vector<Step> path_;
};

}  // namespace
}  // namespace kg

// The agent behind the C ABI: kg_uct_next asks for an order, kg_uct_reward
// credits it.
struct kg_uct {
  kg::JoinGraph g;
  std::unique_ptr<kg::OrderSearch> search;
  bool pending = false;
};

namespace kg {
namespace {

// ------------------------------------------------------------ device steps
__global__ void k_sk_iota(int32_t* out, int64_t n, int32_t start) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = start + static_cast<int32_t>(i);
}

// Cross product with the tuple ids first..first+m-1 of a table that no
// condition joins to the tuples so far (the full scan of a table function
// without an index, permutable_skinner_join_translator.cc:663-832).
__global__ void k_sk_cross(const int32_t* in, int64_t n, int32_t width, int32_t first, int64_t m, int32_t* out) {
  const int64_t total = n * m;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int w1 = width + 1;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < total; j += stride) {
    const int64_t i = j / m;
    for (int q = 0; q < width; q++) out[j * w1 + q] = in[i * width + q];
    out[j * w1 + width] = first + static_cast<int32_t>(j - i * m);
  }
}

// One equality condition between two tuple positions: the key of each side is
// its column at the base row of the side's tuple id.
struct EqCheck {
  int32_t pa, pb;
  kg_col a, b;
  const int32_t* sa;  // tuple id -> base row (NULL: identity)
  const int32_t* sb;
};

struct Filter {
  int32_t width;
  int32_t pos;      // position of the table just joined
  int32_t min_id;   // its tuple ids below this are already complete (offset + 1)
  int32_t nchecks;
  EqCheck chk[kMaxChecks];
};

__device__ __forceinline__ bool key_eq(const EqCheck& c, int32_t ta, int32_t tb) {
  const int64_t ra = c.sa ? c.sa[ta] : ta;
  const int64_t rb = c.sb ? c.sb[tb] : tb;
  // A NULL key joins nothing (the index skips NULL keys,
  // permutable_skinner_join_translator.cc:972-994; NULL = x is not TRUE).
  if ((c.a.nulls && c.a.nulls[ra]) || (c.b.nulls && c.b.nulls[rb])) return false;
  if (c.a.type == KG_REAL || c.b.type == KG_REAL) {
    const double x = c.a.type == KG_REAL ? static_cast<const double*>(c.a.data)[ra]
                                         : static_cast<double>(load_int(c.a.type, c.a.data, ra));
    const double y = c.b.type == KG_REAL ? static_cast<const double*>(c.b.data)[rb]
                                         : static_cast<double>(load_int(c.b.type, c.b.data, rb));
    return x == y;
  }
  return load_int(c.a.type, c.a.data, ra) == load_int(c.b.type, c.b.data, rb);
}

__global__ void k_sk_flags(Filter f, const int32_t* t, int64_t n, int32_t* keep) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t* row = t + i * f.width;
    bool ok = row[f.pos] >= f.min_id;
    for (int c = 0; c < f.nchecks && ok; c++) ok = key_eq(f.chk[c], row[f.chk[c].pa], row[f.chk[c].pb]);
    keep[i] = ok ? 1 : 0;
  }
}

__global__ void k_sk_compact(const int32_t* t, int64_t n, int32_t width, const int32_t* keep, const int32_t* pos,
                             int32_t* out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    if (keep[i])
      for (int q = 0; q < width; q++) out[static_cast<int64_t>(pos[i]) * width + q] = t[i * width + q];
}

// Several conditions link the next table to the tables joined so far: for
// each partial tuple, the index bucket of every condition's key (the
// per-predicate GetBucket of permutable_skinner_join_translator.cc:359-433),
// laid out lists[i * k + j] for kg_intersect_batch (the bucket intersection
// of :435-662).
constexpr int kMaxLinks = 8;

struct BucketSrc {
  int32_t k;
  int32_t width;
  int32_t pos[kMaxLinks];            // tuple position of the probing table
  kg_col key[kMaxLinks];             // its key column
  const int32_t* rows[kMaxLinks];    // its tuple id -> row (NULL: identity)
  IndexDir dir[kMaxLinks];
  const int32_t* tids[kMaxLinks];
};

__global__ void k_sk_buckets(BucketSrc b, const int32_t* t, int64_t n, const int32_t** lists, int32_t* sizes) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    for (int j = 0; j < b.k; j++) {
      const int32_t tid = t[i * b.width + b.pos[j]];
      const int64_t r = b.rows[j] ? b.rows[j][tid] : tid;
      int64_t key = 0, off = 0;
      int32_t cnt = 0;
      if (index_key(b.key[j], r, b.dir[j].enc, key)) index_find(b.dir[j], key, off, cnt);
      lists[i * b.k + j] = b.tids[j] + off;
      sizes[i * b.k + j] = cnt;
    }
  }
}

// New partial tuples from the intersections: output j extends tuple i (the
// last i with offs[i] <= j) by ids[j].
__global__ void k_sk_from_lists(const int32_t* t, int64_t n, int32_t width, const int64_t* offs, const int32_t* ids,
                                int64_t total, int32_t* out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int w1 = width + 1;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < total; j += stride) {
    int64_t lo = 0, hi = n;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (offs[mid] <= j) lo = mid;
      else hi = mid;
    }
    for (int q = 0; q < width; q++) out[j * w1 + q] = t[lo * width + q];
    out[j * w1 + width] = ids[j];
  }
}

__global__ void k_sk_fill(int32_t* p, int32_t v, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

struct Perm {
  int32_t p[kMaxTables];
};

// Partial tuples (join order positions) -> result tuples (table order).
__global__ void k_sk_permute(const int32_t* t, int64_t n, int32_t width, Perm perm, int32_t* out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    for (int q = 0; q < width; q++) out[i * width + q] = t[i * width + perm.p[q]];
}

int sgrid(kg_ctx* ctx, int64_t work) {
  const int64_t g = ceil_div(std::max<int64_t>(work, 1), kSBlock * 4);
  return static_cast<int>(std::min<int64_t>(g, static_cast<int64_t>(ctx->num_cus) * 16));
}

// A grow-only device buffer owned by one join call.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int reserve(kg_ctx* ctx, size_t want) {
    if (want <= bytes) return KG_OK;
    KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (p) KG_HIP_CHECK(ctx, hipFree(p));
    p = nullptr;
    bytes = 0;
    const size_t b = std::max<size_t>(want + want / 2, 1 << 16);
    if (hipMalloc(&p, b) != hipSuccess) {
      (void)hipGetLastError();
      return kg_fail(ctx, KG_ERR_OOM, "skinner join: device allocation of %zu bytes failed", b);
    }
    bytes = b;
    return KG_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

bool fixed_width_key(int32_t type) { return type_width(type) > 0; }

// One condition side: table and key column.
struct Side {
  int32_t table;
  kg_col col;
};

struct Episode {
  int64_t slice = 0;        // leftmost tuples joined
  int64_t results = 0;      // valid tuples found (repeats included)
  int64_t partial = 0;      // partial tuples produced by the join steps
  int64_t new_tuples = 0;   // tuples added to the result set
};

class SkinnerJoin {
 public:
  SkinnerJoin(kg_ctx* ctx, int n, const kg_skinner_table* tables, int nconds, const kg_skinner_cond* conds,
              const kg_skinner_opts& o)
      : ctx_(ctx), n_(n), tables_(tables, tables + n), conds_(conds, conds + nconds), opts_(o),
        offset_(n, -1), index_(nconds * 2, nullptr) {}

  ~SkinnerJoin() {
    for (size_t i = 0; i < index_.size(); i++)
      if (index_[i] && !shared_.count(static_cast<int>(i))) kg_index_free(ctx_, index_[i]);
    for (DevBuf* b : {&a_, &b_, &keep_, &pos_, &cub_, &lists_, &sizes_, &nt_, &offs_, &ids_}) b->release();
  }

  // The device column index of one condition side, built on first use over
  // the table's materialized tuples (the reference builds a MemoryColumnIndex
  // per join column of a filtered child up front,
  // permutable_skinner_join_translator.cc:116-132, 934-994; here only the
  // sides some chosen order probes are built). Keys compare as doubles when
  // either side is REAL. Sides with the same table, column and key encoding
  // share one index.
  int Index(int c, int s, kg_index** out) {
    *out = nullptr;
    const Side sd = side(c, s);
    if (index_[c * 2 + s]) {
      *out = index_[c * 2 + s];
      return KG_OK;
    }
    const int enc = IndexEncOf(c);
    for (size_t e = 0; e < index_.size(); e++) {
      const Side o = side(e / 2, e % 2);
      if (index_[e] && o.table == sd.table && o.col.data == sd.col.data && o.col.type == sd.col.type &&
          index_enc(index_[e]) == enc) {
        index_[c * 2 + s] = *out = index_[e];
        shared_.insert(c * 2 + s);
        return KG_OK;
      }
    }
    int st = index_build(ctx_, &sd.col, tables_[sd.table].d_rows, tables_[sd.table].card, enc, &index_[c * 2 + s]);
    if (st != KG_OK) return st;
    *out = index_[c * 2 + s];
    return KG_OK;
  }

  // Runs one episode of `order`; sets *done when the leftmost table is complete.
  int Run(const std::vector<int>& order, kg_tuple_set** set, int64_t* set_cap, int64_t* set_size, Episode* ep,
          bool* done) {
    const int first = order[0];
    const int64_t card0 = tables_[first].card;
    const int64_t start = offset_[first] + 1;
    const int64_t rem = card0 - start;
    const int64_t budget = opts_.budget_per_episode > 0 ? opts_.budget_per_episode : (1 << 20);
    auto it = per_leftmost_.find(order);
    const double per = it != per_leftmost_.end() ? std::max(it->second, 1.0) : static_cast<double>(n_) * 64.0;
    int64_t slice = static_cast<int64_t>(static_cast<double>(budget) / per);
    slice = std::max<int64_t>(1, std::min<int64_t>(slice, rem));
    ep->slice = slice;
    int st;
    if ((st = a_.reserve(ctx_, slice * 4)) != KG_OK) return st;
    hipLaunchKernelGGL(k_sk_iota, dim3(sgrid(ctx_, slice)), dim3(kSBlock), 0, ctx_->stream, a_.as<int32_t>(), slice,
                       static_cast<int32_t>(start));
    KG_LAUNCH_CHECK(ctx_);
    int64_t cur = slice;
    int width = 1;
    std::vector<int> pos_of(n_, -1);
    pos_of[first] = 0;
    int64_t partial = slice;
    for (int p = 1; p < n_ && cur > 0; p++) {
      const int t = order[p];
      // Conditions between t and the tables joined so far.
      std::vector<std::pair<int, int>> link;  // (condition, side of t)
      for (size_t c = 0; c < conds_.size(); c++)
        for (int s = 0; s < 2; s++)
          if (side(c, s).table == t && pos_of[side(c, 1 - s).table] >= 0 && side(c, 1 - s).table != t)
            link.emplace_back(static_cast<int>(c), s);
      int via = -1;
      kg_index* idx = nullptr;
      for (size_t l = 0; l < link.size() && via < 0; l++) {
        if ((st = Index(link[l].first, link[l].second, &idx)) != KG_OK) return st;
        if (idx) via = static_cast<int>(l);
      }
      int64_t next = 0;
      const int32_t min_id = static_cast<int32_t>(offset_[t] + 1);
      const bool intersect = link.size() >= 2 && link.size() <= static_cast<size_t>(kMaxLinks);
      if (intersect) {
        if ((st = Intersect(link, pos_of, width, cur, min_id, &next)) != KG_OK) return st;
      } else if (via >= 0) {
        const int c = link[via].first, s = link[via].second;
        const Side src = side(c, 1 - s);
        for (;;) {
          const int64_t cap = static_cast<int64_t>(b_.bytes / (4 * (width + 1)));
          st = kg_index_join_expand(ctx_, cur, width, a_.as<int32_t>(), pos_of[src.table], &src.col,
                                    tables_[src.table].d_rows, idx, cap, b_.as<int32_t>(), &next);
          if (st == KG_ERR_CAPACITY && next > cap) {
            if (next > INT32_MAX) return kg_fail(ctx_, KG_ERR_CAPACITY, "skinner join: %lld partial tuples in one "
                                                 "step; lower budget_per_episode", (long long)next);
            if ((st = b_.reserve(ctx_, static_cast<size_t>(next) * 4 * (width + 1))) != KG_OK) return st;
            continue;
          }
          if (st != KG_OK) return st;
          break;
        }
      } else {
        const int64_t m = tables_[t].card - min_id;
        next = cur * std::max<int64_t>(m, 0);
        if (next > INT32_MAX) return kg_fail(ctx_, KG_ERR_CAPACITY, "skinner join: cross product of %lld partial "
                                             "tuples; lower budget_per_episode", (long long)next);
        if ((st = b_.reserve(ctx_, static_cast<size_t>(next) * 4 * (width + 1))) != KG_OK) return st;
        if (next > 0)
          hipLaunchKernelGGL(k_sk_cross, dim3(sgrid(ctx_, next)), dim3(kSBlock), 0, ctx_->stream, a_.as<int32_t>(),
                             cur, width, min_id, m, b_.as<int32_t>());
        KG_LAUNCH_CHECK(ctx_);
      }
      width++;
      pos_of[t] = width - 1;
      partial += next;
      // Filter: tuple ids of t already complete, and every other condition
      // between t and the joined tables (the predicate flags of
      // TogglePredicateFlags, skinner_join_executor.cc:395-430).
      Filter f{};
      f.width = width;
      f.pos = pos_of[t];
      f.min_id = (via >= 0 && !intersect) ? min_id : 0;
      for (size_t l = 0; l < link.size() && !intersect; l++) {
        if (static_cast<int>(l) == via) continue;
        if (f.nchecks == kMaxChecks) return kg_fail(ctx_, KG_ERR_UNSUPPORTED, "skinner join: more than %d "
                                                    "conditions close on one table", kMaxChecks);
        const int c = link[l].first, s = link[l].second;
        const Side x = side(c, s), y = side(c, 1 - s);
        EqCheck& e = f.chk[f.nchecks++];
        e.pa = pos_of[x.table];
        e.pb = pos_of[y.table];
        e.a = x.col;
        e.b = y.col;
        e.sa = tables_[x.table].d_rows;
        e.sb = tables_[y.table].d_rows;
      }
      std::swap(a_, b_);
      cur = next;
      if (cur > 0 && (f.nchecks > 0 || f.min_id > 0)) {
        if ((st = Compact(f, &cur)) != KG_OK) return st;
      }
    }
    int64_t results = 0;
    if (width == n_ && cur > 0) {
      results = cur;
      Perm perm{};
      for (int q = 0; q < n_; q++) perm.p[q] = pos_of[q];
      if ((st = b_.reserve(ctx_, static_cast<size_t>(cur) * 4 * n_)) != KG_OK) return st;
      hipLaunchKernelGGL(k_sk_permute, dim3(sgrid(ctx_, cur)), dim3(kSBlock), 0, ctx_->stream, a_.as<int32_t>(), cur,
                         n_, perm, b_.as<int32_t>());
      KG_LAUNCH_CHECK(ctx_);
      if ((st = Grow(set, set_cap, set_size, cur)) != KG_OK) return st;
      int64_t fresh = 0;
      if ((st = kg_tuple_set_insert(ctx_, *set, b_.as<int32_t>(), cur, nullptr, &fresh)) != KG_OK) return st;
      *set_size += fresh;
      ep->new_tuples = fresh;
    }
    ep->results = results;
    ep->partial = partial;
    per_leftmost_[order] = static_cast<double>(partial) / static_cast<double>(slice);
    offset_[first] = start + slice - 1;
    *done = offset_[first] == card0 - 1;
    return KG_OK;
  }

  // The reference's reward (skinner_join_executor.cc:252-274) for an episode
  // that took the leftmost table from `start` - 1 to `start` + slice - 1 and
  // finished every inner loop: initial state = the offsets (no resume state),
  // final state = (last tuple of the slice, then cardinality - 1 for the rest).
  double Reward(const std::vector<int>& order, const std::vector<int64_t>& off_before, const Episode& ep) const {
    const int64_t budget = opts_.budget_per_episode > 0 ? opts_.budget_per_episode : (1 << 20);
    double progress = 0, weight = 1;
    for (int i = 0; i < n_; i++) {
      const int t = order[i];
      const int64_t card = tables_[t].card;
      const int64_t remaining = card - off_before[t];
      weight *= 1.0 / static_cast<double>(remaining);
      const int64_t init = off_before[t];
      const int64_t fin = i == 0 ? off_before[t] + ep.slice : card - 1;
      const int64_t s = std::max(off_before[t], init), e = std::max(off_before[t], fin);
      progress += static_cast<double>(e - s) * weight;
    }
    const double yield = static_cast<double>(ep.results) / static_cast<double>(std::max(budget, ep.partial));
    return 0.5 * progress + 0.5 * yield;
  }

  const std::vector<int64_t>& offsets() const { return offset_; }

 private:
  // Key encoding of condition c's indexes: REAL comparison when either side
  // is REAL, else the integer width of the wider side.
  int IndexEncOf(int c) const {
    const int32_t a = conds_[c].left.type, b = conds_[c].right.type;
    if (a == KG_REAL || b == KG_REAL) return kEncF64;
    return (a == KG_BIGINT || b == KG_BIGINT) ? kEncI64 : kEncI32;
  }

  Side side(size_t c, int s) const {
    return s == 0 ? Side{conds_[c].left_table, conds_[c].left} : Side{conds_[c].right_table, conds_[c].right};
  }

  // Joins table t through every linking condition at once: the buckets of
  // each condition's index for every partial tuple in a_, their sorted
  // intersection from t's first incomplete tuple id on, and the extended
  // tuples into a_'s successor b_ (swapped by the caller).
  int Intersect(const std::vector<std::pair<int, int>>& link, const std::vector<int>& pos_of, int width, int64_t cur,
                int32_t min_id, int64_t* next) {
    const int k = static_cast<int>(link.size());
    BucketSrc bs{};
    bs.k = k;
    bs.width = width;
    int st;
    for (int j = 0; j < k; j++) {
      const int c = link[j].first, s = link[j].second;
      kg_index* idx = nullptr;
      if ((st = Index(c, s, &idx)) != KG_OK) return st;
      const Side src = side(c, 1 - s);
      bs.pos[j] = pos_of[src.table];
      bs.key[j] = src.col;
      bs.rows[j] = tables_[src.table].d_rows;
      index_view(idx, &bs.dir[j], &bs.tids[j]);
    }
    if ((st = lists_.reserve(ctx_, static_cast<size_t>(cur) * k * 8)) != KG_OK ||
        (st = sizes_.reserve(ctx_, static_cast<size_t>(cur) * k * 4)) != KG_OK ||
        (st = nt_.reserve(ctx_, static_cast<size_t>(cur) * 4)) != KG_OK ||
        (st = offs_.reserve(ctx_, static_cast<size_t>(cur + 1) * 8)) != KG_OK)
      return st;
    hipLaunchKernelGGL(k_sk_buckets, dim3(sgrid(ctx_, cur)), dim3(kSBlock), 0, ctx_->stream, bs, a_.as<int32_t>(), cur,
                       lists_.as<const int32_t*>(), sizes_.as<int32_t>());
    hipLaunchKernelGGL(k_sk_fill, dim3(sgrid(ctx_, cur)), dim3(kSBlock), 0, ctx_->stream, nt_.as<int32_t>(), min_id,
                       cur);
    KG_LAUNCH_CHECK(ctx_);
    int64_t total = 0;
    for (;;) {
      const int64_t cap = static_cast<int64_t>(ids_.bytes / 4);
      st = kg_intersect_batch(ctx_, cur, k, lists_.as<const int32_t*>(), sizes_.as<int32_t>(), nt_.as<int32_t>(),
                              offs_.as<int64_t>(), ids_.as<int32_t>(), cap, &total);
      if (st == KG_ERR_CAPACITY && total > cap) {
        if (total > INT32_MAX) return kg_fail(ctx_, KG_ERR_CAPACITY, "skinner join: %lld partial tuples in one "
                                              "step; lower budget_per_episode", (long long)total);
        if ((st = ids_.reserve(ctx_, static_cast<size_t>(total) * 4)) != KG_OK) return st;
        continue;
      }
      if (st != KG_OK) return st;
      break;
    }
    if ((st = b_.reserve(ctx_, static_cast<size_t>(total) * 4 * (width + 1))) != KG_OK) return st;
    if (total > 0)
      hipLaunchKernelGGL(k_sk_from_lists, dim3(sgrid(ctx_, total)), dim3(kSBlock), 0, ctx_->stream, a_.as<int32_t>(),
                         cur, width, offs_.as<int64_t>(), ids_.as<int32_t>(), total, b_.as<int32_t>());
    KG_LAUNCH_CHECK(ctx_);
    *next = total;
    return KG_OK;
  }

  int Compact(const Filter& f, int64_t* n) {
    const int64_t cnt = *n;
    int st;
    if ((st = keep_.reserve(ctx_, (cnt + 1) * 4)) != KG_OK || (st = pos_.reserve(ctx_, (cnt + 1) * 4)) != KG_OK)
      return st;
    size_t tmp = 0;
    KG_HIP_CHECK(ctx_, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, keep_.as<int32_t>(), pos_.as<int32_t>(),
                                                        static_cast<int>(cnt + 1), ctx_->stream));
    if ((st = cub_.reserve(ctx_, tmp)) != KG_OK) return st;
    KG_HIP_CHECK(ctx_, hipMemsetAsync(keep_.as<int32_t>() + cnt, 0, 4, ctx_->stream));
    hipLaunchKernelGGL(k_sk_flags, dim3(sgrid(ctx_, cnt)), dim3(kSBlock), 0, ctx_->stream, f, a_.as<int32_t>(), cnt,
                       keep_.as<int32_t>());
    KG_LAUNCH_CHECK(ctx_);
    size_t tb = cub_.bytes;
    KG_HIP_CHECK(ctx_, hipcub::DeviceScan::ExclusiveSum(cub_.p, tb, keep_.as<int32_t>(), pos_.as<int32_t>(),
                                                        static_cast<int>(cnt + 1), ctx_->stream));
    int32_t total = 0;
    KG_HIP_CHECK(ctx_, hipMemcpyAsync(&total, pos_.as<int32_t>() + cnt, 4, hipMemcpyDeviceToHost, ctx_->stream));
    KG_HIP_CHECK(ctx_, hipStreamSynchronize(ctx_->stream));
    if ((st = b_.reserve(ctx_, static_cast<size_t>(std::max<int32_t>(total, 1)) * 4 * f.width)) != KG_OK) return st;
    hipLaunchKernelGGL(k_sk_compact, dim3(sgrid(ctx_, cnt)), dim3(kSBlock), 0, ctx_->stream, a_.as<int32_t>(), cnt,
                       f.width, keep_.as<int32_t>(), pos_.as<int32_t>(), b_.as<int32_t>());
    KG_LAUNCH_CHECK(ctx_);
    std::swap(a_, b_);
    *n = total;
    return KG_OK;
  }

  // Makes room for `incoming` more tuples: a fuller set is rebuilt at twice
  // the size with its tuples re-inserted.
  int Grow(kg_tuple_set** set, int64_t* cap, int64_t* size, int64_t incoming) {
    if (*size + incoming <= *cap) return KG_OK;
    const int64_t want = std::min<int64_t>(std::max<int64_t>(*cap * 2, *size + incoming), (1ll << 31) - 1);
    if (want < *size + incoming)
      return kg_fail(ctx_, KG_ERR_CAPACITY, "skinner join: more than 2^31 result tuples");
    kg_tuple_set* fresh = nullptr;
    int st = kg_tuple_set_create(ctx_, n_, want, &fresh);
    if (st != KG_OK) return st;
    if (*size > 0) {
      DevBuf tmp;
      if ((st = tmp.reserve(ctx_, static_cast<size_t>(*size) * 4 * n_)) != KG_OK) {
        kg_tuple_set_free(ctx_, fresh);
        return st;
      }
      int64_t got = 0, added = 0;
      st = kg_tuple_set_extract(ctx_, *set, tmp.as<int32_t>(), *size, &got);
      if (st == KG_OK) st = kg_tuple_set_insert(ctx_, fresh, tmp.as<int32_t>(), got, nullptr, &added);
      KG_HIP_CHECK(ctx_, hipStreamSynchronize(ctx_->stream));
      tmp.release();
      if (st != KG_OK) {
        kg_tuple_set_free(ctx_, fresh);
        return st;
      }
    }
    kg_tuple_set_free(ctx_, *set);
    *set = fresh;
    *cap = want;
    return KG_OK;
  }

  kg_ctx* ctx_;
  int n_;
  std::vector<kg_skinner_table> tables_;
  std::vector<kg_skinner_cond> conds_;
  kg_skinner_opts opts_;
  std::vector<int64_t> offset_;
  std::vector<kg_index*> index_;   // per condition side (shared entries listed in shared_)
  std::set<int> shared_;          // sides that reuse an earlier side's index
  std::map<std::vector<int>, double> per_leftmost_;
  DevBuf a_, b_, keep_, pos_, cub_;
  DevBuf lists_, sizes_, nt_, offs_, ids_;  // bucket intersection (Intersect)
};

}  // namespace
}  // namespace kg

using namespace kg;

extern "C" int kg_uct_create(int32_t ntables, const uint8_t* adjacency, int64_t seed, const int32_t* prefix,
                             int32_t nprefix, kg_uct** out) {
  if (!out || ntables < 1 || ntables > 64 || nprefix < 0 || nprefix > ntables || (nprefix && !prefix))
    return KG_ERR_INVALID;
  std::vector<int> pre;
  uint64_t seen = 0;
  for (int i = 0; i < nprefix; i++) {
    if (prefix[i] < 0 || prefix[i] >= ntables || (seen >> prefix[i] & 1)) return KG_ERR_INVALID;
    seen |= 1ull << prefix[i];
    pre.push_back(prefix[i]);
  }
  kg_uct* u = new (std::nothrow) kg_uct();
  if (!u) return KG_ERR_OOM;
  u->g.n = ntables;
  u->g.nbr.assign(ntables, 0);
  if (adjacency)
    for (int i = 0; i < ntables; i++)
      for (int j = 0; j < ntables; j++)
        if (i != j && adjacency[i * ntables + j]) {
          u->g.nbr[i] |= 1ull << j;
          u->g.nbr[j] |= 1ull << i;
        }
  u->search = std::make_unique<OrderSearch>(u->g, pre, join_seed(seed));
  *out = u;
  return KG_OK;
}

extern "C" int kg_uct_next(kg_uct* u, int32_t* order) {
  if (!u || !order || u->pending) return KG_ERR_INVALID;
  std::vector<int> o;
  u->search->next(o);
  for (int i = 0; i < u->g.n; i++) order[i] = o[i];
  u->pending = true;
  return KG_OK;
}

extern "C" int kg_uct_reward(kg_uct* u, double reward) {
  if (!u || !u->pending || !(reward >= 0)) return KG_ERR_INVALID;
  u->search->reward(reward);
  u->pending = false;
  return KG_OK;
}

extern "C" void kg_uct_free(kg_uct* u) { delete u; }

extern "C" int kg_skinner_join(kg_ctx* ctx, int32_t ntables, const kg_skinner_table* tables, int32_t nconds,
                               const kg_skinner_cond* conds, const kg_skinner_opts* opts, kg_tuple_set** out_set,
                               kg_skinner_stats* stats) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, tables && out_set && (nconds == 0 || conds), "null pointer argument");
  KG_REQUIRE(ctx, ntables >= 1 && ntables <= kMaxTables, "1..%d tables (got %d)", kMaxTables, ntables);
  KG_REQUIRE(ctx, nconds >= 0 && nconds <= 64, "0..64 conditions (got %d)", nconds);
  kg_skinner_opts o{};
  if (opts) o = *opts;
  KG_REQUIRE(ctx, o.budget_per_episode >= 0 && o.max_episodes >= 0, "negative budget / episode limit");
  KG_REQUIRE(ctx, o.nprefix >= 0 && o.nprefix <= ntables, "bad prefix length");
  for (int i = 0; i < ntables; i++) {
    KG_REQUIRE(ctx, tables[i].card >= 0 && tables[i].card < INT32_MAX, "table %d: cardinality out of range", i);
  }
  std::vector<uint8_t> adj(static_cast<size_t>(ntables) * ntables, 0);
  for (int c = 0; c < nconds; c++) {
    const kg_skinner_cond& k = conds[c];
    KG_REQUIRE(ctx, k.left_table >= 0 && k.left_table < ntables && k.right_table >= 0 && k.right_table < ntables,
               "condition %d: table out of range", c);
    KG_REQUIRE(ctx, k.left_table != k.right_table, "condition %d joins a table with itself", c);
    KG_REQUIRE(ctx, (k.left.data || tables[k.left_table].card == 0) && (k.right.data || tables[k.right_table].card == 0),
               "condition %d: key column without data", c);
    if (!fixed_width_key(k.left.type) || !fixed_width_key(k.right.type))
      return kg_fail(ctx, KG_ERR_UNSUPPORTED, "condition %d: keys must be fixed-width (not TEXT)", c);
    adj[k.left_table * ntables + k.right_table] = adj[k.right_table * ntables + k.left_table] = 1;
  }
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  kg_skinner_stats st{};
  *out_set = nullptr;
  kg_tuple_set* set = nullptr;
  int64_t cap = 1024, size = 0;
  int rc = kg_tuple_set_create(ctx, ntables, cap, &set);
  if (rc != KG_OK) return rc;
  bool empty = false;
  for (int i = 0; i < ntables; i++) empty |= tables[i].card == 0;  // one input is empty (:858-863)
  if (!empty) {
    SkinnerJoin sj(ctx, ntables, tables, nconds, conds, o);
    kg_uct* agent = nullptr;
    rc = kg_uct_create(ntables, adj.data(), o.seed, o.prefix, o.nprefix, &agent);
    if (rc != KG_OK) rc = kg_fail(ctx, rc, "skinner join: bad join prefix");
    std::set<std::vector<int>> orders;
    std::vector<int32_t> order(ntables);
    bool done = false;
    while (rc == KG_OK && !done) {
      if (o.max_episodes && st.episodes >= o.max_episodes) {
        rc = kg_fail(ctx, KG_ERR_CAPACITY, "skinner join: not complete after %lld episodes",
                     (long long)st.episodes);
        break;
      }
      kg_uct_next(agent, order.data());
      const std::vector<int> ord(order.begin(), order.end());
      const std::vector<int64_t> before = sj.offsets();
      Episode ep;
      rc = sj.Run(ord, &set, &cap, &size, &ep, &done);
      if (rc != KG_OK) break;
      kg_uct_reward(agent, sj.Reward(ord, before, ep));
      orders.insert(ord);
      st.episodes++;
      st.result_tuples += ep.results;
      st.partial_tuples += ep.partial;
      for (int i = 0; i < ntables; i++) st.last_order[i] = order[i];
    }
    st.distinct_orders = static_cast<int64_t>(orders.size());
    kg_uct_free(agent);
  }
  if (rc != KG_OK) {
    kg_tuple_set_free(ctx, set);
    return rc;
  }
  st.distinct_tuples = size;
  if (stats) *stats = st;
  *out_set = set;
  return KG_OK;
}
