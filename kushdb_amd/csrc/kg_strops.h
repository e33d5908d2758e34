// String operations shared by the TEXT kernels, the expression kernel and
// their host-side set-up: byte-wise comparison and the predicates of
// runtime/string.cc:14-82 (StartsWith = rfind(p, 0) == 0, EndsWith, Contains,
// the string_view comparisons) and LIKE as the planner's anchored pattern
// (plan/planner.cc:284-323: '%' matches any sequence, '_' one byte, every other
// byte itself).
#ifndef KG_STROPS_H_
#define KG_STROPS_H_

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kushgpu.h"

namespace kg {

// a < b on unsigned bytes, a proper prefix sorting first; returns -1/0/1.
__host__ __device__ inline int str_cmp(const uint8_t* a, int64_t na, const uint8_t* b, int64_t nb) {
  const int64_t m = na < nb ? na : nb;
  for (int64_t i = 0; i < m; i++) {
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  }
  return na < nb ? -1 : (na > nb ? 1 : 0);
}

// Anchored LIKE: greedy matching with backtracking to the last '%' (linear in
// practice, O(n*m) worst case), equivalent to the anchored regex full match.
__host__ __device__ inline bool like_match(const uint8_t* s, int64_t n, const uint8_t* p, int32_t m) {
  int64_t i = 0, star_i = -1;
  int32_t j = 0, star_j = -1;
  while (i < n) {
    if (j < m && p[j] == '%') {
      star_j = j++;
      star_i = i;
    } else if (j < m && (p[j] == '_' || p[j] == s[i])) {
      i++;
      j++;
    } else if (star_j >= 0) {
      j = star_j + 1;
      i = ++star_i;
    } else {
      return false;
    }
  }
  while (j < m && p[j] == '%') j++;
  return j == m;
}

// `s op p` for a kg_text_op.
__host__ __device__ inline bool text_match(int32_t op, const uint8_t* s, int64_t n, const uint8_t* p, int32_t m) {
  switch (op) {
    case KG_TEXT_STARTS_WITH:
      if (n < m) return false;
      for (int32_t i = 0; i < m; i++)
        if (s[i] != p[i]) return false;
      return true;
    case KG_TEXT_ENDS_WITH:
      if (n < m) return false;
      for (int32_t i = 0; i < m; i++)
        if (s[n - m + i] != p[i]) return false;
      return true;
    case KG_TEXT_CONTAINS: {
      if (m == 0) return true;
      for (int64_t st = 0; st + m <= n; st++) {
        int32_t i = 0;
        while (i < m && s[st + i] == p[i]) i++;
        if (i == m) return true;
      }
      return false;
    }
    case KG_TEXT_LIKE:
      return like_match(s, n, p, m);
    default: {
      const int c = str_cmp(s, n, p, m);
      switch (op) {
        case KG_TEXT_EQ: return c == 0;
        case KG_TEXT_NEQ: return c != 0;
        case KG_TEXT_LT: return c < 0;
        case KG_TEXT_LEQ: return c <= 0;
        case KG_TEXT_GT: return c > 0;
        default: return c >= 0;
      }
    }
  }
}

}  // namespace kg

#endif  // KG_STROPS_H_
