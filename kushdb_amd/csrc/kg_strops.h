// String operations shared by the TEXT kernels, the expression kernel and
// their host-side set-up: byte-wise comparison and the predicates of
// runtime/string.cc:14-82 (StartsWith = rfind(p, 0) == 0, EndsWith, Contains,
// the string_view comparisons) and LIKE as re2 evaluates the planner's anchored
// regex (plan/planner.cc:284-323 builds "^...$" with '%' -> ".*", '_' -> ".";
// runtime/string.cc:79-82 calls RE2::FullMatch with default options, so '.' is
// one UTF-8 character other than '\n').
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

// One character of re2's default '.' at s[i] (re2 compiles "." over UTF-8
// text as [^\n]: the single bytes 00-09 and 0B-7F, and for code points
// 80-10FFFF the lead/continuation byte ranges [C2-DF][80-BF],
// [E0-EF][80-BF]{2}, [F0-F4][80-BF]{3}). Returns the character's byte length,
// or 0 when no such character starts at i ('\n', a stray continuation or
// C0/C1/F5-FF byte, a truncated sequence).
__host__ __device__ inline int like_any_char(const uint8_t* s, int64_t i, int64_t n) {
  const uint8_t c = s[i];
  if (c < 0x80) return c == '\n' ? 0 : 1;
  int len;
  if (c >= 0xC2 && c <= 0xDF) len = 2;
  else if (c >= 0xE0 && c <= 0xEF) len = 3;
  else if (c >= 0xF0 && c <= 0xF4) len = 4;
  else return 0;
  if (i + len > n) return 0;
  for (int k = 1; k < len; k++)
    if ((s[i + k] & 0xC0) != 0x80) return 0;
  return len;
}

// Whether re2 accepts a pattern's bytes as UTF-8 (its parser decodes every
// literal character and rejects the regex otherwise): no stray continuation
// bytes, no overlong forms, nothing above U+10FFFF, no truncated sequence.
// A rejected regex matches nothing.
__host__ __device__ inline bool like_pattern_ok(const uint8_t* p, int32_t m) {
  int32_t j = 0;
  while (j < m) {
    const uint8_t c = p[j];
    if (c < 0x80) { j++; continue; }
    int len;
    uint32_t cp;
    if (c >= 0xC0 && c <= 0xDF) { len = 2; cp = c & 0x1F; }
    else if (c >= 0xE0 && c <= 0xEF) { len = 3; cp = c & 0x0F; }
    else if (c >= 0xF0 && c <= 0xF7) { len = 4; cp = c & 0x07; }
    else return false;
    if (j + len > m) return false;
    for (int k = 1; k < len; k++) {
      if ((p[j + k] & 0xC0) != 0x80) return false;
      cp = (cp << 6) | (p[j + k] & 0x3F);
    }
    const uint32_t lo = len == 2 ? 0x80u : (len == 3 ? 0x800u : 0x10000u);
    if (cp < lo || cp > 0x10FFFFu) return false;
    j += len;
  }
  return true;
}

// Anchored LIKE with re2's default semantics for the planner's regex: '%' is
// ".*" and '_' is ".", where '.' is one character as like_any_char defines it
// (never '\n'); every other pattern byte is literal. Greedy matching that
// backtracks only to the last '%', which advances one character at a time
// (linear in practice, O(n*m) worst case). The caller checks the pattern with
// like_pattern_ok once.
__host__ __device__ inline bool like_match(const uint8_t* s, int64_t n, const uint8_t* p, int32_t m) {
  int64_t i = 0, star_i = -1;
  int32_t j = 0, star_j = -1;
  while (i < n) {
    if (j < m && p[j] == '%') {
      star_j = j++;
      star_i = i;
      continue;
    }
    if (j < m && p[j] == '_') {
      const int k = like_any_char(s, i, n);
      if (k > 0) {
        i += k;
        j++;
        continue;
      }
    } else if (j < m && p[j] == s[i]) {
      i++;
      j++;
      continue;
    }
    if (star_j < 0) return false;
    const int k = like_any_char(s, star_i, n);
    if (k == 0) return false;  // '%' cannot take this character, and no earlier '%' can either
    star_i += k;
    i = star_i;
    j = star_j + 1;
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
      return like_pattern_ok(p, m) && like_match(s, n, p, m);
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
