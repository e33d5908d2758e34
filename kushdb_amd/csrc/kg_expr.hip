// General expressions on the device: the ExpressionTranslator's node kinds
// (compile/translators/expression_translator.cc:43-360) with the evaluation
// rules of compile/proxy/evaluate.cc:11-338, and the Select operator's filter
// (compile/translators/select_translator.cc:23-37).
//
// Evaluation is fused per row: a workgroup takes a tile of rows, one per
// thread, and evaluates the nodes in order; every node's value (8 bytes:
// integers sign-extended, DATE as its Julian day, BOOLEAN 0/1, ENUM code,
// REAL bits, TEXT the row of its column) and NULL flag goes to LDS, laid out
// node-major so that a node's values for the tile are contiguous and a thread
// only ever touches its own row. Only the referenced columns are read from
// HBM and only the root is written, so an expression costs its input bytes
// plus its output bytes per row.
//
// Operators that compare the decoded strings of an ENUM (evaluate.cc:222-300,
// expression_translator.cc:275-296, 328-360) are reduced on the host, once per
// call, to a table over the dictionary's codes: the truth value for a literal
// operand, or the string rank of each code when two ENUM values meet.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "kg_internal.h"
#include "kg_strops.h"

using namespace kg;

namespace {

constexpr int kXMaxNodes = 64;
constexpr int kXMaxCols = 64;
constexpr int kXLdsBytes = 65536;

struct XNode {
  int32_t op, type, a, b, c, flags;
  int32_t otype;   // operand type (child a's type) of binary and unary nodes
  int32_t slen;    // literal bytes (TEXT literal, LIKE pattern)
  int64_t ilit;
  double flit;
  int64_t soff;    // literal offset in the blob
  int64_t toff;    // ENUM table offset (int32 entries); -1: none
  int32_t tcount;  // table entries
  int32_t tmode;   // 1: truth by code of a; 2: compare the ranks of a and b
  int32_t scol;    // TEXT operands: column of a (or -1: a literal)
  int32_t scol_b;  // TEXT operands: column of b (or -1: a literal)
};

struct XCol {
  int32_t type;
  const void* data;
  const int8_t* nulls;
  const int64_t* offsets;  // TEXT
  const uint8_t* bytes;    // TEXT
};

struct XProg {
  int32_t nn, ncols;
  XNode nodes[kXMaxNodes];
  XCol cols[kXMaxCols];
};

__device__ __forceinline__ double dbl(uint64_t v) { return __longlong_as_double(static_cast<long long>(v)); }
__device__ __forceinline__ uint64_t bits(double d) { return static_cast<uint64_t>(__double_as_longlong(d)); }

// Year of a Julian day number (proleptic Gregorian; the inverse of
// Date::BuildDate, runtime/date.cc:10-25, as Date::ExtractYear returns it).
__device__ __forceinline__ int64_t jd_year(int64_t jd) {
  const int64_t a = jd + 32044;
  const int64_t b = (4 * a + 3) / 146097;
  const int64_t c = a - 146097 * b / 4;
  const int64_t e = (4 * c + 3) / 1461;
  const int64_t f = c - 1461 * e / 4;
  const int64_t g = (5 * f + 2) / 153;
  return 100 * b + e - 4800 + g / 10;
}

__device__ __forceinline__ uint64_t load_col(const XCol& c, int64_t r, bool& null) {
  null = c.nulls && c.nulls[r];
  switch (c.type) {
    case KG_SMALLINT: return static_cast<uint64_t>(static_cast<int64_t>(static_cast<const int16_t*>(c.data)[r]));
    case KG_INT:
    case KG_DATE:
    case KG_ENUM: return static_cast<uint64_t>(static_cast<int64_t>(static_cast<const int32_t*>(c.data)[r]));
    case KG_BIGINT:
    case KG_REAL: return static_cast<const uint64_t*>(c.data)[r];
    case KG_BOOLEAN: return static_cast<const int8_t*>(c.data)[r] ? 1 : 0;
    default: return static_cast<uint64_t>(r);  // TEXT: the row
  }
}

// Integer arithmetic wraps in the operand type (two's complement).
__device__ __forceinline__ uint64_t arith(int op, int32_t t, uint64_t a, uint64_t b) {
  if (t == KG_REAL) {
    const double x = dbl(a), y = dbl(b);
    switch (op) {
      case KG_E_ADD: return bits(x + y);
      case KG_E_SUB: return bits(x - y);
      case KG_E_MUL: return bits(x * y);
      default: return bits(x / y);
    }
  }
  uint64_t r;
  switch (op) {
    case KG_E_ADD: r = a + b; break;
    case KG_E_SUB: r = a - b; break;
    default: r = a * b; break;
  }
  if (t == KG_SMALLINT) return static_cast<uint64_t>(static_cast<int64_t>(static_cast<int16_t>(r)));
  if (t == KG_INT) return static_cast<uint64_t>(static_cast<int64_t>(static_cast<int32_t>(r)));
  return r;
}

__device__ __forceinline__ bool compare(int op, int32_t t, uint64_t a, uint64_t b) {
  if (t == KG_REAL) {
    const double x = dbl(a), y = dbl(b);
    switch (op) {
      case KG_E_EQ: return x == y;
      case KG_E_NEQ: return x != y;
      case KG_E_LT: return x < y;
      case KG_E_LEQ: return x <= y;
      case KG_E_GT: return x > y;
      default: return x >= y;
    }
  }
  const int64_t x = static_cast<int64_t>(a), y = static_cast<int64_t>(b);
  switch (op) {
    case KG_E_EQ: return x == y;
    case KG_E_NEQ: return x != y;
    case KG_E_LT: return x < y;
    case KG_E_LEQ: return x <= y;
    case KG_E_GT: return x > y;
    default: return x >= y;
  }
}

__device__ __forceinline__ int text_op_of(int op) {
  switch (op) {
    case KG_E_STARTS_WITH: return KG_TEXT_STARTS_WITH;
    case KG_E_ENDS_WITH: return KG_TEXT_ENDS_WITH;
    case KG_E_CONTAINS: return KG_TEXT_CONTAINS;
    case KG_E_LIKE: return KG_TEXT_LIKE;
    default: return op - KG_E_EQ;  // EQ .. GEQ
  }
}

// The bytes of TEXT operand `v` (the row of column `col`, or the literal).
__device__ __forceinline__ void text_of(const XProg& P, const uint8_t* blob, int col, const XNode& lit, uint64_t v,
                                        const uint8_t*& s, int64_t& len) {
  if (col >= 0) {
    const XCol& c = P.cols[col];
    const int64_t r = static_cast<int64_t>(v);
    s = c.bytes + c.offsets[r];
    len = c.offsets[r + 1] - c.offsets[r];
  } else {
    s = blob + lit.soff;
    len = lit.slen;
  }
}

// mode 0: write the root (typed) and its NULL flags; mode 1: write flag
// bytes, 1 where the BOOLEAN root is TRUE and not NULL.
__global__ void k_expr(const XProg* __restrict__ P, const uint8_t* __restrict__ blob,
                       const int32_t* __restrict__ tabs, const int32_t* __restrict__ rows, int64_t n, int mode,
                       void* out, int8_t* out_nulls, int32_t out_type) {
  extern __shared__ uint64_t xs[];
  const int T = blockDim.x;
  const int nn = P->nn;
  uint64_t* val = xs;
  int8_t* nul = reinterpret_cast<int8_t*>(xs + static_cast<size_t>(nn) * T);
  const int tid = threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * T;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * T + tid; i - tid < n; i += stride) {
    const bool valid = i < n;
    const int64_t r = valid ? (rows ? rows[i] : i) : 0;
    for (int k = 0; k < nn; k++) {
      const XNode& N = P->nodes[k];
      uint64_t x = 0;
      bool isnull = false;
      if (valid) {
        const uint64_t av = N.a >= 0 && N.op != KG_E_COL ? val[N.a * T + tid] : 0;
        const bool an = N.a >= 0 && N.op != KG_E_COL ? nul[N.a * T + tid] != 0 : false;
        const uint64_t bv = N.b >= 0 ? val[N.b * T + tid] : 0;
        const bool bn = N.b >= 0 ? nul[N.b * T + tid] != 0 : false;
        switch (N.op) {
          case KG_E_COL: x = load_col(P->cols[N.a], r, isnull); break;
          case KG_E_LIT:
            x = N.type == KG_REAL ? bits(N.flit) : static_cast<uint64_t>(N.ilit);
            isnull = (N.flags & KG_E_NULL) != 0;
            break;
          case KG_E_ADD:
          case KG_E_SUB:
          case KG_E_MUL:
          case KG_E_DIV:
            isnull = an || bn;
            x = isnull ? av : arith(N.op, N.otype, av, bv);
            break;
          case KG_E_EQ:
          case KG_E_NEQ:
          case KG_E_LT:
          case KG_E_LEQ:
          case KG_E_GT:
          case KG_E_GEQ:
          case KG_E_STARTS_WITH:
          case KG_E_ENDS_WITH:
          case KG_E_CONTAINS:
            isnull = an || bn;
            if (!isnull) {
              const uint64_t tc = static_cast<uint64_t>(N.tcount);
              if (N.tmode == 1) {  // codes outside the dictionary compare false
                x = (av < tc && tabs[N.toff + av]) ? 1 : 0;
              } else if (N.tmode == 2) {
                x = (av < tc && bv < tc &&
                     compare(N.op, KG_INT, static_cast<uint64_t>(static_cast<int64_t>(tabs[N.toff + av])),
                             static_cast<uint64_t>(static_cast<int64_t>(tabs[N.toff + bv]))))
                        ? 1
                        : 0;
              } else if (N.otype == KG_TEXT) {
                const uint8_t *sa, *sb;
                int64_t la, lb;
                text_of(*P, blob, N.scol, P->nodes[N.a], av, sa, la);
                text_of(*P, blob, N.scol_b, P->nodes[N.b], bv, sb, lb);
                x = text_match(text_op_of(N.op), sa, la, sb, static_cast<int32_t>(lb)) ? 1 : 0;
              } else {
                x = compare(N.op, N.otype, av, bv) ? 1 : 0;
              }
            }
            break;
          case KG_E_AND:
            if (an) {
              isnull = true;
            } else if (av) {
              x = bv;
              isnull = bn;
            }
            break;
          case KG_E_OR:
            if (an || !av) {
              x = bv;
              isnull = bn;
            } else {
              x = 1;
            }
            break;
          case KG_E_NOT:
            isnull = an;
            x = an ? 0 : (av ? 0 : 1);
            break;
          case KG_E_IS_NULL: x = an ? 1 : 0; break;
          case KG_E_CASE: {
            const int pick = (!an && av) ? N.b : N.c;
            x = val[pick * T + tid];
            isnull = nul[pick * T + tid] != 0;
            break;
          }
          case KG_E_INT_TO_REAL: x = an ? bits(0.0) : bits(static_cast<double>(static_cast<int64_t>(av))); break;
          case KG_E_EXTRACT_YEAR:
            isnull = an;
            x = an ? 0 : static_cast<uint64_t>(jd_year(static_cast<int64_t>(av)));
            break;
          case KG_E_LIKE:
            if (!an) {
              if (N.tmode == 1) {
                x = (av < static_cast<uint64_t>(N.tcount) && tabs[N.toff + av]) ? 1 : 0;
              } else {
                const uint8_t* sa;
                int64_t la;
                text_of(*P, blob, N.scol, N, av, sa, la);
                x = like_match(sa, la, blob + N.soff, N.slen) ? 1 : 0;
              }
            }
            break;
          default:  // KG_E_IN
            x = (!an && av < static_cast<uint64_t>(N.tcount) && tabs[N.toff + av]) ? 1 : 0;
            break;
        }
      }
      val[k * T + tid] = x;
      nul[k * T + tid] = isnull ? 1 : 0;
    }
    if (!valid) continue;
    const uint64_t x = val[(nn - 1) * T + tid];
    const bool isnull = nul[(nn - 1) * T + tid] != 0;
    if (mode == 1) {
      static_cast<uint8_t*>(out)[i] = (!isnull && x) ? 1 : 0;
      continue;
    }
    out_nulls[i] = isnull ? 1 : 0;
    switch (out_type) {
      case KG_SMALLINT: static_cast<int16_t*>(out)[i] = static_cast<int16_t>(x); break;
      case KG_INT:
      case KG_DATE:
      case KG_ENUM: static_cast<int32_t*>(out)[i] = static_cast<int32_t>(x); break;
      case KG_BOOLEAN: static_cast<int8_t*>(out)[i] = x ? 1 : 0; break;
      default: static_cast<uint64_t*>(out)[i] = x; break;
    }
  }
}

bool is_int(int32_t t) { return t == KG_SMALLINT || t == KG_INT || t == KG_BIGINT; }

std::string dict_str(const kg_enum_dict& d, int64_t code) {
  return std::string(d.bytes + d.offsets[code], static_cast<size_t>(d.offsets[code + 1] - d.offsets[code]));
}

int host_text_op(int op) {
  switch (op) {
    case KG_E_STARTS_WITH: return KG_TEXT_STARTS_WITH;
    case KG_E_ENDS_WITH: return KG_TEXT_ENDS_WITH;
    case KG_E_CONTAINS: return KG_TEXT_CONTAINS;
    case KG_E_LIKE: return KG_TEXT_LIKE;
    default: return op - KG_E_EQ;
  }
}

// Checks the nodes against the plan's type rules and builds the device
// program, literal blob and ENUM tables (host side).
int plan_expr(kg_ctx* ctx, const kg_col* cols, const kg_text_col* text_cols, const kg_enum_dict* dicts,
              int32_t ncols, const kg_expr_node* nodes, int32_t nn, XProg& P, std::vector<uint8_t>& blob,
              std::vector<int32_t>& tabs) {
  KG_REQUIRE(ctx, nn >= 1 && nn <= kXMaxNodes && nodes, "an expression takes 1..%d nodes", kXMaxNodes);
  KG_REQUIRE(ctx, ncols >= 0 && ncols <= kXMaxCols && (ncols == 0 || cols), "at most %d columns", kXMaxCols);
  std::memset(&P, 0, sizeof(P));
  P.nn = nn;
  P.ncols = ncols;
  for (int i = 0; i < ncols; i++) {
    XCol& c = P.cols[i];
    c.type = cols[i].type;
    if (c.type == KG_TEXT) {
      KG_REQUIRE(ctx, text_cols && text_cols[i].offsets, "TEXT column %d needs its text column", i);
      c.offsets = text_cols[i].offsets;
      c.bytes = text_cols[i].bytes;
      c.nulls = text_cols[i].nulls;
    } else {
      KG_REQUIRE(ctx, type_width(c.type) > 0, "column %d: bad type %d", i, c.type);
      c.data = cols[i].data;
      c.nulls = cols[i].nulls;
    }
  }
  auto col_of = [&](int node) { return nodes[node].op == KG_E_COL ? nodes[node].a : -1; };
  auto dict_of = [&](int node) -> const kg_enum_dict* {
    // The dictionary of an ENUM value: its column's (a literal code takes the
    // other operand's).
    const int c = col_of(node);
    if (c < 0 || !dicts || !dicts[c].offsets) return nullptr;
    return &dicts[c];
  };
  auto add_lit = [&](const kg_expr_node& N, XNode& X) {
    X.soff = static_cast<int64_t>(blob.size());
    X.slen = N.str_len;
    if (N.str_len > 0) blob.insert(blob.end(), N.str, N.str + N.str_len);
  };
  for (int k = 0; k < nn; k++) {
    const kg_expr_node& N = nodes[k];
    XNode& X = P.nodes[k];
    X.op = N.op;
    X.type = N.type;
    X.a = N.a;
    X.b = N.b;
    X.c = N.c;
    X.flags = N.flags;
    X.ilit = N.ilit;
    X.flit = N.flit;
    X.toff = -1;
    X.scol = X.scol_b = -1;
    auto child = [&](int c) { return c >= 0 && c < k; };
    const int op = N.op;
    if (op == KG_E_COL) {
      KG_REQUIRE(ctx, N.a >= 0 && N.a < ncols, "node %d: column %d out of range", k, N.a);
      KG_REQUIRE(ctx, cols[N.a].type == N.type, "node %d: column type %d, node type %d", k, cols[N.a].type, N.type);
      X.b = X.c = -1;
      continue;
    }
    if (op == KG_E_LIT) {
      KG_REQUIRE(ctx, N.type == KG_TEXT || type_width(N.type) > 0, "node %d: bad literal type", k);
      KG_REQUIRE(ctx, N.type != KG_TEXT || N.str_len == 0 || N.str, "node %d: TEXT literal without bytes", k);
      if (N.type == KG_TEXT) add_lit(N, X);
      X.a = X.b = X.c = -1;
      continue;
    }
    KG_REQUIRE(ctx, child(N.a), "node %d: child a must precede it", k);
    const int32_t ta = nodes[N.a].type;
    X.otype = ta;
    const bool binary = (op >= KG_E_ADD && op <= KG_E_CONTAINS);
    if (binary) {
      KG_REQUIRE(ctx, child(N.b), "node %d: child b must precede it", k);
    } else if (op != KG_E_CASE) {
      X.b = -1;
    }
    if (op != KG_E_CASE) X.c = -1;
    const int32_t tb = binary ? nodes[N.b].type : -1;
    switch (op) {
      case KG_E_ADD:
      case KG_E_SUB:
      case KG_E_MUL:
        KG_REQUIRE(ctx, ta == tb && (is_int(ta) || ta == KG_REAL) && N.type == ta,
                   "node %d: arithmetic needs equal numeric operand types and keeps them", k);
        break;
      case KG_E_DIV:
        KG_REQUIRE(ctx, ta == KG_REAL && tb == KG_REAL && N.type == KG_REAL, "node %d: DIV takes REAL", k);
        break;
      case KG_E_AND:
      case KG_E_OR:
        KG_REQUIRE(ctx, ta == KG_BOOLEAN && tb == KG_BOOLEAN && N.type == KG_BOOLEAN, "node %d: AND/OR take BOOLEAN", k);
        break;
      case KG_E_EQ:
      case KG_E_NEQ:
      case KG_E_LT:
      case KG_E_LEQ:
      case KG_E_GT:
      case KG_E_GEQ:
      case KG_E_STARTS_WITH:
      case KG_E_ENDS_WITH:
      case KG_E_CONTAINS: {
        KG_REQUIRE(ctx, N.type == KG_BOOLEAN, "node %d: comparisons are BOOLEAN", k);
        const bool strop = op >= KG_E_STARTS_WITH;
        if (ta == KG_ENUM) {
          const bool eqne = op == KG_E_EQ || op == KG_E_NEQ;
          KG_REQUIRE(ctx, tb == KG_ENUM || (!eqne && tb == KG_TEXT), "node %d: ENUM meets ENUM (or TEXT)", k);
          if (eqne) break;  // codes compare directly
          const kg_enum_dict* d = dict_of(N.a);
          KG_REQUIRE(ctx, d, "node %d: ENUM string comparison needs the column's dictionary", k);
          if (tb == KG_TEXT) {
            const kg_expr_node& L = nodes[N.b];
            KG_REQUIRE(ctx, L.op == KG_E_LIT, "node %d: ENUM against TEXT takes a literal", k);
            X.tmode = 1;
            X.toff = static_cast<int64_t>(tabs.size());
            X.tcount = d->count;
            for (int c = 0; c < d->count; c++) {
              const std::string s = dict_str(*d, c);
              tabs.push_back(text_match(host_text_op(op), reinterpret_cast<const uint8_t*>(s.data()),
                                        static_cast<int64_t>(s.size()), reinterpret_cast<const uint8_t*>(L.str),
                                        L.str_len)
                                 ? 1
                                 : 0);
            }
          } else {
            KG_REQUIRE(ctx, !strop, "node %d: string predicates on ENUM take a TEXT literal", k);
            // Ranks of the codes in string order; both values share the dictionary.
            std::vector<int32_t> idx(static_cast<size_t>(d->count));
            for (int c = 0; c < d->count; c++) idx[c] = c;
            std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return dict_str(*d, x) < dict_str(*d, y); });
            X.tmode = 2;
            X.toff = static_cast<int64_t>(tabs.size());
            X.tcount = d->count;
            tabs.resize(tabs.size() + d->count);
            int rank = 0;
            for (int j = 0; j < d->count; j++) {
              if (j > 0 && dict_str(*d, idx[j]) != dict_str(*d, idx[j - 1])) rank = j;
              tabs[X.toff + idx[j]] = rank;
            }
          }
          break;
        }
        KG_REQUIRE(ctx, ta == tb, "node %d: comparison of types %d and %d", k, ta, tb);
        if (ta == KG_TEXT) {
          X.scol = col_of(N.a);
          X.scol_b = col_of(N.b);
          KG_REQUIRE(ctx, (X.scol >= 0 || nodes[N.a].op == KG_E_LIT) && (X.scol_b >= 0 || nodes[N.b].op == KG_E_LIT),
                     "node %d: TEXT operands are columns or literals", k);
          break;
        }
        KG_REQUIRE(ctx, !strop, "node %d: string predicates take TEXT or ENUM", k);
        KG_REQUIRE(ctx, ta != KG_BOOLEAN || op == KG_E_EQ || op == KG_E_NEQ, "node %d: BOOLEAN compares by = and <>", k);
        break;
      }
      case KG_E_NOT:
        KG_REQUIRE(ctx, ta == KG_BOOLEAN && N.type == KG_BOOLEAN, "node %d: NOT takes BOOLEAN", k);
        break;
      case KG_E_IS_NULL:
        KG_REQUIRE(ctx, N.type == KG_BOOLEAN, "node %d: IS NULL is BOOLEAN", k);
        break;
      case KG_E_CASE:
        KG_REQUIRE(ctx, child(N.b) && child(N.c), "node %d: CASE children must precede it", k);
        KG_REQUIRE(ctx, ta == KG_BOOLEAN, "node %d: CASE condition is BOOLEAN", k);
        KG_REQUIRE(ctx, nodes[N.b].type == N.type && nodes[N.c].type == N.type, "node %d: CASE branches of its type", k);
        if (N.type == KG_TEXT) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "node %d: CASE over TEXT", k);
        break;
      case KG_E_INT_TO_REAL:
        KG_REQUIRE(ctx, is_int(ta) && N.type == KG_REAL, "node %d: conversion of an integer to REAL", k);
        break;
      case KG_E_EXTRACT_YEAR:
        KG_REQUIRE(ctx, ta == KG_DATE && N.type == KG_INT, "node %d: EXTRACT(YEAR) of a DATE is INT", k);
        break;
      case KG_E_LIKE:
        KG_REQUIRE(ctx, N.type == KG_BOOLEAN && (ta == KG_TEXT || ta == KG_ENUM), "node %d: LIKE on TEXT or ENUM", k);
        KG_REQUIRE(ctx, N.str_len == 0 || N.str, "node %d: LIKE without a pattern", k);
        if (ta == KG_ENUM) {
          const kg_enum_dict* d = dict_of(N.a);
          KG_REQUIRE(ctx, d, "node %d: LIKE on ENUM needs the column's dictionary", k);
          X.tmode = 1;
          X.toff = static_cast<int64_t>(tabs.size());
          X.tcount = d->count;
          for (int c = 0; c < d->count; c++) {
            const std::string s = dict_str(*d, c);
            tabs.push_back(text_match(KG_TEXT_LIKE, reinterpret_cast<const uint8_t*>(s.data()),
                                      static_cast<int64_t>(s.size()), reinterpret_cast<const uint8_t*>(N.str),
                                      N.str_len)
                               ? 1
                               : 0);
          }
        } else if (!like_pattern_ok(reinterpret_cast<const uint8_t*>(N.str), N.str_len)) {
          // re2 rejects a pattern that is not UTF-8, and the rejected regex
          // matches nothing: an empty code table gives FALSE for every row.
          X.tmode = 1;
          X.toff = static_cast<int64_t>(tabs.size());
          X.tcount = 0;
        } else {
          X.scol = col_of(N.a);
          KG_REQUIRE(ctx, X.scol >= 0, "node %d: LIKE on a TEXT column", k);
          add_lit(N, X);
        }
        break;
      case KG_E_IN: {
        KG_REQUIRE(ctx, ta == KG_ENUM && N.type == KG_BOOLEAN, "node %d: IN takes an ENUM", k);
        KG_REQUIRE(ctx, N.nvalues >= 0 && (N.nvalues == 0 || N.values), "node %d: IN without values", k);
        int32_t mx = 0;
        for (int j = 0; j < N.nvalues; j++) mx = std::max(mx, N.values[j] + 1);
        X.tmode = 1;
        X.toff = static_cast<int64_t>(tabs.size());
        X.tcount = mx;
        tabs.resize(tabs.size() + mx, 0);
        for (int j = 0; j < N.nvalues; j++)
          if (N.values[j] >= 0) tabs[X.toff + N.values[j]] = 1;
        break;
      }
      default:
        return kg_fail(ctx, KG_ERR_INVALID, "node %d: unknown operator %d", k, op);
    }
  }
  return KG_OK;
}

int run_expr(kg_ctx* ctx, int64_t n, const kg_col* cols, const kg_text_col* text_cols, const kg_enum_dict* dicts,
             int32_t ncols, const kg_expr_node* nodes, int32_t nnodes, const int32_t* d_rows, int mode, void* out,
             int8_t* out_nulls, int32_t out_type, size_t extra, char** extra_ptr) {
  XProg* P = new (std::nothrow) XProg;
  if (!P) return kg_fail(ctx, KG_ERR_OOM, "host allocation failed");
  std::vector<uint8_t> blob;
  std::vector<int32_t> tabs;
  int st = plan_expr(ctx, cols, text_cols, dicts, ncols, nodes, nnodes, *P, blob, tabs);
  if (st != KG_OK) {
    delete P;
    return st;
  }
  if (mode == 0) {
    const int32_t rt = nodes[nnodes - 1].type;
    if (rt != out_type || rt == KG_TEXT) {
      delete P;
      return kg_fail(ctx, KG_ERR_INVALID, "root type %d cannot be written to a column of type %d", rt, out_type);
    }
  } else if (nodes[nnodes - 1].type != KG_BOOLEAN) {
    delete P;
    return kg_fail(ctx, KG_ERR_INVALID, "a filter needs a BOOLEAN expression");
  }
  auto r256 = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t pb = r256(sizeof(XProg)), bb = r256(std::max<size_t>(blob.size(), 1)),
               tb = r256(std::max<size_t>(tabs.size(), 1) * 4);
  st = kg_work_reserve(ctx, pb + bb + tb + extra);
  if (st != KG_OK) {
    delete P;
    return st;
  }
  char* w = static_cast<char*>(ctx->work);
  hipError_t e = hipMemcpyAsync(w, P, sizeof(XProg), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && !blob.empty())
    e = hipMemcpyAsync(w + pb, blob.data(), blob.size(), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && !tabs.empty())
    e = hipMemcpyAsync(w + pb + bb, tabs.data(), tabs.size() * 4, hipMemcpyHostToDevice, ctx->stream);
  // The host staging buffers must live until the copies have run.
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  delete P;
  KG_HIP_CHECK(ctx, e);
  if (extra_ptr) *extra_ptr = w + pb + bb + tb;
  if (mode == 1 && extra_ptr) out = *extra_ptr;  // flags live in the extra area
  int T = 256;
  while (T > kWave && static_cast<size_t>(nnodes) * T * 9 > static_cast<size_t>(kXLdsBytes)) T -= kWave;
  const size_t lds = static_cast<size_t>(nnodes) * T * 9;
  const int64_t tiles = (n + T - 1) / T;
  const int grid = static_cast<int>(std::min<int64_t>(std::max<int64_t>(tiles, 1), int64_t{ctx->num_cus} * 16));
  if (n > 0) {
    kg_time_begin(ctx);
    hipLaunchKernelGGL(k_expr, dim3(grid), dim3(T), lds, ctx->stream, reinterpret_cast<const XProg*>(w),
                       reinterpret_cast<const uint8_t*>(w + pb), reinterpret_cast<const int32_t*>(w + pb + bb), d_rows,
                       n, mode, out, out_nulls, out_type);
    KG_LAUNCH_CHECK(ctx);
    kg_time_end(ctx);
  }
  return KG_OK;
}

}  // namespace

extern "C" int kg_expr_eval(kg_ctx* ctx, int64_t n, const kg_col* cols, const kg_text_col* text_cols,
                            const kg_enum_dict* dicts, int32_t ncols, const kg_expr_node* nodes, int32_t nnodes,
                            const int32_t* d_rows, kg_col* out) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, n >= 0 && out && (n == 0 || (out->data && out->nulls)), "output column needs data and NULL flags");
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const int st = run_expr(ctx, n, cols, text_cols, dicts, ncols, nodes, nnodes, d_rows, 0, const_cast<void*>(out->data),
                          const_cast<int8_t*>(out->nulls), out->type, 0, nullptr);
  if (st != KG_OK) return st;
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return KG_OK;
}

extern "C" int kg_expr_select(kg_ctx* ctx, int64_t n, const kg_col* cols, const kg_text_col* text_cols,
                              const kg_enum_dict* dicts, int32_t ncols, const kg_expr_node* nodes, int32_t nnodes,
                              const int32_t* d_rows, int32_t* d_out_rows, int64_t* out_count) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, n >= 0 && n < INT32_MAX && out_count && (n == 0 || d_out_rows), "bad arguments");
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  *out_count = 0;
  const int ni = static_cast<int>(n);
  size_t cub_bytes = 0;
  hipcub::CountingInputIterator<int32_t> iota(0);
  if (d_rows) {
    KG_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(nullptr, cub_bytes, d_rows, static_cast<uint8_t*>(nullptr),
                                                    d_out_rows, static_cast<int64_t*>(nullptr), ni, ctx->stream));
  } else {
    KG_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(nullptr, cub_bytes, iota, static_cast<uint8_t*>(nullptr),
                                                    d_out_rows, static_cast<int64_t*>(nullptr), ni, ctx->stream));
  }
  const size_t fb = (static_cast<size_t>(n) + 255) / 256 * 256;
  char* extra = nullptr;
  int st = run_expr(ctx, n, cols, text_cols, dicts, ncols, nodes, nnodes, d_rows, 1, nullptr, nullptr, KG_BOOLEAN,
                    fb + 256 + cub_bytes, &extra);
  if (st != KG_OK) return st;
  if (n == 0) return KG_OK;
  uint8_t* flags = reinterpret_cast<uint8_t*>(extra);
  int64_t* d_num = reinterpret_cast<int64_t*>(extra + fb);
  void* cub = extra + fb + 256;
  if (d_rows) {
    KG_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(cub, cub_bytes, d_rows, flags, d_out_rows, d_num, ni, ctx->stream));
  } else {
    KG_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(cub, cub_bytes, iota, flags, d_out_rows, d_num, ni, ctx->stream));
  }
  KG_HIP_CHECK(ctx, hipMemcpyAsync(out_count, d_num, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return KG_OK;
}
