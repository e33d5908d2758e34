// The reference's on-disk column formats, loaded into HBM-resident columns.
//
// FileManager::Open reads a whole file once and keeps it by path
// (runtime/file_manager.cc:22-60); the column readers then view it in place:
//   * typed columns: a plain array of the C type (runtime/column_data.cc:24-116);
//   * NULL flags: "<stem>_null.kdb", one int8 per row, 1 = NULL
//     (util/load.h:70-90);
//   * TEXT: {int32 cardinality, padded to 8 bytes; slot[n] = {int32 length,
//     padded; uint64 offset from the file start}} followed by the
//     NUL-terminated strings (runtime/column_data.cc:175-229);
//   * ENUM dictionaries (.kdbenum): {uint64 entry_offset; uint64 mask;
//     uint64 array[mask + 1]}, then EnumEntry {uint64 str_offset; uint32
//     str_len; uint32 value} by code, then the strings (runtime/enum.cc:26-131);
//   * column indexes (.kdbindex): {uint64 mask; uint64 array[mask + 1]}, then
//     entries {K key, padded to 8 bytes; uint64 size; uint64 next; int32
//     values[size]} packed back to back to the end of the file
//     (runtime/column_index.cc:150-240).
// Here the bytes go to device memory once per path and context and stay there
// until released; TEXT becomes offsets + bytes (the layout of kg_text_col), a
// column index becomes a kg_index equal to kg_index_build over the indexed
// rows.
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "kg_internal.h"

using namespace kg;

void kg_file_col_free(kg_ctx::FileCol& f) {
  if (f.data) (void)hipFree(f.data);
  if (f.nulls) (void)hipFree(f.nulls);
  if (f.offsets) (void)hipFree(f.offsets);
  f.data = nullptr;
  f.nulls = nullptr;
  f.offsets = nullptr;
}

namespace {

bool read_file(const std::string& path, std::vector<char>& out) {
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in) return false;
  const std::streamsize len = in.tellg();
  if (len < 0) return false;
  out.resize(static_cast<size_t>(len));
  in.seekg(0);
  return len == 0 || static_cast<bool>(in.read(out.data(), len));
}

bool file_exists(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  return static_cast<bool>(in);
}

// "<stem>_null.kdb" beside "<stem>.kdb".
std::string null_path(const std::string& path) {
  const std::string ext = ".kdb";
  if (path.size() >= ext.size() && path.compare(path.size() - ext.size(), ext.size(), ext) == 0)
    return path.substr(0, path.size() - ext.size()) + "_null.kdb";
  return path + "_null.kdb";
}

int width_of(int32_t type) {
  switch (type) {
    case KG_SMALLINT: return 2;
    case KG_INT:
    case KG_DATE:
    case KG_ENUM: return 4;
    case KG_BIGINT:
    case KG_REAL: return 8;
    case KG_BOOLEAN: return 1;
    default: return 0;
  }
}

template <typename T>
T read_at(const std::vector<char>& b, size_t off) {
  T v;
  std::memcpy(&v, b.data() + off, sizeof(T));
  return v;
}

int upload(kg_ctx* ctx, const void* src, size_t bytes, void** dst) {
  *dst = nullptr;
  if (hipMalloc(dst, bytes > 0 ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    *dst = nullptr;
    return kg_fail(ctx, KG_ERR_OOM, "device allocation of %zu bytes failed", bytes);
  }
  if (bytes) KG_HIP_CHECK(ctx, hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return KG_OK;
}

kg_ctx::FileCol* find(kg_ctx* ctx, const std::string& path) {
  for (auto& f : ctx->files)
    if (f.path == path) return &f;
  return nullptr;
}

// Loads the NULL file beside `path`, if present, for n rows.
int load_nulls(kg_ctx* ctx, const std::string& path, int64_t n, int8_t** out) {
  *out = nullptr;
  const std::string np = null_path(path);
  if (!file_exists(np)) return KG_OK;
  std::vector<char> b;
  KG_REQUIRE(ctx, read_file(np, b), "cannot read %s", np.c_str());
  KG_REQUIRE(ctx, static_cast<int64_t>(b.size()) == n, "%s has %zu flags for %lld rows", np.c_str(), b.size(),
             (long long)n);
  void* d = nullptr;
  const int st = upload(ctx, b.data(), b.size(), &d);
  *out = static_cast<int8_t*>(d);
  return st;
}

}  // namespace

extern "C" int kg_file_column(kg_ctx* ctx, const char* path, int32_t type, kg_col* out, int64_t* out_n) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, path && out && out_n, "path, out or out_n is NULL");
  const int w = width_of(type);
  KG_REQUIRE(ctx, w > 0, "type %d is not a fixed-width column type (TEXT: kg_file_text)", type);
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const std::string p(path);
  kg_ctx::FileCol* f = find(ctx, p);
  if (f) {
    KG_REQUIRE(ctx, f->type == type, "%s is resident as type %d, not %d", path, f->type, type);
  } else {
    std::vector<char> b;
    KG_REQUIRE(ctx, read_file(p, b), "cannot read %s", path);
    KG_REQUIRE(ctx, b.size() % w == 0, "%s: %zu bytes is not a whole number of %d-byte values", path, b.size(), w);
    kg_ctx::FileCol c;
    c.path = p;
    c.type = type;
    c.n = static_cast<int64_t>(b.size() / w);
    int st = upload(ctx, b.data(), b.size(), &c.data);
    if (st == KG_OK) st = load_nulls(ctx, p, c.n, &c.nulls);
    if (st == KG_OK) st = hipStreamSynchronize(ctx->stream) == hipSuccess ? KG_OK : KG_ERR_HIP;
    if (st != KG_OK) {
      kg_file_col_free(c);
      return st == KG_ERR_HIP ? kg_fail(ctx, KG_ERR_HIP, "upload of %s failed", path) : st;
    }
    ctx->files.push_back(c);
    f = &ctx->files.back();
  }
  out->type = type;
  out->reserved = 0;
  out->data = f->data;
  out->nulls = f->nulls;
  *out_n = f->n;
  return KG_OK;
}

extern "C" int kg_file_text(kg_ctx* ctx, const char* path, kg_text_col* out, int64_t* out_n) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, path && out && out_n, "path, out or out_n is NULL");
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const std::string p(path);
  kg_ctx::FileCol* f = find(ctx, p);
  if (f) {
    KG_REQUIRE(ctx, f->type == KG_TEXT, "%s is resident as type %d, not TEXT", path, f->type);
  } else {
    std::vector<char> b;
    KG_REQUIRE(ctx, read_file(p, b), "cannot read %s", path);
    KG_REQUIRE(ctx, b.size() >= 8, "%s is too short for a TEXT column", path);
    const int64_t n = read_at<int32_t>(b, 0);
    KG_REQUIRE(ctx, n >= 0 && b.size() >= 8 + 16 * static_cast<size_t>(n), "%s: bad cardinality %lld", path,
               (long long)n);
    std::vector<int64_t> offs(static_cast<size_t>(n) + 1, 0);
    std::vector<char> bytes;
    for (int64_t i = 0; i < n; i++) {
      const int32_t len = read_at<int32_t>(b, 8 + 16 * i);
      const uint64_t off = read_at<uint64_t>(b, 8 + 16 * i + 8);
      KG_REQUIRE(ctx, len >= 0 && off + static_cast<uint64_t>(len) <= b.size(), "%s: slot %lld out of the file",
                 path, (long long)i);
      bytes.insert(bytes.end(), b.data() + off, b.data() + off + len);
      offs[i + 1] = offs[i] + len;
    }
    kg_ctx::FileCol c;
    c.path = p;
    c.type = KG_TEXT;
    c.n = n;
    void* d_off = nullptr;
    int st = upload(ctx, bytes.data(), bytes.size(), &c.data);
    if (st == KG_OK) st = upload(ctx, offs.data(), offs.size() * sizeof(int64_t), &d_off);
    c.offsets = static_cast<int64_t*>(d_off);
    if (st == KG_OK) st = load_nulls(ctx, p, n, &c.nulls);
    if (st == KG_OK) st = hipStreamSynchronize(ctx->stream) == hipSuccess ? KG_OK : KG_ERR_HIP;
    if (st != KG_OK) {
      kg_file_col_free(c);
      return st == KG_ERR_HIP ? kg_fail(ctx, KG_ERR_HIP, "upload of %s failed", path) : st;
    }
    ctx->files.push_back(c);
    f = &ctx->files.back();
  }
  out->offsets = f->offsets;
  out->bytes = static_cast<const uint8_t*>(f->data);
  out->nulls = f->nulls;
  *out_n = f->n;
  return KG_OK;
}

extern "C" int kg_file_release(kg_ctx* ctx, const char* path) {
  if (!ctx) return KG_ERR_INVALID;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i < ctx->files.size();) {
    if (!path || ctx->files[i].path == path) {
      kg_file_col_free(ctx->files[i]);
      ctx->files.erase(ctx->files.begin() + static_cast<std::ptrdiff_t>(i));
    } else {
      i++;
    }
  }
  return KG_OK;
}

extern "C" int kg_file_enum(kg_ctx* ctx, const char* path, char* buf, int64_t buf_cap, int64_t* offsets,
                            int32_t max_codes, int32_t* out_count) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, path && out_count, "path or out_count is NULL");
  std::vector<char> b;
  KG_REQUIRE(ctx, read_file(path, b), "cannot read %s", path);
  KG_REQUIRE(ctx, b.size() >= 16, "%s is too short for an enum dictionary", path);
  const uint64_t entry_offset = read_at<uint64_t>(b, 0);
  KG_REQUIRE(ctx, entry_offset >= 16 && entry_offset <= b.size(), "%s: bad entry offset", path);
  // Entries run from entry_offset to the first string, one per code.
  uint64_t end = b.size();
  if (entry_offset + 16 <= b.size()) end = read_at<uint64_t>(b, entry_offset);
  KG_REQUIRE(ctx, end >= entry_offset && end <= b.size() && (end - entry_offset) % 16 == 0,
             "%s: bad string heap offset", path);
  const int64_t count = static_cast<int64_t>((end - entry_offset) / 16);
  KG_REQUIRE(ctx, count <= INT32_MAX, "%s: too many codes", path);
  *out_count = static_cast<int32_t>(count);
  if (!buf) return KG_OK;  // size query
  KG_REQUIRE(ctx, offsets && max_codes >= count, "offsets must hold %lld + 1 entries", (long long)count);
  int64_t pos = 0;
  offsets[0] = 0;
  for (int64_t i = 0; i < count; i++) {
    const uint64_t e = entry_offset + 16 * static_cast<uint64_t>(i);
    const uint64_t so = read_at<uint64_t>(b, e);
    const uint32_t sl = read_at<uint32_t>(b, e + 8);
    const uint32_t code = read_at<uint32_t>(b, e + 12);
    KG_REQUIRE(ctx, code == static_cast<uint32_t>(i), "%s: entry %lld holds code %u", path, (long long)i, code);
    KG_REQUIRE(ctx, so + sl <= b.size(), "%s: string of code %lld out of the file", path, (long long)i);
    KG_REQUIRE(ctx, pos + sl <= buf_cap, "buffer of %lld bytes is too small", (long long)buf_cap);
    std::memcpy(buf + pos, b.data() + so, sl);
    pos += sl;
    offsets[i + 1] = pos;
  }
  return KG_OK;
}

extern "C" int kg_file_index(kg_ctx* ctx, const char* path, int32_t key_type, kg_index** out) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, path && out, "path or out is NULL");
  *out = nullptr;
  const int w = width_of(key_type);
  if (w == 0 || key_type == KG_BOOLEAN)
    return kg_fail(ctx, KG_ERR_UNSUPPORTED, "index key type %d: fixed-width keys (SMALLINT, INT, DATE, ENUM, "
                   "BIGINT, REAL) only", key_type);
  std::vector<char> b;
  KG_REQUIRE(ctx, read_file(path, b), "cannot read %s", path);
  KG_REQUIRE(ctx, b.size() >= 8, "%s is too short for a column index", path);
  const uint64_t mask = read_at<uint64_t>(b, 0);
  KG_REQUIRE(ctx, mask < (uint64_t{1} << 40), "%s: bad mask", path);
  uint64_t off = 8 + 8 * (mask + 1);
  KG_REQUIRE(ctx, off <= b.size(), "%s: bucket array past the end", path);
  // Entries back to back: key (padded to 8), size, next, values.
  std::vector<int32_t> tids;
  std::vector<char> keys;  // key of every listed tuple, in entry order
  while (off < b.size()) {
    KG_REQUIRE(ctx, off + 24 <= b.size(), "%s: truncated entry at %llu", path, (unsigned long long)off);
    const uint64_t size = read_at<uint64_t>(b, off + 8);
    KG_REQUIRE(ctx, size <= (b.size() - off - 24) / 4, "%s: entry at %llu overruns the file", path,
               (unsigned long long)off);
    for (uint64_t j = 0; j < size; j++) {
      tids.push_back(read_at<int32_t>(b, off + 24 + 4 * j));
      keys.insert(keys.end(), b.data() + off, b.data() + off + w);
    }
    off += 24 + 4 * size;
  }
  // A key column over the listed tuples (key[t] at row t), and those rows.
  int64_t nrows = 0;
  for (int32_t t : tids) {
    KG_REQUIRE(ctx, t >= 0, "%s: negative tuple id", path);
    nrows = std::max<int64_t>(nrows, int64_t{t} + 1);
  }
  std::vector<char> keycol(static_cast<size_t>(nrows) * w, 0);
  std::vector<char> seen(static_cast<size_t>(nrows), 0);
  for (size_t i = 0; i < tids.size(); i++) {
    KG_REQUIRE(ctx, !seen[tids[i]], "%s: tuple %d listed twice", path, tids[i]);
    seen[tids[i]] = 1;
    std::memcpy(keycol.data() + static_cast<size_t>(tids[i]) * w, keys.data() + i * w, w);
  }
  std::vector<int32_t> rows;
  rows.reserve(tids.size());
  for (int64_t t = 0; t < nrows; t++)
    if (seen[t]) rows.push_back(static_cast<int32_t>(t));
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  void* d_keys = nullptr;
  void* d_rows = nullptr;
  int st = upload(ctx, keycol.data(), keycol.size(), &d_keys);
  if (st == KG_OK) st = upload(ctx, rows.data(), rows.size() * sizeof(int32_t), &d_rows);
  if (st == KG_OK) {
    kg_col kc{key_type, 0, d_keys, nullptr};
    st = kg_index_build(ctx, &kc, static_cast<const int32_t*>(d_rows), static_cast<int64_t>(rows.size()), out);
  }
  (void)hipStreamSynchronize(ctx->stream);
  if (d_keys) (void)hipFree(d_keys);
  if (d_rows) (void)hipFree(d_rows);
  return st;
}
