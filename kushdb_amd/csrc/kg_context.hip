// Context, errors and workspace for the kushgpu C ABI.
#include <algorithm>
#include <cstdarg>
#include <mutex>
#include <new>

#include "kg_internal.h"

// Device scratch of the entry points that allocate nothing
// (kg_scan_select_noalloc): one array per GPU, held until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host the name is
// only a stand-in); it uses the device address that hipGetSymbolAddress
// returns for the current device, looked up once per device when the first
// context on that device is created.
static __device__ __align__(256) unsigned char scratch[kg::kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The scratch array's device address on ctx->device (current device).
static int scratch_setup(kg_ctx* ctx) {
  if (ctx->device >= kMaxDevices) return kg_fail(ctx, KG_ERR_UNSUPPORTED, "device %d beyond %d", ctx->device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return kg_fail(ctx, KG_ERR_HIP, "scratch symbol address on device %d: %s", ctx->device, hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = ptr;
  }
  ctx->scratch = g_scratch_ptr[ctx->device];
  return KG_OK;
}

int kg_fail(kg_ctx* ctx, int status, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return status;
}

int kg_dev_alloc(kg_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  bytes = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
  // The smallest cached block that holds `bytes` without wasting more than
  // about half of itself.
  auto it = ctx->block_cache.lower_bound(bytes);
  if (it != ctx->block_cache.end() && it->first <= 2 * bytes + (size_t{1} << 20)) {
    *p = it->second;
    ctx->block_live[*p] = it->first;
    ctx->block_cache_bytes -= it->first;
    ctx->block_cache.erase(it);
    return KG_OK;
  }
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    // Give the cached blocks back to the device and try once more.
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& e : ctx->block_cache) (void)hipFree(e.second);
    ctx->block_cache.clear();
    ctx->block_cache_bytes = 0;
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return kg_fail(ctx, KG_ERR_OOM, "device allocation of %zu bytes failed", bytes);
    }
  }
  ctx->block_live[*p] = bytes;
  return KG_OK;
}

void kg_dev_release(kg_ctx* ctx, void* p) {
  if (!p) return;
  auto it = ctx ? ctx->block_live.find(p) : decltype(ctx->block_live.end()){};
  if (!ctx || it == ctx->block_live.end()) {
    (void)hipFree(p);
    return;
  }
  const size_t bytes = it->second;
  ctx->block_live.erase(it);
  if (bytes > kg::kBlockCacheBytes) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(p);
    return;
  }
  // Room for the block: the largest cached blocks go back to the device (the
  // newest release is the likeliest to be asked for again).
  if (ctx->block_cache_bytes + bytes > kg::kBlockCacheBytes) {
    (void)hipStreamSynchronize(ctx->stream);
    while (ctx->block_cache_bytes + bytes > kg::kBlockCacheBytes && !ctx->block_cache.empty()) {
      auto big = std::prev(ctx->block_cache.end());
      (void)hipFree(big->second);
      ctx->block_cache_bytes -= big->first;
      ctx->block_cache.erase(big);
    }
  }
  ctx->block_cache.emplace(bytes, p);
  ctx->block_cache_bytes += bytes;
}

int kg_allow_lds(kg_ctx* ctx, const void* kernel, size_t bytes) {
  if (static_cast<int64_t>(bytes) > ctx->max_lds)
    return kg_fail(ctx, KG_ERR_CAPACITY, "%zu bytes of LDS exceed the device's %lld", bytes,
                   (long long)ctx->max_lds);
  if (bytes <= 64 * 1024) return KG_OK;
  KG_HIP_CHECK(ctx, hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
  return KG_OK;
}

int kg_work_reserve(kg_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->work_bytes) return KG_OK;
  // Outstanding kernels may still read the old buffer.
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->work) KG_HIP_CHECK(ctx, hipFree(ctx->work));
  ctx->work = nullptr;
  ctx->work_bytes = 0;
  size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
  if (hipMalloc(&ctx->work, want) != hipSuccess) {
    (void)hipGetLastError();
    return kg_fail(ctx, KG_ERR_OOM, "workspace allocation of %zu bytes failed", want);
  }
  ctx->work_bytes = want;
  return KG_OK;
}

int kg_pinned_reserve(kg_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->pinned_bytes) return KG_OK;
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->pinned) KG_HIP_CHECK(ctx, hipHostFree(ctx->pinned));
  ctx->pinned = nullptr;
  ctx->pinned_bytes = 0;
  size_t want = bytes < 4096 ? 4096 : bytes;
  KG_HIP_CHECK(ctx, hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
  ctx->pinned_bytes = want;
  return KG_OK;
}

static hipEvent_t take_event(kg_ctx* ctx) {
  if (!ctx->ev_pool.empty()) {
    hipEvent_t e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void kg_time_begin(kg_ctx* ctx) {
  if (!ctx->timing) return;
  hipEvent_t a = take_event(ctx), b = take_event(ctx);
  if (!a || !b) return;
  (void)hipEventRecord(a, ctx->stream);
  ctx->ev_pending.push_back({a, b});
}

void kg_time_end(kg_ctx* ctx) {
  if (!ctx->timing || ctx->ev_pending.empty()) return;
  (void)hipEventRecord(ctx->ev_pending.back().second, ctx->stream);
}

extern "C" {

int kg_ctx_set_timing(kg_ctx* ctx, int on) {
  if (!ctx) return KG_ERR_INVALID;
  ctx->timing = on ? 1 : 0;
  return KG_OK;
}

int kg_ctx_kernel_stats(kg_ctx* ctx, double* total_ms, int64_t* launches, int reset) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, total_ms && launches, "null output pointer");
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  double t = 0;
  for (auto& pr : ctx->ev_pending) {
    float ms = 0;
    KG_HIP_CHECK(ctx, hipEventElapsedTime(&ms, pr.first, pr.second));
    t += ms;
  }
  *total_ms = t;
  *launches = static_cast<int64_t>(ctx->ev_pending.size());
  if (reset) {
    for (auto& pr : ctx->ev_pending) {
      ctx->ev_pool.push_back(pr.first);
      ctx->ev_pool.push_back(pr.second);
    }
    ctx->ev_pending.clear();
  }
  return KG_OK;
}

int kg_ctx_create(int device, kg_ctx** out) {
  if (!out) return KG_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return KG_ERR_HIP;
  }
  if (device < 0 || device >= count) return KG_ERR_INVALID;
  kg_ctx* ctx = new (std::nothrow) kg_ctx();
  if (!ctx) return KG_ERR_OOM;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return KG_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ctx->num_cus = prop.multiProcessorCount;
    ctx->max_lds = static_cast<int64_t>(prop.sharedMemPerBlock);
  }
  if (hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    return KG_ERR_HIP;
  }
  ctx->own_stream = true;
  if (scratch_setup(ctx) != KG_OK) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return KG_ERR_HIP;
  }
  *out = ctx;
  return KG_OK;
}

int kg_ctx_destroy(kg_ctx* ctx) {
  if (!ctx) return KG_OK;
  for (kg_ctx* l : ctx->serve_lanes) kg_ctx_destroy(l);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& pr : ctx->ev_pending) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
  for (auto& f : ctx->files) kg_file_col_free(f);
  for (auto& e : ctx->block_cache) (void)hipFree(e.second);
  if (ctx->work) hipFree(ctx->work);
  if (ctx->pinned) hipHostFree(ctx->pinned);
  if (ctx->own_stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return KG_OK;
}

const char* kg_last_error(const kg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int kg_ctx_set_stream(kg_ctx* ctx, void* hip_stream) {
  if (!ctx) return KG_ERR_INVALID;
  KG_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // Cached device blocks are reused in the order of the context's stream:
  // everything queued on the old stream ends before the new one is used.
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->own_stream) {
    KG_HIP_CHECK(ctx, hipStreamDestroy(ctx->stream));
    ctx->own_stream = false;
    ctx->stream = nullptr;
  }
  ctx->stream = static_cast<hipStream_t>(hip_stream);
  return KG_OK;
}

int kg_ctx_synchronize(kg_ctx* ctx) {
  if (!ctx) return KG_ERR_INVALID;
  KG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return KG_OK;
}

int kg_ctx_set_deterministic(kg_ctx* ctx, int on) {
  if (!ctx) return KG_ERR_INVALID;
  ctx->deterministic = on ? 1 : 0;
  return KG_OK;
}

int kg_ctx_set_probe_regions(kg_ctx* ctx, int32_t regions) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, regions == -1 || regions == 0 || (regions % 8 == 0 && regions <= 1024),
             "probe regions must be -1, 0 or a multiple of 8 up to 1024 (got %d)", regions);
  ctx->probe_regions = regions;
  return KG_OK;
}

const char* kg_ctx_last_path(const kg_ctx* ctx) { return ctx ? ctx->last_path : ""; }

int kg_ctx_set_build_regions(kg_ctx* ctx, int32_t regions) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, regions >= -1 && regions <= 32768, "build regions must be -1, 0 or up to 32768 (got %d)", regions);
  ctx->build_regions = regions;
  return KG_OK;
}

int kg_ctx_set_agg_regions(kg_ctx* ctx, int32_t regions) {
  if (!ctx) return KG_ERR_INVALID;
  KG_REQUIRE(ctx, regions >= -1 && regions <= 65536, "aggregation regions must be -1, 0 or up to 65536 (got %d)",
             regions);
  ctx->agg_regions = regions;
  return KG_OK;
}

}  // extern "C"
