// api.cpp -- C ABI of libswr: contexts, device memory accounting and the primitive entry points.
// The reference's equivalent lifetime is Pathtracer's constructor/destructor
// (src/core/Pathtracer.cpp:369-392); its one cross-thread call is terminate() (:420-423).
#include <algorithm>
#include <cstring>

#include "swr_internal.h"

namespace swr {

void* dev_alloc(swr_context* ctx, size_t bytes) {
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~size_t(255);
  if (ctx->budget && ctx->in_use + bytes > ctx->budget)
    fail(SWR_E_BUDGET, "device memory budget exceeded: need " + std::to_string(bytes) + " B with " +
                           std::to_string(ctx->in_use) + " of " + std::to_string(ctx->budget) + " B in use");
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(e == hipErrorOutOfMemory ? SWR_E_BUDGET : SWR_E_HIP,
         std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
  }
  ctx->allocs[p] = bytes;
  ctx->in_use += bytes;
  return p;
}

void dev_free(swr_context* ctx, void* p) {
  if (!p) return;
  auto it = ctx->allocs.find(p);
  if (it == ctx->allocs.end()) fail(SWR_E_INVALID, "free of a pointer this context does not own");
  ctx->in_use -= it->second;
  ctx->allocs.erase(it);
  SWR_HIP(hipFree(p));
}

void ensure(swr_context* ctx, DeviceBuffer& b, size_t bytes) {
  if (b.bytes >= bytes && b.ptr) return;
  if (b.ptr) {
    SWR_HIP(hipStreamSynchronize(ctx->stream));
    dev_free(ctx, b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  b.ptr = dev_alloc(ctx, bytes);
  b.bytes = bytes;
}

void release(swr_context* ctx, DeviceBuffer& b) {
  if (b.ptr) dev_free(ctx, b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

}  // namespace swr

using swr::Error;

// Runs f, turning library errors into status codes recorded on the context.
template <class F>
static int guarded(swr_context* ctx, F&& f) {
  try {
    f();
    return SWR_OK;
  } catch (const Error& e) {
    if (ctx) ctx->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->err = "host allocation failed";
    return SWR_E_INVALID;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return SWR_E_INVALID;
  }
}

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

extern "C" {

int swr_create(int device, swr_context** out) {
  if (!out) return SWR_E_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return SWR_E_HIP;
  if (device < 0 || device >= count) return SWR_E_INVALID;
  auto* ctx = new swr_context();
  ctx->device = device;
  int rc = guarded(ctx, [&] {
    SWR_HIP(hipSetDevice(device));
    SWR_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    SWR_HIP(hipGetDeviceProperties(&prop, device));
    ctx->num_cus = prop.multiProcessorCount;
    SWR_HIP(hipEventCreate(&ctx->ev0));
    SWR_HIP(hipEventCreate(&ctx->ev1));
    swr::sort_scratch_init(ctx);  // with `device` current
  });
  if (rc != SWR_OK) {
    delete ctx;
    return rc;
  }
  *out = ctx;
  return SWR_OK;
}

void swr_destroy(swr_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (swr_scene* s : ctx->scenes) swr::detach_scene(s);
  ctx->scenes.clear();
  for (swr_batch* b : ctx->batches) swr::detach_batch(b);
  ctx->batches.clear();
  swr::destroy_loop_buffers(ctx);
  for (auto& kv : ctx->allocs) (void)hipFree(kv.first);
  ctx->allocs.clear();
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* swr_last_error(const swr_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int swr_set_budget(swr_context* ctx, uint64_t bytes) {
  if (!ctx) return SWR_E_INVALID;
  ctx->budget = bytes;
  return SWR_OK;
}

uint64_t swr_bytes_in_use(const swr_context* ctx) { return ctx ? ctx->in_use : 0; }

int swr_cancel(swr_context* ctx) {
  if (!ctx) return SWR_E_INVALID;
  ctx->cancel.store(true);
  return SWR_OK;
}

int swr_reset_cancel(swr_context* ctx) {
  if (!ctx) return SWR_E_INVALID;
  ctx->cancel.store(false);
  return SWR_OK;
}

double swr_last_device_ms(const swr_context* ctx) { return ctx ? ctx->last_ms : -1.0; }

int swr_get_tuning(const swr_context* ctx, swr_tuning* out) {
  if (!ctx || !out) return SWR_E_INVALID;
  out->sort_keys_per_thread = ctx->tuning.sort_kpt;
  out->ray_key_bits = ctx->tuning.ray_key_bits;
  out->hit_key_bits = ctx->tuning.hit_key_bits;
  out->trace_stack_entries = ctx->tuning.trace_stack;
  out->sort_algorithm = ctx->tuning.sort_algorithm;
  out->shadow_overlap = ctx->tuning.shadow_overlap;
  return SWR_OK;
}

int swr_set_tuning(swr_context* ctx, const swr_tuning* t) {
  if (!ctx || !t) return SWR_E_INVALID;
  const int k = t->sort_keys_per_thread;
  if (!(k == 8 || k == 12 || k == 16 || k == 20 || k == 24) || t->ray_key_bits < 3 || t->ray_key_bits > 32 ||
      t->hit_key_bits < 0 || t->hit_key_bits > 32 ||
      !(t->trace_stack_entries == 0 || t->trace_stack_entries == 8 || t->trace_stack_entries == 16 ||
        t->trace_stack_entries == 32 || t->trace_stack_entries == 64) ||
      !(t->sort_algorithm == SWR_SORT_ONESWEEP || t->sort_algorithm == SWR_SORT_REDUCE_SCAN ||
        t->sort_algorithm == SWR_SORT_BUCKET) ||
      !(t->shadow_overlap == 0 || t->shadow_overlap == 1)) {
    ctx->err =
        "set_tuning: keys per thread must be 8/12/16/20/24, ray key bits 3..32, hit key bits 0..32, "
        "trace stack entries 0/8/16/32/64, sort algorithm 0/1/2, shadow overlap 0/1";
    return SWR_E_INVALID;
  }
  ctx->tuning.sort_kpt = k;
  ctx->tuning.ray_key_bits = t->ray_key_bits;
  ctx->tuning.hit_key_bits = t->hit_key_bits;
  ctx->tuning.trace_stack = t->trace_stack_entries;
  ctx->tuning.sort_algorithm = t->sort_algorithm;
  ctx->tuning.shadow_overlap = t->shadow_overlap;
  return SWR_OK;
}

int swr_synchronize(swr_context* ctx) {
  if (!ctx) return SWR_E_INVALID;
  return guarded(ctx, [&] { SWR_HIP(hipStreamSynchronize(ctx->stream)); });
}

int swr_device_alloc(swr_context* ctx, size_t bytes, void** out_dev) {
  if (!ctx || !out_dev) return SWR_E_INVALID;
  return guarded(ctx, [&] {
    SWR_HIP(hipSetDevice(ctx->device));
    *out_dev = swr::dev_alloc(ctx, bytes);
  });
}

int swr_device_free(swr_context* ctx, void* dev) {
  if (!ctx) return SWR_E_INVALID;
  return guarded(ctx, [&] {
    SWR_HIP(hipStreamSynchronize(ctx->stream));
    swr::dev_free(ctx, dev);
  });
}

int swr_memcpy_h2d(swr_context* ctx, void* dst_dev, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst_dev || !src))) return SWR_E_INVALID;
  return guarded(ctx, [&] {
    SWR_HIP(hipMemcpyAsync(dst_dev, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    SWR_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int swr_memcpy_d2h(swr_context* ctx, void* dst, const void* src_dev, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src_dev))) return SWR_E_INVALID;
  return guarded(ctx, [&] {
    SWR_HIP(hipMemcpyAsync(dst, src_dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SWR_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int swr_sort_pairs_u32(swr_context* ctx, uint32_t* keys_dev, uint32_t* vals_dev, size_t n, int key_bits,
                       int iota_values) {
  if (!ctx) return SWR_E_INVALID;
  return guarded(ctx, [&] {
    if (n && (!keys_dev || !vals_dev)) swr::fail(SWR_E_INVALID, "sort: null buffer");
    SWR_HIP(hipSetDevice(ctx->device));
    if (n == 0) return;
    swr::ensure(ctx, ctx->sort.alt_keys, n * sizeof(uint32_t));
    swr::ensure(ctx, ctx->sort.alt_vals, n * sizeof(uint32_t));
    swr::sort_reserve(ctx, n);
    uint32_t *ok = nullptr, *ov = nullptr;
    swr::Timer timer(ctx);
    swr::sort_pairs(ctx, keys_dev, vals_dev, static_cast<uint32_t*>(ctx->sort.alt_keys.ptr),
                    static_cast<uint32_t*>(ctx->sort.alt_vals.ptr), n, key_bits, iota_values != 0, &ok, &ov,
                    ctx->stream);
    timer.stop();
    if (ok != keys_dev) {
      SWR_HIP(hipMemcpyAsync(keys_dev, ok, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
      SWR_HIP(hipMemcpyAsync(vals_dev, ov, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    SWR_HIP(hipStreamSynchronize(ctx->stream));
    timer.finish();
  });
}

int swr_sort_pairs_u32_noalloc(swr_context* ctx, uint32_t* keys_dev, uint32_t* vals_dev, size_t n, int key_bits,
                               int iota_values) {
  if (!ctx) return SWR_E_INVALID;
  return guarded(ctx, [&] {
    if (n && (!keys_dev || !vals_dev)) swr::fail(SWR_E_INVALID, "sort: null buffer");
    SWR_HIP(hipSetDevice(ctx->device));
    swr::Timer timer(ctx);
    swr::sort_pairs_noalloc(ctx, keys_dev, vals_dev, n, key_bits, iota_values != 0, ctx->stream);
    timer.stop();
    SWR_HIP(hipStreamSynchronize(ctx->stream));
    timer.finish();
  });
}

int swr_ray_keys(swr_context* ctx, const float* rays_dev, size_t n, const float lo[3], const float scale[3],
                 uint32_t* keys_dev) {
  if (!ctx || !lo || !scale) return SWR_E_INVALID;
  return guarded(ctx, [&] {
    if (n && (!rays_dev || !keys_dev || !aligned16(rays_dev))) swr::fail(SWR_E_INVALID, "ray_keys: rays must be 16-byte aligned");
    swr::Timer timer(ctx);
    swr::launch_ray_keys(reinterpret_cast<const float4*>(rays_dev), n, make_float3(lo[0], lo[1], lo[2]),
                         make_float3(scale[0], scale[1], scale[2]), keys_dev, ctx->stream);
    SWR_HIP(hipGetLastError());
    timer.stop();
    SWR_HIP(hipStreamSynchronize(ctx->stream));
    timer.finish();
  });
}

int swr_hit_keys(swr_context* ctx, const uint32_t* hits_dev, size_t n, uint32_t* keys_dev) {
  if (!ctx) return SWR_E_INVALID;
  return guarded(ctx, [&] {
    if (n && (!hits_dev || !keys_dev || !aligned16(hits_dev))) swr::fail(SWR_E_INVALID, "hit_keys: hits must be 16-byte aligned");
    swr::Timer timer(ctx);
    swr::launch_hit_keys(reinterpret_cast<const uint4*>(hits_dev), n, keys_dev, ctx->stream);
    SWR_HIP(hipGetLastError());
    timer.stop();
    SWR_HIP(hipStreamSynchronize(ctx->stream));
    timer.finish();
  });
}

int swr_mesh_face_keys(swr_context* ctx, const uint32_t* mesh_dev, const uint32_t* face_dev, size_t n,
                       uint32_t* keys_dev) {
  if (!ctx) return SWR_E_INVALID;
  return guarded(ctx, [&] {
    if (n && (!mesh_dev || !face_dev || !keys_dev)) swr::fail(SWR_E_INVALID, "mesh_face_keys: null buffer");
    swr::Timer timer(ctx);
    swr::launch_mesh_face_keys(mesh_dev, face_dev, n, keys_dev, ctx->stream);
    SWR_HIP(hipGetLastError());
    timer.stop();
    SWR_HIP(hipStreamSynchronize(ctx->stream));
    timer.finish();
  });
}

int swr_gather_u32(swr_context* ctx, const uint32_t* in_dev, const uint32_t* perm_dev, size_t n, int rec_words,
                   uint32_t* out_dev) {
  if (!ctx) return SWR_E_INVALID;
  return guarded(ctx, [&] {
    if (n && (!in_dev || !perm_dev || !out_dev)) swr::fail(SWR_E_INVALID, "gather: null buffer");
    if ((rec_words % 4 == 0 && (!aligned16(in_dev) || !aligned16(out_dev))) ||
        (rec_words % 2 == 0 && ((reinterpret_cast<uintptr_t>(in_dev) | reinterpret_cast<uintptr_t>(out_dev)) & 7)))
      swr::fail(SWR_E_INVALID, "gather: records must be aligned to their vector width");
    swr::Timer timer(ctx);
    swr::launch_gather_u32(in_dev, perm_dev, n, rec_words, out_dev, ctx->stream);
    SWR_HIP(hipGetLastError());
    timer.stop();
    SWR_HIP(hipStreamSynchronize(ctx->stream));
    timer.finish();
  });
}

}  // extern "C"
