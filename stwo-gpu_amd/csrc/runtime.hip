// Runtime of libstwo_hip: context, errors, caching pool, timing and the memory
// entry points of the C ABI (cuda/src/utils.cu:10-80 in the reference).
#include "runtime.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "field.hpp"

namespace sh {

namespace {
thread_local int t_status = 0;
thread_local char t_message[512] = {0};
std::recursive_mutex g_mutex;
Context g_ctx;
}  // namespace

void set_error(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_message, sizeof t_message, fmt, ap);
  va_end(ap);
  t_status = status;
  fprintf(stderr, "stwo_hip: %s\n", t_message);
}
int last_status() { return t_status; }

void fail(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_message, sizeof t_message, fmt, ap);
  va_end(ap);
  t_status = status;
  fprintf(stderr, "stwo_hip: %s\n", t_message);
  throw Failure{status};
}

static void init_context(int device) {
  Context& c = g_ctx;
  if (c.device == device) return;
  if (c.device >= 0) fail(STWO_HIP_ERR_ARG, "library already initialised on device %d", c.device);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device)
    fail(STWO_HIP_ERR_NODEV, "no HIP device %d (found %d)", device, count);
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipGetDeviceProperties(&c.props, device), "hipGetDeviceProperties");
  if (strncmp(c.props.gcnArchName, "gfx950", 6) != 0)
    fail(STWO_HIP_ERR_NODEV, "device %d is %s; this library is built for gfx950", device, c.props.gcnArchName);
  // default priority, non-blocking w.r.t. the legacy stream
  hip_check(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate");
  c.cur = c.stream;
  for (int i = 0; i < N_AUX_STREAMS; i++) {
    hip_check(hipStreamCreateWithFlags(&c.aux[i], hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&c.join[i], hipEventDisableTiming), "hipEventCreate");
  }
  hip_check(hipEventCreateWithFlags(&c.fork, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreate(&c.timer_start), "hipEventCreate");
  hip_check(hipEventCreate(&c.timer_stop), "hipEventCreate");
  c.device = device;
}

std::unique_lock<std::recursive_mutex> lock() {
  std::unique_lock<std::recursive_mutex> g(g_mutex);
  if (g_ctx.device < 0) init_context(0);
  hip_check(hipSetDevice(g_ctx.device), "hipSetDevice");
  return g;
}
Context& ctx() { return g_ctx; }

hipStream_t slot_stream(size_t i) {
  Context& c = g_ctx;
  while (c.slots.size() <= i) {
    hipStream_t s;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    c.slots.push_back(s);
  }
  return c.slots[i];
}

static void resolve_timings() {
  Context& c = g_ctx;
  std::vector<TimedLaunch> later;
  for (auto& t : c.pending) {
    if (hipEventQuery(t.stop) != hipSuccess) {  // launched on another stream, still running
      later.push_back(t);
      continue;
    }
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, t.start, t.stop), "hipEventElapsedTime");
    auto& s = c.stats[t.name];
    s.first += 1;
    s.second += ms;
    c.event_pool.push_back(t.start);
    c.event_pool.push_back(t.stop);
  }
  c.pending.swap(later);
}

void sync() {
  hip_check(hipStreamSynchronize(g_ctx.stream), "hipStreamSynchronize");
  if (!g_ctx.pending.empty()) resolve_timings();
}

static hipEvent_t take_event() {
  Context& c = g_ctx;
  if (!c.event_pool.empty()) {
    hipEvent_t e = c.event_pool.back();
    c.event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hip_check(hipEventCreate(&e), "hipEventCreate");
  return e;
}
void launch_begin(const char* name) {
  Context& c = g_ctx;
  if (!c.kernel_timing) return;
  TimedLaunch t{name, take_event(), take_event()};
  hip_check(hipEventRecord(t.start, c.cur), "hipEventRecord");
  c.pending.push_back(t);
}
void launch_end(const char* name) {
  Context& c = g_ctx;
  if (!c.kernel_timing || c.pending.empty()) return;
  hip_check(hipEventRecord(c.pending.back().stop, c.cur), "hipEventRecord");
}

// ---- pool ----
size_t Pool::round(size_t b) {
  if (b == 0) b = 1;
  if (b <= (1u << 20)) return (b + 511) & ~(size_t)511;
  return (b + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
}
void* Pool::alloc(size_t bytes) {
  size_t r = round(bytes);
  const hipStream_t cur = g_ctx.cur;
  for (auto it = cache_.lower_bound(r); it != cache_.end() && it->first <= r + r / 4; ++it) {  // close fit only
    Block& b = it->second;
    if (b.stream != cur && hipEventQuery(b.done) != hipSuccess) continue;
    void* p = b.p;
    g_ctx.event_pool.push_back(b.done);
    cached_ -= it->first;
    live_[p] = it->first;
    cache_.erase(it);
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, r);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    release_cached();
    e = hipMalloc(&p, r);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(STWO_HIP_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", r, hipGetErrorString(e));
    }
  }
  live_[p] = r;
  return p;
}
void Pool::free(void* p) {
  if (!p) return;
  auto it = live_.find(p);
  if (it == live_.end()) fail(STWO_HIP_ERR_ARG, "free of a pointer the library did not allocate (%p)", p);
  Block b{p, g_ctx.cur, take_event()};
  hip_check(hipEventRecord(b.done, b.stream), "hipEventRecord");
  cache_.emplace(it->second, b);
  cached_ += it->second;
  live_.erase(it);
}
void Pool::release_cached() {
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  for (auto& kv : cache_) {
    hip_check(hipFree(kv.second.p), "hipFree");
    g_ctx.event_pool.push_back(kv.second.done);
  }
  cache_.clear();
  cached_ = 0;
}

static Staging& take_staging(size_t bytes) {
  Context& c = g_ctx;
  for (auto& st : c.staging) {
    if (st.busy && hipEventQuery(st.done) == hipSuccess) st.busy = false;
    if (!st.busy && st.cap >= bytes) return st;
  }
  Staging st{};
  st.cap = std::max<size_t>(bytes, 64 << 10);
  // coherent (fine-grained): small uploads are read by a kernel straight from this memory,
  // so the device must not keep lines of an earlier upload cached
  hip_check(hipHostMalloc(&st.host, st.cap, hipHostMallocCoherent), "hipHostMalloc");
  hip_check(hipEventCreateWithFlags(&st.done, hipEventDisableTiming), "hipEventCreate");
  c.staging.push_back(st);
  return c.staging.back();
}

// Small uploads: a kernel on the current stream reads the staging buffer over the bus.  A
// DMA copy would be ordered with the stream's kernels through a cross-engine wait whose
// latency (tens of microseconds) every following kernel sees.
__global__ void upload_kernel(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint32_t n4) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4) dst[i] = src[i];
}
constexpr size_t UPLOAD_KERNEL_MAX = 64 << 10;

void upload_to(void* dev, const void* host, size_t bytes) {
  if (!bytes) return;
  Staging& st = take_staging(bytes);
  memcpy(st.host, host, bytes);
  if (bytes <= UPLOAD_KERNEL_MAX && bytes % 4 == 0 && ((uintptr_t)dev | (uintptr_t)st.host) % 4 == 0) {
    const uint32_t n4 = (uint32_t)(bytes / 4);
    upload_kernel<<<(n4 + 255) / 256, 256, 0, g_ctx.cur>>>((uint32_t*)dev, (const uint32_t*)st.host, n4);
    hip_check(hipGetLastError(), "upload_kernel");
  } else {
    hip_check(hipMemcpyAsync(dev, st.host, bytes, hipMemcpyHostToDevice, g_ctx.cur), "hipMemcpyAsync H2D");
  }
  hip_check(hipEventRecord(st.done, g_ctx.cur), "hipEventRecord");
  st.busy = true;
}

DevBuf upload(const void* host, size_t bytes) {
  DevBuf b(bytes);
  upload_to(b.p, host, bytes);
  return b;
}

// ---- synthetic uniform M31 stream (same definition as the oracle's generator) ----
__global__ void fill_uniform_kernel(uint32_t* dst, uint64_t n, uint64_t seed, uint64_t offset) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = uniform_m31(seed, offset + i);
}

// plain 16-byte copy (column copies; also the reference kernel for HBM counter calibration)
__global__ void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace sh

using namespace sh;

// =========================================================================
// C ABI: library control
// =========================================================================
extern "C" int stwo_hip_init(int device) {
  try {
    std::unique_lock<std::recursive_mutex> g(g_mutex);
    init_context(device);
    return STWO_HIP_OK;
  } catch (const Failure& f) {
    return f.status;
  }
}
extern "C" int stwo_hip_last_status(void) { return t_status; }
extern "C" const char* stwo_hip_last_error(void) { return t_message; }
extern "C" void stwo_hip_clear_error(void) {
  t_status = 0;
  t_message[0] = 0;
}
extern "C" int stwo_hip_synchronize(void) {
  return guarded([] { sync(); });
}
extern "C" void* stwo_hip_stream(void) {
  void* s = nullptr;
  guarded([&] { s = (void*)ctx().stream; });
  return s;
}
extern "C" int stwo_hip_malloc(void** ptr, size_t bytes) {
  return guarded([&] { *ptr = ctx().pool.alloc(bytes); });
}
extern "C" int stwo_hip_free(void* ptr) {
  return guarded([&] { ctx().pool.free(ptr); });
}
extern "C" int stwo_hip_memcpy(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx().stream), "hipMemcpy");
    sync();
  });
}
extern "C" int stwo_hip_release_cached_memory(void) {
  return guarded([] { ctx().pool.release_cached(); });
}
extern "C" int stwo_hip_fill_uniform_m31(uint32_t* dst, uint64_t n, uint64_t seed, uint64_t offset) {
  return guarded([&] {
    if (!n) return;
    uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 256u * 64u);
    timed("fill_uniform", [&] {
      fill_uniform_kernel<<<blocks, 256, 0, ctx().cur>>>(dst, n, seed, offset);
    });
  });
}
extern "C" int stwo_hip_device_copy(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    require(((uintptr_t)dst | (uintptr_t)src | bytes) % 16 == 0, "stwo_hip_device_copy: needs 16-byte alignment");
    uint64_t n16 = bytes / 16;
    if (!n16) return;
    uint32_t blocks = (uint32_t)std::min<uint64_t>((n16 + 255) / 256, 256u * 32u);
    timed("copy16", [&] {
      copy16_kernel<<<blocks, 256, 0, ctx().cur>>>((const uint4*)src, (uint4*)dst, n16);
    });
  });
}
extern "C" int stwo_hip_timer_begin(void) {
  return guarded([] { hip_check(hipEventRecord(ctx().timer_start, ctx().stream), "hipEventRecord"); });
}
extern "C" int stwo_hip_timer_end(float* ms) {
  return guarded([&] {
    hip_check(hipEventRecord(ctx().timer_stop, ctx().stream), "hipEventRecord");
    hip_check(hipEventSynchronize(ctx().timer_stop), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(ms, ctx().timer_start, ctx().timer_stop), "hipEventElapsedTime");
  });
}
extern "C" int stwo_hip_kernel_timing(int enable) {
  return guarded([&] {
    sync();
    ctx().kernel_timing = enable != 0;
  });
}
extern "C" int stwo_hip_kernel_stats(const char* name, uint64_t* launches, double* total_ms) {
  return guarded([&] {
    sync();
    auto it = ctx().stats.find(name);
    *launches = it == ctx().stats.end() ? 0 : it->second.first;
    *total_ms = it == ctx().stats.end() ? 0 : it->second.second;
  });
}
extern "C" int stwo_hip_kernel_stats_reset(void) {
  return guarded([] {
    sync();
    ctx().stats.clear();
  });
}

// =========================================================================
// C ABI: reference memory functions (bindings.rs:77-99, 196-216)
// =========================================================================
static void copy_sync(void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx().stream), "hipMemcpy");
  sync();
}

extern "C" void copy_uint32_t_vec_from_device_to_host(const uint32_t* device_ptr, uint32_t* host_ptr,
                                                      uint32_t size) {
  guarded([&] { copy_sync(host_ptr, device_ptr, (size_t)size * 4); });
}
extern "C" uint32_t* copy_uint32_t_vec_from_host_to_device(const uint32_t* host_ptr, uint32_t size) {
  uint32_t* d = nullptr;
  guarded([&] {
    d = (uint32_t*)ctx().pool.alloc((size_t)size * 4);
    copy_sync(d, host_ptr, (size_t)size * 4);
  });
  return d;
}
extern "C" uint32_t* copy_uint32_t_vec_from_device_to_device(const uint32_t* from, uint32_t* dst,
                                                             uint32_t size) {
  guarded([&] { copy_sync(dst, from, (size_t)size * 4); });
  return dst;
}
extern "C" uint32_t* cuda_malloc_uint32_t(uint32_t size) {
  uint32_t* d = nullptr;
  guarded([&] { d = (uint32_t*)ctx().pool.alloc((size_t)size * 4); });
  return d;
}
extern "C" StwoBlake2sHash* cuda_malloc_blake_2s_hash(size_t size) {
  StwoBlake2sHash* d = nullptr;
  guarded([&] { d = (StwoBlake2sHash*)ctx().pool.alloc(size * 32); });
  return d;
}
extern "C" uint32_t* cuda_alloc_zeroes_uint32_t(uint32_t size) {
  uint32_t* d = nullptr;
  guarded([&] {
    d = (uint32_t*)ctx().pool.alloc((size_t)size * 4);
    hip_check(hipMemsetAsync(d, 0, (size_t)size * 4, ctx().stream), "hipMemsetAsync");
    sync();
  });
  return d;
}
extern "C" StwoBlake2sHash* cuda_alloc_zeroes_blake_2s_hash(size_t size) {
  StwoBlake2sHash* d = nullptr;
  guarded([&] {
    d = (StwoBlake2sHash*)ctx().pool.alloc(size * 32);
    hip_check(hipMemsetAsync(d, 0, size * 32, ctx().stream), "hipMemsetAsync");
    sync();
  });
  return d;
}
extern "C" void cuda_free_memory(void* device_ptr) {
  guarded([&] { ctx().pool.free(device_ptr); });
}
extern "C" StwoBlake2sHash* copy_blake_2s_hash_vec_from_host_to_device(const StwoBlake2sHash* from, size_t size) {
  StwoBlake2sHash* d = nullptr;
  guarded([&] {
    d = (StwoBlake2sHash*)ctx().pool.alloc(size * 32);
    copy_sync(d, from, size * 32);
  });
  return d;
}
extern "C" void copy_blake_2s_hash_vec_from_device_to_host(const StwoBlake2sHash* from, StwoBlake2sHash* to,
                                                           size_t size) {
  guarded([&] { copy_sync(to, from, size * 32); });
}
extern "C" void copy_blake_2s_hash_vec_from_device_to_device(const StwoBlake2sHash* from, StwoBlake2sHash* dst,
                                                             size_t size) {
  guarded([&] { copy_sync(dst, from, size * 32); });
}
extern "C" uint32_t** copy_device_pointer_vec_from_host_to_device(uint32_t* const* from, size_t size) {
  uint32_t** d = nullptr;
  guarded([&] {
    d = (uint32_t**)ctx().pool.alloc(size * sizeof(void*));
    copy_sync(d, from, size * sizeof(void*));
  });
  return d;
}
