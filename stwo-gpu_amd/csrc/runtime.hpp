// Internal runtime of libstwo_hip: per-device context (stream, caching pool,
// twiddle cache), error recording, pointer-table uploads and launch timing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdarg>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/stwo_hip.h"

namespace sh {

// ---- error recording (thread-local, printed once per failure) ----
void set_error(int status, const char* fmt, ...);
int last_status();
struct Failure {  // thrown inside the library, converted to a status at the ABI edge
  int status;
};
[[noreturn]] void fail(int status, const char* fmt, ...);
inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(STWO_HIP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}
inline void require(bool ok, const char* fmt, const char* detail = "") {
  if (!ok) fail(STWO_HIP_ERR_ARG, fmt, detail);
}

// ---- caching device allocator --------------------------------------------
// A freed block remembers the stream it was last used on and an event recorded there
// at the time of the free.  It is handed out again at once to work on the same stream
// (later work on that stream is ordered after its previous users), and to work on
// another stream only once that event has completed.
class Pool {
 public:
  void* alloc(size_t bytes);  // for work launched on ctx().cur
  void free(void* p);         // every user of p is already enqueued on ctx().cur
  void release_cached();      // synchronises the device, then returns cached blocks to HIP
  size_t cached_bytes() const { return cached_; }

 private:
  struct Block {
    void* p;
    hipStream_t stream;
    hipEvent_t done;
  };
  static size_t round(size_t b);
  std::unordered_map<void*, size_t> live_;
  std::multimap<size_t, Block> cache_;
  size_t cached_ = 0;
};

struct TwiddleKey {
  uint32_t initial, step, log;
  bool operator<(const TwiddleKey& o) const {
    return initial != o.initial ? initial < o.initial : step != o.step ? step < o.step : log < o.log;
  }
};
struct TwiddleEntry {
  uint32_t* tw;
  uint32_t* itw;
};

// Pinned host staging buffers for host-to-device uploads of small tables; a buffer is
// reused only after the event recorded behind its copy has completed.
struct Staging {
  void* host;
  size_t cap;
  hipEvent_t done;
  bool busy;
};

struct TimedLaunch {
  const char* name;
  hipEvent_t start, stop;
};

constexpr int N_AUX_STREAMS = 4;

struct Context {
  int device = -1;
  hipStream_t stream = nullptr;  // the library stream: every call is ordered on it
  hipStream_t cur = nullptr;     // where kernels are launched now (the library stream, or an
                                 // auxiliary stream while a plan fans independent work out)
  hipStream_t aux[N_AUX_STREAMS] = {};
  // streams for concurrent jobs (proof server slots); they live as long as the process,
  // since cached pool blocks, staging buffers and timing events keep referring to the
  // stream their last user ran on
  std::vector<hipStream_t> slots;
  hipEvent_t fork = nullptr, join[N_AUX_STREAMS] = {};
  Pool pool;
  std::map<TwiddleKey, TwiddleEntry> twiddles;
  // domain point tables of the circle domains that quotients were accumulated over,
  // keyed like the twiddle trees (half coset initial index, step, domain log size)
  std::map<TwiddleKey, void*> domain_tables;
  bool kernel_timing = false;
  std::vector<TimedLaunch> pending;
  std::map<std::string, std::pair<uint64_t, double>> stats;
  std::vector<hipEvent_t> event_pool;
  std::vector<Staging> staging;
  hipEvent_t timer_start = nullptr, timer_stop = nullptr;
  hipDeviceProp_t props;
};

// Lock the library, make sure the context exists; every ABI function holds this.
std::unique_lock<std::recursive_mutex> lock();
Context& ctx();
void sync();  // synchronise the library stream (resolves pending kernel timings)
hipStream_t slot_stream(size_t i);  // the i-th job slot stream, created on first use

// Device buffer from the pool, returned to the pool on scope exit.
struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { p = ctx().pool.alloc(bytes); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p) ctx().pool.free(p);
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
  ~DevBuf() {
    if (p) ctx().pool.free(p);
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};
// Copy a host array to a fresh pooled device buffer (stream-ordered, through pinned staging).
DevBuf upload(const void* host, size_t bytes);
// Same into existing device memory.
void upload_to(void* dev, const void* host, size_t bytes);

// Launch bookkeeping: when kernel timing is on, bracket the launch with events.
void launch_begin(const char* name);
void launch_end(const char* name);
template <class F>
inline void timed(const char* name, F&& f) {
  // an error left by an earlier call on this thread is reported as such, not as this launch's
  const hipError_t before = hipGetLastError();
  if (before != hipSuccess) fail(STWO_HIP_ERR_HIP, "%s: earlier error pending: %s", name, hipGetErrorString(before));
  launch_begin(name);
  f();
  hip_check(hipGetLastError(), name);
  launch_end(name);
}

// Several same-shape proof jobs in one launch: job z = blockIdx.z keeps its buffers z * js
// bytes after job 0's (equal arena layouts), and tables hold job 0's pointers, so a kernel
// adds the job's offset to every arena pointer it uses (js = 0 and one job elsewhere).
template <class T>
__device__ __forceinline__ T* jptr(T* p, uint64_t jo) {
  return (T*)((char*)p + jo);
}
__device__ __forceinline__ uint64_t job_offset(uint64_t js) { return (uint64_t)blockIdx.z * js; }

// Run `body` converting failures into a status; used at every ABI entry.
template <class F>
inline int guarded(F&& body) {
  try {
    auto g = lock();
    body();
    return STWO_HIP_OK;
  } catch (const Failure& f) {
    return f.status;
  }
}

inline uint32_t log2_exact(uint64_t n) {
  uint32_t l = 0;
  while ((1ull << l) < n) l++;
  return l;
}
inline bool is_pow2(uint64_t n) { return n && !(n & (n - 1)); }
inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace sh
