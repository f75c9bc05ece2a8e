// Twiddle trees: PolyOps::precompute_twiddles (stwo_gpu_backend/src/poly.rs:170-192).
//
// Tree layout (cuda/src/poly/twiddles.cu:29-73): for the coset doubled j times
// (j = 0, 1, ...), the x coordinates of its first half in bit-reversed order,
// one level after another, then a trailing 1 (so the tree has coset.size()
// entries).  The inverse tree is the element-wise inverse (poly.rs:180-185).
//
// Here one launch fills every level (the reference launches once per level), and
// the extension entry point caches trees per coset so a prover computes each once.
#include "field.hpp"
#include "runtime.hpp"

namespace sh {
void batch_inverse_m31(const uint32_t* src, uint32_t* dst, uint64_t n);  // batch_inverse.hip

namespace tw {

__host__ __device__ inline pt pdouble_n(pt p, uint32_t n) {
  for (uint32_t i = 0; i < n; i++) p = pmul(p, p);
  return p;
}
__device__ inline pt ppow(pt p, uint32_t e) {
  pt r = {1, 0};
  while (e) {
    if (e & 1) r = pmul(r, p);
    p = pmul(p, p);
    e >>= 1;
  }
  return r;
}

__global__ void tree_kernel(uint32_t* out, pt initial, pt step, uint32_t log_size) {
  const uint32_t size = 1u << log_size;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= size) return;
  if (t == size - 1) {
    out[t] = 1;
    return;
  }
  // level j occupies [size - (size >> j), size - (size >> (j + 1)))
  uint32_t j = 0;
  while (t >= size - (size >> (j + 1))) j++;
  uint32_t i = t - (size - (size >> j));
  uint32_t lvl_log = log_size - j;  // coset log size at this level
  pt ini = pdouble_n(initial, j), st = pdouble_n(step, j);
  out[t] = pmul(ini, ppow(st, brev(i, lvl_log - 1))).x;
}

void compute_tree(uint32_t* out, pt initial, pt step, uint32_t log_size) {
  uint32_t size = 1u << log_size;
  timed("twiddle_tree", [&] {
    tree_kernel<<<ceil_div(size, 256), 256, 0, ctx().cur>>>(out, initial, step, log_size);
  });
}

}  // namespace tw
}  // namespace sh

using namespace sh;

extern "C" uint32_t* precompute_twiddles(StwoCirclePointM31 initial, StwoCirclePointM31 step, size_t total_size) {
  uint32_t* out = nullptr;
  guarded([&] {
    require(is_pow2(total_size) && total_size >= 1 && total_size <= (1ull << 30),
            "precompute_twiddles: size must be a power of two");
    out = (uint32_t*)ctx().pool.alloc(total_size * 4);
    tw::compute_tree(out, pt{initial.x, initial.y}, pt{step.x, step.y}, log2_exact(total_size));
    sync();
  });
  return out;
}

extern "C" int stwo_hip_twiddles(uint32_t initial_index, uint32_t step_size, uint32_t log_size,
                                 const uint32_t** twiddles, const uint32_t** itwiddles) {
  return guarded([&] {
    require(log_size <= 30, "stwo_hip_twiddles: log size too large");
    TwiddleKey key{initial_index & 0x7fffffffu, step_size & 0x7fffffffu, log_size};
    auto& cache = ctx().twiddles;
    auto it = cache.find(key);
    if (it == cache.end()) {
      uint64_t size = 1ull << log_size;
      TwiddleEntry e;
      e.tw = (uint32_t*)ctx().pool.alloc(size * 4);
      e.itw = (uint32_t*)ctx().pool.alloc(size * 4);
      tw::compute_tree(e.tw, point_at_index(key.initial), point_at_index(key.step), log_size);
      batch_inverse_m31(e.tw, e.itw, size);
      // the cached tree is shared by work on every stream: finish it before handing it out
      hip_check(hipStreamSynchronize(ctx().cur), "hipStreamSynchronize");
      it = cache.emplace(key, e).first;
    }
    *twiddles = it->second.tw;
    *itwiddles = it->second.itw;
  });
}

extern "C" int stwo_hip_release_twiddles(void) {
  return guarded([] {
    sync();
    // proof-server jobs on the slot streams read the cached tables too: wait for every
    // stream, so that no block handed back to the pool here is still being read (new jobs
    // are enqueued under the library lock and fetch the tables again)
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    for (auto& kv : ctx().twiddles) {
      ctx().pool.free(kv.second.tw);
      ctx().pool.free(kv.second.itw);
    }
    ctx().twiddles.clear();
    for (auto& kv : ctx().domain_tables) ctx().pool.free(kv.second);
    ctx().domain_tables.clear();
  });
}
