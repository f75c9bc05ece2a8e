// Proof-of-work grinding: GrindOps<Blake2sChannel>::grind, which the reference backend
// delegates to stwo's SimdBackend on the CPU (stwo_gpu_backend/src/backend.rs:14-18).
//
// The smallest nonce whose BLAKE2s(digest || nonce) has pow_bits trailing zero bits
// (the rule of channel.hpp).  One launch: lane t of the grid tests nonces t, t + W,
// t + 2W, ... (W = grid size) and records hits with a 64-bit atomic minimum.  A lane
// stops once the recorded minimum is below the next nonce it would test, so the
// smallest hit is always tested (every hit recorded before it is larger) and the
// result does not depend on scheduling.  Every lane also stops after max_rounds, so
// the launch always drains; no hit within the bound is reported as an error.
#include <array>
#include <vector>

#include "channel.hpp"
#include "runtime.hpp"

namespace sh {
namespace grind {

constexpr uint32_t NT = 256;
constexpr uint32_t BLOCKS = 2048;  // 8 per CU on 256 CUs
constexpr uint64_t WIDTH = (uint64_t)NT * BLOCKS;

// js: byte stride between the jobs of a batched launch (blockIdx.z = job; 0 for one job)
__global__ __launch_bounds__(NT) void grind_kernel(const uint32_t* __restrict__ digest, uint32_t pow_bits,
                                                   uint32_t max_rounds, unsigned long long* best, uint64_t js) {
  const uint64_t jo = job_offset(js);
  digest = jptr(digest, jo);
  best = jptr(best, jo);
  uint32_t d[8];
#pragma unroll
  for (int i = 0; i < 8; i++) d[i] = digest[i];
  const uint64_t t = blockIdx.x * (uint64_t)NT + threadIdx.x;
  for (uint32_t r = 0; r < max_rounds; r++) {
    const uint64_t nonce = t + r * WIDTH;
    if (__hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nonce) break;
    if (pow_trailing_zeros(d, nonce) >= pow_bits) {
      atomicMin(best, (unsigned long long)nonce);
      break;
    }
  }
}

__global__ void init_kernel(unsigned long long* best, uint64_t js) { *jptr(best, job_offset(js)) = ~0ull; }

}  // namespace grind

// Result written to *d_nonce (device memory, ~0 when no nonce below max_rounds * WIDTH).
// n_jobs / js: same-shape jobs of a batched launch, js bytes apart.
void grind_async(const uint32_t* d_digest, uint32_t pow_bits, uint64_t* d_nonce, uint32_t n_jobs, uint64_t js) {
  require(pow_bits <= 40, "grind: at most 40 bits of work");
  const uint64_t need = (1ull << pow_bits) * 64;  // ~e^-64 chance of no hit
  const uint32_t rounds = (uint32_t)std::max<uint64_t>(1, (need + grind::WIDTH - 1) / grind::WIDTH);
  timed("grind", [&] {
    grind::init_kernel<<<dim3(1, 1, n_jobs), 1, 0, ctx().cur>>>((unsigned long long*)d_nonce, js);
    grind::grind_kernel<<<dim3(grind::BLOCKS, 1, n_jobs), grind::NT, 0, ctx().cur>>>(
        d_digest, pow_bits, rounds, (unsigned long long*)d_nonce, js);
  });
}

}  // namespace sh

using namespace sh;

// The channel's tree hash of a word sequence (blake2s.hpp), on the host: for a Fiat-Shamir
// channel kept on the CPU next to the device backend, as stwo's is.
extern "C" int stwo_hip_host_tree_hash(const uint32_t* words, uint64_t n_words, uint32_t* out) {
  if (!out || (n_words && !words)) return STWO_HIP_ERR_ARG;
  std::vector<std::array<uint32_t, 8>> level((std::max<uint64_t>(n_words, 1) + 15) / 16);
  for (uint64_t i = 0; i < level.size(); i++) {
    uint32_t m[16] = {0};
    for (uint64_t j = 0; j < 16 && 16 * i + j < n_words; j++) m[j] = words[16 * i + j];
    uint32_t h[8];
    b2s::f0(m, h);
    std::copy(h, h + 8, level[i].begin());
  }
  while (level.size() > 1) {
    std::vector<std::array<uint32_t, 8>> up((level.size() + 1) / 2);
    for (size_t i = 0; i < up.size(); i++) {
      if (2 * i + 1 < level.size()) {
        uint32_t m[16], h[8];
        std::copy(level[2 * i].begin(), level[2 * i].end(), m);
        std::copy(level[2 * i + 1].begin(), level[2 * i + 1].end(), m + 8);
        b2s::f0(m, h);
        std::copy(h, h + 8, up[i].begin());
      } else {
        up[i] = level[2 * i];
      }
    }
    level.swap(up);
  }
  std::copy(level[0].begin(), level[0].end(), out);
  return STWO_HIP_OK;
}

extern "C" int stwo_hip_grind(const uint32_t* digest, uint32_t pow_bits, uint64_t* nonce) {
  return guarded([&] {
    require(digest != nullptr && nonce != nullptr, "grind: null argument");
    DevBuf d = upload(digest, 32);
    DevBuf out(8);
    grind_async(d.as<uint32_t>(), pow_bits, out.as<uint64_t>(), 1, 0);
    uint64_t r = 0;
    hip_check(hipMemcpyAsync(&r, out.p, 8, hipMemcpyDeviceToHost, ctx().cur), "hipMemcpy");
    sync();
    require(r != ~0ull, "grind: no nonce found within the search bound");
    *nonce = r;
  });
}
