// Batch inversion: FieldOps::batch_inverse (stwo_gpu_backend/src/field.rs:8-30).
//
// The reference builds a Montgomery product tree per 512/1024-element block in
// shared memory (cuda/src/batch_inverse.cu:24-125).  Here each thread runs
// Montgomery's trick over K strided elements in registers (coalesced loads,
// one field inversion per K elements, no LDS, no barriers).  For QM31 only the
// M31 norm of each element's CM31 denominator is batch-inverted: the QM31
// inverse (x - y u) / (x^2 - R y^2) needs one CM31 inverse, which needs one M31
// inverse (fields.cu:80-84, 138-144).
//
// Zero inputs: inverse(0) is defined as 0 and does not disturb its neighbours
// (the product chain skips it).
#include "field.hpp"
#include "runtime.hpp"

namespace sh {
namespace binv {

constexpr int K_M31 = 16, K_QM31 = 8;

__global__ __launch_bounds__(256) void m31_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                  uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t base = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint32_t v[K_M31], pre[K_M31];
  uint32_t acc = 1;
#pragma unroll
  for (int k = 0; k < K_M31; k++) {
    uint64_t i = base + k * stride;
    v[k] = i < n ? src[i] : 1u;
    pre[k] = acc;
    if (v[k]) acc = mul(acc, v[k]);
  }
  uint32_t inv_acc = inv(acc);
#pragma unroll
  for (int k = K_M31 - 1; k >= 0; k--) {
    uint64_t i = base + k * stride;
    uint32_t r = v[k] ? mul(inv_acc, pre[k]) : 0u;
    if (v[k]) inv_acc = mul(inv_acc, v[k]);
    if (i < n) dst[i] = r;
  }
}

__global__ __launch_bounds__(256) void qm31_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t base = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  qm31 v[K_QM31];
  cm31 den[K_QM31];
  uint32_t norm[K_QM31], pre[K_QM31];
  uint32_t acc = 1;
#pragma unroll
  for (int k = 0; k < K_QM31; k++) {
    uint64_t i = base + k * stride;
    uint4 w = i < n ? src[i] : make_uint4(1, 0, 0, 0);
    v[k] = {{w.x, w.y}, {w.z, w.w}};
    den[k] = csub(cmul(v[k].x, v[k].x), cmul_r(cmul(v[k].y, v[k].y)));  // x^2 - R y^2
    norm[k] = reduce63((uint64_t)den[k].a * den[k].a + (uint64_t)den[k].b * den[k].b);
    pre[k] = acc;
    if (norm[k]) acc = mul(acc, norm[k]);
  }
  uint32_t inv_acc = inv(acc);
#pragma unroll
  for (int k = K_QM31 - 1; k >= 0; k--) {
    uint64_t i = base + k * stride;
    uint32_t ni = norm[k] ? mul(inv_acc, pre[k]) : 0u;
    if (norm[k]) inv_acc = mul(inv_acc, norm[k]);
    cm31 di = {mul(den[k].a, ni), mul(neg(den[k].b), ni)};
    qm31 r = {cmul(v[k].x, di), cneg(cmul(v[k].y, di))};
    if (i < n) dst[i] = make_uint4(r.x.a, r.x.b, r.y.a, r.y.b);
  }
}

// 16 elements per thread as four 16-byte loads of 4 consecutive values (1 KiB per wave and
// load instead of 256 bytes); n a multiple of 4, 16-byte aligned columns
__global__ __launch_bounds__(256) void m31x4_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    uint64_t n4) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t base = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint32_t v[16], pre[16];
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const uint64_t q = base + g * stride;
    const uint4 w = q < n4 ? src[q] : make_uint4(1, 1, 1, 1);
    v[4 * g] = w.x;
    v[4 * g + 1] = w.y;
    v[4 * g + 2] = w.z;
    v[4 * g + 3] = w.w;
  }
  uint32_t acc = 1;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    pre[k] = acc;
    if (v[k]) acc = mul(acc, v[k]);
  }
  uint32_t inv_acc = inv(acc);
#pragma unroll
  for (int g = 3; g >= 0; g--) {
    uint32_t r[4];
#pragma unroll
    for (int j = 3; j >= 0; j--) {
      const int k = 4 * g + j;
      r[j] = v[k] ? mul(inv_acc, pre[k]) : 0u;
      if (v[k]) inv_acc = mul(inv_acc, v[k]);
    }
    const uint64_t q = base + g * stride;
    if (q < n4) dst[q] = make_uint4(r[0], r[1], r[2], r[3]);
  }
}

static uint32_t grid_for(uint64_t n, int k) {
  uint64_t per_block = 256ull * k;
  return (uint32_t)((n + per_block - 1) / per_block);
}

}  // namespace binv

void batch_inverse_m31(const uint32_t* src, uint32_t* dst, uint64_t n) {
  if (!n) return;
  timed("batch_inverse_m31", [&] {
    if (n % 4 == 0 && (((uintptr_t)src | (uintptr_t)dst) & 15) == 0)
      binv::m31x4_kernel<<<binv::grid_for(n / 4, 4), 256, 0, ctx().cur>>>((const uint4*)src, (uint4*)dst, n / 4);
    else
      binv::m31_kernel<<<binv::grid_for(n, binv::K_M31), 256, 0, ctx().cur>>>(src, dst, n);
  });
}
void batch_inverse_qm31(const uint32_t* src, uint32_t* dst, uint64_t n) {
  if (!n) return;
  timed("batch_inverse_qm31", [&] {
    binv::qm31_kernel<<<binv::grid_for(n, binv::K_QM31), 256, 0, ctx().cur>>>((const uint4*)src, (uint4*)dst, n);
  });
}

}  // namespace sh

using namespace sh;

extern "C" void batch_inverse_base_field(const uint32_t* from, uint32_t* dst, size_t size) {
  guarded([&] {
    batch_inverse_m31(from, dst, size);
    sync();
  });
}
extern "C" void batch_inverse_secure_field(const uint32_t* from, uint32_t* dst, size_t size) {
  guarded([&] {
    batch_inverse_qm31(from, dst, size);
    sync();
  });
}
