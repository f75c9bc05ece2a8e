// FRI folding: FriOps::fold_line / fold_circle_into_line (stwo_gpu_backend/src/fri.rs:15-66).
//
// fold_line (cuda/src/fri/fold_line.cu:5-34): for i < n/2,
//   out[i] = (f[2i] + f[2i+1]) + alpha (f[2i] - f[2i+1]) x_i^-1,
//   x_i^-1 = itwiddles[twiddle_offset + i]   (twiddle_offset = tree_len - n, fri.rs:24-26).
// fold_circle_into_line (cuda/src/fri/fold_circle_into_line.cu:7-46): for i < n/2,
//   f1 = (f[2i] - f[2i+1]) y_i^-1 with y_i^-1 = circle twiddle i of the tree part at
//   twiddle_offset (root.size - dst.size, fri.rs:56), and
//   dst[i] = dst[i] alpha^2 + (f[2i] + f[2i+1]) + alpha f1.
// Circle domains of 2 or 4 points (assumed canonic, as FRI builds them) take y from the
// domain directly: their tree part has no (x, y) pair to derive the circle twiddles from.
// Secure columns are SoA (four M31 columns); pointers travel by value.
#include "field.hpp"
#include "runtime.hpp"

namespace sh {
namespace fri {

struct Soa4 {
  uint32_t* c[4];
};

__device__ __forceinline__ qm31 ld(const Soa4& s, size_t i) { return {{s.c[0][i], s.c[1][i]}, {s.c[2][i], s.c[3][i]}}; }
__device__ __forceinline__ void st(const Soa4& s, size_t i, qm31 v) {
  s.c[0][i] = v.x.a;
  s.c[1][i] = v.x.b;
  s.c[2][i] = v.y.a;
  s.c[3][i] = v.y.b;
}

// alpha_dev (optional): read alpha from device memory (a challenge drawn on the device)
__global__ __launch_bounds__(256) void fold_line_kernel(const uint32_t* __restrict__ itw, uint32_t half_n,
                                                        Soa4 src, qm31 alpha, const qm31* alpha_dev, Soa4 dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half_n) return;
  if (alpha_dev) alpha = *alpha_dev;
  uint32_t xi = itw[i];
  qm31 a = ld(src, 2 * (size_t)i), b = ld(src, 2 * (size_t)i + 1);
  qm31 f0 = qadd(a, b), f1 = qscale(qsub(a, b), xi);
  st(dst, i, qadd(f0, qmul(alpha, f1)));
}

// accumulate = false: dst is overwritten with the fold (as if it held zeros)
__global__ __launch_bounds__(256) void fold_circle_kernel(const uint32_t* __restrict__ tw, uint32_t half_n,
                                                          Soa4 src, qm31 alpha, const qm31* alpha_dev,
                                                          bool accumulate, Soa4 dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half_n) return;
  if (alpha_dev) alpha = *alpha_dev;
  uint32_t yi;
  if (half_n >= 4) {
    // circle twiddle i: [y, -y, -x, x] over pairs (x, y) of the first line level
    uint32_t k2 = (i >> 2) << 1, sel = i & 3;
    uint32_t t = tw[k2 + ((sel < 2) ? 1 : 0)];
    yi = (sel == 1 || sel == 2) ? neg(t) : t;
  } else {
    // circle domains of 2 and 4 points: the tree's first line level has fewer than two
    // entries, so take y of the canonic domain point brev(2i) directly
    uint32_t log_n = half_n == 1 ? 1 : 2;
    uint32_t idx = circle_domain_index(1u << (30 - log_n), (1u << (32 - log_n)) & 0x7fffffffu, log_n - 1,
                                       brev(2 * i, log_n));
    yi = inv(point_at_index(idx).y);
  }
  qm31 a = ld(src, 2 * (size_t)i), b = ld(src, 2 * (size_t)i + 1);
  qm31 f0 = qadd(a, b), f1 = qscale(qsub(a, b), yi);
  qm31 r = qadd(f0, qmul(alpha, f1));
  if (accumulate) r = qadd(qmul(ld(dst, i), qsqr(alpha)), r);
  st(dst, i, r);
}

static Soa4 soa(uint32_t* const* p) { return Soa4{{p[0], p[1], p[2], p[3]}}; }

}  // namespace fri

void fold_line_async(const uint32_t* itw, size_t twiddle_offset, size_t n, uint32_t* const* src, qm31 alpha,
                     uint32_t* const* dst, const qm31* alpha_dev) {
  require(n >= 2 && is_pow2(n) && n <= (1ull << 31), "fold_line: n must be a power of two >= 2");
  uint32_t half = (uint32_t)(n >> 1);
  timed("fold_line", [&] {
    fri::fold_line_kernel<<<ceil_div(half, 256), 256, 0, ctx().cur>>>(itw + twiddle_offset, half, fri::soa(src),
                                                                           alpha, alpha_dev, fri::soa(dst));
  });
}
void fold_circle_into_line_async(const uint32_t* itw, size_t twiddle_offset, size_t n, uint32_t* const* src,
                                 qm31 alpha, uint32_t* const* dst, const qm31* alpha_dev, bool accumulate) {
  require(n >= 2 && is_pow2(n) && n <= (1ull << 31), "fold_circle_into_line: n must be a power of two >= 2");
  uint32_t half = (uint32_t)(n >> 1);
  timed("fold_circle_into_line", [&] {
    fri::fold_circle_kernel<<<ceil_div(half, 256), 256, 0, ctx().cur>>>(
        itw + twiddle_offset, half, fri::soa(src), alpha, alpha_dev, accumulate, fri::soa(dst));
  });
}

}  // namespace sh

using namespace sh;

extern "C" int stwo_hip_fold_line_async(const uint32_t* itwiddles, size_t twiddle_offset, size_t n,
                                        uint32_t* const* eval_values, StwoQM31 alpha, uint32_t* const* folded_values) {
  return guarded([&] {
    fold_line_async(itwiddles, twiddle_offset, n, eval_values, qm31{{alpha.a, alpha.b}, {alpha.c, alpha.d}},
                    folded_values, nullptr);
  });
}

extern "C" int stwo_hip_fold_circle_into_line_async(const uint32_t* itwiddles, size_t twiddle_offset, size_t n,
                                                    uint32_t* const* eval_values, StwoQM31 alpha,
                                                    uint32_t* const* folded_values) {
  return guarded([&] {
    fold_circle_into_line_async(itwiddles, twiddle_offset, n, eval_values,
                                qm31{{alpha.a, alpha.b}, {alpha.c, alpha.d}}, folded_values, nullptr, true);
  });
}

extern "C" void fold_line(const uint32_t* gpu_domain, size_t twiddle_offset, size_t n, uint32_t* const* eval_values,
                          StwoQM31 alpha, uint32_t* const* folded_values) {
  guarded([&] {
    fold_line_async(gpu_domain, twiddle_offset, n, eval_values, qm31{{alpha.a, alpha.b}, {alpha.c, alpha.d}},
                    folded_values, nullptr);
    sync();
  });
}

extern "C" void fold_circle_into_line(const uint32_t* gpu_domain, size_t twiddle_offset, size_t n,
                                      uint32_t* const* eval_values, StwoQM31 alpha, uint32_t* const* folded_values) {
  guarded([&] {
    fold_circle_into_line_async(gpu_domain, twiddle_offset, n, eval_values,
                                qm31{{alpha.a, alpha.b}, {alpha.c, alpha.d}}, folded_values, nullptr, true);
    sync();
  });
}
