// PolyOps::eval_at_point (stwo_gpu_backend/src/poly.rs:84-94).
//
// For coefficients c[0 .. 2^L) and a QM31 point (x, y), the factors are
// f = [pi^(L-2)(x), ..., pi(x), x, y] with pi(x) = 2x^2 - 1
// (cuda/src/poly/eval_at_point.cu:99-110), and the value is the fold
//   fold(c, f) = fold(lo half, f[1:]) + fold(hi half, f[1:]) * f[0],
// i.e. bit b of a coefficient index selects factor f[L-1-b].
//
// Two launches: blocks reduce 2^11 coefficients each in LDS (bits 0..10), then
// one block reduces the block results (remaining bits).  The reference's shared
// memory aliasing (eval_at_point.cu:17-18) has no counterpart here: separate
// arrays, and every level reads into registers before a barrier and writes after.
#include <utility>
#include <vector>

#include "field.hpp"
#include "runtime.hpp"

namespace sh {
namespace eap {

constexpr int CB = 11;  // coefficients per first-pass block = 2^CB
constexpr int NT = 256;

struct Factors {
  qm31 f[31];
};

// Reduce `cnt` level values in s (cnt power of two) with factors f[fi], f[fi-1], ...
__device__ void tree_reduce(qm31* s, uint32_t cnt, const Factors& F, int fi) {
  while (cnt > 1) {
    uint32_t half = cnt >> 1;
    qm31 r[4];
    uint32_t k = 0;
    for (uint32_t i = threadIdx.x; i < half; i += NT, k++) r[k & 3] = qadd(s[2 * i], qmul(s[2 * i + 1], F.f[fi]));
    __syncthreads();
    k = 0;
    for (uint32_t i = threadIdx.x; i < half; i += NT, k++) s[i] = r[k & 3];
    __syncthreads();
    cnt = half;
    fi--;
  }
}

// First pass: block b folds coefficients [b 2^cb, (b+1) 2^cb) with f[L-1], ..., f[L-cb].
__global__ __launch_bounds__(NT) void first_pass(const uint32_t* __restrict__ c, uint32_t L, uint32_t cb, Factors F,
                                                 qm31* partial) {
  __shared__ qm31 s[1 << (CB - 1)];
  const uint32_t pairs = 1u << (cb - 1);
  const size_t base = (size_t)blockIdx.x << cb;
  const qm31 fy = F.f[L - 1];
  for (uint32_t i = threadIdx.x; i < pairs; i += NT) {
    uint32_t c0 = c[base + 2 * i], c1 = c[base + 2 * i + 1];
    qm31 v = qscale(fy, c1);
    v.x.a = add(v.x.a, c0);
    s[i] = v;
  }
  __syncthreads();
  tree_reduce(s, pairs, F, (int)L - 2);
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// Later passes: block b folds partials [b 2^g, (b+1) 2^g) with f[fi], f[fi-1], ...
__global__ __launch_bounds__(NT) void reduce_pass(const qm31* in, uint32_t g, Factors F, int fi, qm31* out) {
  __shared__ qm31 s[1 << (CB - 1)];
  const uint32_t cnt = 1u << g;
  for (uint32_t i = threadIdx.x; i < cnt; i += NT) s[i] = in[((size_t)blockIdx.x << g) + i];
  __syncthreads();
  tree_reduce(s, cnt, F, fi);
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// ---- many columns of one size at one point held in device memory ----
__global__ void factors_kernel(const qm31* __restrict__ point, uint32_t L, Factors* F) {
  qm31 x = point[0];
  F->f[L - 1] = point[1];
  for (uint32_t i = 2; i <= L; i++) {
    F->f[L - i] = x;
    x = qsub(qadd(qsqr(x), qsqr(x)), qone());
  }
}

// block (b, col) folds coefficients [b 2^cb, (b+1) 2^cb) of column col
__global__ __launch_bounds__(NT) void first_pass_cols(const uint32_t* const* __restrict__ cols, uint32_t L,
                                                      uint32_t cb, const Factors* __restrict__ F, qm31* partial) {
  __shared__ qm31 s[1 << (CB - 1)];
  const uint32_t* c = cols[blockIdx.y];
  const uint32_t pairs = 1u << (cb - 1);
  const size_t base = (size_t)blockIdx.x << cb;
  const qm31 fy = F->f[L - 1];
  for (uint32_t i = threadIdx.x; i < pairs; i += NT) {
    uint32_t c0 = c[base + 2 * i], c1 = c[base + 2 * i + 1];
    qm31 v = qscale(fy, c1);
    v.x.a = add(v.x.a, c0);
    s[i] = v;
  }
  __syncthreads();
  tree_reduce(s, pairs, *F, (int)L - 2);
  if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = s[0];
}

// Columns of log >= 12: block (b, col) folds coefficients [b 2^12, (b+1) 2^12) of column
// col; each lane folds 16 consecutive coefficients in registers (bits 0..3: factors
// f[L-1] .. f[L-4]) from four 16-byte loads, then the 256 lane values are folded in LDS.
__global__ __launch_bounds__(NT) void first_pass_cols16(const uint32_t* const* __restrict__ cols, uint32_t L,
                                                        const Factors* __restrict__ F, qm31* partial) {
  __shared__ qm31 s[NT];
  const uint4* c = reinterpret_cast<const uint4*>(cols[blockIdx.y] + ((size_t)blockIdx.x << 12) + 16 * threadIdx.x);
  const uint4 q0 = c[0], q1 = c[1], q2 = c[2], q3 = c[3];
  const uint32_t v[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                          q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
  const qm31 f0 = F->f[L - 1], f1 = F->f[L - 2], f2 = F->f[L - 3], f3 = F->f[L - 4];
  qm31 a[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    a[j] = qscale(f0, v[2 * j + 1]);
    a[j].x.a = add(a[j].x.a, v[2 * j]);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) a[j] = qadd(a[2 * j], qmul(a[2 * j + 1], f1));
#pragma unroll
  for (int j = 0; j < 2; j++) a[j] = qadd(a[2 * j], qmul(a[2 * j + 1], f2));
  s[threadIdx.x] = qadd(a[0], qmul(a[1], f3));
  __syncthreads();
  tree_reduce(s, NT, *F, (int)L - 5);
  if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = s[0];
}

__global__ __launch_bounds__(NT) void reduce_pass_dev(const qm31* in, uint32_t g, const Factors* __restrict__ F,
                                                      int fi, qm31* out) {
  __shared__ qm31 s[1 << (CB - 1)];
  const uint32_t cnt = 1u << g;
  for (uint32_t i = threadIdx.x; i < cnt; i += NT) s[i] = in[((size_t)blockIdx.x << g) + i];
  __syncthreads();
  tree_reduce(s, cnt, *F, fi);
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

}  // namespace eap

size_t eval_at_point_columns_scratch_bytes(uint32_t n_cols, uint32_t L) {
  const uint32_t cb = L >= 12 ? 12 : std::min<uint32_t>(L, eap::CB);
  return 512 + 2 * sizeof(qm31) * (size_t)n_cols * (1ull << (L - cb));
}

// d_out[c] = value at d_point (x, y) of the polynomial with coefficients d_cols[c] (2^L each).
void eval_at_point_columns_async(const uint32_t* const* d_cols, uint32_t n_cols, uint32_t L, const qm31* d_point,
                                 qm31* d_out, void* d_scratch) {
  require(L >= 1 && L <= 31 && n_cols >= 1 && n_cols <= 65535, "eval_at_point_columns: bad sizes");
  eap::Factors* F = (eap::Factors*)d_scratch;
  const uint32_t cb = L >= 12 ? 12 : std::min<uint32_t>(L, eap::CB);
  uint32_t blocks = 1u << (L - cb);
  qm31* pa = (qm31*)((char*)d_scratch + 512);
  qm31* pb = pa + (size_t)n_cols * blocks;
  timed("eval_at_point", [&] {
    eap::factors_kernel<<<1, 1, 0, ctx().cur>>>(d_point, L, F);
    if (cb == 12)
      eap::first_pass_cols16<<<dim3(blocks, n_cols), eap::NT, 0, ctx().cur>>>(d_cols, L, F, blocks == 1 ? d_out : pa);
    else
      eap::first_pass_cols<<<dim3(blocks, n_cols), eap::NT, 0, ctx().cur>>>(d_cols, L, cb, F,
                                                                             blocks == 1 ? d_out : pa);
  });
  int fi = (int)(L - cb) - 1;
  while (blocks > 1) {
    const uint32_t g = std::min<uint32_t>(log2_exact(blocks), eap::CB - 1);
    const uint32_t nb = blocks >> g;
    qm31* dst = nb == 1 ? d_out : pb;
    timed("eval_at_point", [&] {
      eap::reduce_pass_dev<<<nb * n_cols, eap::NT, 0, ctx().cur>>>(pa, g, F, fi, dst);
    });
    std::swap(pa, pb);
    fi -= (int)g;
    blocks = nb;
  }
}

qm31 eval_at_point_sync(const uint32_t* coeffs, uint64_t n, qm31 px, qm31 py) {
  require(is_pow2(n) && n <= (1ull << 31), "eval_at_point: size must be a power of two");
  uint32_t L = log2_exact(n);
  if (L == 0) {
    uint32_t c0;
    hip_check(hipMemcpyAsync(&c0, coeffs, 4, hipMemcpyDeviceToHost, ctx().stream), "hipMemcpy");
    sync();
    return qm31{{c0, 0}, {0, 0}};
  }
  eap::Factors F{};
  F.f[L - 1] = py;
  qm31 x = px;
  for (uint32_t i = 2; i <= L; i++) {
    F.f[L - i] = x;
    x = qsub(qadd(qsqr(x), qsqr(x)), qone());
  }
  uint32_t cb = std::min<uint32_t>(L, eap::CB);
  uint32_t blocks = 1u << (L - cb);
  DevBuf pa(sizeof(qm31) * blocks), pb(sizeof(qm31) * blocks);
  timed("eval_at_point", [&] {
    eap::first_pass<<<blocks, eap::NT, 0, ctx().cur>>>(coeffs, L, cb, F, pa.as<qm31>());
  });
  int fi = (int)(L - cb) - 1;
  while (blocks > 1) {
    uint32_t lb = log2_exact(blocks);
    uint32_t g = std::min<uint32_t>(lb, eap::CB - 1);
    uint32_t nb = blocks >> g;
    timed("eval_at_point", [&] {
      eap::reduce_pass<<<nb, eap::NT, 0, ctx().cur>>>(pa.as<qm31>(), g, F, fi, pb.as<qm31>());
    });
    std::swap(pa.p, pb.p);
    fi -= (int)g;
    blocks = nb;
  }
  qm31 r;
  hip_check(hipMemcpyAsync(&r, pa.p, sizeof(qm31), hipMemcpyDeviceToHost, ctx().stream), "hipMemcpy");
  sync();
  return r;
}

}  // namespace sh

using namespace sh;

extern "C" StwoQM31 eval_at_point(const uint32_t* coeffs, uint32_t coeffs_size, StwoQM31 point_x, StwoQM31 point_y) {
  StwoQM31 out{0, 0, 0, 0};
  guarded([&] {
    qm31 r = eval_at_point_sync(coeffs, coeffs_size, qm31{{point_x.a, point_x.b}, {point_x.c, point_x.d}},
                                qm31{{point_y.a, point_y.b}, {point_y.c, point_y.d}});
    out = StwoQM31{r.x.a, r.x.b, r.y.a, r.y.b};
  });
  return out;
}
