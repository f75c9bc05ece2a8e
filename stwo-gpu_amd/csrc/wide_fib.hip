// Wide-Fibonacci composition on the extended domain:
// ComponentProver<CudaBackend>::evaluate_constraint_quotients_on_domain
// (stwo_gpu_backend/src/examples/wide_fibonacci.rs:90-141) via
// cuda/src/examples/wide_fibonacci.cu:3-66.
//
// For each row: sum_{i=0}^{n-3} (t_i^2 + t_{i+1}^2 - t_{i+2}) * powers[n-3-i], times the
// row's vanishing-polynomial inverse, written to the four coordinate columns.
// Unlike the reference launch (grid and block swapped, no bounds check,
// wide_fibonacci.cu:54), every row is covered exactly once; the constraint sum is
// accumulated in 64-bit lazy sums per coordinate.
#include <vector>

#include "field.hpp"
#include "gmem.hpp"
#include "runtime.hpp"

namespace sh {
namespace wfib {

__device__ __forceinline__ uint64_t fold64(uint64_t s) { return (s & P) + (s >> 31); }

constexpr int WG = 6;  // columns per load group (a multiple of the fold period 3)

// js: byte stride between the jobs of a batched launch (blockIdx.z = job; 0 for one job)
__global__ __launch_bounds__(256) void quotients_kernel(uint32_t n_rows, uint32_t n_cols,
                                                        const uint32_t* const* __restrict__ trace,
                                                        const uint4* __restrict__ powers,
                                                        const uint32_t* __restrict__ dinv, uint32_t* __restrict__ q0,
                                                        uint32_t* __restrict__ q1, uint32_t* __restrict__ q2,
                                                        uint32_t* __restrict__ q3, uint64_t js) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  const uint64_t jo = job_offset(js);
  powers = jptr(powers, jo);
  dinv = jptr(dinv, jo);
  q0 = jptr(q0, jo);
  q1 = jptr(q1, jo);
  q2 = jptr(q2, jo);
  q3 = jptr(q3, jo);
  uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  uint32_t sa = sqr(at(as_global(jptr(trace[0], jo)), r)), sb = sqr(at(as_global(jptr(trace[1], jo)), r));
  // constraint i reads column i + 2; columns are loaded one group ahead of the arithmetic,
  // unconditionally (indices past the last column re-read it and are never used), so the
  // arithmetic waits only for the older group
  const uint32_t n_cons = n_cols - 2, last = n_cols - 1;
  uint32_t cv[WG], cn[WG];
#pragma unroll
  for (int g = 0; g < WG; g++) cv[g] = at(as_global(jptr(trace[min(2u + g, last)], jo)), r);
  for (uint32_t i = 0; i < n_cons; i += WG) {
#pragma unroll
    for (int g = 0; g < WG; g++) cn[g] = at(as_global(jptr(trace[min(2u + i + WG + g, last)], jo)), r);
#pragma unroll
    for (int g = 0; g < WG; g++) {
      if (i + g >= n_cons) break;
      const uint32_t c = cv[g];
      const uint64_t cst = sub(add(sa, sb), c);
      const uint4 pw = powers[n_cons - 1 - (i + g)];
      s0 += pw.x * cst;
      s1 += pw.y * cst;
      s2 += pw.z * cst;
      s3 += pw.w * cst;
      if (g % 3 == 2) {  // at most three products between folds
        s0 = fold64(s0);
        s1 = fold64(s1);
        s2 = fold64(s2);
        s3 = fold64(s3);
      }
      sa = sb;
      sb = sqr(c);
    }
#pragma unroll
    for (int g = 0; g < WG; g++) cv[g] = cn[g];
  }
  uint32_t d = dinv[r];
  q0[r] = mul(reduce63(s0), d);
  q1[r] = mul(reduce63(s1), d);
  q2[r] = mul(reduce63(s2), d);
  q3[r] = mul(reduce63(s3), d);
}

// Trace of a wide-Fibonacci instance (the AIR of examples/wide_fibonacci.rs): row r has
// col[0] = U(seed, r), col[1] = U(seed, n + r) (uniform_m31) and
// col[i] = col[i-2]^2 + col[i-1]^2.  One lane per row; every column store is coalesced.
__global__ __launch_bounds__(256) void trace_kernel(uint32_t* const* __restrict__ cols, uint32_t n_cols,
                                                    uint32_t log_n, const uint64_t* __restrict__ seed_dev, uint64_t js) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = 1u << log_n;
  if (r >= n) return;
  const uint64_t jo = job_offset(js);
  const uint64_t seed = *jptr(seed_dev, jo);
  uint32_t a = uniform_m31(seed, r), b = uniform_m31(seed, (uint64_t)n + r);
  at(as_global(jptr(cols[0], jo)), r) = a;
  at(as_global(jptr(cols[1], jo)), r) = b;
  for (uint32_t i = 2; i < n_cols; i++) {
    const uint32_t c = add(sqr(a), sqr(b));
    at(as_global(jptr(cols[i], jo)), r) = c;
    a = b;
    b = c;
  }
}

// 1 / Z(p) on the rows of CanonicCoset(log_ext) (bit-reversed order), where
// Z(p) = pi^(log_n - 1)(p.x), pi(x) = 2x^2 - 1, vanishes exactly on CanonicCoset(log_n).
__global__ __launch_bounds__(256) void vanishing_inv_kernel(uint32_t* __restrict__ out, uint32_t log_ext,
                                                            uint32_t log_n, uint64_t js) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (1u << log_ext)) return;
  out = jptr(out, job_offset(js));
  const uint32_t hi = 1u << (30 - log_ext), hs = (1u << (32 - log_ext)) & 0x7fffffffu;
  uint32_t x = point_at_index(circle_domain_index(hi, hs, log_ext - 1, brev(r, log_ext))).x;
  for (uint32_t i = 1; i < log_n; i++) x = sub(add(sqr(x), sqr(x)), 1);
  out[r] = inv(x);
}

// powers[k] = alpha^k (QM31 AoS), alpha read from device memory
__global__ __launch_bounds__(256) void powers_kernel(const qm31* __restrict__ alpha, uint32_t n, uint4* powers,
                                                     uint64_t js) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t jo = job_offset(js);
  const qm31 v = qpow(*jptr(alpha, jo), k);
  jptr(powers, jo)[k] = make_uint4(v.x.a, v.x.b, v.y.a, v.y.b);
}

}  // namespace wfib

// n_jobs / js: same-shape jobs of a batched launch, js bytes apart (see jptr)
void wide_fib_trace_async(uint32_t* const* d_cols, uint32_t n_cols, uint32_t log_n, const uint64_t* d_seed,
                          uint32_t n_jobs, uint64_t js) {
  require(n_cols >= 2 && log_n >= 1 && log_n <= 30, "wide fibonacci trace: bad sizes");
  timed("wide_fib_trace", [&] {
    wfib::trace_kernel<<<dim3(ceil_div(1ull << log_n, 256), 1, n_jobs), 256, 0, ctx().cur>>>(d_cols, n_cols, log_n,
                                                                                            d_seed, js);
  });
}
void vanishing_inverse_async(uint32_t* d_out, uint32_t log_ext, uint32_t log_n, uint32_t n_jobs, uint64_t js) {
  require(log_n >= 1 && log_ext > log_n && log_ext <= 30, "vanishing inverse: bad sizes");
  timed("vanishing_inverse", [&] {
    wfib::vanishing_inv_kernel<<<dim3(ceil_div(1ull << log_ext, 256), 1, n_jobs), 256, 0, ctx().cur>>>(
        d_out, log_ext, log_n, js);
  });
}
void qm31_powers_async(const qm31* d_alpha, uint32_t n, uint32_t* d_powers, uint32_t n_jobs, uint64_t js) {
  timed("qm31_powers", [&] {
    wfib::powers_kernel<<<dim3(ceil_div(n, 256), 1, n_jobs), 256, 0, ctx().cur>>>(d_alpha, n, (uint4*)d_powers, js);
  });
}

void wide_fib_quotients_async(uint32_t* const q[4], const uint32_t* const* d_trace, const uint32_t* powers,
                              const uint32_t* dinv, uint32_t n_rows, uint32_t n_cols, uint32_t n_jobs, uint64_t js) {
  require(n_cols >= 3, "wide fibonacci: needs at least 3 columns");
  timed("wide_fib_quotients", [&] {
    wfib::quotients_kernel<<<dim3(ceil_div(n_rows, 256), 1, n_jobs), 256, 0, ctx().cur>>>(
        n_rows, n_cols, d_trace, (const uint4*)powers, dinv, q[0], q[1], q[2], q[3], js);
  });
}

}  // namespace sh

using namespace sh;

extern "C" void evaluate_wide_fibonacci_constraint_quotients_on_domain(
    uint32_t* quotients_0, uint32_t* quotients_1, uint32_t* quotients_2, uint32_t* quotients_3,
    uint32_t* const* trace_evaluations, const uint32_t* random_coeff_powers, const uint32_t* denominator_inverses,
    uint32_t extended_domain_size, uint32_t number_of_columns) {
  guarded([&] {
    if (!extended_domain_size) return;
    DevBuf table = upload(trace_evaluations, (size_t)number_of_columns * sizeof(void*));
    uint32_t* q[4] = {quotients_0, quotients_1, quotients_2, quotients_3};
    wide_fib_quotients_async(q, table.as<const uint32_t*>(), random_coeff_powers, denominator_inverses,
                             extended_domain_size, number_of_columns, 1, 0);
    sync();
  });
}

extern "C" int stwo_hip_wide_fib_trace(uint32_t* const* columns, uint32_t n_cols, uint32_t log_n, uint64_t seed) {
  return guarded([&] {
    require(columns != nullptr, "wide fibonacci trace: null columns");
    DevBuf table = upload(columns, (size_t)n_cols * sizeof(void*));
    DevBuf s = upload(&seed, sizeof(seed));
    wide_fib_trace_async(table.as<uint32_t*>(), n_cols, log_n, s.as<uint64_t>(), 1, 0);
  });
}

extern "C" int stwo_hip_vanishing_inverse(uint32_t* out, uint32_t log_ext, uint32_t log_n) {
  return guarded([&] { vanishing_inverse_async(out, log_ext, log_n, 1, 0); });
}
