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
#include "runtime.hpp"

namespace sh {
namespace wfib {

__device__ __forceinline__ uint64_t fold64(uint64_t s) { return (s & P) + (s >> 31); }

__global__ __launch_bounds__(256) void quotients_kernel(uint32_t n_rows, uint32_t n_cols,
                                                        const uint32_t* const* __restrict__ trace,
                                                        const uint4* __restrict__ powers,
                                                        const uint32_t* __restrict__ dinv, uint32_t* __restrict__ q0,
                                                        uint32_t* __restrict__ q1, uint32_t* __restrict__ q2,
                                                        uint32_t* __restrict__ q3) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  uint32_t a = trace[0][r], b = trace[1][r];
  uint32_t sa = mul(a, a), sb = mul(b, b);
  for (uint32_t i = 0; i + 2 < n_cols; i++) {
    uint32_t c = trace[i + 2][r];
    uint64_t cst = sub(add(sa, sb), c);
    uint4 pw = powers[n_cols - 3 - i];
    s0 += pw.x * cst;
    s1 += pw.y * cst;
    s2 += pw.z * cst;
    s3 += pw.w * cst;
    s0 = fold64(s0);
    s1 = fold64(s1);
    s2 = fold64(s2);
    s3 = fold64(s3);
    sa = sb;
    sb = mul(c, c);
  }
  uint32_t d = dinv[r];
  q0[r] = mul(reduce63(s0), d);
  q1[r] = mul(reduce63(s1), d);
  q2[r] = mul(reduce63(s2), d);
  q3[r] = mul(reduce63(s3), d);
}

}  // namespace wfib

void wide_fib_quotients_async(uint32_t* const q[4], const uint32_t* const* d_trace, const uint32_t* powers,
                              const uint32_t* dinv, uint32_t n_rows, uint32_t n_cols) {
  require(n_cols >= 3, "wide fibonacci: needs at least 3 columns");
  timed("wide_fib_quotients", [&] {
    wfib::quotients_kernel<<<ceil_div(n_rows, 256), 256, 0, ctx().cur>>>(
        n_rows, n_cols, d_trace, (const uint4*)powers, dinv, q[0], q[1], q[2], q[3]);
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
                             extended_domain_size, number_of_columns);
    sync();
  });
}
