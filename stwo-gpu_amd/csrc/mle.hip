// Multilinear-extension ops the reference crate binds for its lookups:
// GkrOps::gen_eq_evals (stwo_gpu_backend/src/lookups/gkr.rs:12-27 ->
// cuda/src/gkr.cu:4-38) and MleOps::fix_first_variable for M31 and QM31
// evaluations (lookups/mle.rs:12-57 -> cuda/src/mle.cu:4-30).
//
// Variable order: variable 0 is the most significant bit of the evaluation
// index, so fixing the first variable pairs evals[i] with evals[i + n/2], and
// gen_eq_evals(y, v)[x] = v * prod_k (y_k if bit (n-1-k) of x else 1 - y_k).
//
// gen_eq_evals: the reference multiplies y_size factors per output.  Here a
// block writes 2^LO consecutive outputs that share their high index bits: the
// product over the high variables is formed once per block (v times one factor
// per high bit), the low 2^LO products are built in LDS by doubling (one QM31
// product per table entry), and each output is one product of the two.  The
// kernel is bound by its 16-byte-per-output stores.
#include "field.hpp"
#include "runtime.hpp"

namespace sh {
namespace mle {

constexpr int LO = 10;  // low variables per block table (1024 entries, 16 KiB of LDS)

struct EqFactors {
  qm31w y[32];  // y_0 .. y_{n-1}
};

__global__ __launch_bounds__(256) void gen_eq_kernel(EqFactors f, uint32_t n_vars, qm31 v, qm31w* __restrict__ out) {
  __shared__ qm31 lo[1 << LO];
  const uint32_t lo_vars = n_vars < (uint32_t)LO ? n_vars : (uint32_t)LO;
  const uint32_t hi_vars = n_vars - lo_vars;
  // high product: variables 0 .. hi_vars-1 from the bits of the block index (MSB = y_0)
  qm31 hi = v;
  for (uint32_t k = 0; k < hi_vars; k++) {
    const qm31 y = from_w(f.y[k]);
    const bool bit = (blockIdx.x >> (hi_vars - 1 - k)) & 1u;
    hi = qmul(hi, bit ? y : qsub(qone(), y));
  }
  // low table by doubling: after step s the table holds 2^s products over the low
  // variables hi_vars .. hi_vars+s-1, index bits in the same MSB-first order
  if (threadIdx.x == 0) lo[0] = hi;
  __syncthreads();
  for (uint32_t s = 0; s < lo_vars; s++) {
    const qm31 y = from_w(f.y[hi_vars + s]);
    const qm31 y0 = qsub(qone(), y);
    const uint32_t size = 1u << s;
    // read phase then write phase: entry j becomes entries 2j (bit 0) and 2j+1 (bit 1)
    qm31 a[4];
    const uint32_t per = (size + blockDim.x - 1) / blockDim.x;  // <= 4 for LO = 10 and 256 threads
    for (uint32_t r = 0; r < per; r++) {
      const uint32_t j = threadIdx.x + r * blockDim.x;
      if (j < size) a[r] = lo[j];
    }
    __syncthreads();
    for (uint32_t r = 0; r < per; r++) {
      const uint32_t j = threadIdx.x + r * blockDim.x;
      if (j < size) {
        lo[2 * j] = qmul(a[r], y0);
        lo[2 * j + 1] = qmul(a[r], y);
      }
    }
    __syncthreads();
  }
  const uint64_t base = (uint64_t)blockIdx.x << lo_vars;
  for (uint32_t j = threadIdx.x; j < (1u << lo_vars); j += blockDim.x) out[base + j] = to_w(lo[j]);
}

// out[i] = lhs + (rhs - lhs) a with lhs = evals[i], rhs = evals[i + half] (mle.cu:4-12)
__global__ __launch_bounds__(256) void fix_first_base_kernel(const uint32_t* __restrict__ evals, uint64_t half,
                                                             qm31 a, qm31w* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half) return;
  const uint32_t l = evals[i], r = evals[i + half];
  qm31 t = qscale(a, sub(r, l));
  t.x.a = add(t.x.a, l);
  out[i] = to_w(t);
}
__global__ __launch_bounds__(256) void fix_first_secure_kernel(const qm31w* __restrict__ evals, uint64_t half,
                                                               qm31 a, qm31w* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half) return;
  const qm31 l = from_w(evals[i]), r = from_w(evals[i + half]);
  out[i] = to_w(qadd(l, qmul(qsub(r, l), a)));
}

}  // namespace mle
}  // namespace sh

using namespace sh;

extern "C" void gen_eq_evals(StwoQM31 v, const StwoQM31* y, uint32_t y_size, StwoQM31* evals, uint32_t evals_size) {
  guarded([&] {
    require(y_size <= 31, "gen_eq_evals: at most 31 variables");
    require(evals_size == (1ull << y_size), "gen_eq_evals: evals_size must be 2^y_size");
    require(evals != nullptr && (y_size == 0 || y != nullptr), "gen_eq_evals: null argument");
    mle::EqFactors f{};
    for (uint32_t k = 0; k < y_size; k++) f.y[k] = {y[k].a % P, y[k].b % P, y[k].c % P, y[k].d % P};
    const qm31 vv = {{v.a % P, v.b % P}, {v.c % P, v.d % P}};
    const uint32_t lo_vars = y_size < (uint32_t)mle::LO ? y_size : (uint32_t)mle::LO;
    const uint32_t blocks = 1u << (y_size - lo_vars);
    timed("gen_eq_evals", [&] {
      mle::gen_eq_kernel<<<blocks, 256, 0, ctx().cur>>>(f, y_size, vv, reinterpret_cast<qm31w*>(evals));
    });
    sync();
  });
}

extern "C" void fix_first_variable_base_field(const uint32_t* evals, size_t evals_size, StwoQM31 assignment,
                                              uint32_t* output_evals) {
  guarded([&] {
    require(evals_size >= 2 && evals_size % 2 == 0 && evals_size <= (1ull << 32),
            "fix_first_variable: evals_size must be even, 2 .. 2^32");
    require(evals != nullptr && output_evals != nullptr, "fix_first_variable: null argument");
    const uint64_t half = evals_size / 2;
    const qm31 a = {{assignment.a % P, assignment.b % P}, {assignment.c % P, assignment.d % P}};
    timed("fix_first_variable", [&] {
      mle::fix_first_base_kernel<<<ceil_div(half, 256), 256, 0, ctx().cur>>>(
          evals, half, a, reinterpret_cast<qm31w*>(output_evals));
    });
    sync();
  });
}

extern "C" void fix_first_variable_secure_field(const uint32_t* evals, size_t evals_size, StwoQM31 assignment,
                                                uint32_t* output_evals) {
  guarded([&] {
    require(evals_size >= 2 && evals_size % 2 == 0 && evals_size <= (1ull << 32),
            "fix_first_variable: evals_size must be even, 2 .. 2^32");
    require(evals != nullptr && output_evals != nullptr, "fix_first_variable: null argument");
    const uint64_t half = evals_size / 2;
    const qm31 a = {{assignment.a % P, assignment.b % P}, {assignment.c % P, assignment.d % P}};
    timed("fix_first_variable", [&] {
      mle::fix_first_secure_kernel<<<ceil_div(half, 256), 256, 0, ctx().cur>>>(
          reinterpret_cast<const qm31w*>(evals), half, a, reinterpret_cast<qm31w*>(output_evals));
    });
    sync();
  });
}
