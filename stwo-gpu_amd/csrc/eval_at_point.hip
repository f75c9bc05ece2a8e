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
#include "gmem.hpp"
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
// js (here and below): byte stride between the jobs of a batched launch (blockIdx.z = job)
__global__ void factors_kernel(const qm31* __restrict__ point, uint32_t L, Factors* F, uint64_t js) {
  const uint64_t jo = job_offset(js);
  point = jptr(point, jo);
  F = jptr(F, jo);
  qm31 x = point[0];
  F->f[L - 1] = point[1];
  for (uint32_t i = 2; i <= L; i++) {
    F->f[L - i] = x;
    x = qsub(qadd(qsqr(x), qsqr(x)), qone());
  }
}

// block (b, col) folds coefficients [b 2^cb, (b+1) 2^cb) of column col
__global__ __launch_bounds__(NT) void first_pass_cols(const uint32_t* const* __restrict__ cols, uint32_t L,
                                                      uint32_t cb, const Factors* __restrict__ F, qm31* partial,
                                                      uint64_t js) {
  __shared__ qm31 s[1 << (CB - 1)];
  const uint64_t jo = job_offset(js);
  F = jptr(F, jo);
  partial = jptr(partial, jo);
  const uint32_t* c = jptr(cols[blockIdx.y], jo);
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

// ---- columns of log >= 12 as a matrix-vector product -------------------------------
// The value at the point is sum_i c_i B_i with B_i the product of the factors selected by
// the bits of i.  Split i = 4096 h + j: B_i = Lo[j] Hi[h], Lo over the low 12 bits and Hi
// over the rest (two small tables per point).  A block takes 4096 coefficients (one h) of
// a group of columns; a lane multiplies its 16 coefficients by its 16 Lo values held in
// registers (M31 x QM31 products accumulated in 64 bits), the block sums the lanes, and
// the sum times Hi[h] is the block's partial; partials are summed per column afterwards.
// Every coefficient is read once; about 12 VALU instructions per coefficient.
constexpr uint32_t LO_BITS = 12, GEMV_COLS = 8;

__global__ __launch_bounds__(256) void basis_tables_kernel(const Factors* __restrict__ F, uint32_t L, qm31* lo,
                                                           qm31* hi, uint64_t js) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t jo = job_offset(js);
  F = jptr(F, jo);
  lo = jptr(lo, jo);
  hi = jptr(hi, jo);
  const uint32_t n_lo = 1u << LO_BITS, n_hi = 1u << (L - LO_BITS);
  if (t < n_lo) {
    qm31 v = qone();
    for (uint32_t b = 0; b < LO_BITS; b++)
      if ((t >> b) & 1) v = qmul(v, F->f[L - 1 - b]);
    lo[t] = v;
  } else if (t < n_lo + n_hi) {
    const uint32_t h = t - n_lo;
    qm31 v = qone();
    for (uint32_t b = 0; b < L - LO_BITS; b++)
      if ((h >> b) & 1) v = qmul(v, F->f[L - 1 - LO_BITS - b]);
    hi[h] = v;
  }
}

__device__ __forceinline__ uint32_t addm(uint32_t a, uint32_t b) { return add(a, b); }
__device__ __forceinline__ qm31 wave_sum(qm31 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v.x.a = addm(v.x.a, (uint32_t)__shfl_xor((int)v.x.a, o));
    v.x.b = addm(v.x.b, (uint32_t)__shfl_xor((int)v.x.b, o));
    v.y.a = addm(v.y.a, (uint32_t)__shfl_xor((int)v.y.a, o));
    v.y.b = addm(v.y.b, (uint32_t)__shfl_xor((int)v.y.b, o));
  }
  return v;
}

__global__ __launch_bounds__(256) void gemv_kernel(const uint32_t* const* __restrict__ cols, uint32_t n_cols,
                                                   const qm31* __restrict__ lo, const qm31* __restrict__ hi,
                                                   qm31* __restrict__ partial, uint64_t js) {
  __shared__ qm31 wsum[GEMV_COLS][4];
  const uint64_t jo = job_offset(js);
  lo = jptr(lo, jo);
  hi = jptr(hi, jo);
  partial = jptr(partial, jo);
  const uint32_t h = blockIdx.x, c0 = blockIdx.y * GEMV_COLS, t = threadIdx.x;
  const uint32_t n_hi = gridDim.x;
  qm31 l[16];
#pragma unroll
  for (int j = 0; j < 16; j++) l[j] = lo[16 * t + j];
  // the 16 coefficients of column c0 + g for this lane: global loads (not flat, so waiting
  // for the wave shuffles does not wait for them), the next column's issued before the
  // current one is reduced; the column index is clamped so the prefetch needs no branch
  const uint32_t off4 = (uint32_t)((((size_t)h << LO_BITS) + 16 * t) >> 2);
  const uint32_t n_here = min(GEMV_COLS, n_cols - c0);
  auto load16 = [&](uint32_t g, uint4 (&q)[4]) {
    const SH_GLOBAL uint4* p =
        reinterpret_cast<const SH_GLOBAL uint4*>(as_global(jptr(cols[c0 + min(g, n_here - 1)], jo)));
#pragma unroll
    for (int k = 0; k < 4; k++) q[k] = at(p, off4 + k);
  };
  uint4 qc[4], qn[4];
  load16(0, qc);
  for (uint32_t g = 0; g < n_here; g++) {
    load16(g + 1, qn);
    const uint32_t v[16] = {qc[0].x, qc[0].y, qc[0].z, qc[0].w, qc[1].x, qc[1].y, qc[1].z, qc[1].w,
                            qc[2].x, qc[2].y, qc[2].z, qc[2].w, qc[3].x, qc[3].y, qc[3].z, qc[3].w};
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      s0 += (uint64_t)l[j].x.a * v[j];
      s1 += (uint64_t)l[j].x.b * v[j];
      s2 += (uint64_t)l[j].y.a * v[j];
      s3 += (uint64_t)l[j].y.b * v[j];
      if (j % 4 == 3) {  // four products of canonical values fit on top of a folded sum
        s0 = (s0 & P) + (s0 >> 31);
        s1 = (s1 & P) + (s1 >> 31);
        s2 = (s2 & P) + (s2 >> 31);
        s3 = (s3 & P) + (s3 >> 31);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) qc[k] = qn[k];
    qm31 r = {{reduce63(s0), reduce63(s1)}, {reduce63(s2), reduce63(s3)}};
    r = wave_sum(r);
    if ((t & 63) == 0) wsum[g][t >> 6] = r;
  }
  __syncthreads();
  if (t < GEMV_COLS && c0 + t < n_cols) {
    qm31 r = qadd(qadd(wsum[t][0], wsum[t][1]), qadd(wsum[t][2], wsum[t][3]));
    partial[(size_t)(c0 + t) * n_hi + h] = qmul(r, hi[h]);
  }
}

// out[c] = sum of the n_hi partials of column c (one block per column)
__global__ __launch_bounds__(256) void sum_partials_kernel(const qm31* __restrict__ partial, uint32_t n_hi,
                                                           qm31* __restrict__ out, uint64_t js) {
  __shared__ qm31 ws[4];
  const uint64_t jo = job_offset(js);
  partial = jptr(partial, jo);
  out = jptr(out, jo);
  const uint32_t c = blockIdx.x, t = threadIdx.x;
  qm31 r = qzero();
  for (uint32_t h = t; h < n_hi; h += 256) r = qadd(r, partial[(size_t)c * n_hi + h]);
  r = wave_sum(r);
  if ((t & 63) == 0) ws[t >> 6] = r;
  __syncthreads();
  if (t == 0) out[c] = qadd(qadd(ws[0], ws[1]), qadd(ws[2], ws[3]));
}

__global__ __launch_bounds__(NT) void reduce_pass_dev(const qm31* in, uint32_t g, const Factors* __restrict__ F,
                                                      int fi, qm31* out, uint64_t js) {
  __shared__ qm31 s[1 << (CB - 1)];
  const uint64_t jo = job_offset(js);
  in = jptr(in, jo);
  F = jptr(F, jo);
  out = jptr(out, jo);
  const uint32_t cnt = 1u << g;
  for (uint32_t i = threadIdx.x; i < cnt; i += NT) s[i] = in[((size_t)blockIdx.x << g) + i];
  __syncthreads();
  tree_reduce(s, cnt, *F, fi);
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

}  // namespace eap

size_t eval_at_point_columns_scratch_bytes(uint32_t n_cols, uint32_t L) {
  if (L >= eap::LO_BITS)  // factors | Lo | Hi | partials
    return 512 + sizeof(qm31) * ((1ull << eap::LO_BITS) + (1ull << (L - eap::LO_BITS)) * (1 + (size_t)n_cols));
  const uint32_t cb = std::min<uint32_t>(L, eap::CB);
  return 512 + 2 * sizeof(qm31) * (size_t)n_cols * (1ull << (L - cb));
}

// d_out[c] = value at d_point (x, y) of the polynomial with coefficients d_cols[c] (2^L each).
// n_jobs / js: same-shape jobs of a batched launch, js bytes apart (point, output, scratch and
// columns all move by the job's offset).
void eval_at_point_columns_async(const uint32_t* const* d_cols, uint32_t n_cols, uint32_t L, const qm31* d_point,
                                 qm31* d_out, void* d_scratch, uint32_t n_jobs, uint64_t js) {
  // coefficient offsets are formed in 32 bits (gmem.hpp): at most 2^30 coefficients
  require(L >= 1 && L <= 30 && n_cols >= 1 && n_cols <= 65535, "eval_at_point_columns: bad sizes");
  eap::Factors* F = (eap::Factors*)d_scratch;
  if (L >= eap::LO_BITS) {
    const uint32_t n_hi = 1u << (L - eap::LO_BITS);
    qm31* lo = (qm31*)((char*)d_scratch + 512);
    qm31* hi = lo + (1u << eap::LO_BITS);
    qm31* partial = hi + n_hi;
    timed("eval_at_point", [&] {
      eap::factors_kernel<<<dim3(1, 1, n_jobs), 1, 0, ctx().cur>>>(d_point, L, F, js);
      eap::basis_tables_kernel<<<dim3(ceil_div((1u << eap::LO_BITS) + n_hi, 256), 1, n_jobs), 256, 0, ctx().cur>>>(
          F, L, lo, hi, js);
      eap::gemv_kernel<<<dim3(n_hi, ceil_div(n_cols, eap::GEMV_COLS), n_jobs), 256, 0, ctx().cur>>>(
          d_cols, n_cols, lo, hi, partial, js);
      eap::sum_partials_kernel<<<dim3(n_cols, 1, n_jobs), 256, 0, ctx().cur>>>(partial, n_hi, d_out, js);
    });
    return;
  }
  // small columns (log < 12): fold in LDS, block results reduced by later passes
  const uint32_t cb = std::min<uint32_t>(L, eap::CB);
  uint32_t blocks = 1u << (L - cb);
  qm31* pa = (qm31*)((char*)d_scratch + 512);
  qm31* pb = pa + (size_t)n_cols * blocks;
  timed("eval_at_point", [&] {
    eap::factors_kernel<<<dim3(1, 1, n_jobs), 1, 0, ctx().cur>>>(d_point, L, F, js);
    eap::first_pass_cols<<<dim3(blocks, n_cols, n_jobs), eap::NT, 0, ctx().cur>>>(d_cols, L, cb, F,
                                                                                   blocks == 1 ? d_out : pa, js);
  });
  int fi = (int)(L - cb) - 1;
  while (blocks > 1) {
    const uint32_t g = std::min<uint32_t>(log2_exact(blocks), eap::CB - 1);
    const uint32_t nb = blocks >> g;
    qm31* dst = nb == 1 ? d_out : pb;
    timed("eval_at_point", [&] {
      eap::reduce_pass_dev<<<dim3(nb * n_cols, 1, n_jobs), eap::NT, 0, ctx().cur>>>(pa, g, F, fi, dst, js);
    });
    std::swap(pa, pb);
    fi -= (int)g;
    blocks = nb;
  }
}

qm31 eval_at_point_sync(const uint32_t* coeffs, uint64_t n, qm31 px, qm31 py) {
  require(is_pow2(n) && n <= (1ull << 30), "eval_at_point: size must be a power of two <= 2^30");
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

extern "C" int stwo_hip_eval_at_points_columns(const uint32_t* const* columns, uint32_t n_cols, uint32_t log_size,
                                               const StwoCirclePointQM31* points, uint32_t n_points,
                                               StwoQM31* values) {
  return guarded([&] {
    require(columns != nullptr && values != nullptr && points != nullptr && n_cols >= 1 && n_points >= 1 &&
                log_size >= 1 && log_size <= 30,
            "eval_at_points_columns: bad arguments");
    std::vector<qm31> pts(2 * (size_t)n_points);
    for (uint32_t k = 0; k < n_points; k++) {
      const StwoCirclePointQM31& q = points[k];
      pts[2 * k] = qm31{{q.x.a, q.x.b}, {q.x.c, q.x.d}};
      pts[2 * k + 1] = qm31{{q.y.a, q.y.b}, {q.y.c, q.y.d}};
    }
    DevBuf table = upload(columns, (size_t)n_cols * sizeof(void*));
    DevBuf dpt = upload(pts.data(), pts.size() * sizeof(qm31));
    DevBuf out(sizeof(qm31) * n_cols * (size_t)n_points);
    DevBuf scratch(eval_at_point_columns_scratch_bytes(n_cols, log_size));
    // one batched evaluation per point (the scratch is reused in stream order)
    for (uint32_t k = 0; k < n_points; k++)
      eval_at_point_columns_async(table.as<const uint32_t*>(), n_cols, log_size, dpt.as<qm31>() + 2 * k,
                                  out.as<qm31>() + (size_t)k * n_cols, scratch.p, 1, 0);
    hip_check(hipMemcpyAsync(values, out.p, sizeof(qm31) * n_cols * (size_t)n_points, hipMemcpyDeviceToHost,
                             ctx().cur),
              "hipMemcpyAsync D2H");
    sync();
  });
}

// ---- the batched evaluation with no device allocation ----------------------------------
// For callers that cannot allocate (code under graph capture, serving loops, harnesses that
// forbid hipMalloc): the pointer table, the points, the values and the partial sums live in
// this fixed array instead of pooled buffers.  Each device holds its own copy until the
// process ends.  Every call on a device shares that copy, so calls that overlap on one device
// race; nothing in it has to start at zero (every partial is written before it is read).
namespace {
constexpr size_t NOALLOC_BYTES = 32u << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
constexpr int MAX_DEVICES = 64;
void* g_scratch[MAX_DEVICES];  // device address of `scratch` per device, under the library lock

// The device address of this device's `scratch` (the array's name is only a host-side
// stand-in); looked up on the first call on each device, with that device current.
void* noalloc_scratch() {
  const int dev = ctx().device;
  require(dev >= 0 && dev < MAX_DEVICES, "eval_at_points_columns_noalloc: device index out of range");
  if (!g_scratch[dev]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    require(p != nullptr, "eval_at_points_columns_noalloc: no device address for the scratch array");
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

extern "C" int stwo_hip_eval_at_points_columns_noalloc(const uint32_t* const* columns, uint32_t n_cols,
                                                       uint32_t log_size, const StwoCirclePointQM31* points,
                                                       uint32_t n_points, StwoQM31* values) {
  return guarded([&] {
    require(columns != nullptr && values != nullptr && points != nullptr && n_cols >= 1 && n_points >= 1 &&
                log_size >= 1 && log_size <= 30 && n_cols <= 65535,
            "eval_at_points_columns_noalloc: bad arguments");
    const size_t t_bytes = align256((size_t)n_cols * sizeof(void*));
    const size_t p_bytes = align256(2 * sizeof(qm31) * (size_t)n_points);
    const size_t o_bytes = align256(sizeof(qm31) * (size_t)n_cols * n_points);
    const size_t need = t_bytes + p_bytes + o_bytes + eval_at_point_columns_scratch_bytes(n_cols, log_size);
    if (need > NOALLOC_BYTES)
      fail(STWO_HIP_ERR_ARG,
           "eval_at_points_columns_noalloc: %u columns of log %u at %u points need %zu scratch bytes; the "
           "fixed scratch holds %zu",
           n_cols, log_size, n_points, need, NOALLOC_BYTES);
    char* base = (char*)noalloc_scratch();
    const uint32_t** table = (const uint32_t**)base;
    qm31* dpt = (qm31*)(base + t_bytes);
    qm31* out = (qm31*)(base + t_bytes + p_bytes);
    void* work = base + t_bytes + p_bytes + o_bytes;
    // points in the kernels' (x, y) layout; QM31 words are (a, b, c, d) in both
    static_assert(sizeof(StwoCirclePointQM31) == 2 * sizeof(qm31), "point layout");
    hipStream_t s = ctx().cur;
    hip_check(hipMemcpyAsync(table, columns, (size_t)n_cols * sizeof(void*), hipMemcpyHostToDevice, s),
              "hipMemcpyAsync H2D");
    hip_check(hipMemcpyAsync(dpt, points, 2 * sizeof(qm31) * (size_t)n_points, hipMemcpyHostToDevice, s),
              "hipMemcpyAsync H2D");
    for (uint32_t k = 0; k < n_points; k++)
      eval_at_point_columns_async(table, n_cols, log_size, dpt + 2 * k, out + (size_t)k * n_cols, work, 1, 0);
    hip_check(hipMemcpyAsync(values, out, sizeof(qm31) * n_cols * (size_t)n_points, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync D2H");
    sync();
  });
}

extern "C" int stwo_hip_eval_at_point_columns(const uint32_t* const* columns, uint32_t n_cols, uint32_t log_size,
                                              StwoQM31 point_x, StwoQM31 point_y, StwoQM31* values) {
  const StwoCirclePointQM31 p{point_x, point_y};
  return stwo_hip_eval_at_points_columns(columns, n_cols, log_size, &p, 1, values);
}
