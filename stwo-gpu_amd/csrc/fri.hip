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
#include <algorithm>

#include "field.hpp"
#include "gmem.hpp"
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

// moves the four coordinate columns by a job offset (see jptr)
__device__ __forceinline__ Soa4 jsoa(Soa4 s, uint64_t jo) {
  return Soa4{{jptr(s.c[0], jo), jptr(s.c[1], jo), jptr(s.c[2], jo), jptr(s.c[3], jo)}};
}

// alpha_dev (optional): read alpha from device memory (a challenge drawn on the device).
// js: byte stride between the jobs of a batched launch (blockIdx.z = job; 0 for one job).
__global__ __launch_bounds__(256) void fold_line_kernel(const uint32_t* __restrict__ itw, uint32_t half_n,
                                                        Soa4 src, qm31 alpha, const qm31* alpha_dev, Soa4 dst,
                                                        uint64_t js) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half_n) return;
  const uint64_t jo = job_offset(js);
  src = jsoa(src, jo);
  dst = jsoa(dst, jo);
  if (alpha_dev) alpha = *jptr(alpha_dev, jo);
  const uint32_t xi = itw[i];
  const qm31 a = ld(src, 2 * (size_t)i), b = ld(src, 2 * (size_t)i + 1);
  st(dst, i, fold_mat(a, b, xi, qmat_of(alpha)));
}

// accumulate = false: dst is overwritten with the fold (as if it held zeros)
__global__ __launch_bounds__(256) void fold_circle_kernel(const uint32_t* __restrict__ tw, uint32_t half_n,
                                                          Soa4 src, qm31 alpha, const qm31* alpha_dev,
                                                          bool accumulate, Soa4 dst, uint64_t js) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half_n) return;
  const uint64_t jo = job_offset(js);
  src = jsoa(src, jo);
  dst = jsoa(dst, jo);
  if (alpha_dev) alpha = *jptr(alpha_dev, jo);
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
  const qm31 a = ld(src, 2 * (size_t)i), b = ld(src, 2 * (size_t)i + 1);
  const qmat M = qmat_of(alpha);
  st(dst, i, accumulate ? fold_mat_acc(a, b, yi, M, ld(dst, i), qmat_of(qsqr(alpha))) : fold_mat(a, b, yi, M));
}

static Soa4 soa(uint32_t* const* p) { return Soa4{{p[0], p[1], p[2], p[3]}}; }

// ---- fold chains with every coefficient known up front --------------------------------
// out[i] depends only on in[2i] and in[2i + 1], so a block that owns 2^S consecutive inputs
// folds them S times with no other block's data: the first fold reads HBM, the later ones
// read the block's previous outputs from LDS, and every layer is written out.  A log-22
// circle evaluation folded to 2^4 values takes two launches instead of eighteen.
constexpr uint32_t CHAIN_MAX_S = 13;  // 2^13 inputs per block: 5 x 2^12 words of LDS (80 KiB)
constexpr int CHAIN_NT = 1024;  // two blocks (32 waves) per CU, 64 VGPRs at most
constexpr int CHAIN_ROUND = 1;  // first-fold outputs per thread whose loads are in flight together
constexpr uint32_t CHAIN_SMALL_S = 10;  // tiles of at most 2^10 inputs: 256 threads
constexpr int CHAIN_SMALL_NT = 256;

struct ChainArgs {
  const uint32_t* itw;  // the whole inverse tree
  uint64_t tree_len;
  Soa4 src;             // input layer of this launch
  uint32_t log_in;      // its log size
  uint32_t circle;      // 1: the first fold is the circle fold, accumulated into dst[0]
  uint32_t n_fold;      // folds in this launch (<= S)
  uint32_t S;           // each block folds inputs [b 2^S, (b + 1) 2^S)
  qmat mat[CHAIN_MAX_S];  // fold f multiplies by its alpha through this matrix
  qmat acc_mat;           // the circle fold's alpha^2 (the factor of the layer it accumulates into)
  Soa4 dst[CHAIN_MAX_S];
};

// circle twiddle i at the tree part tw: [y, -y, -x, x] over pairs (x, y) of the first line
// level: the word loaded, and its sign applied once the load is needed (applying it at once
// would wait for every load in flight)
__device__ __forceinline__ uint32_t circle_twiddle_word(const uint32_t* tw, uint32_t i) {
  return tw[((i >> 2) << 1) + ((i & 3) < 2 ? 1 : 0)];
}
__device__ __forceinline__ uint32_t circle_twiddle_sign(uint32_t t, uint32_t i) {
  const uint32_t sel = i & 3;
  return (sel == 1 || sel == 2) ? neg(t) : t;
}
// NT threads for tiles of at most 2^SMAX inputs (the wide launches: 1024 and 2^13; a chain's
// last layers, at most 2^10 values in one block: 256, which keeps the barriers of its
// short folds cheap)
template <int NT, int SMAX>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void fold_chain_kernel(ChainArgs A) {
  extern __shared__ uint32_t lds[];
  const uint32_t h0 = 1u << (A.S - 1);  // outputs of the block's first fold
  uint32_t* const pl[4] = {lds, lds + h0, lds + 2 * h0, lds + 3 * h0};
  // twiddles of the later folds, read ahead: fold j >= 1 has cnt_j = h0 >> j outputs in this
  // block, its twiddles at twl[h0 - 2 cnt_j, h0 - cnt_j); the loads are issued first and stored
  // to LDS after the first fold, so their latency overlaps that fold's loads
  uint32_t* const twl = lds + 4 * h0;
  constexpr int PRE = (1 << (SMAX - 1)) / NT;
  uint32_t v[PRE];
  {
#pragma unroll
    for (int q = 0; q < PRE; q++) {
      const uint32_t u = threadIdx.x + q * NT;
      if (u + 1 < h0) {
        const uint32_t r = h0 - u;                     // in (h0 >> j, h0 >> (j - 1)]
        const uint32_t j = A.S - (32 - __clz(r - 1));  // so that ceil(log2 r) = S - j
        if (j < A.n_fold) {
          const uint32_t k = u - (h0 - (h0 >> (j - 1)));
          v[q] = A.itw[(A.tree_len - (1ull << (A.log_in - j))) + blockIdx.x * (h0 >> j) + k];
        }
      }
    }
  }
  {
    const uint64_t n_in = 1ull << A.log_in;
    const uint32_t* tw = A.itw + (A.tree_len - (A.circle ? n_in >> 1 : n_in));
    const uint32_t g0 = blockIdx.x * h0;
    // CHAIN_ROUND outputs per thread per round: every load of a round is issued before the
    // first store (the stores could alias the loads, so they would otherwise serialise)
    for (uint32_t k0 = 0; k0 < h0; k0 += CHAIN_ROUND * NT) {
      uint2 w[CHAIN_ROUND][4];
      uint32_t t[CHAIN_ROUND];
      qm31 d[CHAIN_ROUND];
#pragma unroll
      for (int q = 0; q < CHAIN_ROUND; q++) {
        const uint32_t k = k0 + threadIdx.x + q * NT, g = g0 + k;
        if (k < h0) {  // pair g of the input: values 2g, 2g + 1 (one 8-byte load per coordinate)
#pragma unroll
          for (int c = 0; c < 4; c++)
            w[q][c] = at(reinterpret_cast<const SH_GLOBAL uint2*>(as_global(A.src.c[c])), g);
          t[q] = A.circle ? circle_twiddle_word(tw, g) : tw[g];
          if (A.circle) d[q] = ld(A.dst[0], g);
        }
      }
#pragma unroll
      for (int q = 0; q < CHAIN_ROUND; q++) {
        const uint32_t k = k0 + threadIdx.x + q * NT, g = g0 + k;
        if (k < h0) {
          const qm31 a = {{w[q][0].x, w[q][1].x}, {w[q][2].x, w[q][3].x}};
          const qm31 b = {{w[q][0].y, w[q][1].y}, {w[q][2].y, w[q][3].y}};
          const qm31 r = A.circle ? fold_mat_acc(a, b, circle_twiddle_sign(t[q], g), A.mat[0], d[q], A.acc_mat)
                                  : fold_mat(a, b, t[q], A.mat[0]);
          st(A.dst[0], g, r);
          pl[0][k] = r.x.a;
          pl[1][k] = r.x.b;
          pl[2][k] = r.y.a;
          pl[3][k] = r.y.b;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PRE; q++) {
    const uint32_t u = threadIdx.x + q * NT;
    if (u + 1 < h0 && A.S - (32 - __clz(h0 - u - 1)) < A.n_fold) twl[u] = v[q];
  }
  __syncthreads();
  // the later folds: the block's previous outputs from LDS, twiddles from the read-ahead
  constexpr int PER = (1 << (SMAX - 2)) / NT;  // outputs per thread of the second fold
  for (uint32_t j = 1; j < A.n_fold; j++) {
    const uint32_t cnt = h0 >> j;
    const uint32_t* tw = twl + (h0 - 2 * cnt);
    const uint32_t g0 = blockIdx.x * cnt;
    const qmat& M = A.mat[j];
    qm31 r[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const uint32_t k = threadIdx.x + q * NT;
      if (k < cnt) {
        const qm31 a = {{pl[0][2 * k], pl[1][2 * k]}, {pl[2][2 * k], pl[3][2 * k]}};
        const qm31 b = {{pl[0][2 * k + 1], pl[1][2 * k + 1]}, {pl[2][2 * k + 1], pl[3][2 * k + 1]}};
        r[q] = fold_mat(a, b, tw[k], M);
      }
    }
    __syncthreads();  // every read of the previous layer is done
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const uint32_t k = threadIdx.x + q * NT;
      if (k < cnt) {
        st(A.dst[j], g0 + k, r[q]);
        pl[0][k] = r[q].x.a;
        pl[1][k] = r[q].x.b;
        pl[2][k] = r[q].y.a;
        pl[3][k] = r[q].y.b;
      }
    }
    __syncthreads();
  }
}

}  // namespace fri

void fold_circle_into_line_async(const uint32_t* itw, size_t twiddle_offset, size_t n, uint32_t* const* src,
                                 qm31 alpha, uint32_t* const* dst, const qm31* alpha_dev, bool accumulate,
                                 uint32_t n_jobs, uint64_t js);

// layers: n_layers x 4 device pointers (host array), layer j of 2^(log_circle - 1 - j) values.
void fri_fold_chain_async(const uint32_t* itw, uint64_t tree_len, uint32_t log_circle, uint32_t* const* circle,
                          uint32_t n_layers, const qm31* alphas, uint32_t* const* layers) {
  // pair offsets are formed in 32 bits (8-byte pairs): at most 2^30 values per layer
  require(log_circle >= 1 && log_circle <= 30 && n_layers >= 1 && n_layers <= log_circle,
          "fri_fold_chain: need 1 <= n_layers <= log_circle <= 30");
  require(tree_len >= (1ull << (log_circle - 1)), "fri_fold_chain: twiddle tree smaller than the circle domain");
  uint32_t j = 0, log_cur = log_circle;
  uint32_t* const* cur = circle;
  if (log_circle <= 2) {  // circle domains of 2 and 4 points: their twiddles come from the domain itself
    fold_circle_into_line_async(itw, tree_len - (1ull << (log_circle - 1)), 1ull << log_circle, circle, alphas[0],
                                layers, nullptr, true, 1, 0);
    j = 1;
    log_cur--;
    cur = layers;
  }
  while (j < n_layers) {
    // 2^S inputs per block, S <= 13: a layer of at most 2^13 values is folded by one block to
    // the end; a block's later folds cost latency rather than bandwidth, so wide layers take
    // the largest ranges (fewest block-folds)
    const uint32_t S = std::min<uint32_t>(fri::CHAIN_MAX_S, log_cur);
    const uint32_t nf = std::min<uint32_t>(S, n_layers - j);
    fri::ChainArgs A{};
    A.itw = itw;
    A.tree_len = tree_len;
    A.src = fri::soa(cur);
    A.log_in = log_cur;
    A.circle = j == 0;
    if (A.circle) A.acc_mat = qmat_of(qsqr(alphas[0]));
    A.n_fold = nf;
    A.S = S;
    for (uint32_t f = 0; f < nf; f++) {
      A.mat[f] = qmat_of(alphas[j + f]);
      A.dst[f] = fri::soa(layers + 4 * (size_t)(j + f));
    }
    timed(log_cur > S ? "fri_fold_chain" : "fri_fold_chain_tail", [&] {
      if (S <= fri::CHAIN_SMALL_S)
        fri::fold_chain_kernel<fri::CHAIN_SMALL_NT, fri::CHAIN_SMALL_S>
            <<<1u << (log_cur - S), fri::CHAIN_SMALL_NT, 20u << (S - 1), ctx().cur>>>(A);
      else
        fri::fold_chain_kernel<fri::CHAIN_NT, fri::CHAIN_MAX_S>
            <<<1u << (log_cur - S), fri::CHAIN_NT, 20u << (S - 1), ctx().cur>>>(A);
    });
    j += nf;
    log_cur -= nf;
    cur = layers + 4 * (size_t)(j - 1);
  }
}

// n_jobs / js: same-shape jobs of a batched launch, js bytes apart (source, destination and
// alpha_dev move by the job's offset; the twiddles are shared)
void fold_line_async(const uint32_t* itw, size_t twiddle_offset, size_t n, uint32_t* const* src, qm31 alpha,
                     uint32_t* const* dst, const qm31* alpha_dev, uint32_t n_jobs, uint64_t js) {
  require(n >= 2 && is_pow2(n) && n <= (1ull << 31), "fold_line: n must be a power of two >= 2");
  uint32_t half = (uint32_t)(n >> 1);
  timed("fold_line", [&] {
    fri::fold_line_kernel<<<dim3(ceil_div(half, 256), 1, n_jobs), 256, 0, ctx().cur>>>(
        itw + twiddle_offset, half, fri::soa(src), alpha, alpha_dev, fri::soa(dst), js);
  });
}
void fold_circle_into_line_async(const uint32_t* itw, size_t twiddle_offset, size_t n, uint32_t* const* src,
                                 qm31 alpha, uint32_t* const* dst, const qm31* alpha_dev, bool accumulate,
                                 uint32_t n_jobs, uint64_t js) {
  require(n >= 2 && is_pow2(n) && n <= (1ull << 31), "fold_circle_into_line: n must be a power of two >= 2");
  uint32_t half = (uint32_t)(n >> 1);
  timed("fold_circle_into_line", [&] {
    fri::fold_circle_kernel<<<dim3(ceil_div(half, 256), 1, n_jobs), 256, 0, ctx().cur>>>(
        itw + twiddle_offset, half, fri::soa(src), alpha, alpha_dev, accumulate, fri::soa(dst), js);
  });
}

}  // namespace sh

using namespace sh;

extern "C" int stwo_hip_fold_line_async(const uint32_t* itwiddles, size_t twiddle_offset, size_t n,
                                        uint32_t* const* eval_values, StwoQM31 alpha, uint32_t* const* folded_values) {
  return guarded([&] {
    fold_line_async(itwiddles, twiddle_offset, n, eval_values, qm31{{alpha.a, alpha.b}, {alpha.c, alpha.d}},
                    folded_values, nullptr, 1, 0);
  });
}

extern "C" int stwo_hip_fold_circle_into_line_async(const uint32_t* itwiddles, size_t twiddle_offset, size_t n,
                                                    uint32_t* const* eval_values, StwoQM31 alpha,
                                                    uint32_t* const* folded_values) {
  return guarded([&] {
    fold_circle_into_line_async(itwiddles, twiddle_offset, n, eval_values,
                                qm31{{alpha.a, alpha.b}, {alpha.c, alpha.d}}, folded_values, nullptr, true, 1, 0);
  });
}

extern "C" int stwo_hip_fri_fold_chain_async(const uint32_t* itwiddles, uint64_t tree_len, uint32_t log_circle,
                                             uint32_t* const* circle, uint32_t n_layers, const StwoQM31* alphas,
                                             uint32_t* const* layers) {
  return guarded([&] {
    require(itwiddles && circle && alphas && layers, "fri_fold_chain: null argument");
    require(n_layers >= 1 && n_layers <= 30, "fri_fold_chain: bad layer count");
    qm31 a[30];
    for (uint32_t j = 0; j < n_layers; j++) a[j] = qm31{{alphas[j].a, alphas[j].b}, {alphas[j].c, alphas[j].d}};
    fri_fold_chain_async(itwiddles, tree_len, log_circle, circle, n_layers, a, layers);
  });
}

extern "C" void fold_line(const uint32_t* gpu_domain, size_t twiddle_offset, size_t n, uint32_t* const* eval_values,
                          StwoQM31 alpha, uint32_t* const* folded_values) {
  guarded([&] {
    fold_line_async(gpu_domain, twiddle_offset, n, eval_values, qm31{{alpha.a, alpha.b}, {alpha.c, alpha.d}},
                    folded_values, nullptr, 1, 0);
    sync();
  });
}

extern "C" void fold_circle_into_line(const uint32_t* gpu_domain, size_t twiddle_offset, size_t n,
                                      uint32_t* const* eval_values, StwoQM31 alpha, uint32_t* const* folded_values) {
  guarded([&] {
    fold_circle_into_line_async(gpu_domain, twiddle_offset, n, eval_values,
                                qm31{{alpha.a, alpha.b}, {alpha.c, alpha.d}}, folded_values, nullptr, true, 1, 0);
    sync();
  });
}
