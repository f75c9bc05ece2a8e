// Circle FFT over M31: PolyOps::interpolate / interpolate_columns / evaluate /
// evaluate_polynomials (stwo_gpu_backend/src/poly.rs:35-168).
//
// Butterflies and layer indexing follow cuda/src/poly/ifft.cu:3-39 (inverse:
// v0 + v1, (v0 - v1) t) and cuda/src/poly/rfft.cu:5-41 (forward: v0 + v1 t,
// v0 - v1 t).  Layer 0 is the circle layer (twiddles [y, -y, -x, x] from the
// first line level, cuda/src/poly/utils.cu:3-14); layer i >= 1 pairs indices
// that differ in bit i and uses line level i-1 of the twiddle tree at index
// h = idx >> (i + 1).  The domain's tree part starts at tree_len - N/2
// (ifft.cu:137, rfft.cu:46).
//
// MI355X design (not the reference's one-launch-per-layer structure):
//   * One launch ("pass") covers up to k layers of a tile held in registers.
//     A tile is 2^B elements of one column (or several small columns) laid out
//     as [t low index bits | k layer bits | hb high bits]; the t low bits are
//     contiguous in HBM, so every access moves >= 128-byte runs.
//   * Within a pass the layers are done radix-8 in registers; between radix
//     stages the tile is exchanged through LDS.
//   * interpolate: the 2^-n scale is fused into the store of the last pass.
//   * evaluate from coefficients (zero padded to the domain): the top layers on
//     zeros only duplicate values, so the first pass reads coeffs[idx mod 2^c]
//     and skips those layers (no extend copy, no wasted butterflies).
#include <algorithm>

#include "field.hpp"
#include "runtime.hpp"

namespace sh {
namespace fft {

struct PassArgs {
  uint32_t* const* dst;        // device table: destination column pointers
  const uint32_t* const* src;  // device table: source column pointers (may equal dst)
  const uint32_t* tw;          // tree base of this domain (line level 0)
  uint32_t n_cols;
  uint32_t log_n;    // domain log size
  uint32_t src_log;  // log size used to index the source (< log_n only for the first forward pass)
  uint32_t act_log;  // layers at or above act_log are skipped (forward from 2^act_log coefficients)
  uint32_t a, k;     // this pass applies layers [a, a + k)
  uint32_t t, hb;    // tile bits: t low (contiguous) bits, hb high bits (or column bits)
  uint32_t multi;    // 1: the hb bits select a column (whole small columns per tile)
  uint32_t scale;    // 0: none; else multiply outputs by it
  uint32_t n_groups;     // column groups (grid x = n_groups * tiles_per_group)
  uint32_t lob_count_log;  // log2 of the number of low blocks per column (a - t)
};

// Twiddle of layer `layer` (global) at butterfly-block index h.
template <bool INV>
__device__ __forceinline__ uint32_t twiddle(const uint32_t* __restrict__ tw, uint32_t log_n, uint32_t layer,
                                            uint32_t h) {
  if (layer == 0) {
    uint32_t k2 = (h >> 2) << 1;
    uint32_t sel = h & 3;  // [y, -y, -x, x] with (x, y) = (tw[2k], tw[2k+1])
    uint32_t v = tw[k2 + ((sel == 0 || sel == 1) ? 1 : 0)];
    return (sel == 1 || sel == 2) ? neg(v) : v;
  }
  uint32_t half = 1u << (log_n - 1);
  uint32_t off = half - (half >> (layer - 1));
  return tw[off + h];
}

template <int B, int NT>
struct Tile {
  static constexpr int E = (1 << B) / NT;
};

// p (tile position) with W zero bits inserted at bit `at`
__device__ __forceinline__ uint32_t spread(uint32_t gamma, uint32_t at, uint32_t W) {
  uint32_t lowmask = (1u << at) - 1;
  return (gamma & lowmask) | ((gamma & ~lowmask) << W);
}

struct Geo {
  uint32_t t, k, hb, a, log_n, multi;
  uint32_t hi_tile, lob, col_group;
  // global element index (within its column) and column of tile position p
  __device__ __forceinline__ void locate(uint32_t p, uint32_t& col, uint32_t& idx) const {
    uint32_t j = p & ((1u << t) - 1);
    uint32_t m = (p >> t) & ((1u << k) - 1);
    uint32_t r = p >> (t + k);
    if (multi) {
      col = (col_group << hb) + r;
      idx = m;
    } else {
      col = col_group;
      idx = ((((hi_tile << hb) + r) << (a + k)) | (m << a) | (lob << t) | j);
    }
  }
  // butterfly-block index h for layer bit b (global layer a + b) of position p
  __device__ __forceinline__ uint32_t h_of(uint32_t p, uint32_t b) const {
    uint32_t m = (p >> t) & ((1u << k) - 1);
    uint32_t r = p >> (t + k);
    if (multi) return m >> (b + 1);
    return (((hi_tile << hb) + r) << (k - b - 1)) | (m >> (b + 1));
  }
};

// One radix-2^W stage on the thread's groups: layer bits [lo, lo + W) of the k
// layer bits, active only below `limit` (k, or src_log - a for a forward first pass).
template <bool INV, int B, int NT, int W>
__device__ __forceinline__ void stage(uint32_t (&v)[(1 << B) / NT], const Geo& g, const uint32_t* __restrict__ tw,
                                      uint32_t lo, uint32_t limit) {
  constexpr int E = Tile<B, NT>::E;
  constexpr int R = 1 << W;
  constexpr int G = E / R;
  const uint32_t tid = threadIdx.x;
#pragma unroll
  for (int gi = 0; gi < G; gi++) {
    uint32_t gamma = tid + NT * gi;
    uint32_t pbase = spread(gamma, g.t + lo, W);
#pragma unroll
    for (int s = 0; s < W; s++) {
      const int bl = INV ? s : (W - 1 - s);
      const uint32_t b = lo + bl;
      if (b >= limit) continue;
      const uint32_t layer = g.a + b;
#pragma unroll
      for (int rh = 0; rh < (R >> (bl + 1)); rh++) {
        uint32_t p0 = pbase | ((uint32_t)(rh << (bl + 1)) << (g.t + lo));
        uint32_t tw_v = twiddle<INV>(tw, g.log_n, layer, g.h_of(p0, b));
#pragma unroll
        for (int rl = 0; rl < (1 << bl); rl++) {
          const int e0 = gi * R + (rh << (bl + 1)) + rl;
          const int e1 = e0 + (1 << bl);
          uint32_t x0 = v[e0], x1 = v[e1];
          if (INV) {
            v[e0] = add(x0, x1);
            v[e1] = mul(sub(x0, x1), tw_v);
          } else {
            uint32_t tmp = mul(x1, tw_v);
            v[e0] = add(x0, tmp);
            v[e1] = sub(x0, tmp);
          }
        }
      }
    }
  }
}

template <int B, int NT, int W>
__device__ __forceinline__ void lds_write(uint32_t* s, const uint32_t (&v)[(1 << B) / NT], const Geo& g,
                                          uint32_t lo) {
  constexpr int E = Tile<B, NT>::E;
  constexpr int R = 1 << W;
#pragma unroll
  for (int gi = 0; gi < E / R; gi++) {
    uint32_t pbase = spread(threadIdx.x + NT * gi, g.t + lo, W);
#pragma unroll
    for (int r = 0; r < R; r++) s[pbase | ((uint32_t)r << (g.t + lo))] = v[gi * R + r];
  }
}
template <int B, int NT, int W>
__device__ __forceinline__ void lds_read(const uint32_t* s, uint32_t (&v)[(1 << B) / NT], const Geo& g,
                                         uint32_t lo) {
  constexpr int E = Tile<B, NT>::E;
  constexpr int R = 1 << W;
#pragma unroll
  for (int gi = 0; gi < E / R; gi++) {
    uint32_t pbase = spread(threadIdx.x + NT * gi, g.t + lo, W);
#pragma unroll
    for (int r = 0; r < R; r++) v[gi * R + r] = s[pbase | ((uint32_t)r << (g.t + lo))];
  }
}

template <int B, int NT, int W>
__device__ __forceinline__ void gmem_load(uint32_t (&v)[(1 << B) / NT], const Geo& g, const PassArgs& A,
                                          uint32_t lo) {
  constexpr int E = Tile<B, NT>::E;
  constexpr int R = 1 << W;
  const uint32_t smask = (1u << A.src_log) - 1;
  const uint32_t* col0 = g.multi ? nullptr : A.src[g.col_group];
#pragma unroll
  for (int gi = 0; gi < E / R; gi++) {
    uint32_t pbase = spread(threadIdx.x + NT * gi, g.t + lo, W);
#pragma unroll
    for (int r = 0; r < R; r++) {
      uint32_t col, idx;
      g.locate(pbase | ((uint32_t)r << (g.t + lo)), col, idx);
      const uint32_t* c = g.multi ? (col < A.n_cols ? A.src[col] : nullptr) : col0;
      v[gi * R + r] = c ? c[idx & smask] : 0u;
    }
  }
}
template <int B, int NT, int W>
__device__ __forceinline__ void gmem_store(const uint32_t (&v)[(1 << B) / NT], const Geo& g, const PassArgs& A,
                                           uint32_t lo) {
  constexpr int E = Tile<B, NT>::E;
  constexpr int R = 1 << W;
  uint32_t* col0 = g.multi ? nullptr : A.dst[g.col_group];
#pragma unroll
  for (int gi = 0; gi < E / R; gi++) {
    uint32_t pbase = spread(threadIdx.x + NT * gi, g.t + lo, W);
#pragma unroll
    for (int r = 0; r < R; r++) {
      uint32_t col, idx;
      g.locate(pbase | ((uint32_t)r << (g.t + lo)), col, idx);
      uint32_t* c = g.multi ? (col < A.n_cols ? A.dst[col] : nullptr) : col0;
      uint32_t x = v[gi * R + r];
      if (A.scale) x = mul(x, A.scale);
      if (c) c[idx] = x;
    }
  }
}

// Dispatch on the (runtime) width of a stage: 3 except possibly the last chunk.
#define SH_W_DISPATCH(w, CALL) \
  do {                         \
    if ((w) == 3) {            \
      constexpr int W = 3;     \
      CALL;                    \
    } else if ((w) == 2) {     \
      constexpr int W = 2;     \
      CALL;                    \
    } else {                   \
      constexpr int W = 1;     \
      CALL;                    \
    }                          \
  } while (0)

template <bool INV, int B, int NT>
__global__ __launch_bounds__(NT) void fft_pass_kernel(PassArgs A) {
  constexpr int E = Tile<B, NT>::E;
  __shared__ uint32_t s[1 << B];
  uint32_t v[E];

  Geo g;
  g.t = A.t;
  g.k = A.k;
  g.hb = A.hb;
  g.a = A.a;
  g.log_n = A.log_n;
  g.multi = A.multi;
  g.col_group = blockIdx.x % A.n_groups;
  uint32_t tile = blockIdx.x / A.n_groups;
  g.lob = tile & ((1u << A.lob_count_log) - 1);
  g.hi_tile = tile >> A.lob_count_log;

  // layer chunks of width <= 3 over [0, k); inverse runs them upward, forward downward
  const uint32_t n_chunks = (A.k + 2) / 3;
  const uint32_t limit = INV ? A.k : (A.act_log > A.a ? std::min(A.k, A.act_log - A.a) : 0u);
  auto chunk_lo = [&](uint32_t c) { return 3 * c; };
  auto chunk_w = [&](uint32_t c) { return std::min(3u, A.k - 3 * c); };

  uint32_t c = INV ? 0 : n_chunks - 1;
  {
    uint32_t lo = chunk_lo(c), w = chunk_w(c);
    SH_W_DISPATCH(w, (gmem_load<B, NT, W>(v, g, A, lo)));
  }
  for (uint32_t it = 0; it < n_chunks; it++) {
    uint32_t lo = chunk_lo(c), w = chunk_w(c);
    SH_W_DISPATCH(w, (stage<INV, B, NT, W>(v, g, A.tw, lo, limit)));
    if (it + 1 == n_chunks) break;
    uint32_t cn = INV ? c + 1 : c - 1;
    uint32_t lon = chunk_lo(cn), wn = chunk_w(cn);
    if (it) __syncthreads();
    SH_W_DISPATCH(w, (lds_write<B, NT, W>(s, v, g, lo)));
    __syncthreads();
    SH_W_DISPATCH(wn, (lds_read<B, NT, W>(s, v, g, lon)));
    c = cn;
  }
  {
    uint32_t lo = chunk_lo(c), w = chunk_w(c);
    SH_W_DISPATCH(w, (gmem_store<B, NT, W>(v, g, A, lo)));
  }
}

// ---- circle domains of log size 1 and 2 (stwo CpuBackend special cases) ----
// CanonicCoset(n).half_coset.initial = G^(2^(30-n)).
template <bool INV>
__global__ void fft_tiny_kernel(uint32_t* const* dst, const uint32_t* const* src, uint32_t n_cols,
                                uint32_t log_n, uint32_t src_log) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  const uint32_t n = 1u << log_n;
  uint32_t v[4];
  for (uint32_t i = 0; i < n; i++) v[i] = (i < (1u << src_log)) ? src[c][i] : 0u;
  pt p = point_at_index(1u << (30 - log_n));
  if (INV) {
    uint32_t yi = inv(p.y);
    if (log_n == 1) {
      uint32_t a = add(v[0], v[1]), b = mul(sub(v[0], v[1]), yi);
      uint32_t h = inv(2);
      v[0] = mul(a, h);
      v[1] = mul(b, h);
    } else {
      uint32_t xi = inv(p.x), q = inv(4);
      uint32_t a0 = add(v[0], v[1]), a1 = mul(sub(v[0], v[1]), yi);
      uint32_t a2 = add(v[2], v[3]), a3 = mul(sub(v[2], v[3]), neg(yi));
      v[0] = mul(add(a0, a2), q);
      v[2] = mul(mul(sub(a0, a2), xi), q);
      v[1] = mul(add(a1, a3), q);
      v[3] = mul(mul(sub(a1, a3), xi), q);
    }
  } else {
    auto fb = [](uint32_t& x0, uint32_t& x1, uint32_t t) {
      uint32_t tmp = mul(x1, t);
      uint32_t y0 = add(x0, tmp), y1 = sub(x0, tmp);
      x0 = y0;
      x1 = y1;
    };
    if (log_n == 1) {
      fb(v[0], v[1], p.y);
    } else {
      fb(v[0], v[2], p.x);
      fb(v[1], v[3], p.x);
      fb(v[0], v[1], p.y);
      fb(v[2], v[3], neg(p.y));
    }
  }
  for (uint32_t i = 0; i < n; i++) dst[c][i] = v[i];
}

// ---------------------------------------------------------------------------
// Host planning.  Contiguous passes (layers [0, kc)) use tiles of 2^12 or 2^13
// elements; strided passes use 2^14-element tiles whose low index bits give
// >= 32 contiguous words per run.  A tile never exceeds its column: columns of
// log size <= 12 are grouped several to a tile.
// ---------------------------------------------------------------------------
constexpr int BS = 14;  // strided tile bits (64 KiB of LDS)
constexpr uint32_t KS_MAX = BS - 5;

static int threads_for(int B) { return B == 12 ? 256 : 512; }

struct Pass {
  uint32_t a, k, t, hb, multi, B;
};

static std::vector<Pass> plan(uint32_t log_n) {
  std::vector<Pass> ps;
  if (log_n <= 12) {
    ps.push_back({0, log_n, 0, 12 - log_n, 1, 12});
    return ps;
  }
  if (log_n == 13) {
    ps.push_back({0, 13, 0, 0, 0, 13});
    return ps;
  }
  uint32_t kc = (log_n - 12 > KS_MAX) ? 13 : 12;
  ps.push_back({0, kc, 0, 0, 0, kc});
  uint32_t a = kc;
  while (a < log_n) {
    uint32_t rem = log_n - a;
    uint32_t n_left = (rem + KS_MAX - 1) / KS_MAX;
    uint32_t k = (rem + n_left - 1) / n_left;
    uint32_t t = std::min<uint32_t>(a, BS - k);
    uint32_t hb = BS - k - t;  // > 0 only when a < BS - k; fits because log_n >= BS
    ps.push_back({a, k, t, hb, 0, (uint32_t)BS});
    a += k;
  }
  return ps;
}

static const char* pass_name(bool inv, const Pass& p) {
  if (p.multi) return inv ? "fft_inv_tile" : "fft_fwd_tile";
  if (p.a == 0) return inv ? "fft_inv_contig" : "fft_fwd_contig";
  return inv ? "fft_inv_strided" : "fft_fwd_strided";
}

template <bool INV>
static void launch_pass(const Pass& p, const PassArgs& base) {
  const char* name = pass_name(INV, p);
  PassArgs A = base;
  A.a = p.a;
  A.k = p.k;
  A.t = p.t;
  A.hb = p.hb;
  A.multi = p.multi;
  uint32_t blocks;
  if (p.multi) {
    A.n_groups = ceil_div(A.n_cols, 1u << p.hb);
    A.lob_count_log = 0;
    blocks = A.n_groups;
  } else {
    A.n_groups = A.n_cols;
    A.lob_count_log = p.a - p.t;
    uint32_t hi_log = A.log_n - p.a - p.k - p.hb;
    blocks = A.n_cols << (A.lob_count_log + hi_log);
  }
  hipStream_t st = ctx().stream;
  timed(name, [&] {
    if (p.B == 12)
      fft_pass_kernel<INV, 12, 256><<<blocks, 256, 0, st>>>(A);
    else if (p.B == 13)
      fft_pass_kernel<INV, 13, 512><<<blocks, 512, 0, st>>>(A);
    else
      fft_pass_kernel<INV, BS, 512><<<blocks, 512, 0, st>>>(A);
  });
}

// Inverse FFT of n_cols columns of one log size (device pointer tables): the first pass
// reads d_src, every pass writes d_cols (d_src == d_cols: in place).
void interpolate_same_size(uint32_t* const* d_cols, uint32_t n_cols, uint32_t log_n, const uint32_t* itw,
                           uint64_t tree_len, const uint32_t* const* d_src = nullptr) {
  if (!d_src) d_src = (const uint32_t* const*)d_cols;
  require(log_n >= 1 && log_n <= 30, "interpolate: log size out of range");
  require(tree_len >= (1ull << (log_n - 1)), "interpolate: twiddle tree smaller than the domain");
  if (log_n <= 2) {
    timed("fft_tiny_inv", [&] {
      fft_tiny_kernel<true><<<ceil_div(n_cols, 64), 64, 0, ctx().stream>>>(d_cols, d_src, n_cols, log_n, log_n);
    });
    return;
  }
  PassArgs A{};
  A.dst = d_cols;
  A.tw = itw + (tree_len - (1ull << (log_n - 1)));
  A.n_cols = n_cols;
  A.log_n = log_n;
  A.src_log = log_n;
  A.act_log = log_n;
  auto ps = plan(log_n);
  for (size_t i = 0; i < ps.size(); i++) {
    A.src = i == 0 ? d_src : (const uint32_t* const*)d_cols;
    // scale by 2^-log_n in the last pass: 2^-n = 2^(31 - n) since 2^31 = 1
    A.scale = (i + 1 == ps.size()) ? (1u << (31 - log_n)) % P : 0u;
    launch_pass<true>(ps[i], A);
  }
}

// Forward FFT: src columns (2^src_log coefficients) evaluated on the canonic domain
// of log_n >= src_log, written to dst (device pointer tables).
void evaluate_same_size(uint32_t* const* d_dst, const uint32_t* const* d_src, uint32_t n_cols, uint32_t src_log,
                        uint32_t log_n, const uint32_t* tw, uint64_t tree_len) {
  require(log_n >= 1 && log_n <= 30 && src_log <= log_n, "evaluate: log sizes out of range");
  require(tree_len >= (1ull << (log_n - 1)), "evaluate: twiddle tree smaller than the domain");
  if (log_n <= 2) {
    timed("fft_tiny_fwd", [&] {
      fft_tiny_kernel<false><<<ceil_div(n_cols, 64), 64, 0, ctx().stream>>>(d_dst, d_src, n_cols, log_n, src_log);
    });
    return;
  }
  PassArgs A{};
  A.dst = d_dst;
  A.tw = tw + (tree_len - (1ull << (log_n - 1)));
  A.n_cols = n_cols;
  A.log_n = log_n;
  A.scale = 0;
  auto ps = plan(log_n);
  for (size_t i = ps.size(); i-- > 0;) {
    bool first = (i + 1 == ps.size());
    A.src = first ? d_src : (const uint32_t* const*)d_dst;
    A.src_log = first ? src_log : log_n;
    A.act_log = src_log;
    launch_pass<false>(ps[i], A);
  }
}

// Group columns by size and run each class.  cols/logs are host arrays; src (optional)
// holds the inputs when they are not transformed in place.
void interpolate_mixed(uint32_t* const* cols, const uint32_t* logs, uint32_t n_cols, const uint32_t* itw,
                       uint64_t tree_len, const uint32_t* const* src = nullptr) {
  std::map<uint32_t, std::pair<std::vector<uint32_t*>, std::vector<const uint32_t*>>> by;
  for (uint32_t i = 0; i < n_cols; i++) {
    by[logs[i]].first.push_back(cols[i]);
    by[logs[i]].second.push_back(src ? src[i] : cols[i]);
  }
  for (auto& kv : by) {
    size_t n = kv.second.first.size();
    DevBuf table = upload(kv.second.first.data(), n * sizeof(void*));
    DevBuf stable = upload(kv.second.second.data(), n * sizeof(void*));
    interpolate_same_size(table.as<uint32_t*>(), (uint32_t)n, kv.first, itw, tree_len,
                          stable.as<const uint32_t*>());
  }
}
void evaluate_mixed(const uint32_t* const* src, const uint32_t* src_logs, uint32_t* const* dst,
                    const uint32_t* dst_logs, uint32_t n_cols, const uint32_t* tw, uint64_t tree_len) {
  std::map<std::pair<uint32_t, uint32_t>, std::pair<std::vector<const uint32_t*>, std::vector<uint32_t*>>> by;
  for (uint32_t i = 0; i < n_cols; i++) {
    auto& e = by[{src_logs[i], dst_logs[i]}];
    e.first.push_back(src[i]);
    e.second.push_back(dst[i]);
  }
  for (auto& kv : by) {
    size_t n = kv.second.first.size();
    DevBuf ts = upload(kv.second.first.data(), n * sizeof(void*));
    DevBuf td = upload(kv.second.second.data(), n * sizeof(void*));
    evaluate_same_size(td.as<uint32_t*>(), ts.as<const uint32_t*>(), (uint32_t)n, kv.first.first, kv.first.second,
                       tw, tree_len);
  }
}

}  // namespace fft
}  // namespace sh

using namespace sh;

// =========================================================================
// C ABI
// =========================================================================
extern "C" void interpolate(uint32_t eval_domain_size, uint32_t* values, const uint32_t* inverse_twiddles_tree,
                            uint32_t inverse_twiddle_tree_size, uint32_t values_size) {
  guarded([&] {
    require(is_pow2(values_size) && values_size == 2ull * eval_domain_size, "interpolate: bad sizes");
    uint32_t* cols[1] = {values};
    uint32_t logs[1] = {log2_exact(values_size)};
    fft::interpolate_mixed(cols, logs, 1, inverse_twiddles_tree, inverse_twiddle_tree_size);
    sync();
  });
}

extern "C" void interpolate_columns(uint32_t eval_domain_size, uint32_t* const* values,
                                    const uint32_t* inverse_twiddles_tree, uint32_t inverse_twiddle_tree_size,
                                    uint32_t values_size, uint32_t number_of_rows) {
  guarded([&] {
    if (!values_size) return;
    require(is_pow2(number_of_rows) && number_of_rows == 2ull * eval_domain_size, "interpolate_columns: bad sizes");
    std::vector<uint32_t> logs(values_size, log2_exact(number_of_rows));
    fft::interpolate_mixed(values, logs.data(), values_size, inverse_twiddles_tree, inverse_twiddle_tree_size);
    sync();
  });
}

extern "C" void evaluate(uint32_t eval_domain_size, uint32_t* values, const uint32_t* twiddles_tree,
                         uint32_t twiddle_tree_size, uint32_t values_size) {
  guarded([&] {
    require(is_pow2(values_size) && values_size == 2ull * eval_domain_size, "evaluate: bad sizes");
    uint32_t l = log2_exact(values_size);
    const uint32_t* src[1] = {values};
    uint32_t* dst[1] = {values};
    fft::evaluate_mixed(src, &l, dst, &l, 1, twiddles_tree, twiddle_tree_size);
    sync();
  });
}

extern "C" void evaluate_columns(const uint32_t* eval_domain_sizes, uint32_t* const* values,
                                 const uint32_t* twiddles_tree, uint32_t twiddle_tree_size,
                                 uint32_t number_of_columns, const uint32_t* column_sizes) {
  guarded([&] {
    if (!number_of_columns) return;
    std::vector<uint32_t> logs(number_of_columns);
    for (uint32_t i = 0; i < number_of_columns; i++) {
      require(is_pow2(column_sizes[i]) && column_sizes[i] == 2ull * eval_domain_sizes[i],
              "evaluate_columns: bad sizes");
      logs[i] = log2_exact(column_sizes[i]);
    }
    fft::evaluate_mixed((const uint32_t* const*)values, logs.data(), values, logs.data(), number_of_columns,
                        twiddles_tree, twiddle_tree_size);
    sync();
  });
}

extern "C" int stwo_hip_interpolate_columns_async(uint32_t* const* cols, const uint32_t* log_sizes, uint32_t n_cols,
                                                  const uint32_t* itwiddles, uint64_t tree_len) {
  return guarded([&] { fft::interpolate_mixed(cols, log_sizes, n_cols, itwiddles, tree_len); });
}

extern "C" int stwo_hip_interpolate_columns_oop_async(const uint32_t* const* src, uint32_t* const* dst,
                                                      const uint32_t* log_sizes, uint32_t n_cols,
                                                      const uint32_t* itwiddles, uint64_t tree_len) {
  return guarded([&] { fft::interpolate_mixed(dst, log_sizes, n_cols, itwiddles, tree_len, src); });
}

extern "C" int stwo_hip_evaluate_columns_async(const uint32_t* const* coeffs, const uint32_t* coeff_log,
                                               uint32_t* const* dst, const uint32_t* eval_log, uint32_t n_cols,
                                               const uint32_t* twiddles, uint64_t tree_len) {
  return guarded([&] {
    for (uint32_t i = 0; i < n_cols; i++) {
      require(coeff_log[i] <= eval_log[i], "evaluate: coefficient log size exceeds the domain");
      require(coeff_log[i] == eval_log[i] || (const void*)coeffs[i] != (const void*)dst[i],
              "evaluate: dst aliases coeffs with a larger domain");
    }
    fft::evaluate_mixed(coeffs, coeff_log, dst, eval_log, n_cols, twiddles, tree_len);
  });
}
