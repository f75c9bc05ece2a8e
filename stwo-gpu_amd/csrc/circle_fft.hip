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
//     stages the tile is exchanged through LDS.  The next stage's twiddles are
//     loaded into registers before each exchange, so their latency overlaps the
//     barriers.  Where a wave's 64 lanes differ only in index bits below the stage's
//     twiddle bits (strided tiles with >= 64-word runs, and the upper stages of the
//     contiguous passes) the twiddles are wave-uniform and come through the scalar
//     cache into SGPRs: no vector-memory instructions or VGPRs for them, and their
//     doubling runs on the scalar ALU (256 columns of log 20: fused LDE pass 1.35 ->
//     1.20 ms).
//   * Butterflies multiply by twice the twiddle (mul_dbl in field.hpp), one
//     instruction fewer per product mod P than a plain multiplication; the passes
//     are VALU-bound (about 12 instructions per butterfly).
//   * Global accesses use a scalar column base and a 32-bit vector offset.
//   * interpolate: the 2^-n scale is fused into the store of the last pass (a
//     31-bit rotation, since 2^-n = 2^(31-n) mod P).
//   * evaluate from coefficients (zero padded to the domain): the top layers on
//     zeros only duplicate values, so the first pass reads coeffs[idx mod 2^c]
//     and skips those layers (no extend copy, no wasted butterflies).
#include <algorithm>
#include <type_traits>

#include "fft.hpp"
#include "field.hpp"
#include "gmem.hpp"
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
  uint32_t scale;    // 0: none; else multiply outputs by 2^scale (1 <= scale <= 30)
  uint32_t n_groups;     // column groups (grid x = n_groups * tiles_per_group)
  uint32_t band;         // consecutive blocks walk `band` columns at one tile, then the next tile
  uint32_t lob_count_log;  // log2 of the number of low blocks per column (a - t)
  // block order (CONTIG / STRIDED, set by set_block_order): tiles per column = 2^tiles_log;
  // the last band holds band_last columns; *_magic: ceil(2^32 / d) for the divisions by
  // band and band_last (0 for d = 1), exact for every numerator the order forms
  uint32_t tiles_log, band_last, band_magic, band_last_magic;
};

// m = ceil(2^32 / d) for 2 <= d, or 0 for d = 1: divu(n, m) = n / d while n (m d - 2^32) < 2^32
inline uint32_t div_magic(uint32_t d) { return d == 1 ? 0u : (uint32_t)(((1ull << 32) + d - 1) / d); }
__device__ __forceinline__ uint32_t divu(uint32_t n, uint32_t m) { return m ? __umulhi(n, m) : n; }

// accelerator complex dies (XCDs) of gfx950, each with its own L2
constexpr uint32_t N_XCD = 8;

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// p (tile position) with W zero bits inserted at bit `at`
template <int AT, int W>
__device__ __forceinline__ uint32_t spread(uint32_t gamma) {
  constexpr uint32_t lowmask = (1u << AT) - 1;
  return (gamma & lowmask) | ((gamma & ~lowmask) << W);
}

// Pass shape, fixed at compile time.  A tile holds 2^B positions p = [T low index
// bits | K layer bits | HB high bits]:
//   CONTIG  (a = 0):            T = 0, K = B, HB = 0 (one column, 2^B consecutive values)
//   STRIDED (a > 0):            T = B - K, HB = 0 (2^T-word contiguous runs)
//   MULTI   (a = 0, log_n = K): T = 0, HB = B - K bits selecting the column
enum Kind { CONTIG = 0, STRIDED = 1, MULTI = 2 };

template <bool INV, int B, int NT, int K, int KIND, bool MASKED = false, int ACT = -1>
struct Pass {
  static constexpr int T = KIND == STRIDED ? B - K : 0;
  static constexpr int HB = KIND == MULTI ? B - K : 0;
  static constexpr int E = (1 << B) / NT;
  static constexpr int NCH = (K + 2) / 3;
  // layers at or above act() are skipped (compile-time when ACT >= 0)
  __device__ __forceinline__ uint32_t act() const {
    if constexpr (ACT >= 0)
      return ACT;
    else
      return A.act_log;
  }

  const PassArgs& A;
  uint32_t* s;
  uint32_t tile, col_group, lob;
  uint32_t block_base;  // index of tile position 0 (CONTIG / STRIDED); columns hold < 2^31 values
  uint32_t hi_tile;     // tile's high index bits (twiddle block index)
  const SH_GLOBAL uint32_t* src;  // CONTIG / STRIDED: this block's column
  SH_GLOBAL uint32_t* dst;

  __device__ __forceinline__ Pass(const PassArgs& a, uint32_t* lds) : A(a), s(lds) {
    if constexpr (KIND == MULTI) {
      col_group = blockIdx.x % A.n_groups;
      tile = blockIdx.x / A.n_groups;
    } else {
      // bands of `band` columns: the blocks in flight share each tile's twiddles across the
      // band and touch few columns (pages) at a time.  Blocks go round-robin to the 8 XCDs
      // (block i to XCD i % 8), each with its own L2, so each XCD takes a contiguous range of
      // this order: the band members of one tile then run on one XCD and read its twiddles
      // from that XCD's L2 once.
      // (The divisions by band sizes multiply by host-computed reciprocals, on the scalar
      // unit; a division by a runtime value costs ~5 VALU instructions per wave.)
      const uint32_t G = gridDim.x, q = G / N_XCD, rem = G % N_XCD;
      const uint32_t x = blockIdx.x % N_XCD;
      const uint32_t v = x * q + min(x, rem) + blockIdx.x / N_XCD;
      const uint32_t bi = divu(v >> A.tiles_log, A.band_magic);
      const uint32_t r = v - ((bi * A.band) << A.tiles_log);
      const bool last = (bi + 1) * A.band >= A.n_groups;
      const uint32_t w = last ? A.band_last : A.band;
      tile = divu(r, last ? A.band_last_magic : A.band_magic);
      col_group = bi * A.band + (r - tile * w);
    }
    if constexpr (KIND == MULTI) {
      block_base = 0;
      hi_tile = 0;
      lob = 0;
      src = nullptr;
      dst = nullptr;
    } else {
      lob = tile & ((1u << A.lob_count_log) - 1);
      set_hi(tile >> A.lob_count_log);
      src = as_global(A.src[col_group]);
      dst = as_global(A.dst[col_group]);
    }
  }
  // place the tile in high block `hi` (the bits above a + K) of its column
  __device__ __forceinline__ void set_hi(uint32_t hi) {
    hi_tile = hi;
    block_base = (hi_tile << (A.a + K)) | (lob << T);
  }

  // global index (within its column) of tile position p
  __device__ __forceinline__ uint32_t index_of(uint32_t p) const {
    if constexpr (KIND == MULTI) return p & ((1u << K) - 1);
    else if constexpr (T == 0) return block_base + p;
    else return block_base + ((p >> T) << A.a) + (p & ((1u << T) - 1));
  }
  __device__ __forceinline__ uint32_t col_of(uint32_t p) const { return (col_group << HB) + (p >> K); }

  // twiddle of layer bit b at the butterfly block containing position p: the word loaded
  // from its level (tw_index) and the sign applied after the load (tw_fix)
  template <int BIT>
  __device__ __forceinline__ uint32_t tw_h(uint32_t p) const {
    const uint32_t m = (p >> T) & ((1u << K) - 1);
    return (hi_tile << (K - BIT - 1)) | (m >> (BIT + 1));
  }
  template <int BIT>
  __device__ __forceinline__ uint32_t tw_index(uint32_t p) const {
    const uint32_t h = tw_h<BIT>(p);
    if constexpr (KIND != STRIDED && BIT == 0) {
      // circle layer: [y, -y, -x, x] over pairs (x, y) of the first line level
      return ((h >> 2) << 1) + ((h & 3) < 2 ? 1 : 0);
    } else {
      return h;
    }
  }
  // the butterflies multiply by twice the twiddle (mul_dbl); a negated circle twiddle
  // -t is represented by 2P - 2t (that is 2(P - t), with t = 0 giving the factor 2P)
  template <int BIT>
  __device__ __forceinline__ uint32_t tw_fix(uint32_t v, uint32_t p) const {
    const uint32_t v2 = v << 1;
    if constexpr (KIND != STRIDED && BIT == 0) {
      const uint32_t sel = tw_h<BIT>(p) & 3;
      return (sel == 1 || sel == 2) ? 2 * P - v2 : v2;
    } else {
      return v2;
    }
  }

  // Twiddles of one radix-2^W stage held in registers (at most 7 per 8 values): slot of
  // layer bit BL, value group gi, butterfly block rh.
  static constexpr int NTW = E - E / 8;
  template <int W, int BL>
  static constexpr int tw_slot(int gi, int rh) {
    return (E >> W) * ((1 << W) - ((1 << W) >> BL)) + gi * ((1 << W) >> (BL + 1)) + rh;
  }
  template <int LO, int W>
  __device__ __forceinline__ void load_tw(uint32_t (&tw)[NTW]) const {
    constexpr int R = 1 << W;
    // the 64 lanes of a wave differ in the low 6 bits of threadIdx.x; when those land below the
    // twiddle's index bits (T + LO >= 6) every lane of the wave needs the same twiddles, so they
    // are loaded once per wave through the scalar cache into SGPRs
    constexpr bool UNIFORM = T + LO >= 6;
    const uint32_t tid = UNIFORM ? (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x & ~63u) : threadIdx.x;
    static_for<0, W>([&](auto si) {
      constexpr int BL = si.value;
      constexpr int BIT = LO + BL;
      if constexpr (MASKED) {
        if (A.a + BIT >= act()) return;  // layer skipped by stage()
      }
      const SH_GLOBAL uint32_t* lvl = level(BIT);
#pragma unroll
      for (int gi = 0; gi < E / R; gi++) {
        const uint32_t pb = spread<T + LO, W>(tid + NT * gi);
#pragma unroll
        for (int rh = 0; rh < (R >> (BL + 1)); rh++) {
          const uint32_t i = tw_index<BIT>(pb | ((uint32_t)(rh << (BL + 1)) << (T + LO)));
          if constexpr (UNIFORM)
            tw[tw_slot<W, BL>(gi, rh)] = ((const __attribute__((address_space(4))) uint32_t*)lvl)[i];
          else
            tw[tw_slot<W, BL>(gi, rh)] = at(lvl, i);
        }
      }
    });
  }
  // line level used by layer bit b (global layer a + b >= 1); level 0 for the circle layer
  __device__ __forceinline__ const SH_GLOBAL uint32_t* level(uint32_t b) const {
    const uint32_t layer = A.a + b;
    if (layer == 0) return as_global(A.tw);
    const uint32_t half = 1u << (A.log_n - 1);
    return as_global(A.tw) + (half - (half >> (layer - 1)));
  }

  template <int LO, int W>
  __device__ __forceinline__ void load(uint32_t (&v)[E]) const {
    constexpr int R = 1 << W;
    // only masked forward passes read from fewer coefficients than the domain (src_log <
    // log_n, the first pass of an evaluation from 2^act_log coefficients); every other
    // CONTIG / STRIDED pass reads whole columns, and without the mask its addresses are one
    // add per load.  MULTI tiles always mask: it also drops the column bits of p.
    const uint32_t smask = (MASKED || KIND == MULTI) ? (1u << A.src_log) - 1 : ~0u;
#pragma unroll
    for (int gi = 0; gi < E / R; gi++) {
      const uint32_t pb = spread<T + LO, W>(threadIdx.x + NT * gi);
      if constexpr (KIND == MULTI) {
        const uint32_t col = col_of(pb);
        const SH_GLOBAL uint32_t* c = col < A.n_cols ? as_global(A.src[col]) : nullptr;
        if constexpr (!MASKED && LO == 0 && W == 3) {
          // 8 consecutive values of a whole column (K >= 3): two 16-byte loads
          const uint32_t i0 = pb & ((1u << K) - 1);
          uint4 x0 = {0, 0, 0, 0}, x1 = {0, 0, 0, 0};
          if (c) {
            x0 = at(reinterpret_cast<const SH_GLOBAL uint4*>(c), i0 >> 2);
            x1 = at(reinterpret_cast<const SH_GLOBAL uint4*>(c), (i0 >> 2) + 1);
          }
          v[gi * 8 + 0] = x0.x; v[gi * 8 + 1] = x0.y; v[gi * 8 + 2] = x0.z; v[gi * 8 + 3] = x0.w;
          v[gi * 8 + 4] = x1.x; v[gi * 8 + 5] = x1.y; v[gi * 8 + 6] = x1.z; v[gi * 8 + 7] = x1.w;
        } else {
          // with the coefficient count known at compile time only the distinct values are
          // loaded: r and r + 2^D read the same coefficient when D = ACT - LO < W
          constexpr int D = ACT < 0 ? W : (ACT - LO <= 0 ? 0 : (ACT - LO < W ? ACT - LO : W));
#pragma unroll
          for (int r = 0; r < R; r++) {
            if (r < (1 << D))
              v[gi * R + r] = c ? c[(pb + (r << LO)) & smask] : 0u;
            else
              v[gi * R + r] = v[gi * R + (r & ((1 << D) - 1))];
          }
        }
      } else if constexpr (LO == 0 && T == 0 && W == 3) {
        const uint32_t i0 = index_of(pb);
        // unmasked passes read whole domains (src_log = log_n >= 3): no per-element fallback,
        // so the compiler keeps the two 16-byte loads instead of merging both paths
        if (!MASKED || A.src_log >= 3) {
          const SH_GLOBAL uint4* q = reinterpret_cast<const SH_GLOBAL uint4*>(src);
          const uint32_t qi = (i0 & smask) >> 2;
          uint4 x0 = at(q, qi), x1 = at(q, qi + 1);
          v[gi * 8 + 0] = x0.x; v[gi * 8 + 1] = x0.y; v[gi * 8 + 2] = x0.z; v[gi * 8 + 3] = x0.w;
          v[gi * 8 + 4] = x1.x; v[gi * 8 + 5] = x1.y; v[gi * 8 + 6] = x1.z; v[gi * 8 + 7] = x1.w;
        } else {
#pragma unroll
          for (int r = 0; r < 8; r++) v[gi * 8 + r] = at(src, (i0 + r) & smask);
        }
      } else if constexpr (!MASKED) {
        // value r of the group sits r << (LO + a) words after value 0 (pb has zeros in the
        // group's bits): one vector offset for the group and a scalar one per r
        const BufCol b(src + block_base);
        const uint32_t i0 = index_of(pb) - block_base;
#pragma unroll
        for (int r = 0; r < R; r++) v[gi * R + r] = b.load(i0, r_offset<LO>(r));
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) v[gi * R + r] = at(src, index_of(pb | (r << (T + LO))) & smask);
      }
    }
  }
  // word offset of value r of a radix group at stage LO from the group's first value (uniform)
  template <int LO>
  __device__ __forceinline__ uint32_t r_offset(int r) const {
    return (uint32_t)r << (LO + (KIND == STRIDED ? A.a : 0u));
  }

  template <int LO, int W>
  __device__ __forceinline__ void store(uint32_t (&v)[E]) const {
    constexpr int R = 1 << W;
    if (A.scale) {
      // x 2^s mod 2^31 - 1 is a rotation of the 31-bit value: canonical in, canonical out
      const uint32_t sl = A.scale, sr = 31 - A.scale;
#pragma unroll
      for (int e = 0; e < E; e++) v[e] = ((v[e] << sl) & P) | (v[e] >> sr);
    }
#pragma unroll
    for (int gi = 0; gi < E / R; gi++) {
      const uint32_t pb = spread<T + LO, W>(threadIdx.x + NT * gi);
      if constexpr (KIND == MULTI) {
        const uint32_t col = col_of(pb);
        if (col < A.n_cols) {
          SH_GLOBAL uint32_t* c = as_global(A.dst[col]);
          if constexpr (LO == 0 && W == 3) {
            SH_GLOBAL uint4* q = reinterpret_cast<SH_GLOBAL uint4*>(c);
            const uint32_t qi = (pb & ((1u << K) - 1)) >> 2;
            at(q, qi) = make_uint4(v[gi * 8 + 0], v[gi * 8 + 1], v[gi * 8 + 2], v[gi * 8 + 3]);
            at(q, qi + 1) = make_uint4(v[gi * 8 + 4], v[gi * 8 + 5], v[gi * 8 + 6], v[gi * 8 + 7]);
          } else {
#pragma unroll
            for (int r = 0; r < R; r++) c[(pb + (r << LO)) & ((1u << K) - 1)] = v[gi * R + r];
          }
        }
      } else if constexpr (LO == 0 && T == 0 && W == 3) {
        SH_GLOBAL uint4* q = reinterpret_cast<SH_GLOBAL uint4*>(dst);
        const uint32_t qi = index_of(pb) >> 2;
        at(q, qi) = make_uint4(v[gi * 8 + 0], v[gi * 8 + 1], v[gi * 8 + 2], v[gi * 8 + 3]);
        at(q, qi + 1) = make_uint4(v[gi * 8 + 4], v[gi * 8 + 5], v[gi * 8 + 6], v[gi * 8 + 7]);
      } else {
        const BufCol b(dst + block_base);
        const uint32_t i0 = index_of(pb) - block_base;
#pragma unroll
        for (int r = 0; r < R; r++) b.store(v[gi * R + r], i0, r_offset<LO>(r));
      }
    }
  }

  // LDS exchanges use the plain position as address.  Chunks 0 and 1 of the contiguous
  // passes are 8-way bank conflicted; an XOR swizzle that removes every conflict costs
  // ~150 VALU instructions per thread and measured slower (these passes are VALU-bound),
  // and 8 words of padding after every 64 (~20 instructions per thread, conflict free)
  // measured no faster: the conflicted exchanges overlap other waves' butterflies.
  template <int LO, int W>
  __device__ __forceinline__ void lds_write(const uint32_t (&v)[E]) const {
    constexpr int R = 1 << W;
#pragma unroll
    for (int gi = 0; gi < E / R; gi++) {
      const uint32_t pb = spread<T + LO, W>(threadIdx.x + NT * gi);
#pragma unroll
      for (int r = 0; r < R; r++) s[pb | (r << (T + LO))] = v[gi * R + r];
    }
  }
  template <int LO, int W>
  __device__ __forceinline__ void lds_read(uint32_t (&v)[E]) const {
    constexpr int R = 1 << W;
#pragma unroll
    for (int gi = 0; gi < E / R; gi++) {
      const uint32_t pb = spread<T + LO, W>(threadIdx.x + NT * gi);
#pragma unroll
      for (int r = 0; r < R; r++) v[gi * R + r] = s[pb | (r << (T + LO))];
    }
  }

  // radix-2^W butterflies on layer bits [LO, LO + W); forward passes skip layers >= act_log.
  // pre: the stage's twiddles, loaded by load_tw<LO, W>.
  template <int LO, int W>
  __device__ __forceinline__ void stage(uint32_t (&v)[E], const uint32_t (&pre)[NTW]) const {
    constexpr int R = 1 << W;
    static_for<0, W>([&](auto si) {
      constexpr int BL = INV ? si.value : (W - 1 - si.value);
      constexpr int BIT = LO + BL;
      if constexpr (MASKED) {
        if (A.a + BIT >= act()) return;  // uniform: layer over zero padding
      }
#pragma unroll
      for (int gi = 0; gi < E / R; gi++) {
        const uint32_t pb = spread<T + LO, W>(threadIdx.x + NT * gi);
#pragma unroll
        for (int rh = 0; rh < (R >> (BL + 1)); rh++) {
          const uint32_t p = pb | ((uint32_t)(rh << (BL + 1)) << (T + LO));
          const uint32_t tw = tw_fix<BIT>(pre[tw_slot<W, BL>(gi, rh)], p);
#pragma unroll
          for (int rl = 0; rl < (1 << BL); rl++) {
            const int e0 = gi * R + (rh << (BL + 1)) + rl;
            const int e1 = e0 + (1 << BL);
            const uint32_t x0 = v[e0], x1 = v[e1];
            if constexpr (INV) {
              v[e0] = add(x0, x1);
              v[e1] = mul_dbl(sub(x0, x1), tw);
            } else {
              const uint32_t t = mul_dbl(x1, tw);
              v[e0] = add(x0, t);
              v[e1] = sub(x0, t);
            }
          }
        }
      }
    });
  }

  // The twiddles of the next stage are loaded before the LDS exchange, so their latency
  // overlaps the barriers instead of following them.
  __device__ __forceinline__ void run() {
    uint32_t v[E], tw[NTW];
    static_for<0, NCH>([&](auto it) {
      constexpr int c = INV ? it.value : NCH - 1 - it.value;
      constexpr int LO = 3 * c;
      constexpr int W = (K - LO) < 3 ? (K - LO) : 3;
      if constexpr (it.value == 0) {
        load_tw<LO, W>(tw);
        load<LO, W>(v);
      } else {
        lds_read<LO, W>(v);
      }
      stage<LO, W>(v, tw);
      if constexpr (it.value + 1 < NCH) {
        constexpr int cn = INV ? it.value + 1 : NCH - 2 - it.value;
        constexpr int LON = 3 * cn;
        constexpr int WN = (K - LON) < 3 ? (K - LON) : 3;
        load_tw<LON, WN>(tw);
        if constexpr (it.value > 0) __syncthreads();  // previous round's reads are done
        lds_write<LO, W>(v);
        __syncthreads();
      } else {
        store<LO, W>(v);
      }
    });
  }
};

template <bool INV, int B, int NT, int K, int KIND, bool MASKED, int ACT = -1>
__global__ __launch_bounds__(NT) void fft_pass_kernel(PassArgs A) {
  __shared__ uint32_t s[1 << B];
  Pass<INV, B, NT, K, KIND, MASKED, ACT> ps(A, s);
  ps.run();
}

// Low-degree extension middle pass.  For columns of log n (14 <= n <= 21), after the
// contiguous inverse pass (layers 0..11): inverse layers 12..n-1 on a strided tile,
// scale by 2^-n and store the coefficients; then, from the same registers (the
// inverse's last chunk mapping is the forward's first), forward layers n-1..12 of the
// evaluation domain (log m = n + blowup) for each of its 2^(m-n) blocks -- the top
// m-n forward layers only duplicate the zero-padded coefficients -- and store the
// partial LDE.  The forward contiguous pass (layers 0..11) then finishes the LDE.
// One HBM read of the evaluations' partial transform replaces the separate inverse
// and forward strided passes (which read the coefficients back).
constexpr int BS_LDE = 14;
// The fused middle pass saves one HBM read of the coefficients, but it holds the
// coefficients and the forward tile in registers.
constexpr int NT_S = 1024;  // threads per 2^14-element tile: 16 values per thread
// Tile bits and threads of the fused pass for K layers: while K <= 8, 2^13-element tiles
// (32-word runs) on 512 threads, 16 values per thread (two radix-8 groups per stage), so
// the kept coefficients and the forward copy fit in 90 VGPRs and two blocks share a CU
// (256 columns of log 20: 1.33 ms, against 1.36 ms on 1024 threads with 8 values each
// and 1.63 ms for the separate strided passes; the default for log 14..20); 2^14 on 1024
// threads for K = 9 (98 VGPRs, one block per CU: a variant for the autotuner only).
template <int K>
constexpr int lde_tile_bits() {
  return K <= 8 ? 13 : BS_LDE;
}
template <int K>
constexpr int lde_threads() {
  return K <= 8 ? 512 : NT_S;
}
template <int K>
__global__ __launch_bounds__(lde_threads<K>()) __attribute__((amdgpu_waves_per_eu(K <= 8 ? 5 : 4)))
void fft_lde_strided_kernel(PassArgs Ai, PassArgs Af, uint32_t n_blocks) {
  constexpr int BL = lde_tile_bits<K>();
  __shared__ uint32_t s[1 << BL];
  using PI = Pass<true, BL, lde_threads<K>(), K, STRIDED>;
  using PF = Pass<false, BL, lde_threads<K>(), K, STRIDED, false>;
  constexpr int E = PI::E, NCH = PI::NCH;
  constexpr int LO_TOP = 3 * (NCH - 1), W_TOP = K - LO_TOP;
  // each stage's twiddles are loaded before the exchange that precedes it (as in Pass::run)
  PI pi(Ai, s);
  uint32_t v[E], tw[PI::NTW];
  static_for<0, NCH>([&](auto it) {
    constexpr int c = it.value;
    constexpr int LO = 3 * c;
    constexpr int W = (K - LO) < 3 ? (K - LO) : 3;
    if constexpr (c == 0) {
      pi.template load_tw<LO, W>(tw);
      pi.template load<LO, W>(v);
    } else {
      pi.template lds_read<LO, W>(v);
    }
    pi.template stage<LO, W>(v, tw);
    if constexpr (c + 1 < NCH) {
      constexpr int LON = 3 * (c + 1);
      constexpr int WN = (K - LON) < 3 ? (K - LON) : 3;
      pi.template load_tw<LON, WN>(tw);
      if constexpr (c > 0) __syncthreads();
      pi.template lds_write<LO, W>(v);
      __syncthreads();
    }
  });
  pi.template store<LO_TOP, W_TOP>(v);  // scales v in place: v now holds the coefficients
  PF pf(Af, s);
  for (uint32_t b = 0; b < n_blocks; b++) {
    pf.set_hi(b);
    uint32_t w[E];
#pragma unroll
    for (int e = 0; e < E; e++) w[e] = v[e];
    static_for<0, NCH>([&](auto it) {
      constexpr int c = NCH - 1 - it.value;
      constexpr int LO = 3 * c;
      constexpr int W = (K - LO) < 3 ? (K - LO) : 3;
      if constexpr (it.value == 0) pf.template load_tw<LO, W>(tw);
      if constexpr (it.value > 0) pf.template lds_read<LO, W>(w);
      pf.template stage<LO, W>(w, tw);
      if constexpr (c > 0) {
        constexpr int LON = 3 * (c - 1);
        constexpr int WN = (K - LON) < 3 ? (K - LON) : 3;
        pf.template load_tw<LON, WN>(tw);
        __syncthreads();  // earlier readers of s are done
        pf.template lds_write<LO, W>(w);
        __syncthreads();
      } else {
        pf.template store<LO, W>(w);
      }
    });
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
// columns per block band (8 measured ~1 % faster than one band over all 256 columns of
// log 20)
constexpr uint32_t FFT_BAND = 8;
constexpr uint32_t KS_MAX = BS - 5;

struct PassPlan {
  uint32_t a, k, t, hb, multi, B;
};

// Block order of a CONTIG / STRIDED launch over A.n_cols columns of 2^tiles_log tiles each.
static void set_block_order(PassArgs& A, uint32_t tiles_log) {
  A.n_groups = A.n_cols;
  A.band = std::min(A.n_cols, FFT_BAND);
  A.tiles_log = tiles_log;
  const uint32_t n_bands = ceil_div(A.n_cols, A.band);
  A.band_last = A.n_cols - (n_bands - 1) * A.band;
  A.band_magic = div_magic(A.band);
  A.band_last_magic = div_magic(A.band_last);
  // exactness of divu for the numerators formed: column indexes (< n_cols) and positions
  // within a band (< band 2^tiles_log)
  const uint64_t max_r = ((uint64_t)A.band << tiles_log);
  auto exact = [](uint64_t n_max, uint32_t d, uint32_t m) {
    return m == 0 || n_max * ((uint64_t)m * d - (1ull << 32)) < (1ull << 32);
  };
  require(tiles_log < 31 && max_r < (1ull << 31) && exact(A.n_cols, A.band, A.band_magic) &&
              exact(max_r, A.band, A.band_magic) && exact(max_r, A.band_last, A.band_last_magic),
          "fft: block order out of range");
}

// kc_pref: layers in the contiguous pass (12 or 13; 0 = the fewest passes, then 12).
static std::vector<PassPlan> plan(uint32_t log_n, uint32_t kc_pref = 0) {
  std::vector<PassPlan> ps;
  if (log_n <= 12) {
    ps.push_back({0, log_n, 0, 12 - log_n, 1, 12});
    return ps;
  }
  if (log_n == 13) {
    ps.push_back({0, 13, 0, 0, 0, 13});
    return ps;
  }
  uint32_t kc = kc_pref ? kc_pref : ((log_n - 12 > KS_MAX) ? 13 : 12);
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

static const char* pass_name(bool inv, const PassPlan& p) {
  if (p.multi) return inv ? "fft_inv_tile" : "fft_fwd_tile";
  if (p.a == 0) return inv ? "fft_inv_contig" : "fft_fwd_contig";
  return inv ? "fft_inv_strided" : "fft_fwd_strided";
}

template <bool INV, int B, int NT, int K, int KIND>
static void launch_k(uint32_t blocks, const PassArgs& A) {
  // MULTI tiles evaluated from 2^(K - 1) .. 2^(K - 3) coefficients (blowup 1..3): the
  // coefficient count is a template argument, so the skipped layers cost no branches and only
  // distinct coefficients are loaded (2^16 columns of 2^10 at blowup 4: 0.61 -> 0.57 ms)
  if constexpr (KIND == MULTI && !INV) {
    const uint32_t skip = K - A.act_log;
    if (A.act_log < K && skip <= 3) {
      if (skip == 1)
        fft_pass_kernel<false, B, NT, K, KIND, true, K - 1><<<blocks, NT, 0, ctx().cur>>>(A);
      else if (skip == 2)
        fft_pass_kernel<false, B, NT, K, KIND, true, K - 2><<<blocks, NT, 0, ctx().cur>>>(A);
      else
        fft_pass_kernel<false, B, NT, K, KIND, true, K - 3><<<blocks, NT, 0, ctx().cur>>>(A);
      return;
    }
  }
  // forward passes that reach layers at or above act_log (zero-padded coefficients) skip them
  if (!INV && A.act_log < A.a + K)
    fft_pass_kernel<false, B, NT, K, KIND, true><<<blocks, NT, 0, ctx().cur>>>(A);
  else
    fft_pass_kernel<INV, B, NT, K, KIND, false><<<blocks, NT, 0, ctx().cur>>>(A);
}

template <bool INV, int K>
static void launch_multi(uint32_t blocks, const PassArgs& A, uint32_t k) {
  if constexpr (K >= 3) {
    if (k == K) return launch_k<INV, 12, 256, K, MULTI>(blocks, A);
    launch_multi<INV, K - 1>(blocks, A, k);
  }
}
template <bool INV, int K>
static void launch_strided(uint32_t blocks, const PassArgs& A, uint32_t k) {
  if constexpr (K >= 1) {
    if (k == K) return launch_k<INV, BS, NT_S, K, STRIDED>(blocks, A);
    launch_strided<INV, K - 1>(blocks, A, k);
  }
}

template <bool INV>
static void launch_pass(const PassPlan& p, const PassArgs& base) {
  const char* name = pass_name(INV, p);
  PassArgs A = base;
  A.a = p.a;
  A.k = p.k;
  A.t = p.t;
  A.hb = p.hb;
  A.multi = p.multi;
  uint32_t blocks;
  require(p.multi || p.a == 0 || (p.t == BS - p.k && p.hb == 0), "fft: unsupported pass shape");
  if (p.multi) {
    A.n_groups = ceil_div(A.n_cols, 1u << p.hb);
    A.lob_count_log = 0;
    blocks = A.n_groups;
  } else {
    A.lob_count_log = p.a - p.t;
    uint32_t hi_log = A.log_n - p.a - p.k - p.hb;
    set_block_order(A, A.lob_count_log + hi_log);
    blocks = A.n_cols << (A.lob_count_log + hi_log);
  }
  timed(name, [&] {
    if (p.multi)
      launch_multi<INV, 12>(blocks, A, p.k);
    else if (p.a == 0 && p.B == 12)
      launch_k<INV, 12, 256, 12, CONTIG>(blocks, A);
    else if (p.a == 0)
      launch_k<INV, 13, 512, 13, CONTIG>(blocks, A);
    else
      launch_strided<INV, KS_MAX>(blocks, A, p.k);
  });
}

// Inverse FFT of n_cols columns of one log size (device pointer tables): the first pass
// reads d_src, every pass writes d_cols (d_src == d_cols: in place).
void interpolate_same_size(uint32_t* const* d_cols, uint32_t n_cols, uint32_t log_n, const uint32_t* itw,
                           uint64_t tree_len, const uint32_t* const* d_src = nullptr, uint32_t kc = 0) {
  if (!d_src) d_src = (const uint32_t* const*)d_cols;
  require(log_n >= 1 && log_n <= 30, "interpolate: log size out of range");
  require(tree_len >= (1ull << (log_n - 1)), "interpolate: twiddle tree smaller than the domain");
  if (log_n <= 2) {
    timed("fft_tiny_inv", [&] {
      fft_tiny_kernel<true><<<ceil_div(n_cols, 64), 64, 0, ctx().cur>>>(d_cols, d_src, n_cols, log_n, log_n);
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
  auto ps = plan(log_n, kc);
  for (size_t i = 0; i < ps.size(); i++) {
    A.src = i == 0 ? d_src : (const uint32_t* const*)d_cols;
    // scale by 2^-log_n in the last pass: 2^-n = 2^(31 - n) since 2^31 = 1 (log_n >= 3 here)
    A.scale = (i + 1 == ps.size()) ? 31 - log_n : 0u;
    launch_pass<true>(ps[i], A);
  }
}

// Forward FFT: src columns (2^src_log coefficients) evaluated on the canonic domain
// of log_n >= src_log, written to dst (device pointer tables).
void evaluate_same_size(uint32_t* const* d_dst, const uint32_t* const* d_src, uint32_t n_cols, uint32_t src_log,
                        uint32_t log_n, const uint32_t* tw, uint64_t tree_len, uint32_t kc = 0) {
  require(log_n >= 1 && log_n <= 30 && src_log <= log_n, "evaluate: log sizes out of range");
  require(tree_len >= (1ull << (log_n - 1)), "evaluate: twiddle tree smaller than the domain");
  if (log_n <= 2) {
    timed("fft_tiny_fwd", [&] {
      fft_tiny_kernel<false><<<ceil_div(n_cols, 64), 64, 0, ctx().cur>>>(d_dst, d_src, n_cols, log_n, src_log);
    });
    return;
  }
  PassArgs A{};
  A.dst = d_dst;
  A.tw = tw + (tree_len - (1ull << (log_n - 1)));
  A.n_cols = n_cols;
  A.log_n = log_n;
  A.scale = 0;
  auto ps = plan(log_n, kc);
  for (size_t i = ps.size(); i-- > 0;) {
    bool first = (i + 1 == ps.size());
    A.src = first ? d_src : (const uint32_t* const*)d_dst;
    A.src_log = first ? src_log : log_n;
    A.act_log = src_log;
    launch_pass<false>(ps[i], A);
  }
}

template <int K>
static void launch_lde_strided(uint32_t blocks, const PassArgs& Ai, const PassArgs& Af, uint32_t n_blocks, uint32_t k) {
  if constexpr (K >= 2) {
    if (k == K) {
      fft_lde_strided_kernel<K><<<blocks, lde_threads<K>(), 0, ctx().cur>>>(Ai, Af, n_blocks);
      return;
    }
    launch_lde_strided<K - 1>(blocks, Ai, Af, n_blocks, k);
  }
}

// Interpolate d_src (2^n evaluations) into d_coeffs and evaluate on the canonic domain of
// log n + log_blowup into d_lde.  d_src may equal d_coeffs.
void lde_same_size(const uint32_t* const* d_src, uint32_t* const* d_coeffs, uint32_t* const* d_lde, uint32_t n_cols,
                   uint32_t n, uint32_t log_blowup, const uint32_t* tw, const uint32_t* itw, uint64_t tree_len,
                   LdeVariant var) {
  const uint32_t m = n + log_blowup;
  require(m <= 30, "lde: log size out of range");
  require(tree_len >= (1ull << (m - 1)), "lde: twiddle tree smaller than the evaluation domain");
  if (!var.fused || n < 14 || n > 12 + KS_MAX) {
    const uint32_t kci = (var.kc == 13 && n >= 14 && n - 13 <= KS_MAX) ? 13 : 0;
    const uint32_t kcf = (var.kc == 13 && m >= 14 && m - 13 <= KS_MAX) ? 13 : 0;
    interpolate_same_size(d_coeffs, n_cols, n, itw, tree_len, d_src, kci);
    evaluate_same_size(d_lde, (const uint32_t* const*)d_coeffs, n_cols, n, m, tw, tree_len, kcf);
    return;
  }
  const uint32_t k = n - 12, t = (k <= 8 ? 13 : BS_LDE) - k;
  PassArgs Ai{};
  Ai.src = d_src;
  Ai.dst = d_coeffs;
  Ai.tw = itw + (tree_len - (1ull << (n - 1)));
  Ai.n_cols = n_cols;
  Ai.log_n = n;
  Ai.src_log = n;
  Ai.act_log = n;
  // inverse layers 0..11, evaluations -> d_coeffs
  launch_pass<true>(PassPlan{0, 12, 0, 0, 0, 12}, Ai);
  // fused middle pass
  Ai.src = (const uint32_t* const*)d_coeffs;
  Ai.a = 12;
  Ai.k = k;
  Ai.t = t;
  Ai.hb = 0;
  Ai.multi = 0;
  Ai.scale = 31 - n;
  Ai.lob_count_log = 12 - t;
  set_block_order(Ai, 12 - t);
  PassArgs Af = Ai;
  Af.src = (const uint32_t* const*)d_lde;
  Af.dst = d_lde;
  Af.tw = tw + (tree_len - (1ull << (m - 1)));
  Af.log_n = m;
  Af.src_log = m;
  Af.act_log = n;
  Af.scale = 0;
  const uint32_t blocks = n_cols << (12 - t);
  timed("fft_lde_strided", [&] { launch_lde_strided<KS_MAX>(blocks, Ai, Af, 1u << log_blowup, k); });
  // forward layers 0..11 on the LDE
  PassArgs Ac{};
  Ac.src = (const uint32_t* const*)d_lde;
  Ac.dst = d_lde;
  Ac.tw = tw + (tree_len - (1ull << (m - 1)));
  Ac.n_cols = n_cols;
  Ac.log_n = m;
  Ac.src_log = m;
  Ac.act_log = m;
  launch_pass<false>(PassPlan{0, 12, 0, 0, 0, 12}, Ac);
}

void lde_mixed(const uint32_t* const* src, uint32_t* const* coeffs, uint32_t* const* lde, const uint32_t* logs,
               uint32_t n_cols, uint32_t log_blowup, const uint32_t* tw, const uint32_t* itw, uint64_t tree_len) {
  std::map<uint32_t, std::vector<const void*>> by;
  for (uint32_t i = 0; i < n_cols; i++) {
    auto& e = by[logs[i]];
    e.push_back(src[i]);
    e.push_back(coeffs[i]);
    e.push_back(lde[i]);
  }
  for (auto& kv : by) {
    size_t cnt = kv.second.size() / 3;
    std::vector<const void*> t(3 * cnt);
    for (size_t i = 0; i < cnt; i++) {
      t[i] = kv.second[3 * i];
      t[cnt + i] = kv.second[3 * i + 1];
      t[2 * cnt + i] = kv.second[3 * i + 2];
    }
    DevBuf d = upload(t.data(), t.size() * sizeof(void*));
    const void* const* base = d.as<const void*>();
    lde_same_size((const uint32_t* const*)base, (uint32_t* const*)(base + cnt), (uint32_t* const*)(base + 2 * cnt),
                  (uint32_t)cnt, kv.first, log_blowup, tw, itw, tree_len, lde_variant_for(kv.first, log_blowup, cnt));
  }
}

// Group columns by size and run each class.  cols/logs are host arrays; src (optional)
// holds the inputs when they are not transformed in place.
void interpolate_mixed(uint32_t* const* cols, const uint32_t* logs, uint32_t n_cols, const uint32_t* itw,
                       uint64_t tree_len, const uint32_t* const* src = nullptr) {
  // one size (the common case, e.g. 2^16 columns of one trace): the caller's pointer table
  // is uploaded as it is, once when the columns are transformed in place
  if (n_cols && std::all_of(logs, logs + n_cols, [&](uint32_t l) { return l == logs[0]; })) {
    DevBuf table = upload(cols, (size_t)n_cols * sizeof(void*));
    DevBuf stable = src ? upload(src, (size_t)n_cols * sizeof(void*)) : DevBuf();
    interpolate_same_size(table.as<uint32_t*>(), n_cols, logs[0], itw, tree_len,
                          src ? stable.as<const uint32_t*>() : table.as<const uint32_t*>());
    return;
  }
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
  // one (source, destination) size pair: the caller's tables as they are, one upload when
  // the columns are evaluated in place
  if (n_cols && std::all_of(src_logs, src_logs + n_cols, [&](uint32_t l) { return l == src_logs[0]; }) &&
      std::all_of(dst_logs, dst_logs + n_cols, [&](uint32_t l) { return l == dst_logs[0]; })) {
    const bool in_place = std::equal(src, src + n_cols, (const uint32_t* const*)dst);
    DevBuf td = upload(dst, (size_t)n_cols * sizeof(void*));
    DevBuf ts = in_place ? DevBuf() : upload(src, (size_t)n_cols * sizeof(void*));
    evaluate_same_size(td.as<uint32_t*>(), in_place ? td.as<const uint32_t*>() : ts.as<const uint32_t*>(), n_cols,
                       src_logs[0], dst_logs[0], tw, tree_len);
    return;
  }
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

extern "C" int stwo_hip_lde_columns_async(const uint32_t* const* evals, uint32_t* const* coeffs, uint32_t* const* lde,
                                          const uint32_t* log_sizes, uint32_t n_cols, uint32_t log_blowup,
                                          const uint32_t* twiddles, const uint32_t* itwiddles, uint64_t tree_len) {
  return guarded([&] {
    for (uint32_t i = 0; i < n_cols; i++) {
      require(log_sizes[i] >= 1, "lde: empty column");
      require((const void*)lde[i] != (const void*)coeffs[i] && (const void*)lde[i] != (const void*)evals[i],
              "lde: the LDE buffer must not alias its inputs");
    }
    fft::lde_mixed(evals, coeffs, lde, log_sizes, n_cols, log_blowup, twiddles, itwiddles, tree_len);
  });
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
