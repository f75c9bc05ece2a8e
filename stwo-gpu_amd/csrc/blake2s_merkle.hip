// Blake2s Merkle commitment: MerkleOps<Blake2sMerkleHasher>::commit_on_layer
// (stwo_gpu_backend/src/blake2s.rs:19-97) and the whole-tree commit that stwo's
// MerkleProver::commit drives one layer at a time.
//
// Hash of a node (cuda/src/blake2s.cu:97-151): state = 0^8; if the node has
// children, state = F(state, left || right); then for each chunk of 16 columns
// of this layer (zero-padded remainder) state = F(state, chunk).  F is the
// BLAKE2s compression with the IV in v[8..15] and zero counter and flags
// (blake2s.cu:56-95).
//
// MI355X design: one lane per node; a wavefront reads 64 consecutive rows of a
// column (256-byte coalesced loads); the tree below the wide layers is built by
// subtree launches over runs of light layers in LDS; a run topped by at most 2^12
// nodes takes one launch: blocks build 128-node subtrees and the last block to finish
// builds the levels above their roots; wider runs take one launch per 8-9 layers.  Columns of any
// layer (mixed-degree commitments) are injected as the layers are built.
#include <algorithm>
#include <cstring>

#include "field.hpp"
#include "gmem.hpp"
#include "blake2s_quad.hpp"
#include "merkle_plan.hpp"
#include "runtime.hpp"

namespace sh {
namespace merkle {

__device__ __forceinline__ uint32_t rotr(uint32_t x, uint32_t c) { return __builtin_amdgcn_alignbit(x, x, c); }

#define B2_G(a, b, c, d, x, y) \
  a = a + b + (x);             \
  d = rotr(d ^ a, 16);         \
  c = c + d;                   \
  b = rotr(b ^ c, 12);         \
  a = a + b + (y);             \
  d = rotr(d ^ a, 8);          \
  c = c + d;                   \
  b = rotr(b ^ c, 7);

#define B2_ROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
  B2_G(v0, v4, v8, v12, m[s0], m[s1]);                                                  \
  B2_G(v1, v5, v9, v13, m[s2], m[s3]);                                                  \
  B2_G(v2, v6, v10, v14, m[s4], m[s5]);                                                 \
  B2_G(v3, v7, v11, v15, m[s6], m[s7]);                                                 \
  B2_G(v0, v5, v10, v15, m[s8], m[s9]);                                                 \
  B2_G(v1, v6, v11, v12, m[s10], m[s11]);                                               \
  B2_G(v2, v7, v8, v13, m[s12], m[s13]);                                                \
  B2_G(v3, v4, v9, v14, m[s14], m[s15]);

// h <- F(h, m) with zero counter and flags
__device__ __forceinline__ void compress(uint32_t (&h)[8], const uint32_t (&m)[16]) {
  uint32_t v0 = h[0], v1 = h[1], v2 = h[2], v3 = h[3], v4 = h[4], v5 = h[5], v6 = h[6], v7 = h[7];
  uint32_t v8 = 0x6A09E667u, v9 = 0xBB67AE85u, v10 = 0x3C6EF372u, v11 = 0xA54FF53Au;
  uint32_t v12 = 0x510E527Fu, v13 = 0x9B05688Cu, v14 = 0x1F83D9ABu, v15 = 0x5BE0CD19u;
  B2_ROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
  B2_ROUND(14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3)
  B2_ROUND(11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4)
  B2_ROUND(7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8)
  B2_ROUND(9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13)
  B2_ROUND(2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9)
  B2_ROUND(12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11)
  B2_ROUND(13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10)
  B2_ROUND(6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5)
  B2_ROUND(10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0)
  h[0] ^= v0 ^ v8;
  h[1] ^= v1 ^ v9;
  h[2] ^= v2 ^ v10;
  h[3] ^= v3 ^ v11;
  h[4] ^= v4 ^ v12;
  h[5] ^= v5 ^ v13;
  h[6] ^= v6 ^ v14;
  h[7] ^= v7 ^ v15;
}

// Absorb the columns of one layer at `row` (16-column chunks, zero-padded remainder).
// jo: the job offset of the column pointers (see jptr).  The column table is read through
// the scalar cache and each column becomes a scalar buffer descriptor with row * 4 as the
// lane's offset (global loads of pointer + jo + row * 4 compile to a 64-bit vector add per
// load, and generic-pointer reads to flat loads).
// Row row_base + row_off: row_base uniform (the descriptor starts there, so the lane offsets
// stay small), row_off per lane.
__device__ __forceinline__ uint32_t column_word(const uint32_t* const* cols, uint32_t c, uint32_t row_base,
                                                uint32_t row_off, uint64_t jo) {
  return BufCol(jptr(ld_uniform(cols, c), jo) + row_base).load(row_off, 0);
}
__device__ __forceinline__ void absorb_columns_simple(uint32_t (&h)[8], const uint32_t* const* cols, uint32_t n_cols,
                                                      uint32_t row_base, uint32_t row_off, uint64_t jo) {
  uint32_t c = 0;
  for (; c + 16 <= n_cols; c += 16) {
    uint32_t m[16];
#pragma unroll
    for (int j = 0; j < 16; j++) m[j] = column_word(cols, c + j, row_base, row_off, jo);
    compress(h, m);
  }
  if (c < n_cols) {
    uint32_t m[16];
#pragma unroll
    for (int j = 0; j < 16; j++) m[j] = (c + j < n_cols) ? column_word(cols, c + j, row_base, row_off, jo) : 0u;
    compress(h, m);
  }
}

__device__ __forceinline__ void store_hash(SH_GLOBAL Hash* out, uint32_t i, const uint32_t (&h)[8]) {
  SH_GLOBAL uint4* o = reinterpret_cast<SH_GLOBAL uint4*>(out + i);
  o[0] = make_uint4(h[0], h[1], h[2], h[3]);
  o[1] = make_uint4(h[4], h[5], h[6], h[7]);
}
__device__ __forceinline__ void load_children(const SH_GLOBAL Hash* prev, uint32_t i, uint32_t (&m)[16]) {
  const SH_GLOBAL uint4* p = reinterpret_cast<const SH_GLOBAL uint4*>(prev + 2 * (size_t)i);
  uint4 a = p[0], b = p[1], c = p[2], d = p[3];
  m[0] = a.x; m[1] = a.y; m[2] = a.z; m[3] = a.w;
  m[4] = b.x; m[5] = b.y; m[6] = b.z; m[7] = b.w;
  m[8] = c.x; m[9] = c.y; m[10] = c.z; m[11] = c.w;
  m[12] = d.x; m[13] = d.y; m[14] = d.z; m[15] = d.w;
}

// One layer, one lane per node (commit_on_first_layer / commit_on_layer_with_previous).
// js: byte stride between the jobs of a batched launch (blockIdx.z = job; 0 for one job)
__global__ __launch_bounds__(256) void layer_kernel(uint32_t size, uint32_t n_cols, const uint32_t* const* cols,
                                                    const Hash* prev, Hash* out, uint64_t js) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= size) return;
  const uint64_t jo = job_offset(js);
  uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (prev) {
    uint32_t m[16];
    load_children(as_global(jptr(prev, jo)), i, m);
    compress(h, m);
  }
  absorb_columns_simple(h, cols, n_cols, blockIdx.x * blockDim.x, threadIdx.x, jo);
  store_hash(as_global(jptr(out, jo)), i, h);
}


// (the quad compression, quad_compress, is in blake2s_quad.hpp)

// Narrow layer, one quad per node (4 x size lanes): children compress, then the 16-column
// chunks, each staged through the quad's 16 LDS words.  With STAGED the block first copies
// the column-pointer table to LDS: a lane's pointer then comes from an LDS read instead of
// a global load, so waiting for it does not drain the column loads already in flight
// (vector-memory loads complete in order).
constexpr uint32_t QUAD_STAGED_MAX_COLS = 4096;  // 32 KiB of pointers
template <bool STAGED>
__global__ __launch_bounds__(256) void quad_layer_kernel(uint32_t size, uint32_t n_cols, const uint32_t* const* cols,
                                                         const Hash* prev, Hash* out, uint64_t js) {
  __shared__ uint32_t msg[64][16];
  extern __shared__ const uint32_t* sp[];  // STAGED: the n_cols column pointers
  if constexpr (STAGED) {
    for (uint32_t j = threadIdx.x; j < n_cols; j += blockDim.x) sp[j] = cols[j];
    __syncthreads();
  }
  const uint64_t jo = job_offset(js);
  auto col_ptr = [&](uint32_t col) -> const uint32_t* {
    if constexpr (STAGED) return jptr(sp[col], jo);
    else return jptr(cols[col], jo);
  };
  if (prev) prev = jptr(prev, jo);
  out = jptr(out, jo);
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = t >> 2, q = t & 3, lq = threadIdx.x >> 2;
  if (i >= size) return;  // whole quads leave together
  uint32_t* mb = msg[lq];
  uint32_t h_lo = 0, h_hi = 0;
  if (prev) {
    const uint4 w = reinterpret_cast<const uint4*>(prev + 2 * (size_t)i)[q];
    mb[4 * q] = w.x;
    mb[4 * q + 1] = w.y;
    mb[4 * q + 2] = w.z;
    mb[4 * q + 3] = w.w;
    __builtin_amdgcn_wave_barrier();
    quad_compress(h_lo, h_hi, mb, q);
  }
  // column chunks: the loads of the next PF chunks are in flight while one is hashed (a
  // narrow layer has few waves, so a chunk's HBM latency is hidden only by distance)
  constexpr int PF = 4;
  uint32_t nx[PF][4];
#pragma unroll
  for (int d = 0; d < PF; d++)
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t col = 16 * d + 4 * q + k;
      nx[d][k] = col < n_cols ? at(as_global(col_ptr(col)), i) : 0u;
    }
  for (uint32_t c0 = 0; c0 < n_cols; c0 += 16 * PF) {
#pragma unroll
    for (int d = 0; d < PF; d++) {
      const uint32_t c = c0 + 16 * d;
      if (c >= n_cols) break;  // uniform
      uint32_t cur[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        cur[k] = nx[d][k];
        const uint32_t col = c + 16 * PF + 4 * q + k;
        nx[d][k] = col < n_cols ? at(as_global(col_ptr(col)), i) : 0u;
      }
      __builtin_amdgcn_wave_barrier();  // the quad's reads of the previous message are done
#pragma unroll
      for (int k = 0; k < 4; k++) mb[4 * q + k] = cur[k];
      __builtin_amdgcn_wave_barrier();
      quad_compress(h_lo, h_hi, mb, q);
    }
  }
  out[i].s[q] = h_lo;
  out[i].s[4 + q] = h_hi;
}

// Subtree levels: layers top, top-1, ..., top-depth+1 (each halving), built by one block.
// The block reads 2^(w+1) children hashes of layer top+1 from `prev` (global memory)
// starting at child 2 base0 (none when `prev` is null: layer `top` is the leaf layer),
// produces the nodes0 = 2^w nodes of layer `top` starting at
// base0 and halves them in LDS down to 2^(w-depth+1) nodes of layer top-depth+1, i.e. it
// builds 2^(w-depth+1) adjacent subtrees side by side (w >= depth-1).  Wide levels take one
// lane per node; levels of at most NT/4 nodes one lane quad per node (the four G
// functions of each half-round side by side).  outs[k] and inj[k] are indexed by layer k;
// s holds nodes0 hashes.
// Leaf r of a LeafFold (lf.src[0] non-null): computed from the folded layer, written to
// the leaf layer's four columns (cols, job offset jo) by lane `writer` (all lanes when it
// is 4), and returned.
__device__ __forceinline__ qm31 leaf_fold_value(const LeafFold& lf, const uint32_t* const* cols, uint32_t r,
                                                uint64_t jo, uint32_t writer, uint32_t lane) {
  const qm31 al = *jptr(lf.alpha, jo);
  qm31 a, b;
  uint32_t* ac[4] = {&a.x.a, &a.x.b, &a.y.a, &a.y.b};
  uint32_t* bc[4] = {&b.x.a, &b.x.b, &b.y.a, &b.y.b};
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const uint2 w = at(reinterpret_cast<const SH_GLOBAL uint2*>(as_global(jptr(lf.src[c], jo))), r);
    *ac[c] = w.x;
    *bc[c] = w.y;
  }
  const qm31 f = fold_mat(a, b, at(as_global(lf.tw), r), qmat_of(al));
  const uint32_t fw[4] = {f.x.a, f.x.b, f.y.a, f.y.b};
#pragma unroll
  for (int c = 0; c < 4; c++)
    if (writer == 4 || (uint32_t)c == lane)
      at(as_global(jptr(const_cast<uint32_t*>(ld_uniform(cols, c)), jo)), r) = fw[c];  // the plan's leaf columns
  return f;
}

template <int NT>
__device__ __forceinline__ void build_levels(const Hash* prev, uint32_t top, uint32_t depth, uint32_t nodes0,
                                             uint32_t base0, Hash* const* outs, const LayerCols* inj, Hash* s,
                                             uint32_t (*qmsg)[16], uint64_t jo, const LeafFold& lf) {
  for (uint32_t d = 0; d < depth; d++) {
    const uint32_t layer = top - d;
    const uint32_t cnt = nodes0 >> d;
    const uint32_t base = base0 >> d;
    const LayerCols lc = ld_uniform(inj, layer);
    SH_GLOBAL Hash* const out = as_global(jptr(outs[layer], jo));
    // the leaf level of a LeafFold launch: four columns, computed here (prev is null)
    const bool fold = d == 0 && lf.src[0] != nullptr;
    if ((d > 0 || !prev) && cnt <= NT / 4) {
      // narrow level: one quad per node, children read from s in place (a leaf level of at
      // most NT / 4 nodes too: it has no children, only its columns)
      const uint32_t i = threadIdx.x >> 2, q = threadIdx.x & 3;
      uint32_t h_lo = 0, h_hi = 0;
      if (i < cnt) {
        if (d > 0) quad_compress(h_lo, h_hi, s[2 * i].s, q);
        uint32_t* mb = qmsg[i];
        if (fold) {
          // each lane of the quad folds the leaf, stores its own coordinate and stages it
          const qm31 f = leaf_fold_value(lf, lc.cols, base + i, jo, q, q);
          const uint32_t fw[4] = {f.x.a, f.x.b, f.y.a, f.y.b};
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int k = 0; k < 4; k++) mb[4 * q + k] = q == 0 ? fw[k] : 0u;
          __builtin_amdgcn_wave_barrier();
          quad_compress(h_lo, h_hi, mb, q);
        }
        for (uint32_t c = fold ? lc.n_cols : 0; c < lc.n_cols; c += 16) {
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const uint32_t col = c + 4 * q + k;
            mb[4 * q + k] = col < lc.n_cols ? at(as_global(jptr(at(as_global(lc.cols), col), jo)), base + i) : 0u;
          }
          __builtin_amdgcn_wave_barrier();
          quad_compress(h_lo, h_hi, mb, q);
        }
        out[base + i].s[q] = h_lo;
        out[base + i].s[4 + q] = h_hi;
      }
      __syncthreads();  // all reads of s for this layer are done
      if (i < cnt) {
        s[i].s[q] = h_lo;
        s[i].s[4 + q] = h_hi;
      }
      __syncthreads();
      continue;
    }
    // one node per lane per round; a round's results are written back after a barrier, and
    // later rounds read children at indices above everything written so far
    for (uint32_t r0 = 0; r0 < cnt; r0 += NT) {
      const uint32_t i = r0 + threadIdx.x;
      uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (i < cnt) {
        if (d > 0 || prev) {  // leaves (no prev) have no children to compress
          uint32_t m[16];
          if (d == 0) {
            load_children(as_global(prev), base + i, m);
          } else {
#pragma unroll
            for (int w = 0; w < 8; w++) {
              m[w] = s[2 * i].s[w];
              m[8 + w] = s[2 * i + 1].s[w];
            }
          }
          compress(h, m);
        }
        if (fold) {
          const qm31 f = leaf_fold_value(lf, lc.cols, base + i, jo, 4, 0);
          uint32_t m[16] = {f.x.a, f.x.b, f.y.a, f.y.b, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
          compress(h, m);
        } else {
          absorb_columns_simple(h, lc.cols, lc.n_cols, base + r0, threadIdx.x, jo);
        }
        store_hash(out, base + i, h);
      }
      __syncthreads();  // this round's reads of s are done
      if (i < cnt) {
#pragma unroll
        for (int w = 0; w < 8; w++) s[i].s[w] = h[w];
      }
      __syncthreads();
    }
  }
}

// Subtree launch over layers top .. top-depth+1.  Block b builds the levels of its
// 2^w_log nodes of layer `top` (build_levels).  With `arrive` (a counter that is zero
// at launch), each block builds w_log + 1 levels, down to one root per block, and the
// last block to finish builds the remaining depth - w_log - 1 levels from all block
// roots, so a whole tree top needs one launch whose wide levels still spread over the
// chip.  No block waits for another: each block publishes its nodes (device-scope fence)
// and counts itself in; the block that counts last continues, the others exit.  The last
// block resets the counter for the next launch.
constexpr uint32_t SUB_BLOCK_LOG = 7;  // nodes per block at the top layer (one lane quad each below)
// Runs topped by at most 2^12 nodes (at most 32 blocks) finish in their last block.  The
// device-scope fence of every block writes its XCD's L2 back, which costs more than a
// second launch once there are hundreds of blocks: a wider run is cut into at most
// 2^SUB_WIDE_BLOCKS_LOG blocks (one per CU; 149 VGPRs keep one 512-thread block per CU)
// of 2^7..2^8 nodes, each reduced to its root, and the layers above the roots form the
// next run.
constexpr uint32_t TAIL_MAX_TOP = 12;
constexpr uint32_t SUB_WIDE_BLOCKS_LOG = 8;
static uint32_t wide_w_log(uint32_t top) { return std::max<uint32_t>(top - SUB_WIDE_BLOCKS_LOG, SUB_BLOCK_LOG); }
constexpr int SUB_NT = 512;  // threads per subtree block (measured: 512 ahead of 256; 1024 spills)
template <int NT>
__global__ __launch_bounds__(NT) void subtree_kernel(const Hash* prev, uint32_t top, uint32_t depth, uint32_t w_log,
                                                     Hash* const* outs, const LayerCols* inj, RootDraw rd,
                                                     uint32_t* arrive, uint64_t js, LeafFold lf) {
  extern __shared__ Hash s[];  // max(2^w_log, block count / 2) hashes
  __shared__ uint32_t qmsg[NT / 4][16];  // quad mode: staging of injected column chunks
  __shared__ uint32_t is_last;
  // batched jobs (blockIdx.z): every arena pointer moves by the job's offset; one arrival
  // counter per job
  const uint64_t jo = job_offset(js);
  if (prev) prev = jptr(prev, jo);
  if (arrive) arrive += blockIdx.z;
  if (rd.ch) {
    rd.ch = jptr(rd.ch, jo);
    rd.dst = jptr(rd.dst, jo);
  }
  const uint32_t bd = arrive ? w_log + 1 : depth;  // levels built by every block
  build_levels<NT>(prev, top, bd, 1u << w_log, blockIdx.x << w_log, outs, inj, s, qmsg, jo, lf);
  if (arrive) {
    __threadfence();  // this block's nodes are visible device-wide before it counts in
    __syncthreads();
    if (threadIdx.x == 0) is_last = atomicAdd(arrive, 1u) == gridDim.x - 1 ? 1u : 0u;
    __syncthreads();
    if (!is_last) return;  // uniform over the block
    __threadfence();       // every other block's nodes are visible to the loads below
    if (threadIdx.x == 0) *arrive = 0u;
    const uint32_t t_top = top - bd;  // the block roots are layer t_top + 1
    build_levels<NT>(jptr(outs[t_top + 1], jo), t_top, depth - bd, gridDim.x >> 1, 0, outs, inj, s, qmsg, jo,
                     LeafFold{});
  }
  // the launch that wrote the root (layer 0, held in s[0] after the last barrier) runs
  // the channel step on one lane quad: Channel::mix8 and Channel::draw_qm31 / draw_point
  if (rd.ch && top + 1 == depth && threadIdx.x < 4) channel_step_quad(rd, s[0].s, qmsg[0], threadIdx.x);
}

static uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

// A layer with many injected columns is a long per-node hash chain (1 + ceil(cols/16)
// compressions); it gets its own launch with one lane per node, so all of its nodes
// run in parallel.  Runs of layers with few columns are built by subtree launches.
constexpr uint32_t HEAVY_CHUNKS = 2;
constexpr uint32_t QUAD_LAYER_MAX = 14;  // layers of at most 2^14 nodes are hashed by quads
constexpr int32_t WIDE_LAYER = 16;

size_t MerklePlan::table_bytes(uint32_t n_cols, uint32_t max_log, uint32_t n_jobs) {
  const size_t cols = ((size_t)std::max<uint32_t>(n_cols, 1) * sizeof(void*) + 15) & ~(size_t)15;
  const size_t lcs = ((size_t)(max_log + 1) * sizeof(LayerCols) + 15) & ~(size_t)15;
  const size_t outs = ((size_t)(max_log + 1) * sizeof(Hash*) + 15) & ~(size_t)15;
  return cols + lcs + outs + (size_t)(max_log + 1) * n_jobs * sizeof(uint32_t);
}

MerklePlan::MerklePlan(uint32_t n_cols, const uint32_t* const* cols, const uint32_t* logs, Hash* const* layers,
                       uint32_t max_log_)
    : max_log(max_log_) {
  require(max_log <= 31, "merkle: log size too large");
  const size_t bytes = table_bytes(n_cols, max_log);
  own = DevBuf(bytes);
  std::vector<char> host(bytes);
  build(n_cols, cols, logs, layers, own.as<char>(), host.data());
  upload_to(own.p, host.data(), bytes);
}

MerklePlan::MerklePlan(uint32_t n_cols, const uint32_t* const* cols, const uint32_t* logs, Hash* const* layers,
                       uint32_t max_log_, void* dev, void* host_image, uint32_t n_jobs_, uint64_t jstride_)
    : max_log(max_log_), n_jobs(n_jobs_), jstride(jstride_) {
  require(max_log <= 31, "merkle: log size too large");
  require(n_jobs >= 1 && n_jobs <= 65535, "merkle: bad job count");
  build(n_cols, cols, logs, layers, (char*)dev, (char*)host_image);
}

void MerklePlan::build(uint32_t n_cols, const uint32_t* const* cols, const uint32_t* logs, Hash* const* layers,
                       char* dev, char* host) {
  // stable sort by decreasing size (MerkleProver::commit)
  std::vector<std::vector<const uint32_t*>> by_layer(max_log + 1);
  for (uint32_t i = 0; i < n_cols; i++) {
    require(logs[i] <= max_log, "merkle: column larger than the tree");
    require(logs[i] <= 30, "merkle: columns hold at most 2^30 values");
    by_layer[logs[i]].push_back(cols[i]);
  }
  std::vector<const uint32_t*> flat;
  std::vector<size_t> off(max_log + 1);
  for (uint32_t k = 0; k <= max_log; k++) {
    off[k] = flat.size();
    flat.insert(flat.end(), by_layer[k].begin(), by_layer[k].end());
  }
  if (flat.empty()) flat.push_back(nullptr);
  const size_t o_lcs = (flat.size() * sizeof(void*) + 15) & ~(size_t)15;
  const size_t o_outs = o_lcs + (((size_t)(max_log + 1) * sizeof(LayerCols) + 15) & ~(size_t)15);
  const uint32_t* const* dcols = (const uint32_t* const*)dev;
  lcs.resize(max_log + 1);
  for (uint32_t k = 0; k <= max_log; k++) lcs[k] = {dcols + off[k], (uint32_t)by_layer[k].size(), 0};
  outs.assign(layers, layers + max_log + 1);
  memcpy(host, flat.data(), flat.size() * sizeof(void*));
  memcpy(host + o_lcs, lcs.data(), lcs.size() * sizeof(LayerCols));
  memcpy(host + o_outs, outs.data(), outs.size() * sizeof(Hash*));
  // arrival counters of the subtree launches that finish in their last block, by top layer;
  // zero at upload, and each launch leaves its counter at zero
  const size_t o_arrive = o_outs + (((size_t)(max_log + 1) * sizeof(Hash*) + 15) & ~(size_t)15);
  memset(host + o_arrive, 0, (size_t)(max_log + 1) * n_jobs * sizeof(uint32_t));
  dlcs = (const LayerCols*)(dev + o_lcs);
  douts = (Hash* const*)(dev + o_outs);
  darrive = (uint32_t*)(dev + o_arrive);
  // launch schedule, top-down: wide layers (more than 2^16 nodes, enough lanes to fill
  // the chip) and injection-heavy layers get one launch each, runs of the remaining
  // layers are built by subtree launches
  // a small tree whose leaves take one compression is built from its leaves by one
  // subtree launch (no separate leaf launch)
  const bool leaves_in_subtree = max_log <= TAIL_MAX_TOP && lcs[max_log].n_cols <= 16;
  if (!leaves_in_subtree) steps.push_back({Step::LAYER, max_log, 1});
  int32_t top = (int32_t)max_log - (leaves_in_subtree ? 0 : 1);
  while (top >= 0) {
    if (top > WIDE_LAYER || ((lcs[top].n_cols + 15) / 16 >= HEAVY_CHUNKS && top >= 8)) {
      steps.push_back({Step::LAYER, (uint32_t)top, 1});
      top--;
      continue;
    }
    const uint32_t max_depth = (uint32_t)top <= TAIL_MAX_TOP ? (uint32_t)top + 1 : wide_w_log((uint32_t)top) + 1;
    uint32_t depth = 1;
    while (depth < max_depth && !((lcs[top - depth].n_cols + 15) / 16 >= HEAVY_CHUNKS && top - (int32_t)depth >= 8))
      depth++;
    steps.push_back({Step::SUBTREE, (uint32_t)top, depth});
    top -= (int32_t)depth;
  }
}

bool MerklePlan::enqueue(const std::vector<hipEvent_t>* layer_ready, const RootDraw* root_draw,
                         const LeafFold* leaf_fold) const {
  size_t next = 0;
  return enqueue_from(&next, 0, layer_ready, root_draw, leaf_fold);
}

bool MerklePlan::leaf_fold_ok() const {
  return !steps.empty() && steps[0].kind == Step::SUBTREE && steps[0].top == max_log && lcs[max_log].n_cols == 4;
}

bool MerklePlan::enqueue_from(size_t* next, uint32_t min_layer, const std::vector<hipEvent_t>* layer_ready,
                              const RootDraw* root_draw, const LeafFold* leaf_fold) const {
  require(!leaf_fold || (leaf_fold_ok() && *next == 0), "merkle: leaf fold on a plan that cannot take it");
  hipStream_t st = ctx().cur;
  bool drawn = false;
  for (; *next < steps.size() && lowest_layer(steps[*next]) >= min_layer; ++*next) {
    const Step& s = steps[*next];
    if (layer_ready) {  // the columns injected into the layers of this step must be written
      const uint32_t lo = lowest_layer(s);
      for (uint32_t k = lo; k <= s.top; k++)
        if ((*layer_ready)[k]) hip_check(hipStreamWaitEvent(st, (*layer_ready)[k], 0), "hipStreamWaitEvent");
    }
    if (s.kind == Step::LAYER) {
      const uint32_t k = s.top;
      const Hash* prev = k == max_log ? nullptr : outs[k + 1];
      timed(k == max_log ? "merkle_leaf" : "merkle_layer", [&] {
        if (k <= QUAD_LAYER_MAX)
        {
          const uint32_t nc = lcs[k].n_cols, blocks = blocks_for(4ull << k, 256);
          if (nc >= 32 && nc <= QUAD_STAGED_MAX_COLS)
            quad_layer_kernel<true><<<dim3(blocks, 1, n_jobs), 256, nc * sizeof(void*), st>>>(
                1u << k, nc, lcs[k].cols, prev, outs[k], jstride);
          else
            quad_layer_kernel<false><<<dim3(blocks, 1, n_jobs), 256, 0, st>>>(1u << k, nc, lcs[k].cols, prev,
                                                                              outs[k], jstride);
        }
        else
          layer_kernel<<<dim3(blocks_for(1ull << k, 256), 1, n_jobs), 256, 0, st>>>(1u << k, lcs[k].n_cols,
                                                                                     lcs[k].cols, prev, outs[k],
                                                                                     jstride);
      });
    } else {
      // narrow runs: 2^SUB_BLOCK_LOG nodes of layer top per block (the whole layer when it
      // is smaller), and runs deeper than one block's subtree finish in the last block;
      // wide runs: one subtree per block, at most 2^SUB_WIDE_BLOCKS_LOG blocks
      const bool narrow = s.top <= TAIL_MAX_TOP;
      const uint32_t w_log = narrow ? std::min<uint32_t>(SUB_BLOCK_LOG, s.top) : wide_w_log(s.top);
      const uint32_t blocks = 1u << (s.top - w_log);
      const bool tail = s.depth > w_log + 1;
      require(!tail || (narrow && blocks >= 2), "merkle: bad subtree step");
      const size_t lds = sizeof(Hash) * std::max<size_t>(1u << w_log, tail ? blocks / 2 : 0);
      // the launch that writes the root also runs the channel step
      RootDraw rd{nullptr, nullptr, 0};
      if (root_draw && s.top + 1 == s.depth) {
        rd = *root_draw;
        drawn = true;
      }
      const LeafFold lf = (leaf_fold && s.top == max_log) ? *leaf_fold : LeafFold{};
      timed("merkle_subtree", [&] {
        subtree_kernel<SUB_NT><<<dim3(blocks, 1, n_jobs), SUB_NT, lds, st>>>(
            s.top == max_log ? nullptr : outs[s.top + 1], s.top, s.depth, w_log, douts, dlcs, rd,
            tail ? darrive + (size_t)s.top * n_jobs : nullptr, jstride, lf);
      });
    }
  }
  return drawn;
}

// Whole tree.  cols/logs are host arrays; layers[k] (host array of device buffers) gets 2^k hashes.
void commit_tree(uint32_t n_cols, const uint32_t* const* cols, const uint32_t* logs, Hash* const* layers,
                 uint32_t max_log) {
  MerklePlan plan(n_cols, cols, logs, layers, max_log);
  plan.enqueue();
}

}  // namespace merkle
}  // namespace sh

using namespace sh;
using sh::merkle::Hash;

extern "C" void commit_on_first_layer(size_t size, size_t amount_of_columns, uint32_t* const* columns,
                                      StwoBlake2sHash* result) {
  guarded([&] {
    // column offsets are formed in 32 bits (gmem.hpp): injected columns hold at most 2^30 values
    require(size <= (amount_of_columns ? (1ull << 30) : 0xffffffffull), "commit_on_first_layer: size too large");
    if (!size) return;
    timed("merkle_leaf", [&] {
      merkle::layer_kernel<<<merkle::blocks_for(size, 256), 256, 0, ctx().cur>>>(
          (uint32_t)size, (uint32_t)amount_of_columns, (const uint32_t* const*)columns, nullptr, (Hash*)result, 0);
    });
    sync();
  });
}

extern "C" void commit_on_layer_with_previous(size_t size, size_t amount_of_columns, uint32_t* const* columns,
                                              const StwoBlake2sHash* previous_layer, StwoBlake2sHash* result) {
  guarded([&] {
    require(size <= (amount_of_columns ? (1ull << 30) : 0xffffffffull),
            "commit_on_layer_with_previous: size too large");
    if (!size) return;
    timed("merkle_node", [&] {
      merkle::layer_kernel<<<merkle::blocks_for(size, 256), 256, 0, ctx().cur>>>(
          (uint32_t)size, (uint32_t)amount_of_columns, (const uint32_t* const*)columns, (const Hash*)previous_layer,
          (Hash*)result, 0);
    });
    sync();
  });
}

extern "C" int stwo_hip_merkle_commit_async(uint32_t n_cols, const uint32_t* const* cols, const uint32_t* log_sizes,
                                            StwoBlake2sHash* const* layers, uint32_t max_log) {
  return guarded([&] { merkle::commit_tree(n_cols, cols, log_sizes, (Hash* const*)layers, max_log); });
}

extern "C" int stwo_hip_merkle_root(const StwoBlake2sHash* root_layer, StwoBlake2sHash* root_out) {
  return guarded([&] {
    hip_check(hipMemcpyAsync(root_out, root_layer, 32, hipMemcpyDefault, ctx().cur), "hipMemcpy");
    sync();
  });
}
