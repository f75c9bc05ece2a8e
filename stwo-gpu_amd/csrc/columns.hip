// Column operations: ColumnOps::bit_reverse_column (stwo_gpu_backend/src/column.rs:10-34),
// PolyOps::new_canonical_ordered (poly.rs:23-33) and AccumulationOps::accumulate
// (accumulation.rs:9-19).
//
// bit_reverse: the reference swaps a[i] <-> a[brev(i)] one element per thread
// (cuda/src/bit_reverse.cu:6-16), which scatters every access.  Here an index
// is split into [hi a bits | mid | lo a bits]; brev maps (hi, mid, lo) to
// (brev lo, brev mid, brev hi), so the 2^a x 2^a tile at `mid` and the tile at
// brev(mid) exchange places as a bit-reversed transpose through LDS, with
// 2^a-element contiguous runs on both the read and the write side.
#include "field.hpp"
#include "runtime.hpp"

namespace sh {
namespace cols {

// tile = 2^TA x 2^TA elements for the smaller columns: 64 x 64 for M31 (256-byte runs) below
// 2^14 values, 32 x 32 for QM31 (512-byte runs) below 2^12 values
template <class T>
constexpr int tile_bits() {
  return sizeof(T) == 4 ? 6 : 5;
}

constexpr uint32_t BR_NT = 256;
// 4-byte elements: 64 x 64 tiles, every load of a thread in flight before its LDS stores
template <class T>
__global__ __launch_bounds__(BR_NT) void bitrev_tiled_kernel(T* a, uint32_t log_n) {
  constexpr int TA = tile_bits<T>();
  // LDS rows padded by one element, so the transposed reads hit distinct banks
  constexpr uint32_t W = 1u << TA, RW = W + 1;
  constexpr int PER = (int)(W * W / BR_NT);     // elements per thread and tile
  __shared__ T ta[W * RW];
  __shared__ T tb[W * RW];
  const uint32_t mid_bits = log_n - 2 * TA;
  const uint32_t m = blockIdx.x;
  const uint32_t mr = brev(m, mid_bits);
  if (mr < m) return;  // the pair is handled by block mr
  const bool pair = mr != m;
  auto at = [&](uint32_t hi, uint32_t mid, uint32_t lo) -> size_t {
    return ((size_t)hi << (log_n - TA)) | ((size_t)mid << TA) | lo;
  };
  // every load of the thread is issued before the first LDS store
  T va[PER], vb[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + BR_NT * k, hi = e >> TA, lo = e & (W - 1);
    va[k] = a[at(hi, m, lo)];
    if (pair) vb[k] = a[at(hi, mr, lo)];
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + BR_NT * k, hi = e >> TA, lo = e & (W - 1);
    ta[hi * RW + lo] = va[k];
    if (pair) tb[hi * RW + lo] = vb[k];
  }
  __syncthreads();
  // element (hi, m, lo) goes to (brev lo, mr, brev hi)
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + BR_NT * k, hi2 = e >> TA, lo2 = e & (W - 1);
    const uint32_t src = brev(lo2, TA) * RW + brev(hi2, TA);
    a[at(hi2, mr, lo2)] = ta[src];
    if (pair) a[at(hi2, m, lo2)] = tb[src];
  }
}

// 16-byte elements: 32 x 32 tiles (512-byte runs), unpadded LDS; measured faster with the
// loads interleaved with the LDS stores than with the 4-byte kernel's load-all-first loop
template <class T>
__global__ __launch_bounds__(BR_NT) void bitrev_tiled16_kernel(T* a, uint32_t log_n) {
  constexpr int TA = 5;
  constexpr uint32_t W = 1u << TA;
  __shared__ T ta[W * W];
  __shared__ T tb[W * W];
  const uint32_t mid_bits = log_n - 2 * TA;
  const uint32_t m = blockIdx.x;
  const uint32_t mr = brev(m, mid_bits);
  if (mr < m) return;  // the pair is handled by block mr
  auto at = [&](uint32_t hi, uint32_t mid, uint32_t lo) -> size_t {
    return ((size_t)hi << (log_n - TA)) | ((size_t)mid << TA) | lo;
  };
  for (uint32_t e = threadIdx.x; e < W * W; e += blockDim.x) {
    uint32_t hi = e >> TA, lo = e & (W - 1);
    ta[e] = a[at(hi, m, lo)];
    if (mr != m) tb[e] = a[at(hi, mr, lo)];
  }
  __syncthreads();
  // element (hi, m, lo) goes to (brev lo, mr, brev hi)
  for (uint32_t e = threadIdx.x; e < W * W; e += blockDim.x) {
    uint32_t hi2 = e >> TA, lo2 = e & (W - 1);
    uint32_t src = (brev(lo2, TA) << TA) | brev(hi2, TA);
    a[at(hi2, mr, lo2)] = ta[src];
    if (mr != m) a[at(hi2, m, lo2)] = tb[src];
  }
}

// 4-byte elements in columns of at least 2^14: 128 x 128 tiles (512-byte runs; 2^28 values
// 0.70 -> 0.61 ms against 64 x 64 tiles, 0.58 ms with two elements per 8-byte access), both
// tiles of the pair in registers and one padded LDS tile for the two transposes
template <class T, int TA, int NT>
__global__ __launch_bounds__(NT) void bitrev_big_kernel(T* a, uint32_t log_n) {
  constexpr uint32_t W = 1u << TA, RW = W + 1;
  constexpr int PER = (int)(W * W / NT) / 2;  // element pairs per thread and tile
  __shared__ T t[W * RW];
  const uint32_t mid_bits = log_n - 2 * TA;
  const uint32_t m = blockIdx.x;
  const uint32_t mr = brev(m, mid_bits);
  if (mr < m) return;
  const bool pair = mr != m;
  // pair e of a tile: row e >> (TA - 1), elements lo, lo + 1 with lo = 2 (e mod W / 2); one
  // 8-byte access per pair, so a wave covers a whole 512-byte row
  auto at2 = [&](uint32_t hi, uint32_t mid, uint32_t lo) -> size_t {
    return (((size_t)hi << (log_n - TA)) | ((size_t)mid << TA) | lo) >> 1;
  };
  uint2* const a2 = reinterpret_cast<uint2*>(a);
  uint2 va[PER], vb[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + NT * k, hi = e >> (TA - 1), lo = 2 * (e & (W / 2 - 1));
    va[k] = a2[at2(hi, m, lo)];
    if (pair) vb[k] = a2[at2(hi, mr, lo)];
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + NT * k, hi = e >> (TA - 1), lo = 2 * (e & (W / 2 - 1));
    t[hi * RW + lo] = va[k].x;
    t[hi * RW + lo + 1] = va[k].y;
  }
  __syncthreads();
  // element (hi, m, lo) goes to (brev lo, mr, brev hi)
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + NT * k, hi2 = e >> (TA - 1), lo2 = 2 * (e & (W / 2 - 1));
    a2[at2(hi2, mr, lo2)] = make_uint2(t[brev(lo2, TA) * RW + brev(hi2, TA)], t[brev(lo2 + 1, TA) * RW + brev(hi2, TA)]);
  }
  if (!pair) return;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + NT * k, hi = e >> (TA - 1), lo = 2 * (e & (W / 2 - 1));
    t[hi * RW + lo] = vb[k].x;
    t[hi * RW + lo + 1] = vb[k].y;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + NT * k, hi2 = e >> (TA - 1), lo2 = 2 * (e & (W / 2 - 1));
    a2[at2(hi2, m, lo2)] = make_uint2(t[brev(lo2, TA) * RW + brev(hi2, TA)], t[brev(lo2 + 1, TA) * RW + brev(hi2, TA)]);
  }
}

// 16-byte elements in columns of at least 2^12: 64 x 64 tiles (1 KiB runs), both tiles of the
// pair in registers; LDS row r starts at element 64 r + r / 8, so the transposed reads of 8
// consecutive lanes (rows 8 apart) fall in distinct bank groups (with one element of padding
// per row they all hit one group: 3.1 ms for 2^28, against 2.5 ms for 32 x 32 tiles and 2.17
// ms here)
__global__ __launch_bounds__(1024) void bitrev_big16_kernel(uint4* a, uint32_t log_n) {
  constexpr uint32_t TA = 6, W = 1u << TA, NT = 1024;
  constexpr int PER = (int)(W * W / NT);
  __shared__ uint4 t[W * W + W / 8];
  const uint32_t mid_bits = log_n - 2 * TA;
  const uint32_t m = blockIdx.x;
  const uint32_t mr = brev(m, mid_bits);
  if (mr < m) return;
  const bool pair = mr != m;
  auto at = [&](uint32_t hi, uint32_t mid, uint32_t lo) -> size_t {
    return ((size_t)hi << (log_n - TA)) | ((size_t)mid << TA) | lo;
  };
  auto row = [](uint32_t r) { return r * W + (r >> 3); };
  // plain words in registers (arrays of the vector type would be kept in scratch memory)
  uint32_t va[PER][4], vb[PER][4];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + NT * k, hi = e >> TA, lo = e & (W - 1);
    const uint4 x = a[at(hi, m, lo)];
    va[k][0] = x.x; va[k][1] = x.y; va[k][2] = x.z; va[k][3] = x.w;
    if (pair) {
      const uint4 y = a[at(hi, mr, lo)];
      vb[k][0] = y.x; vb[k][1] = y.y; vb[k][2] = y.z; vb[k][3] = y.w;
    }
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + NT * k, hi = e >> TA, lo = e & (W - 1);
    t[row(hi) + lo] = make_uint4(va[k][0], va[k][1], va[k][2], va[k][3]);
  }
  __syncthreads();
  // element (hi, m, lo) goes to (brev lo, mr, brev hi)
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + NT * k, hi2 = e >> TA, lo2 = e & (W - 1);
    a[at(hi2, mr, lo2)] = t[row(brev(lo2, TA)) + brev(hi2, TA)];
  }
  if (!pair) return;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + NT * k, hi = e >> TA, lo = e & (W - 1);
    t[row(hi) + lo] = make_uint4(vb[k][0], vb[k][1], vb[k][2], vb[k][3]);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const uint32_t e = threadIdx.x + NT * k, hi2 = e >> TA, lo2 = e & (W - 1);
    a[at(hi2, m, lo2)] = t[row(brev(lo2, TA)) + brev(hi2, TA)];
  }
}

template <class T>
__global__ void bitrev_simple_kernel(T* a, uint32_t n, uint32_t log_n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t j = brev(i, log_n);
  if (j > i) {
    T t = a[i];
    a[i] = a[j];
    a[j] = t;
  }
}

template <class T>
void bit_reverse(T* a, uint64_t n) {
  require(is_pow2(n) && n <= (1ull << 31), "bit_reverse: size must be a power of two <= 2^31");
  uint32_t log_n = log2_exact(n);
  constexpr int TA = tile_bits<T>();
  if (log_n >= 2 * TA) {
    timed("bit_reverse", [&] {
      if constexpr (sizeof(T) == 4) {
        if (log_n >= 14 && ((uintptr_t)a & 7) == 0)  // 8-byte accesses
          bitrev_big_kernel<T, 7, 1024><<<1u << (log_n - 14), 1024, 0, ctx().cur>>>(a, log_n);
        else
          bitrev_tiled_kernel<T><<<1u << (log_n - 2 * TA), cols::BR_NT, 0, ctx().cur>>>(a, log_n);
      } else if (log_n >= 12)
        bitrev_big16_kernel<<<1u << (log_n - 12), 1024, 0, ctx().cur>>>(a, log_n);
      else
        bitrev_tiled16_kernel<T><<<1u << (log_n - 2 * TA), cols::BR_NT, 0, ctx().cur>>>(a, log_n);
    });
  } else if (n > 1) {
    timed("bit_reverse", [&] {
      bitrev_simple_kernel<T><<<ceil_div(n, 256), 256, 0, ctx().cur>>>(a, (uint32_t)n, log_n);
    });
  }
}

// canonic order -> circle-domain order: evens, then odds reversed (cuda/src/poly/twiddles.cu:4-15)
__global__ void sort_values_kernel(const uint32_t* from, uint32_t* dst, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t half = n >> 1;
  dst[i] = i < half ? from[2 * i] : from[n - 1 - 2 * (i - half)];
}

struct Soa4 {
  uint32_t* c[4];
};
__global__ void accumulate_kernel(uint32_t n, Soa4 l, Soa4 r) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 4; k++) l.c[k][i] = add(l.c[k][i], r.c[k][i]);
}

// ---- gathers for Column::at and Merkle decommitment ----
// values[i * n_cols + c] = cols[c][rows[i]]: every queried row of every column in one
// launch (the reference copies a whole column to the host per element, column.rs:49-52).
__global__ __launch_bounds__(256) void gather_rows_kernel(const uint32_t* const* __restrict__ cols, uint32_t n_cols,
                                                          const uint32_t* __restrict__ rows, uint64_t total,
                                                          uint32_t* __restrict__ out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint32_t i = (uint32_t)(t / n_cols), c = (uint32_t)(t % n_cols);
  out[t] = cols[c][rows[i]];
}
struct Hash8 {
  uint32_t s[8];
};
// paths[q * depth + j] = the sibling, at layer depth - j, of the node on leaf q's path
// (j = 0 is the leaf's own sibling): the authentication path of MerkleProver::decommit.
__global__ __launch_bounds__(256) void auth_paths_kernel(const Hash8* const* __restrict__ layers, uint32_t depth,
                                                         const uint32_t* __restrict__ leaves, uint32_t total,
                                                         Hash8* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint32_t q = t / depth, j = t % depth;
  out[t] = layers[depth - j][(leaves[q] >> j) ^ 1u];
}

}  // namespace cols
}  // namespace sh

using namespace sh;

extern "C" int stwo_hip_gather_rows(const uint32_t* const* columns, uint32_t n_cols, uint64_t column_size,
                                    const uint32_t* rows, uint32_t n_rows, uint32_t* values) {
  return guarded([&] {
    if (!n_cols || !n_rows) return;
    require(columns != nullptr && rows != nullptr && values != nullptr, "gather_rows: null argument");
    for (uint32_t i = 0; i < n_rows; i++) require(rows[i] < column_size, "gather_rows: row index out of range");
    const uint64_t total = (uint64_t)n_rows * n_cols;
    require(total <= (1ull << 31), "gather_rows: too many values");
    DevBuf table = upload(columns, n_cols * sizeof(void*));
    DevBuf drows = upload(rows, n_rows * sizeof(uint32_t));
    DevBuf out(total * sizeof(uint32_t));
    timed("gather_rows", [&] {
      cols::gather_rows_kernel<<<ceil_div(total, 256), 256, 0, ctx().cur>>>(table.as<const uint32_t*>(), n_cols,
                                                                            drows.as<uint32_t>(), total,
                                                                            out.as<uint32_t>());
    });
    hip_check(hipMemcpyAsync(values, out.p, total * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx().cur),
              "hipMemcpyAsync D2H");
    sync();
  });
}

extern "C" int stwo_hip_merkle_auth_paths(StwoBlake2sHash* const* layers, uint32_t max_log, const uint32_t* leaves,
                                          uint32_t n_leaves, StwoBlake2sHash* paths) {
  return guarded([&] {
    if (!n_leaves || !max_log) return;
    require(layers != nullptr && leaves != nullptr && paths != nullptr, "merkle_auth_paths: null argument");
    require(max_log <= 31, "merkle_auth_paths: at most 31 layers");
    for (uint32_t q = 0; q < n_leaves; q++)
      require(leaves[q] < (1u << max_log), "merkle_auth_paths: leaf index out of range");
    const uint64_t total = (uint64_t)n_leaves * max_log;
    require(total < (1ull << 31), "merkle_auth_paths: too many hashes");
    DevBuf table = upload(layers, (max_log + 1) * sizeof(void*));
    DevBuf dleaves = upload(leaves, n_leaves * sizeof(uint32_t));
    DevBuf out(total * sizeof(StwoBlake2sHash));
    timed("merkle_auth_paths", [&] {
      cols::auth_paths_kernel<<<ceil_div(total, 256), 256, 0, ctx().cur>>>(
          table.as<const cols::Hash8*>(), max_log, dleaves.as<uint32_t>(), (uint32_t)total, out.as<cols::Hash8>());
    });
    hip_check(hipMemcpyAsync(paths, out.p, total * sizeof(StwoBlake2sHash), hipMemcpyDeviceToHost, ctx().cur),
              "hipMemcpyAsync D2H");
    sync();
  });
}

extern "C" void bit_reverse_base_field(uint32_t* array, size_t size) {
  guarded([&] {
    cols::bit_reverse(array, size);
    sync();
  });
}
extern "C" void bit_reverse_secure_field(uint32_t* array, size_t size) {
  guarded([&] {
    cols::bit_reverse(reinterpret_cast<uint4*>(array), size);
    sync();
  });
}
extern "C" uint32_t* sort_values_and_permute_with_bit_reverse_order(const uint32_t* from, size_t size) {
  uint32_t* dst = nullptr;
  guarded([&] {
    require(is_pow2(size) && size <= (1ull << 31), "sort_values: size must be a power of two");
    dst = (uint32_t*)ctx().pool.alloc(size * 4);
    timed("sort_values", [&] {
      cols::sort_values_kernel<<<ceil_div(size, 256), 256, 0, ctx().cur>>>(from, dst, (uint32_t)size);
    });
    cols::bit_reverse(dst, size);
    sync();
  });
  return dst;
}
extern "C" void accumulate(uint32_t size, uint32_t* const* left_columns, uint32_t* const* right_columns) {
  guarded([&] {
    if (!size) return;
    cols::Soa4 l{{left_columns[0], left_columns[1], left_columns[2], left_columns[3]}};
    cols::Soa4 r{{right_columns[0], right_columns[1], right_columns[2], right_columns[3]}};
    timed("accumulate", [&] {
      cols::accumulate_kernel<<<ceil_div(size, 256), 256, 0, ctx().cur>>>(size, l, r);
    });
    sync();
  });
}
