// DEEP quotients: QuotientOps::accumulate_quotients (stwo_gpu_backend/src/quotient.rs:17-115).
//
// For row r of the (bit-reversed) evaluation domain, p = domain.at(brev(r)), and
// for each sample batch b (point P_b, samples (column j, value v_j)) in order
// (cuda/src/quotients.cu:111-178):
//   acc = acc * rc^|b| + (sum_j c_j col_j[r] - (a_j p.y + b_j)) * den_b(p)^-1
// with line coefficients a = conj(v) - v, c = conj(P.y) - P.y, b = v c - a P.y,
// each times alpha_j = rc^(j+1) (quotients.cu:49-88), and
//   den_b(p) = (Re P.x - p.x) Im P.y - (Re P.y - p.y) Im P.x   in CM31 (quotients.cu:91-109).
// The line coefficients use the cumulative offset of quotients.cu:147-168 (the
// reference's quotients.cu:74 offset is wrong for batches of different sizes).
//
// MI355X design: the per-batch sums of a_j and b_j and the factor c (shared by
// every sample of a batch, since c depends only on P) are folded on the host, so
// a row costs one QM31 x M31 multiply-accumulate per sample, done on 64-bit
// lazy sums (v_mad_u64_u32) and reduced once per batch.  Sample tables are
// uniform across the wavefront and are read with scalar loads.
#include <cstring>
#include <map>
#include <vector>

#include "field.hpp"
#include "gmem.hpp"
#include "runtime.hpp"

namespace sh {
namespace quot {

struct Sample {
  uint32_t col;
  uint32_t pad;
  qm31 alpha;  // rc^(j+1)
};
struct Batch {
  uint32_t first, count;
  qm31 c;      // conj(P.y) - P.y
  qm31 a_sum;  // sum_j alpha_j (conj(v_j) - v_j)
  qm31 b_sum;  // sum_j alpha_j (v_j c - (conj(v_j) - v_j) P.y)
  qm31 coeff;  // rc^count
  cm31 prx, pry, pix, piy;
};

__device__ __forceinline__ uint64_t fold64(uint64_t s) { return (s & P) + (s >> 31); }

// Points of CircleDomain(half coset (hi, hs) of log half_log) from two small tables
// instead of a 31-step power per row: lo[j] = G^(hi + hs j) for j < 2^k and
// up[j] = G^(hs j 2^k) for j < 2^(half_log - k), so position i < 2^half_log is
// lo[i mod 2^k] up[i >> k] and position i + 2^half_log its conjugate
// (CircleDomain::index_at, cuda/src/quotients.cu:17-29).
struct DomainTables {
  const pt* lo;
  const pt* up;
  uint32_t k, half_log;
};
constexpr uint32_t DT_LO_BITS = 11;
inline uint32_t dt_k(uint32_t half_log) { return half_log < DT_LO_BITS ? half_log : DT_LO_BITS; }
inline size_t dt_bytes(uint32_t half_log) {
  const uint32_t k = dt_k(half_log);
  return (sizeof(pt) * ((1ull << k) + (1ull << (half_log - k))) + 255) & ~(size_t)255;
}
__global__ void domain_tables_kernel(uint32_t hi, uint32_t hs, uint32_t half_log, uint32_t k, pt* lo, pt* up) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (1u << k)) lo[t] = point_at_index((uint32_t)((hi + (uint64_t)hs * t) & 0x7fffffffu));
  if (t < (1u << (half_log - k))) up[t] = point_at_index((uint32_t)(((uint64_t)hs * ((uint64_t)t << k)) & 0x7fffffffu));
}
__device__ __forceinline__ pt domain_point(const DomainTables& T, uint32_t i) {
  const uint32_t half = 1u << T.half_log;
  const bool conj = i >= half;
  const uint32_t j = conj ? i - half : i;
  pt q = pmul(T.lo[j & ((1u << T.k) - 1)], T.up[j >> T.k]);
  if (conj) q.y = neg(q.y);
  return q;
}

__global__ __launch_bounds__(256) void accumulate_kernel(DomainTables T, uint32_t log_size,
                                                         const uint32_t* const* __restrict__ cols,
                                                         const Sample* __restrict__ samples,
                                                         const Batch* __restrict__ batches, uint32_t n_batches,
                                                         uint32_t* __restrict__ o0, uint32_t* __restrict__ o1,
                                                         uint32_t* __restrict__ o2, uint32_t* __restrict__ o3) {
  uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= (1u << log_size)) return;
  pt p = domain_point(T, brev(row, log_size));
  qm31 acc = qzero();
  for (uint32_t b = 0; b < n_batches; b++) {
    const Batch B = batches[b];
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    uint32_t j = 0;
    for (; j + 2 <= B.count; j += 2) {
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const Sample S = samples[B.first + j + u];
        uint64_t v = at(as_global(cols[S.col]), row);
        s0 += S.alpha.x.a * v;
        s1 += S.alpha.x.b * v;
        s2 += S.alpha.y.a * v;
        s3 += S.alpha.y.b * v;
      }
      s0 = fold64(s0);
      s1 = fold64(s1);
      s2 = fold64(s2);
      s3 = fold64(s3);
    }
    if (j < B.count) {
      const Sample S = samples[B.first + j];
      uint64_t v = at(as_global(cols[S.col]), row);
      s0 += S.alpha.x.a * v;
      s1 += S.alpha.x.b * v;
      s2 += S.alpha.y.a * v;
      s3 += S.alpha.y.b * v;
    }
    qm31 sum = {{reduce63(s0), reduce63(s1)}, {reduce63(s2), reduce63(s3)}};
    qm31 num = qsub(qmul(B.c, sum), qadd(qscale(B.a_sum, p.y), B.b_sum));
    cm31 d = csub(cmul(cm31{sub(B.prx.a, p.x), B.prx.b}, B.piy), cmul(cm31{sub(B.pry.a, p.y), B.pry.b}, B.pix));
    acc = qadd(qmul(acc, B.coeff), qmul_cm(num, cinv(d)));
  }
  o0[row] = acc.x.a;
  o1[row] = acc.x.b;
  o2[row] = acc.y.a;
  o3[row] = acc.y.b;
}

// Column-major form for up to 4 batches: each sampled column value of a row is loaded
// once and multiplied into the sums of all NB batches, with alpha = 0 for a batch that
// does not sample the column.  Per column the host provides NB alphas (one uniform
// 16 * NB-byte record), so the loop has no data-dependent branches; loads run one group
// of QG columns ahead of the arithmetic.  A batch's sums are folded after every fourth
// product taken into them: a product of two values below P is at most (P - 1)^2 =
// 2^62 - 2^33 + 1, and a folded sum is below 2^34, so four products on top of it stay
// below 2^64.
template <int NB>
struct ColSlot {
  const uint32_t* col;  // the column (device pointer)
  uint32_t mask;        // bit b: batch b samples the column (alpha[b] may be nonzero)
  uint32_t pad;
  qm31 alpha[NB];
};

__device__ __forceinline__ cm31 batch_denominator(const Batch& B, pt p) {
  return csub(cmul(cm31{sub(B.prx.a, p.x), B.prx.b}, B.piy), cmul(cm31{sub(B.pry.a, p.y), B.pry.b}, B.pix));
}
__device__ __forceinline__ qm31 finish_batch(const Batch& B, const uint64_t (&s)[4], pt p, cm31 dinv, qm31 acc) {
  qm31 sum = {{reduce63(s[0]), reduce63(s[1])}, {reduce63(s[2]), reduce63(s[3])}};
  qm31 num = qsub(qmul(B.c, sum), qadd(qscale(B.a_sum, p.y), B.b_sum));
  return qadd(qmul(acc, B.coeff), qmul_cm(num, dinv));
}

constexpr int QG = 16;  // columns per load group (a multiple of the fold period 4; 8: 6 % slower at 200 x 2^21)

// js: byte stride between the jobs of a batched launch (blockIdx.z = job): the slot and batch
// tables and the outputs move by the job's offset (the slots already hold that job's columns)
template <int NB>
__global__ __launch_bounds__(256) void accumulate_cols_kernel(DomainTables T, uint32_t log_size,
                                                              const ColSlot<NB>* __restrict__ slots,
                                                              uint32_t n_slots, const Batch* __restrict__ batches,
                                                              uint32_t* __restrict__ o0, uint32_t* __restrict__ o1,
                                                              uint32_t* __restrict__ o2, uint32_t* __restrict__ o3,
                                                              uint64_t js) {
  uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= (1u << log_size)) return;
  const uint64_t jo = job_offset(js);
  slots = jptr(slots, jo);
  batches = jptr(batches, jo);
  o0 = jptr(o0, jo);
  o1 = jptr(o1, jo);
  o2 = jptr(o2, jo);
  o3 = jptr(o3, jo);
  uint64_t s[NB][4];
#pragma unroll
  for (int b = 0; b < NB; b++) s[b][0] = s[b][1] = s[b][2] = s[b][3] = 0;
  // n_slots is a multiple of QG (zero-alpha slots of column 0 pad it), and QG more such
  // slots follow, so the next group's loads are issued unconditionally: a conditional
  // prefetch makes the compiler wait for every load in flight before the arithmetic
  uint32_t v[QG], vn[QG];
  // products taken into each batch's sums since their last fold (uniform): a batch's sums
  // are folded after its fourth product, not after every fourth column, so batches that
  // sample few of the columns fold rarely
  uint32_t since[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) since[b] = 0;
#pragma unroll
  for (int i = 0; i < QG; i++) v[i] = at(as_global(slots[i].col), row);
  for (uint32_t h = 0; h < n_slots; h += QG) {
#pragma unroll
    for (int i = 0; i < QG; i++) vn[i] = at(as_global(slots[h + QG + i].col), row);
#pragma unroll
    for (int i = 0; i < QG; i++) {
      const ColSlot<NB>& S = slots[h + i];
#pragma unroll
      for (int b = 0; b < NB; b++) {
        if (NB > 1 && !(S.mask & (1u << b))) continue;  // uniform: the slot is the same for every lane
        s[b][0] += (uint64_t)S.alpha[b].x.a * v[i];
        s[b][1] += (uint64_t)S.alpha[b].x.b * v[i];
        s[b][2] += (uint64_t)S.alpha[b].y.a * v[i];
        s[b][3] += (uint64_t)S.alpha[b].y.b * v[i];
        if (++since[b] == 4) {
          since[b] = 0;
          s[b][0] = fold64(s[b][0]);
          s[b][1] = fold64(s[b][1]);
          s[b][2] = fold64(s[b][2]);
          s[b][3] = fold64(s[b][3]);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < QG; i++) v[i] = vn[i];
  }
  const pt p = domain_point(T, brev(row, log_size));
  // one CM31 inversion for the NB denominators (prefix products, Montgomery's trick)
  cm31 d[NB], pre[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    d[b] = batch_denominator(batches[b], p);
    pre[b] = b == 0 ? d[0] : cmul(pre[b - 1], d[b]);
  }
  cm31 inv_all = cinv(pre[NB - 1]);
  cm31 dinv[NB];
#pragma unroll
  for (int b = NB - 1; b > 0; b--) {
    dinv[b] = cmul(inv_all, pre[b - 1]);
    inv_all = cmul(inv_all, d[b]);
  }
  dinv[0] = inv_all;
  qm31 acc = qzero();
#pragma unroll
  for (int b = 0; b < NB; b++) acc = finish_batch(batches[b], s[b], p, dinv[b], acc);
  o0[row] = acc.x.a;
  o1[row] = acc.x.b;
  o2[row] = acc.y.a;
  o3[row] = acc.y.b;
}

// One batch (every column sampled once at one point) built on the device from a
// point, sample values and random coefficient that live in device memory: lane j
// takes samples j, j + 256, ...; the per-batch sums are reduced in LDS.
constexpr uint32_t SB_NT = 256;
__global__ __launch_bounds__(SB_NT) void single_batch_kernel(const uint32_t* const* __restrict__ cols,
                                                             const qm31* __restrict__ point,
                                                             const qm31* __restrict__ values, uint32_t count,
                                                             const qm31* __restrict__ rc_dev, ColSlot<1>* ss,
                                                             uint32_t n_slots, Batch* bs, uint64_t js) {
  __shared__ qm31 sa[SB_NT], sb[SB_NT];
  // batched jobs (blockIdx.z): the slots get the job's own column pointers
  const uint64_t jo = job_offset(js);
  point = jptr(point, jo);
  values = jptr(values, jo);
  rc_dev = jptr(rc_dev, jo);
  ss = jptr(ss, jo);
  bs = jptr(bs, jo);
  const qm31 px = point[0], py = point[1], rc = *rc_dev;
  const qm31 c = qsub(qm31{py.x, cneg(py.y)}, py);
  qm31 la = qzero(), lb = qzero();
  for (uint32_t j = threadIdx.x; j < count; j += SB_NT) {
    const qm31 alpha = qpow(rc, j + 1);
    const qm31 v = values[j];
    const qm31 a = qsub(qm31{v.x, cneg(v.y)}, v);
    const qm31 bb = qsub(qmul(v, c), qmul(a, py));
    la = qadd(la, qmul(alpha, a));
    lb = qadd(lb, qmul(alpha, bb));
    ss[j] = ColSlot<1>{jptr(cols[j], jo), 1, 0, {alpha}};
  }
  for (uint32_t j = count + threadIdx.x; j < n_slots; j += SB_NT) ss[j] = ColSlot<1>{jptr(cols[0], jo), 0, 0, {qzero()}};
  sa[threadIdx.x] = la;
  sb[threadIdx.x] = lb;
  __syncthreads();
  for (uint32_t h = SB_NT / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      sa[threadIdx.x] = qadd(sa[threadIdx.x], sa[threadIdx.x + h]);
      sb[threadIdx.x] = qadd(sb[threadIdx.x], sb[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) bs[0] = Batch{0, count, c, sa[0], sb[0], qpow(rc, count), px.x, py.x, px.y, py.y};
}

// the quotient of an empty sample set: all four coordinates zero
__global__ __launch_bounds__(256) void zero_kernel(uint32_t n, uint32_t* o0, uint32_t* o1, uint32_t* o2,
                                                   uint32_t* o3) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  o0[i] = 0;
  o1[i] = 0;
  o2[i] = 0;
  o3[i] = 0;
}

}  // namespace quot

static uint32_t padded_slots(uint32_t n_cols) { return (n_cols + quot::QG - 1) / quot::QG * quot::QG; }
// slots: the padded loop range plus QG trailing pad slots
static size_t slot_bytes(uint32_t n_slots) {
  return (sizeof(quot::ColSlot<1>) * (n_slots + quot::QG) + 255) & ~(size_t)255;
}
static constexpr size_t BATCH_BYTES = (sizeof(quot::Batch) + 255) & ~(size_t)255;
// scratch of accumulate_quotients_single_point_async: slots | batch (the domain tables
// come from the library's per-domain cache)
size_t quotient_single_point_scratch_bytes(uint32_t n_cols, uint32_t max_log) {
  (void)max_log;
  return slot_bytes(padded_slots(std::max<uint32_t>(n_cols, 1))) + BATCH_BYTES;
}

// Fill the domain tables of CircleDomain(half coset (hi, hs), log log_size) at `dev`
// (dt_bytes(log_size - 1) bytes) on the current stream.
static quot::DomainTables domain_tables_async(uint32_t hi, uint32_t hs, uint32_t log_size, void* dev) {
  const uint32_t hl = log_size - 1, k = quot::dt_k(hl);
  pt* lo = (pt*)dev;
  pt* up = lo + (1u << k);
  const uint32_t n = std::max(1u << k, 1u << (hl - k));
  timed("quotient_domain_tables", [&] {
    quot::domain_tables_kernel<<<ceil_div(n, 256), 256, 0, ctx().cur>>>(hi, hs, hl, k, lo, up);
  });
  return quot::DomainTables{lo, up, k, hl};
}

// The domain tables of CircleDomain(half coset (hi, hs), log log_size), computed on first
// use and kept in the library context (until stwo_hip_release_twiddles), so repeated
// accumulations over one domain launch no table kernel.
static quot::DomainTables cached_domain_tables(uint32_t hi, uint32_t hs, uint32_t log_size) {
  const uint32_t hl = log_size - 1, k = quot::dt_k(hl);
  const TwiddleKey key{hi, hs, log_size};
  auto& cache = ctx().domain_tables;
  auto it = cache.find(key);
  if (it == cache.end()) {
    void* dev = ctx().pool.alloc(quot::dt_bytes(hl));
    domain_tables_async(hi, hs, log_size, dev);
    // the cached tables are read by work on every stream: finish them before sharing
    hip_check(hipStreamSynchronize(ctx().cur), "hipStreamSynchronize");
    it = cache.emplace(key, dev).first;
  }
  pt* lo = (pt*)it->second;
  return quot::DomainTables{lo, lo + (1u << k), k, hl};
}

// accumulate_quotients for one sample point shared by columns 0 .. n_cols-1 of d_cols, with
// the point (x, y), the sample values and the random coefficient read from device memory.
// n_jobs / js: same-shape jobs of a batched launch, js bytes apart (columns, point, values,
// coefficient, scratch and outputs all move by the job's offset; the domain is shared).
void accumulate_quotients_single_point_async(uint32_t hi, uint32_t hs, uint32_t log_size,
                                             const uint32_t* const* d_cols, uint32_t n_cols, const qm31* d_point,
                                             const qm31* d_values, const qm31* d_rc, void* d_scratch,
                                             uint32_t* const out[4], uint32_t n_jobs, uint64_t js) {
  require(log_size >= 1 && log_size <= 30 && n_cols >= 1, "accumulate_quotients: bad sizes");
  const uint32_t n_slots = padded_slots(n_cols);
  quot::ColSlot<1>* ss = (quot::ColSlot<1>*)d_scratch;
  quot::Batch* bs = (quot::Batch*)((char*)d_scratch + slot_bytes(n_slots));
  const quot::DomainTables T = cached_domain_tables(hi, hs, log_size);
  timed("quotient_constants", [&] {
    quot::single_batch_kernel<<<dim3(1, 1, n_jobs), quot::SB_NT, 0, ctx().cur>>>(d_cols, d_point, d_values, n_cols,
                                                                                d_rc, ss, n_slots + quot::QG, bs, js);
  });
  timed("accumulate_quotients", [&] {
    quot::accumulate_cols_kernel<1><<<dim3(ceil_div(1ull << log_size, 256), 1, n_jobs), 256, 0, ctx().cur>>>(
        T, log_size, ss, n_slots, bs, out[0], out[1], out[2], out[3], js);
  });
}

namespace {
// column-major launch: per sampled column one slot with the alpha of every batch (0 when
// the batch does not sample it; a column sampled twice in one batch gets the sum)
template <int NB>
void launch_cols_nb(uint32_t hi, uint32_t hs, uint32_t log_size, const uint32_t* const* h_cols,
                    const std::vector<quot::Sample>& ss, const std::vector<quot::Batch>& bs, uint32_t* const out[4]) {
  std::map<uint32_t, quot::ColSlot<NB>> by_col;
  for (uint32_t b = 0; b < (uint32_t)NB; b++)
    for (uint32_t j = 0; j < bs[b].count; j++) {
      const quot::Sample& S = ss[bs[b].first + j];
      auto it = by_col.find(S.col);
      if (it == by_col.end()) {
        quot::ColSlot<NB> z{};
        z.col = h_cols[S.col];
        for (int k = 0; k < NB; k++) z.alpha[k] = qzero();
        it = by_col.emplace(S.col, z).first;
      }
      it->second.alpha[b] = qadd(it->second.alpha[b], S.alpha);
      it->second.mask |= 1u << b;
    }
  std::vector<quot::ColSlot<NB>> slots;
  for (auto& kv : by_col) slots.push_back(kv.second);
  quot::ColSlot<NB> pad{};
  pad.col = slots.empty() ? h_cols[0] : slots[0].col;
  for (int k = 0; k < NB; k++) pad.alpha[k] = qzero();
  while (slots.empty() || slots.size() % quot::QG) slots.push_back(pad);
  const uint32_t n_loop = (uint32_t)slots.size();
  for (int i = 0; i < quot::QG; i++) slots.push_back(pad);  // read by the last group's prefetch
  // one upload: slots, then the batches at a 256-byte boundary
  const size_t slots_bytes = (slots.size() * sizeof(quot::ColSlot<NB>) + 255) & ~(size_t)255;
  std::vector<unsigned char> host(slots_bytes + bs.size() * sizeof(quot::Batch));
  memcpy(host.data(), slots.data(), slots.size() * sizeof(quot::ColSlot<NB>));
  memcpy(host.data() + slots_bytes, bs.data(), bs.size() * sizeof(quot::Batch));
  DevBuf d = upload(host.data(), host.size());
  const quot::ColSlot<NB>* dsl = d.as<quot::ColSlot<NB>>();
  const quot::Batch* dbs = (const quot::Batch*)((const unsigned char*)d.p + slots_bytes);
  const quot::DomainTables T = cached_domain_tables(hi, hs, log_size);
  timed("accumulate_quotients", [&] {
    quot::accumulate_cols_kernel<NB><<<ceil_div(1ull << log_size, 256), 256, 0, ctx().cur>>>(
        T, log_size, dsl, n_loop, dbs, out[0], out[1], out[2], out[3], 0);
  });
}
void launch_cols(uint32_t hi, uint32_t hs, uint32_t log_size, const uint32_t* const* h_cols,
                 const std::vector<quot::Sample>& ss, const std::vector<quot::Batch>& bs, uint32_t n_batches,
                 uint32_t* const out[4]) {
  switch (n_batches) {
    case 1: launch_cols_nb<1>(hi, hs, log_size, h_cols, ss, bs, out); break;
    case 2: launch_cols_nb<2>(hi, hs, log_size, h_cols, ss, bs, out); break;
    case 3: launch_cols_nb<3>(hi, hs, log_size, h_cols, ss, bs, out); break;
    default: launch_cols_nb<4>(hi, hs, log_size, h_cols, ss, bs, out);
  }
}
}  // namespace

// points: n_batches x StwoCirclePointQM31 (host); indexes/values flattened in batch order (host).
// h_cols: the n_cols column pointers in host memory; d_cols: the same table in device memory,
// or null (uploaded here when the batch-major kernel needs it).
void accumulate_quotients_async(uint32_t hi, uint32_t hs, uint32_t domain_size, const uint32_t* const* d_cols,
                                const uint32_t* const* h_cols, uint32_t n_cols,
                                qm31 rc, uint32_t n_batches, const StwoCirclePointQM31* points,
                                const uint32_t* sizes, const uint32_t* col_idx, const StwoQM31* values,
                                uint32_t* const out[4]) {
  // column offsets are formed in 32 bits (gmem.hpp): at most 2^30 rows
  require(is_pow2(domain_size) && domain_size >= 2 && domain_size <= (1u << 30),
          "accumulate_quotients: domain size must be a power of two in 2 .. 2^30");
  uint32_t log_size = log2_exact(domain_size);
  uint64_t n_samples = 0;
  for (uint32_t b = 0; b < n_batches; b++) n_samples += sizes[b];
  if (n_samples == 0) {
    // every numerator is zero, so is the accumulation: no column is read
    timed("quotients_zero", [&] {
      quot::zero_kernel<<<ceil_div(domain_size, 256), 256, 0, ctx().cur>>>(domain_size, out[0], out[1], out[2],
                                                                             out[3]);
    });
    return;
  }
  require(n_cols >= 1 && h_cols != nullptr, "accumulate_quotients: samples name columns but none are given");
  std::vector<quot::Batch> bs(n_batches);
  std::vector<quot::Sample> ss;
  uint32_t off = 0;
  for (uint32_t b = 0; b < n_batches; b++) {
    qm31 px = from_w(qm31w{points[b].x.a, points[b].x.b, points[b].x.c, points[b].x.d});
    qm31 py = from_w(qm31w{points[b].y.a, points[b].y.b, points[b].y.c, points[b].y.d});
    qm31 c = qsub(qm31{py.x, cneg(py.y)}, py);
    quot::Batch& B = bs[b];
    B.first = (uint32_t)ss.size();
    B.count = sizes[b];
    B.c = c;
    B.a_sum = qzero();
    B.b_sum = qzero();
    B.coeff = qpow(rc, sizes[b]);
    B.prx = px.x;
    B.pix = px.y;
    B.pry = py.x;
    B.piy = py.y;
    qm31 alpha = qone();
    for (uint32_t j = 0; j < sizes[b]; j++) {
      const StwoQM31& w = values[off + j];
      qm31 v = {{w.a, w.b}, {w.c, w.d}};
      alpha = qmul(alpha, rc);
      qm31 a = qsub(qm31{v.x, cneg(v.y)}, v);
      qm31 bb = qsub(qmul(v, c), qmul(a, py));
      B.a_sum = qadd(B.a_sum, qmul(alpha, a));
      B.b_sum = qadd(B.b_sum, qmul(alpha, bb));
      ss.push_back(quot::Sample{col_idx[off + j], 0, alpha});
    }
    off += sizes[b];
  }
  if (n_batches >= 1 && n_batches <= 4) {
    launch_cols(hi, hs, log_size, h_cols, ss, bs, n_batches, out);
    return;
  }
  if (ss.empty()) ss.push_back(quot::Sample{0, 0, qzero()});
  if (bs.empty()) bs.push_back(quot::Batch{});
  // one upload: samples, batches and (when the caller has no device table) the columns
  const size_t ss_bytes = (ss.size() * sizeof(quot::Sample) + 255) & ~(size_t)255;
  const size_t bs_bytes = (bs.size() * sizeof(quot::Batch) + 255) & ~(size_t)255;
  const size_t cols_bytes = d_cols ? 0 : (size_t)n_cols * sizeof(void*);
  std::vector<unsigned char> host(ss_bytes + bs_bytes + cols_bytes);
  memcpy(host.data(), ss.data(), ss.size() * sizeof(quot::Sample));
  memcpy(host.data() + ss_bytes, bs.data(), bs.size() * sizeof(quot::Batch));
  if (cols_bytes) memcpy(host.data() + ss_bytes + bs_bytes, h_cols, cols_bytes);
  DevBuf d = upload(host.data(), host.size());
  const unsigned char* base = (const unsigned char*)d.p;
  if (!d_cols) d_cols = (const uint32_t* const*)(base + ss_bytes + bs_bytes);
  const quot::DomainTables T = cached_domain_tables(hi, hs, log_size);
  timed("accumulate_quotients", [&] {
    quot::accumulate_kernel<<<ceil_div(domain_size, 256), 256, 0, ctx().cur>>>(
        T, log_size, d_cols, (const quot::Sample*)base, (const quot::Batch*)(base + ss_bytes), n_batches, out[0],
        out[1], out[2], out[3]);
  });
}

}  // namespace sh

using namespace sh;

extern "C" int stwo_hip_accumulate_quotients_async(uint32_t half_coset_initial_index, uint32_t half_coset_step_size,
                                                   uint32_t domain_size, const uint32_t* const* columns,
                                                   uint32_t n_columns, StwoQM31 random_coeff, uint32_t n_batches,
                                                   const StwoCirclePointQM31* points, const uint32_t* batch_sizes,
                                                   const uint32_t* column_indexes, const StwoQM31* values,
                                                   uint32_t* const* out) {
  return guarded([&] {
    uint64_t total = 0;
    for (uint32_t b = 0; b < n_batches; b++) total += batch_sizes[b];
    for (uint64_t i = 0; i < total; i++)
      require(column_indexes[i] < n_columns, "accumulate_quotients: column index out of range");
    uint32_t* o[4] = {out[0], out[1], out[2], out[3]};
    accumulate_quotients_async(half_coset_initial_index, half_coset_step_size, domain_size, nullptr, columns,
                               n_columns, qm31{{random_coeff.a, random_coeff.b},
                                                                 {random_coeff.c, random_coeff.d}},
                               n_batches, points, batch_sizes, column_indexes, values, o);
  });
}

extern "C" void accumulate_quotients(uint32_t half_coset_initial_index, uint32_t half_coset_step_size,
                                     uint32_t domain_size, uint32_t* const* columns, size_t number_of_columns,
                                     StwoQM31 random_coeff, const uint32_t* sample_points,
                                     const uint32_t* sample_column_indexes, uint32_t sample_column_indexes_size,
                                     const StwoQM31* sample_column_values,
                                     const uint32_t* sample_column_and_values_sizes, uint32_t sample_size,
                                     uint32_t* result_column_0, uint32_t* result_column_1, uint32_t* result_column_2,
                                     uint32_t* result_column_3, uint32_t flattened_line_coeffs_size) {
  (void)flattened_line_coeffs_size;
  guarded([&] {
    uint64_t total = 0;
    for (uint32_t b = 0; b < sample_size; b++) total += sample_column_and_values_sizes[b];
    require(total == sample_column_indexes_size, "accumulate_quotients: batch sizes do not sum to the index count");
    for (uint64_t i = 0; i < total; i++)
      require(sample_column_indexes[i] < number_of_columns, "accumulate_quotients: column index out of range");
    uint32_t* out[4] = {result_column_0, result_column_1, result_column_2, result_column_3};
    // `columns` is a device table (copy_device_pointer_vec_from_host_to_device); keep a host copy
    std::vector<const uint32_t*> host_cols(std::max<size_t>(number_of_columns, 1));
    if (number_of_columns) {
      hip_check(hipMemcpyAsync(host_cols.data(), columns, number_of_columns * sizeof(void*), hipMemcpyDeviceToHost,
                               ctx().cur),
                "hipMemcpyAsync D2H");
      sync();
    }
    accumulate_quotients_async(half_coset_initial_index, half_coset_step_size, domain_size,
                               (const uint32_t* const*)columns, host_cols.data(), (uint32_t)number_of_columns,
                               qm31{{random_coeff.a, random_coeff.b},
                                                                     {random_coeff.c, random_coeff.d}},
                               sample_size, (const StwoCirclePointQM31*)sample_points, sample_column_and_values_sizes,
                               sample_column_indexes, sample_column_values, out);
    sync();
  });
}
