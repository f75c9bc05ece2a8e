// Wide-Fibonacci STARK prover built from the backend operations, in the shape of stwo's
// prove() for the WideFibComponent (stwo_gpu_backend/src/examples/wide_fibonacci.rs:90-141
// gives the composition step; trace LDE, commitments, OODS sampling, DEEP quotients and
// FRI are the PolyOps / MerkleOps / QuotientOps / FriOps calls of poly.rs, blake2s.rs,
// quotient.rs and fri.rs).  Statement: n_cols columns of 2^log_n rows with
// col[i] = col[i-2]^2 + col[i-1]^2.
//
// Protocol (every challenge comes from the channel of channel.hpp):
//  1. mix the statement (n_cols, log_n, n_queries, pow_bits, last_log, tag);
//  2. trace T_i on CanonicCoset(log_n); interpolate, evaluate on CanonicCoset(log_n + 1)
//     (blowup 2), commit the LDE columns (tree T), mix its root;
//  3. alpha <- draw; composition E(p) = sum_i (T_i^2 + T_{i+1}^2 - T_{i+2}) alpha^(C-3-i) / Z(p)
//     on the rows of CanonicCoset(log_n + 1), Z(p) = pi^(log_n - 1)(p.x); its four
//     coordinate columns are interpolated there and evaluated on CanonicCoset(log_n + 2),
//     committed (tree Q), root mixed;
//  4. OODS point z <- draw_point; samples T_i(z) and E_k(z) (from the coefficients), mixed;
//  5. beta <- draw, a_0 <- draw; F1 = quotients of the trace LDE at z (domain log_n + 1),
//     F2 = quotients of the composition LDE at z (domain log_n + 2), random coefficient beta;
//  6. FRI: line_1 = fold_circle(F2, a_0); for j = 1 .. n_line: commit line_j (tree L_j),
//     mix its root, a_j <- draw, line_{j+1} = fold_line(line_j, a_j), and for j = 1 also
//     fold_circle_into_line(line_2, F1, a_1); the last line (2^last_log values) is mixed
//     and sent in the clear;
//  7. proof of work (pow_bits), nonce mixed; n_queries positions P in [0, 2^(log_n+2));
//  8. decommitments per query: trace rows P>>1 and its sibling, composition rows P and
//     P^1, line j rows P>>j and its sibling, each with the Merkle siblings of every layer
//     above the leaves.
// Nothing returns to the host before the proof is complete: the channel runs in
// single-lane kernels on device-resident state, and one gather kernel assembles the
// proof in device memory before a single copy to pinned host memory.
#include <array>
#include <cstring>
#include <map>

#include "blake2s_quad.hpp"
#include "channel.hpp"
#include "fft.hpp"
#include "prover.hpp"

namespace sh {

// other modules
// (n_jobs, js): same-shape jobs of a batched launch whose buffers lie js bytes apart
void wide_fib_trace_async(uint32_t* const* d_cols, uint32_t n_cols, uint32_t log_n, const uint64_t* d_seed,
                          uint32_t n_jobs, uint64_t js);
void vanishing_inverse_async(uint32_t* d_out, uint32_t log_ext, uint32_t log_n, uint32_t n_jobs, uint64_t js);
void qm31_powers_async(const qm31* d_alpha, uint32_t n, uint32_t* d_powers, uint32_t n_jobs, uint64_t js);
void wide_fib_quotients_async(uint32_t* const q[4], const uint32_t* const* d_trace, const uint32_t* powers,
                              const uint32_t* dinv, uint32_t n_rows, uint32_t n_cols, uint32_t n_jobs, uint64_t js);
size_t eval_at_point_columns_scratch_bytes(uint32_t n_cols, uint32_t L);
void eval_at_point_columns_async(const uint32_t* const* d_cols, uint32_t n_cols, uint32_t L, const qm31* d_point,
                                 qm31* d_out, void* d_scratch, uint32_t n_jobs, uint64_t js);
size_t quotient_single_point_scratch_bytes(uint32_t n_cols, uint32_t max_log);
void accumulate_quotients_single_point_async(uint32_t hi, uint32_t hs, uint32_t log_size,
                                             const uint32_t* const* d_cols, uint32_t n_cols, const qm31* d_point,
                                             const qm31* d_values, const qm31* d_rc, void* d_scratch,
                                             uint32_t* const out[4], uint32_t n_jobs, uint64_t js);
void fold_line_async(const uint32_t* itw, size_t twiddle_offset, size_t n, uint32_t* const* src, qm31 alpha,
                     uint32_t* const* dst, const qm31* alpha_dev, uint32_t n_jobs, uint64_t js);
void fold_circle_into_line_async(const uint32_t* itw, size_t twiddle_offset, size_t n, uint32_t* const* src,
                                 qm31 alpha, uint32_t* const* dst, const qm31* alpha_dev, bool accumulate,
                                 uint32_t n_jobs, uint64_t js);
void grind_async(const uint32_t* d_digest, uint32_t pow_bits, uint64_t* d_nonce, uint32_t n_jobs, uint64_t js);

namespace prover {

constexpr uint32_t MAGIC = 0x50575453u;  // "STWP"
constexpr uint32_t VERSION = 1;
constexpr uint32_t STATEMENT_TAG = 0x57464942u;  // "BIFW"
constexpr uint32_t MAX_QUERIES = 64;
constexpr uint32_t MAX_ALPHAS = 32;

struct JobBlock {
  Channel ch;
  uint32_t header[8];
  uint64_t seed;
  uint64_t nonce;
  qm31 alpha;
  qm31 z[2];
  qm31 beta;
  qm31 fri_alpha[MAX_ALPHAS];
  uint32_t queries[MAX_QUERIES];
};

// the trace seeds of a batch, by job
struct Seeds {
  uint64_t s[MAX_BATCH];
};

struct GatherOp {
  const uint32_t* src;
  uint32_t dst;
  uint16_t words;
  int16_t q;  // query whose position indexes src, or -1 for a fixed source
  uint8_t shift, xr;
  uint16_t pad;
};

// ---- channel steps (one lane per job; blockIdx.z = job, js = byte stride between jobs) ----
__global__ void k_begin(JobBlock* jb, Seeds seeds, Params p, uint32_t n_line, uint64_t js) {
  jb = jptr(jb, job_offset(js));
  jb->seed = seeds.s[blockIdx.z];
  const uint32_t h[8] = {MAGIC, VERSION, p.n_cols, p.log_n, p.n_queries, p.pow_bits, p.last_log, n_line};
  for (int i = 0; i < 8; i++) jb->header[i] = h[i];
  jb->ch.reset();
  const uint32_t st[6] = {p.n_cols, p.log_n, p.n_queries, p.pow_bits, p.last_log, STATEMENT_TAG};
  jb->ch.mix(st, 6);
}
__global__ void k_mix_root_draw(JobBlock* jb, const merkle::Hash* root, qm31* dst, uint64_t js) {
  const uint64_t jo = job_offset(js);
  jb = jptr(jb, jo);
  jb->ch.mix8(jptr(root, jo)->s);
  *jptr(dst, jo) = jb->ch.draw_qm31();
}
__global__ void k_mix_root_draw_point(JobBlock* jb, const merkle::Hash* root, uint64_t js) {
  const uint64_t jo = job_offset(js);
  jb = jptr(jb, jo);
  jb->ch.mix8(jptr(root, jo)->s);
  jb->ch.draw_point(jb->z[0], jb->z[1]);
}
// Tree root (blake2s.hpp) of a QM31 vector held as AoS (soa == false: v[i] = 4 words) or as
// four coordinate columns, by one block: one lane quad per node (quad_compress), leaves
// from global memory into buf, levels alternating between buf2 and buf; returns the root's
// LDS slot.  qm: 16 LDS words per quad.
constexpr uint32_t TREE_NT = 256;
constexpr uint32_t TREE_NQ = TREE_NT / 4;
constexpr uint32_t TREE_MAX_LEAVES = 1040;  // 4 words per element: up to 4160 elements
struct Soa4c {
  const uint32_t* c[4];
};
__device__ const merkle::Hash* block_tree_root(const qm31* aos, Soa4c soa, bool use_soa, uint32_t n,
                                               merkle::Hash* buf, merkle::Hash* buf2, uint32_t (*qm)[16]) {
  const uint32_t nb = (4 * n + 15) / 16;
  const uint32_t qi = threadIdx.x >> 2, q = threadIdx.x & 3;
  uint32_t* mb = qm[qi];
  // leaf i holds elements 4i .. 4i+3; lane q stages element 4i + q (words 4q .. 4q+3)
#pragma unroll 1
  for (uint32_t i = qi; i < nb; i += TREE_NQ) {  // uniform over the quad
    const uint32_t e = 4 * i + q;
    uint32_t x[4] = {0, 0, 0, 0};
    if (e < n) {
      if (use_soa) {
        for (int k = 0; k < 4; k++) x[k] = soa.c[k][e];
      } else {
        const qm31 v = aos[e];
        x[0] = v.x.a;
        x[1] = v.x.b;
        x[2] = v.y.a;
        x[3] = v.y.b;
      }
    }
    __builtin_amdgcn_wave_barrier();  // the quad's reads of its previous message are done
    for (int k = 0; k < 4; k++) mb[4 * q + k] = x[k];
    __builtin_amdgcn_wave_barrier();
    uint32_t h_lo = 0, h_hi = 0;
    merkle::quad_compress(h_lo, h_hi, mb, q);
    buf[i].s[q] = h_lo;
    buf[i].s[4 + q] = h_hi;
  }
  __syncthreads();
  // each level reads one buffer and writes the other (an unpaired last node moves up as is)
  merkle::Hash* in = buf;
  merkle::Hash* out = buf2;
  uint32_t cnt = nb;
  while (cnt > 1) {
    const uint32_t half = (cnt + 1) / 2;
#pragma unroll 1
    for (uint32_t i = qi; i < half; i += TREE_NQ) {  // uniform over the quad
      uint32_t lo = 0, hi = 0;
      if (2 * i + 1 < cnt) {
        merkle::quad_compress(lo, hi, in[2 * i].s, q);  // children 2i, 2i+1: 16 words
      } else {
        lo = in[2 * i].s[q];
        hi = in[2 * i].s[4 + q];
      }
      out[i].s[q] = lo;
      out[i].s[4 + q] = hi;
    }
    __syncthreads();
    merkle::Hash* t = in;
    in = out;
    out = t;
    cnt = half;
  }
  return in;
}

// mix the tree root of the OODS samples, draw the quotient coefficient and the first FRI
// folding coefficient (Channel::mix8, two Channel::draw_qm31), on one quad
__global__ __launch_bounds__(TREE_NT) void k_mix_samples_draw(JobBlock* jb, const qm31* samples, uint32_t n,
                                                             uint64_t js) {
  __shared__ merkle::Hash buf[TREE_MAX_LEAVES], buf2[TREE_MAX_LEAVES / 2];
  __shared__ uint32_t qm[TREE_NQ][16];
  const uint64_t jo = job_offset(js);
  jb = jptr(jb, jo);
  samples = jptr(samples, jo);
  const merkle::Hash* root = block_tree_root(samples, Soa4c{}, false, n, buf, buf2, qm);
  if (threadIdx.x < 4) {
    const uint32_t q = threadIdx.x;
    uint32_t g_lo, g_hi, k = 0;
    merkle::quad_mix8(&jb->ch, root->s, qm[0], q, g_lo, g_hi);
    const qm31 beta = merkle::quad_draw_qm31(g_lo, g_hi, k, qm[0], q);
    const qm31 alpha0 = merkle::quad_draw_qm31(g_lo, g_hi, k, qm[0], q);
    jb->ch.digest[q] = g_lo;
    jb->ch.digest[4 + q] = g_hi;
    if (q == 0) {
      jb->ch.n_draws = k;
      jb->beta = beta;
      jb->fri_alpha[0] = alpha0;
    }
  }
}
// mix the tree root of the last FRI layer (Channel::mix8), on one quad
__global__ __launch_bounds__(TREE_NT) void k_mix_last(JobBlock* jb, Soa4c last, uint32_t n, uint64_t js) {
  __shared__ merkle::Hash buf[TREE_MAX_LEAVES], buf2[TREE_MAX_LEAVES / 2];
  __shared__ uint32_t qm[TREE_NQ][16];
  const uint64_t jo = job_offset(js);
  jb = jptr(jb, jo);
  for (int k = 0; k < 4; k++) last.c[k] = jptr(last.c[k], jo);
  const merkle::Hash* root = block_tree_root(nullptr, last, true, n, buf, buf2, qm);
  if (threadIdx.x < 4) {
    const uint32_t q = threadIdx.x;
    uint32_t g_lo, g_hi;
    merkle::quad_mix8(&jb->ch, root->s, qm[0], q, g_lo, g_hi);
    jb->ch.digest[q] = g_lo;
    jb->ch.digest[4 + q] = g_hi;
    if (q == 0) jb->ch.n_draws = 0;
  }
}
__global__ void k_mix_nonce_queries(JobBlock* jb, uint32_t n_queries, uint32_t log_size, uint64_t js) {
  jb = jptr(jb, job_offset(js));
  const uint32_t w[2] = {(uint32_t)jb->nonce, (uint32_t)(jb->nonce >> 32)};
  jb->ch.mix(w, 2);
  jb->ch.draw_queries(jb->queries, n_queries, log_size);
}

__global__ __launch_bounds__(256) void k_gather(const GatherOp* __restrict__ ops, uint32_t n_ops,
                                                const JobBlock* __restrict__ jb, uint32_t* __restrict__ proof,
                                                uint64_t js) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_ops) return;
  const uint64_t jo = job_offset(js);
  jb = jptr(jb, jo);
  proof = jptr(proof, jo);
  const GatherOp op = ops[i];
  const uint32_t idx = op.q < 0 ? 0u : ((jb->queries[op.q] >> op.shift) ^ op.xr);
  const uint32_t* src = jptr(op.src, jo) + (size_t)idx * op.words;
  for (uint32_t w = 0; w < op.words; w++) proof[op.dst + w] = src[w];
}

// ---- layout ----
void validate(const Params& p) {
  require(p.n_cols >= 3 && p.n_cols <= 4096, "wide fibonacci prover: n_cols must be in [3, 4096]");
  require(p.last_log >= 2 && p.last_log <= 10, "wide fibonacci prover: last_log must be in [2, 10]");
  require(p.log_n >= 4 && p.log_n >= p.last_log && p.log_n <= 26, "wide fibonacci prover: log_n out of range");
  require(p.n_queries >= 1 && p.n_queries <= MAX_QUERIES, "wide fibonacci prover: n_queries must be in [1, 64]");
  require(p.pow_bits <= 32, "wide fibonacci prover: pow_bits must be <= 32");
}

Layout::Layout(const Params& p) {
  const uint64_t C = p.n_cols;
  n_line = p.log_n + 1 - p.last_log;
  uint64_t off = 8;
  roots = off;
  off += 8ull * (2 + n_line);
  samples = off;
  off += 4 * (C + 4);
  last = off;
  off += 4ull << p.last_log;
  nonce = off;
  off += 2;
  queries = off;
  off += p.n_queries;
  per_query = off;
  uint64_t o = 0;
  q_trace_vals = o;
  o += 2 * C;
  q_trace_path = o;
  o += 8ull * p.log_n;
  q_comp_vals = o;
  o += 8;
  q_comp_path = o;
  o += 8ull * (p.log_n + 1);
  for (uint32_t j = 1; j <= n_line; j++) {
    const uint32_t m = p.log_n + 2 - j;
    q_line.push_back(o);
    o += 8 + 8ull * (m - 1);
  }
  query_words = o;
  total = per_query + p.n_queries * query_words;
}

namespace {

uint64_t align256(uint64_t b) { return (b + 255) & ~255ull; }

// Column stride (words) inside the arena.  Columns of one table are read together at the
// same row (Merkle leaves, composition, quotients); an exact power-of-two stride would
// send all of those reads to the same HBM channels, so each column is padded by
// 4 KiB + 256 B (keeping 16-byte alignment).
constexpr uint64_t COL_PAD = 1088;
uint64_t cs(uint64_t len) { return len + COL_PAD; }

// gather operations of one proof: fixed fields in chunks of 16 words, then per query
// one operation per value word and per path hash
uint64_t gather_ops_upper_bound(const Params& p, const Layout& L) {
  const uint64_t fixed = L.per_query / 16 + 8 + 4 * (1ull << p.last_log) + 3 + L.n_line + (p.n_queries + 15) / 16;
  return fixed + p.n_queries * L.query_words;
}

// arena layout of one job (byte offsets)
struct Arena {
  uint64_t coeffs, lde, tree_t, dinv, powers, comp, comp_lde, tree_q, f1, f2, eap, quot, samples, jb, proof, end;
  std::vector<uint64_t> line, tree_l;
  explicit Arena(const Params& p, const Layout& L) {
    const uint64_t C = p.n_cols, n = 1ull << p.log_n;
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
      uint64_t at = o;
      o += align256(bytes);
      return at;
    };
    coeffs = take(4 * C * cs(n));
    lde = take(4 * C * cs(2 * n));
    tree_t = take(32 * ((4 * n) - 1));  // layers 0 .. log_n + 1
    dinv = take(8 * n);
    powers = take(16 * (C - 2));
    comp = take(4 * 4 * cs(2 * n));
    comp_lde = take(4 * 4 * cs(4 * n));
    tree_q = take(32 * ((8 * n) - 1));
    f1 = take(4 * 4 * cs(2 * n));
    f2 = take(4 * 4 * cs(4 * n));
    for (uint32_t j = 1; j <= L.n_line + 1; j++) line.push_back(take(4 * 4 * cs(1ull << (p.log_n + 2 - j))));
    for (uint32_t j = 1; j <= L.n_line; j++) tree_l.push_back(take(32 * ((2ull << (p.log_n + 2 - j)) - 1)));
    eap = take(std::max(eval_at_point_columns_scratch_bytes(p.n_cols, p.log_n),
                        eval_at_point_columns_scratch_bytes(4, p.log_n + 1)));
    quot = take(quotient_single_point_scratch_bytes(std::max<uint32_t>(p.n_cols, 4), p.log_n + 2));
    samples = take(16 * (C + 4));
    jb = take(sizeof(JobBlock));
    proof = take(4 * L.total);
    end = o;
  }
};

// layers[k] of a tree whose layers 0 .. max_log are packed from `base`
std::vector<merkle::Hash*> tree_layers(char* base, uint32_t max_log) {
  std::vector<merkle::Hash*> v(max_log + 1);
  uint64_t o = 0;
  for (uint32_t k = 0; k <= max_log; k++) {
    v[k] = (merkle::Hash*)(base + o);
    o += 32ull << k;
  }
  return v;
}

// pinned host buffers for finished proofs, reused by capacity
std::multimap<size_t, void*> g_pinned_free;
std::map<void*, size_t> g_pinned_cap;
void* take_pinned(size_t bytes) {
  auto it = g_pinned_free.lower_bound(bytes);
  if (it != g_pinned_free.end() && it->first <= 2 * bytes + 4096) {
    void* p = it->second;
    g_pinned_free.erase(it);
    return p;
  }
  void* p = nullptr;
  hip_check(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc");
  g_pinned_cap[p] = bytes;
  return p;
}
void give_pinned(void* p) { g_pinned_free.emplace(g_pinned_cap.at(p), p); }

struct StreamScope {
  Context& x;
  hipStream_t saved;
  StreamScope(hipStream_t s) : x(ctx()), saved(ctx().cur) { x.cur = s; }
  ~StreamScope() { x.cur = saved; }
};

uint32_t half_initial(uint32_t log) { return 1u << (30 - log); }
uint32_t half_step(uint32_t log) { return (1u << (32 - log)) & 0x7fffffffu; }

}  // namespace

// byte stride between the arenas of a batch's jobs
uint64_t job_stride(const Params& p, const Layout& L) { return (Arena(p, L).end + 4095) & ~4095ull; }

uint64_t footprint_bytes(const Params& p, uint32_t n_jobs) {
  Layout L(p);
  return job_stride(p, L) * n_jobs + 64 * 1024 * (uint64_t)(3 + L.n_line) +  // plus plan tables
         (uint64_t)n_jobs * 16 * ((uint64_t)p.n_cols + 4);                      // batch pointer tables
}

Job::Job(const Params& p_, hipStream_t s, uint32_t n_jobs_)
    : p((validate(p_), p_)), L(p_), stream(s), n_jobs(n_jobs_) {
  require(n_jobs >= 1 && n_jobs <= MAX_BATCH, "wide fibonacci prover: batch of 1 .. 16 jobs");
  StreamScope scope(stream);
  const uint32_t C = p.n_cols, ln = p.log_n;
  const uint64_t n = 1ull << ln;

  // one twiddle tree (CanonicCoset(log_n + 2)) serves every domain of the proof
  {
    const uint32_t lt = ln + 2;
    int st = stwo_hip_twiddles(half_initial(lt), half_step(lt), lt - 1, &tw, &itw);
    if (st != STWO_HIP_OK) fail(st, "wide fibonacci prover: twiddles");
    tree_len = 1ull << (lt - 1);
  }

  Arena A(p, L);
  jstride = n_jobs > 1 ? job_stride(p, L) : 0;
  arena = DevBuf(n_jobs > 1 ? jstride * n_jobs : A.end);
  char* base = arena.as<char>();
  coeffs = (uint32_t*)(base + A.coeffs);
  lde = (uint32_t*)(base + A.lde);
  dinv = (uint32_t*)(base + A.dinv);
  powers = (uint32_t*)(base + A.powers);
  comp = (uint32_t*)(base + A.comp);
  comp_lde = (uint32_t*)(base + A.comp_lde);
  f1 = (uint32_t*)(base + A.f1);
  f2 = (uint32_t*)(base + A.f2);
  for (uint64_t o : A.line) line.push_back((uint32_t*)(base + o));
  eap_scratch = base + A.eap;
  quot_scratch = base + A.quot;
  jb = (JobBlock*)(base + A.jb);
  proof = (uint32_t*)(base + A.proof);
  samples = (qm31*)(base + A.samples);

  // device pointer tables: coeffs | lde | comp | comp_lde of job 0, then (a batch of several
  // jobs) the same four tables over every job, for the LDE launches
  std::vector<const uint32_t*> t;
  for (uint32_t i = 0; i < C; i++) t.push_back(coeffs + i * cs(n));
  for (uint32_t i = 0; i < C; i++) t.push_back(lde + i * cs(2 * n));
  for (uint32_t k = 0; k < 4; k++) t.push_back(comp + k * cs(2 * n));
  for (uint32_t k = 0; k < 4; k++) t.push_back(comp_lde + k * cs(4 * n));
  const size_t n_job0 = t.size();
  if (n_jobs > 1) {
    auto all = [&](size_t first, size_t count) {
      for (uint32_t j = 0; j < n_jobs; j++)
        for (size_t i = 0; i < count; i++) t.push_back((const uint32_t*)((const char*)t[first + i] + j * jstride));
    };
    all(0, C);
    all(C, C);
    all(2 * C, 4);
    all(2 * C + 4, 4);
  }
  // one device region for every table of the job: pointer tables, the Merkle plans'
  // tables and the gather table, filled as one host image and uploaded once
  const uint32_t NL = L.n_line;
  std::vector<uint64_t> o_tree(3 + NL);
  uint64_t tb_bytes = align256(t.size() * sizeof(void*));
  o_tree[0] = tb_bytes;
  tb_bytes += align256(merkle::MerklePlan::table_bytes(C, ln + 1, n_jobs));
  o_tree[1] = tb_bytes;
  tb_bytes += align256(merkle::MerklePlan::table_bytes(4, ln + 2, n_jobs));
  for (uint32_t j = 1; j <= NL; j++) {
    o_tree[1 + j] = tb_bytes;
    tb_bytes += align256(merkle::MerklePlan::table_bytes(4, ln + 2 - j, n_jobs));
  }
  const uint64_t o_ops = tb_bytes;
  tb_bytes += align256(sizeof(GatherOp) * gather_ops_upper_bound(p, L));
  tables_ = DevBuf(tb_bytes);
  char* dtab = tables_.as<char>();
  std::vector<char> image(tb_bytes);
  memcpy(image.data(), t.data(), t.size() * sizeof(void*));
  const uint32_t* const* tb = (const uint32_t* const*)dtab;
  t_coeffs = tb;
  t_lde = tb + C;
  t_comp = tb + 2 * C;
  t_comp_lde = tb + 2 * C + 4;
  if (n_jobs > 1) {
    const uint32_t* const* bt = tb + n_job0;
    b_coeffs = bt;
    b_lde = bt + (size_t)n_jobs * C;
    b_comp = bt + (size_t)2 * n_jobs * C;
    b_comp_lde = bt + (size_t)2 * n_jobs * C + 4 * n_jobs;
  } else {
    b_coeffs = t_coeffs;
    b_lde = t_lde;
    b_comp = t_comp;
    b_comp_lde = t_comp_lde;
  }

  // Merkle plans
  layers_t = tree_layers(base + A.tree_t, ln + 1);
  layers_q = tree_layers(base + A.tree_q, ln + 2);
  {
    std::vector<uint32_t> logs(C, ln + 1);
    tree_t = std::make_unique<merkle::MerklePlan>(C, t.data() + C, logs.data(), layers_t.data(), ln + 1,
                                                  dtab + o_tree[0], image.data() + o_tree[0], n_jobs, jstride);
    std::vector<uint32_t> l4(4, ln + 2);
    tree_q = std::make_unique<merkle::MerklePlan>(4, t.data() + 2 * C + 4, l4.data(), layers_q.data(), ln + 2,
                                                  dtab + o_tree[1], image.data() + o_tree[1], n_jobs, jstride);
  }
  for (uint32_t j = 1; j <= NL; j++) {
    const uint32_t m = ln + 2 - j;
    layers_l.push_back(tree_layers(base + A.tree_l[j - 1], m));
    const uint32_t* cols[4];
    for (int k = 0; k < 4; k++) cols[k] = line[j - 1] + k * cs(1ull << m);
    const uint32_t l4[4] = {m, m, m, m};
    tree_l.push_back(std::make_unique<merkle::MerklePlan>(4, cols, l4, layers_l.back().data(), m,
                                                          dtab + o_tree[1 + j], image.data() + o_tree[1 + j],
                                                          n_jobs, jstride));
  }

  // proof assembly table
  std::vector<GatherOp> g;
  auto fixed = [&](const void* src, uint64_t dst, uint32_t words) {
    for (uint32_t w = 0; w < words; w += 16)
      g.push_back(GatherOp{(const uint32_t*)src + w, (uint32_t)(dst + w), (uint16_t)std::min<uint32_t>(16, words - w),
                           -1, 0, 0, 0});
  };
  auto indexed = [&](const void* src, uint64_t dst, uint32_t words, uint32_t q, uint32_t shift, uint32_t xr) {
    g.push_back(GatherOp{(const uint32_t*)src, (uint32_t)dst, (uint16_t)words, (int16_t)q, (uint8_t)shift,
                         (uint8_t)xr, 0});
  };
  // siblings of layers leaf_log - 1 .. 1 for a leaf position (query position >> shift0)
  auto path = [&](const std::vector<merkle::Hash*>& layers, uint32_t leaf_log, uint64_t dst, uint32_t q,
                  uint32_t shift0) {
    for (uint32_t k = leaf_log - 1, i = 0; k >= 1; k--, i++) indexed(layers[k], dst + 8 * i, 8, q, shift0 + leaf_log - k, 1);
  };
  fixed(jb->header, 0, 8);
  fixed(layers_t[0], L.roots, 8);
  fixed(layers_q[0], L.roots + 8, 8);
  for (uint32_t j = 1; j <= L.n_line; j++) fixed(layers_l[j - 1][0], L.roots + 8 * (1 + j), 8);
  fixed(base + A.samples, L.samples, 4 * (C + 4));
  {
    const uint32_t* last = line[L.n_line];
    const uint32_t nl = 1u << p.last_log;
    for (uint32_t v = 0; v < nl; v++)
      for (uint32_t k = 0; k < 4; k++) fixed(last + k * cs(nl) + v, L.last + 4 * v + k, 1);
  }
  fixed(&jb->nonce, L.nonce, 2);
  fixed(jb->queries, L.queries, p.n_queries);
  for (uint32_t q = 0; q < p.n_queries; q++) {
    const uint64_t qb = L.per_query + q * L.query_words;
    for (uint32_t s = 0; s < 2; s++)
      for (uint32_t i = 0; i < C; i++) indexed(lde + i * cs(2 * n), qb + L.q_trace_vals + s * C + i, 1, q, 1, s);
    path(layers_t, ln + 1, qb + L.q_trace_path, q, 1);
    for (uint32_t s = 0; s < 2; s++)
      for (uint32_t k = 0; k < 4; k++) indexed(comp_lde + k * cs(4 * n), qb + L.q_comp_vals + 4 * s + k, 1, q, 0, s);
    path(layers_q, ln + 2, qb + L.q_comp_path, q, 0);
    for (uint32_t j = 1; j <= L.n_line; j++) {
      const uint32_t m = ln + 2 - j;
      for (uint32_t s = 0; s < 2; s++)
        for (uint32_t k = 0; k < 4; k++)
          indexed(line[j - 1] + k * cs(1ull << m), qb + L.q_line[j - 1] + 4 * s + k, 1, q, j, s);
      path(layers_l[j - 1], m, qb + L.q_line[j - 1] + 8, q, j);
    }
  }
  n_ops = (uint32_t)g.size();
  require(n_ops <= gather_ops_upper_bound(p, L), "wide fibonacci prover: gather table overflow");
  memcpy(image.data() + o_ops, g.data(), g.size() * sizeof(GatherOp));
  ops_ = (const GatherOp*)(dtab + o_ops);
  upload_to(tables_.p, image.data(), tb_bytes);

  host_proof = take_pinned(4 * L.total * n_jobs);
  hip_check(hipEventCreateWithFlags(&done, hipEventDisableTiming), "hipEventCreate");
}

Job::~Job() {
  // the caller has waited for `done`
  if (done) (void)hipEventDestroy(done);
  if (host_proof) give_pinned(host_proof);
  StreamScope scope(stream);
  tree_t.reset();
  tree_q.reset();
  tree_l.clear();
  tables_ = DevBuf();
  arena = DevBuf();
}

void Job::enqueue(const uint64_t* seeds) {
  StreamScope scope(stream);
  const uint32_t C = p.n_cols, ln = p.log_n, K = n_jobs;
  const uint64_t n = 1ull << ln, js = jstride;
  const dim3 one(1, 1, K);  // one block per job

  Seeds sd{};
  for (uint32_t k = 0; k < K; k++) sd.s[k] = seeds[k];
  timed("channel_begin", [&] { k_begin<<<one, 1, 0, stream>>>(jb, sd, p, L.n_line, js); });
  wide_fib_trace_async((uint32_t* const*)t_coeffs, C, ln, &jb->seed, K, js);
  fft::lde_same_size(b_coeffs, (uint32_t* const*)b_coeffs, (uint32_t* const*)b_lde, K * C, ln, 1, tw, itw, tree_len,
                     fft::lde_variant_for(ln, 1, (uint64_t)K * C));
  {
    const merkle::RootDraw rd{&jb->ch, &jb->alpha, 0};
    if (!tree_t->enqueue(nullptr, &rd))
      timed("channel_trace_root",
            [&] { k_mix_root_draw<<<one, 1, 0, stream>>>(jb, layers_t[0], &jb->alpha, js); });
  }

  qm31_powers_async(&jb->alpha, C - 2, powers, K, js);
  vanishing_inverse_async(dinv, ln + 1, ln, K, js);
  uint32_t* q4[4] = {comp, comp + cs(2 * n), comp + 2 * cs(2 * n), comp + 3 * cs(2 * n)};
  wide_fib_quotients_async(q4, t_lde, powers, dinv, (uint32_t)(2 * n), C, K, js);
  fft::lde_same_size(b_comp, (uint32_t* const*)b_comp, (uint32_t* const*)b_comp_lde, 4 * K, ln + 1, 1, tw, itw,
                     tree_len, fft::lde_variant_for(ln + 1, 1, 4ull * K));
  {
    const merkle::RootDraw rd{&jb->ch, jb->z, 1};
    if (!tree_q->enqueue(nullptr, &rd))
      timed("channel_oods_point", [&] { k_mix_root_draw_point<<<one, 1, 0, stream>>>(jb, layers_q[0], js); });
  }

  eval_at_point_columns_async(t_coeffs, C, ln, jb->z, samples, eap_scratch, K, js);
  eval_at_point_columns_async(t_comp, 4, ln + 1, jb->z, samples + C, eap_scratch, K, js);
  timed("channel_samples", [&] { k_mix_samples_draw<<<one, TREE_NT, 0, stream>>>(jb, samples, C + 4, js); });

  uint32_t* f1c[4] = {f1, f1 + cs(2 * n), f1 + 2 * cs(2 * n), f1 + 3 * cs(2 * n)};
  uint32_t* f2c[4] = {f2, f2 + cs(4 * n), f2 + 2 * cs(4 * n), f2 + 3 * cs(4 * n)};
  accumulate_quotients_single_point_async(half_initial(ln + 1), half_step(ln + 1), ln + 1, t_lde, C, jb->z, samples,
                                          &jb->beta, quot_scratch, f1c, K, js);
  accumulate_quotients_single_point_async(half_initial(ln + 2), half_step(ln + 2), ln + 2, t_comp_lde, 4, jb->z,
                                          samples + C, &jb->beta, quot_scratch, f2c, K, js);

  auto coords = [&](uint32_t* c, uint32_t m) {
    std::array<uint32_t*, 4> r;
    for (int k = 0; k < 4; k++) r[k] = c + k * cs(1ull << m);
    return r;
  };
  auto l1 = coords(line[0], ln + 1);
  fold_circle_into_line_async(itw, tree_len - (1ull << (ln + 1)), 4 * n, f2c, qzero(), l1.data(), &jb->fri_alpha[0],
                              false, K, js);
  // line[k] = fold(line[k - 1], alpha_k) for k >= 2 is computed by the launch that hashes
  // line[k]'s leaves when its tree is one subtree launch (LeafFold), saving a fold launch
  // per layer; line[1] also accumulates the circle fold of f1, and line[n_line] has no tree
  auto fused = [&](uint32_t k) { return k >= 2 && k < L.n_line && tree_l[k]->leaf_fold_ok(); };
  for (uint32_t j = 1; j <= L.n_line; j++) {
    const uint32_t m = ln + 2 - j;
    const merkle::RootDraw rd{&jb->ch, &jb->fri_alpha[j], 0};
    merkle::LeafFold lf{};
    if (fused(j - 1)) {
      auto prev_src = coords(line[j - 2], m + 1);
      for (int k = 0; k < 4; k++) lf.src[k] = prev_src[k];
      lf.tw = itw + (tree_len - (1ull << (m + 1)));
      lf.alpha = &jb->fri_alpha[j - 1];
    }
    if (!tree_l[j - 1]->enqueue(nullptr, &rd, fused(j - 1) ? &lf : nullptr))
      timed("channel_fri_root",
            [&] { k_mix_root_draw<<<one, 1, 0, stream>>>(jb, layers_l[j - 1][0], &jb->fri_alpha[j], js); });
    auto src = coords(line[j - 1], m), dst = coords(line[j], m - 1);
    if (!fused(j))
      fold_line_async(itw, tree_len - (1ull << m), 1ull << m, src.data(), qzero(), dst.data(), &jb->fri_alpha[j],
                      K, js);
    if (j == 1)
      fold_circle_into_line_async(itw, tree_len - (1ull << ln), 2 * n, f1c, qzero(), dst.data(), &jb->fri_alpha[1],
                                  true, K, js);
  }
  {
    const uint32_t nl = 1u << p.last_log;
    const uint32_t* last = line[L.n_line];
    Soa4c s{{last, last + cs(nl), last + 2 * cs(nl), last + 3 * cs(nl)}};
    timed("channel_last_layer", [&] { k_mix_last<<<one, TREE_NT, 0, stream>>>(jb, s, nl, js); });
  }
  grind_async(jb->ch.digest, p.pow_bits, &jb->nonce, K, js);
  timed("channel_queries", [&] { k_mix_nonce_queries<<<one, 1, 0, stream>>>(jb, p.n_queries, ln + 2, js); });
  timed("proof_gather", [&] {
    k_gather<<<dim3(ceil_div(n_ops, 256), 1, K), 256, 0, stream>>>(ops_, n_ops, jb, proof, js);
  });
  // every job's proof words, one row per job
  hip_check(hipMemcpy2DAsync(host_proof, 4 * L.total, proof, K > 1 ? js : 4 * L.total, 4 * L.total, K,
                             hipMemcpyDeviceToHost, stream),
            "hipMemcpy2DAsync D2H");
  hip_check(hipEventRecord(done, stream), "hipEventRecord");
}

}  // namespace prover
}  // namespace sh
