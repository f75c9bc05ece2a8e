// Wide-Fibonacci prover on the device: one proof is one stream's worth of kernels with
// no host round trip until the finished proof is copied back.
#pragma once
#include <stdint.h>

#include <memory>
#include <vector>

#include "field.hpp"
#include "merkle_plan.hpp"
#include "runtime.hpp"

namespace sh {
namespace prover {

struct Params {
  uint32_t n_cols, log_n, n_queries, pow_bits, last_log;
  bool operator==(const Params& o) const {
    return n_cols == o.n_cols && log_n == o.log_n && n_queries == o.n_queries && pow_bits == o.pow_bits &&
           last_log == o.last_log;
  }
};
// same-shape jobs proved together (every phase one launch for all of them)
constexpr uint32_t MAX_BATCH = 16;
void validate(const Params& p);

// Word offsets of the proof (the layout the verifier parses).
struct Layout {
  uint32_t n_line;  // committed FRI line layers
  uint64_t roots, samples, last, nonce, queries, per_query, query_words, total;
  // inside one query block
  uint64_t q_trace_vals, q_trace_path, q_comp_vals, q_comp_path;
  std::vector<uint64_t> q_line;  // per line layer: values (8 words) then path
  explicit Layout(const Params& p);
};

uint64_t footprint_bytes(const Params& p, uint32_t n_jobs = 1);  // device bytes of a batch of n_jobs

// A batch of n_jobs same-shape proofs (n_jobs = 1: one proof).  Job k's buffers sit
// k * jstride bytes after job 0's in one arena; the pointers below are job 0's, and every
// launch covers all jobs of the batch.
struct Job {
  Params p;
  Layout L;
  hipStream_t stream;
  uint32_t n_jobs;
  uint64_t jstride;
  DevBuf arena;  // every per-job device buffer, n_jobs arenas of jstride bytes
  std::unique_ptr<merkle::MerklePlan> tree_t, tree_q;
  std::vector<std::unique_ptr<merkle::MerklePlan>> tree_l;
  DevBuf tables_;  // pointer tables, Merkle plan tables and the gather table
  const struct GatherOp* ops_ = nullptr;
  uint32_t n_ops = 0;
  void* host_proof = nullptr;  // pinned
  hipEvent_t done = nullptr;

  // device pointers into the arena
  uint32_t* coeffs;  // n_cols x n (trace, then its coefficients)
  uint32_t* lde;     // n_cols x 2n
  uint32_t* comp;    // 4 x 2n (composition on the extended domain, then its coefficients)
  uint32_t* comp_lde;  // 4 x 4n
  uint32_t* f1;        // 4 x 2n
  uint32_t* f2;        // 4 x 4n
  std::vector<uint32_t*> line;  // line layers 1 .. n_line + 1, 4 x 2^m each
  uint32_t *dinv, *powers, *proof;
  qm31* samples;  // n_cols trace samples, then 4 composition samples
  void *eap_scratch, *quot_scratch;
  struct JobBlock* jb;
  const uint32_t* const* t_coeffs;  // device pointer tables (job 0)
  const uint32_t* const* t_lde;
  const uint32_t* const* t_comp;
  const uint32_t* const* t_comp_lde;
  const uint32_t* const* b_coeffs;  // the same over every job of the batch (the LDE launches)
  const uint32_t* const* b_lde;
  const uint32_t* const* b_comp;
  const uint32_t* const* b_comp_lde;
  std::vector<merkle::Hash*> layers_t, layers_q;
  std::vector<std::vector<merkle::Hash*>> layers_l;
  const uint32_t *tw, *itw;
  uint64_t tree_len;

  // Allocates everything and builds the plans; work is enqueued on `stream`.
  Job(const Params& p, hipStream_t stream, uint32_t n_jobs = 1);
  ~Job();
  // the whole proof of every job (job k from seeds[k]), then the copy to host_proof and `done`
  void enqueue(const uint64_t* seeds);
  void enqueue(uint64_t seed) { enqueue(&seed); }
  const uint32_t* result(uint32_t k = 0) const { return (const uint32_t*)host_proof + k * L.total; }
};

}  // namespace prover
}  // namespace sh
