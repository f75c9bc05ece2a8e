// Merkle commitment plan shared by the whole-tree entry point and batched commits:
// device tables of the columns injected into each layer, built once, and the launch
// schedule (leaf layer, per-layer launches for injection-heavy layers, subtree
// launches covering runs of light layers).
#pragma once
#include <stdint.h>

#include <vector>

#include "channel.hpp"
#include "runtime.hpp"

namespace sh {
namespace merkle {

struct Hash {
  uint32_t s[8];
};

// After the root is written: mix it into a Fiat-Shamir channel (device memory) and draw a
// QM31 into *dst, or a circle point into dst[0], dst[1] (`point`).  Run by the launch that
// produces the root, it saves a single-lane kernel launch after every commitment.
struct RootDraw {
  Channel* ch;
  qm31* dst;
  uint32_t point;
};

// A leaf layer that is one FRI line fold of another layer (a proof's FRI layer trees): the
// launch that hashes the leaves computes them, leaf r = (a + b) + alpha (a - b) tw[r] with
// a, b = src[2r], src[2r + 1] (QM31 in four coordinate columns), writes them to the leaf
// layer's four columns and hashes them, so the fold needs no launch of its own.  src and
// alpha are job 0's (they move by the job offset), alpha is read from device memory (the
// previous commitment's channel step drew it).
struct LeafFold {
  const uint32_t* src[4];
  const uint32_t* tw;  // the folded layer's line twiddles, tw[r] for leaf r
  const qm31* alpha;
};

// Columns injected into a layer: device array of column pointers.
struct LayerCols {
  const uint32_t* const* cols;
  uint32_t n_cols;
  uint32_t pad;
};

struct MerklePlan {
  struct Step {
    enum Kind { LAYER, SUBTREE } kind;
    uint32_t top, depth;
  };
  uint32_t max_log;
  // jobs hashed by every launch (blockIdx.z): n_jobs trees of the same shape whose buffers lie
  // jstride bytes apart (the proof jobs of a batch); the tables hold job 0's pointers
  uint32_t n_jobs = 1;
  uint64_t jstride = 0;
  DevBuf own;                     // the device tables, unless they live in caller memory
  const LayerCols* dlcs = nullptr;  // device: per-layer column tables
  Hash* const* douts = nullptr;     // device: layer output pointers
  uint32_t* darrive = nullptr;      // device: arrival counters of subtree launches, [layer][job]
  std::vector<LayerCols> lcs;       // host copies (device column-table pointers)
  std::vector<Hash*> outs;
  std::vector<Step> steps;

  // Bytes of device tables for a tree with n_cols columns and layers 0 .. max_log.
  static size_t table_bytes(uint32_t n_cols, uint32_t max_log, uint32_t n_jobs = 1);

  // cols/logs: host arrays (device column pointers); layers: host array of max_log + 1
  // device buffers, layers[k] holding 2^k hashes.  The tables are uploaded to a buffer
  // of the plan's own.
  MerklePlan(uint32_t n_cols, const uint32_t* const* cols, const uint32_t* logs, Hash* const* layers,
             uint32_t max_log);
  // Same, with the tables placed at device address `dev` (table_bytes long); their bytes
  // are written to `host_image` and the caller uploads them to `dev` before enqueue().
  // n_jobs / jstride: see above (table_bytes(n_cols, max_log, n_jobs) bytes at dev).
  MerklePlan(uint32_t n_cols, const uint32_t* const* cols, const uint32_t* logs, Hash* const* layers,
             uint32_t max_log, void* dev, void* host_image, uint32_t n_jobs = 1, uint64_t jstride = 0);
  // On ctx().cur; allocates nothing.  layer_ready (optional, indexed by layer): an event
  // per layer (or null) that marks its injected columns written, e.g. on another stream;
  // each step waits only for the layers it reads, so hashing starts while other
  // columns are still being produced.
  // root_draw (optional): run by the subtree launch that writes the root; returns whether
  // it ran (false when the root is written by a layer launch: the caller mixes it itself).
  // leaf_fold (optional, leaf_fold_ok()): the leaves are computed by that fold
  bool enqueue(const std::vector<hipEvent_t>* layer_ready = nullptr, const RootDraw* root_draw = nullptr,
               const LeafFold* leaf_fold = nullptr) const;
  // whether enqueue can compute the leaves by a LeafFold: the leaf layer holds exactly four
  // columns (one QM31 column) and one subtree launch hashes it
  bool leaf_fold_ok() const;
  // Enqueue steps from *next on while every layer they read is >= min_layer (the caller
  // has enqueued the producers of those layers); advances *next.  Interleaving these
  // calls with the producers' launches hands the hashing to the device as early as the
  // host submits its inputs.
  bool enqueue_from(size_t* next, uint32_t min_layer, const std::vector<hipEvent_t>* layer_ready,
                    const RootDraw* root_draw = nullptr, const LeafFold* leaf_fold = nullptr) const;
  // lowest layer a step reads injected columns of
  static uint32_t lowest_layer(const Step& s) { return s.kind == Step::LAYER ? s.top : s.top + 1 - s.depth; }

 private:
  void build(uint32_t n_cols, const uint32_t* const* cols, const uint32_t* logs, Hash* const* layers, char* dev,
             char* host);
};

}  // namespace merkle
}  // namespace sh
