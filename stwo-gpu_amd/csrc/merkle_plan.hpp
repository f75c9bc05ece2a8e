// Merkle commitment plan shared by the whole-tree entry point and batched commits:
// device tables of the columns injected into each layer, built once, and the launch
// schedule (leaf layer, per-layer launches for injection-heavy layers, subtree
// launches covering runs of light layers).
#pragma once
#include <stdint.h>

#include <vector>

#include "runtime.hpp"

namespace sh {
namespace merkle {

struct Hash {
  uint32_t s[8];
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
  DevBuf dcols, dlcs, douts;
  std::vector<LayerCols> lcs;
  std::vector<Hash*> outs;
  std::vector<Step> steps;

  // cols/logs: host arrays (device column pointers); layers: host array of max_log + 1
  // device buffers, layers[k] holding 2^k hashes.
  MerklePlan(uint32_t n_cols, const uint32_t* const* cols, const uint32_t* logs, Hash* const* layers,
             uint32_t max_log);
  void enqueue() const;  // on ctx().cur; allocates nothing
};

}  // namespace merkle
}  // namespace sh
