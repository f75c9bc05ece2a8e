// Host-side interface of the circle FFT module used by the batched commit plan.
#pragma once
#include <stdint.h>

namespace sh {
namespace fft {

// Pass schedule of one LDE size class, chosen by the autotuner: the number of layers
// in the contiguous passes (12 or 13) and whether the strided inverse and forward
// passes are fused into one launch.
struct LdeVariant {
  uint8_t kc = 12;
  bool fused = false;
};

// Interpolate d_src into d_coeffs (may alias) and evaluate on the canonic domain of log
// n + log_blowup into d_lde; device pointer tables of n_cols columns; allocates nothing.
void lde_same_size(const uint32_t* const* d_src, uint32_t* const* d_coeffs, uint32_t* const* d_lde, uint32_t n_cols,
                   uint32_t n, uint32_t log_blowup, const uint32_t* tw, const uint32_t* itw, uint64_t tree_len,
                   LdeVariant var);

// Autotuner results (per log size, blowup and column-count bucket); defaults otherwise.
LdeVariant lde_variant_for(uint32_t n, uint32_t log_blowup, uint64_t n_cols);
void set_lde_variant(uint32_t n, uint32_t log_blowup, uint64_t n_cols, LdeVariant v);

}  // namespace fft
}  // namespace sh
