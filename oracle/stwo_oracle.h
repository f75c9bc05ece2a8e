/*
 * stwo_oracle.h -- CPU restatement of the Stwo circle-STARK backend operations.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP backend:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load it.  The product path (stwo-gpu_amd/) never links or calls it.
 *
 * Semantics follow the reference backend's CUDA kernels and the Rust trait
 * contracts they are tested against (stwo CpuBackend, which the reference's
 * tests use as the expected value).  Every function cites the reference
 * file:line it restates.  All arithmetic is exact modulo P = 2^31 - 1 and every
 * stored word is canonical (in [0, P)).
 *
 * Layouts:
 *   M31 column         : uint32_t[n]
 *   QM31 (AoS)         : uint32_t[4*n], element = (a, b, c, d) = (a + b i) + (c + d i) u
 *   QM31 SoA column    : four uint32_t[n] coordinate columns
 *   Blake2s hash       : uint32_t[8] (little-endian words of the 32-byte digest)
 *   circle point (M31) : (x, y); QM31 point: x[4], y[4]
 *   coset              : (initial_index, step_size, log_size), indices mod 2^31
 */
#ifndef STWO_ORACLE_H
#define STWO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_P 2147483647u

/* ---- fields (cuda/src/fields.cu:3-160, cuda/include/fields.cuh:18-19) ---- */
uint32_t or_m31_add(uint32_t a, uint32_t b);
uint32_t or_m31_sub(uint32_t a, uint32_t b);
uint32_t or_m31_neg(uint32_t a);
uint32_t or_m31_mul(uint32_t a, uint32_t b);
uint32_t or_m31_inv(uint32_t a);
void or_qm31_add(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void or_qm31_sub(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void or_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void or_qm31_inv(const uint32_t a[4], uint32_t out[4]);
void or_qm31_pow(const uint32_t a[4], uint64_t e, uint32_t out[4]);
void or_cm31_inv(const uint32_t a[2], uint32_t out[2]);

/* ---- circle group (cuda/include/point.cuh:17-65) ---- */
void or_point_at_index(uint32_t index, uint32_t out[2]);
void or_coset_at(uint32_t initial_index, uint32_t step_size, uint32_t i, uint32_t out[2]);
/* circle domain point i: half_coset[i] for i < N/2, else conj(half_coset[i - N/2]) */
void or_circle_domain_at(uint32_t half_initial_index, uint32_t half_step_size,
                         uint32_t half_log_size, uint64_t i, uint32_t out[2]);

/* ---- columns ---- */
void or_bit_reverse_u32(uint32_t *a, uint64_t n);
void or_bit_reverse_qm31(uint32_t *a, uint64_t n);
void or_new_canonical_ordered(const uint32_t *in, uint32_t *out, uint64_t n);

/* ---- twiddles (cuda/src/poly/twiddles.cu:29-73, stwo_gpu_backend/src/poly.rs:170-192) ---- */
void or_precompute_twiddles(uint32_t initial_index, uint32_t step_size, uint32_t log_size,
                            uint32_t *twiddles, uint32_t *itwiddles);

/* ---- circle FFT (cuda/src/poly/ifft.cu:136-160, rfft.cu:45-64) ----
 * The domain is CircleDomain(half_coset); half_log = log2(N) - 1.  The tree
 * must belong to a coset of which half_coset is a repeated doubling. */
void or_interpolate(uint32_t *values, uint32_t half_initial_index, uint32_t half_log,
                    const uint32_t *itwiddles, uint64_t tree_len);
void or_evaluate(uint32_t *values, uint32_t half_initial_index, uint32_t half_log,
                 const uint32_t *twiddles, uint64_t tree_len);
/* interpolate then extend + evaluate on CanonicCoset(log_n + log_blowup); many columns. */
void or_lde_columns(const uint32_t *const *evals, uint32_t *const *coeffs_out,
                    uint32_t *const *lde_out, uint32_t n_cols, uint32_t log_n,
                    uint32_t log_blowup, const uint32_t *twiddles,
                    const uint32_t *itwiddles, uint64_t tree_len);

/* ---- eval_at_point (cuda/src/poly/eval_at_point.cu:99-155) ---- */
void or_eval_at_point(const uint32_t *coeffs, uint32_t log_n, const uint32_t px[4],
                      const uint32_t py[4], uint32_t out[4]);

/* ---- batch inverse (cuda/src/batch_inverse.cu:24-177) ---- */
void or_batch_inverse_m31(const uint32_t *src, uint32_t *dst, uint64_t n);
void or_batch_inverse_qm31(const uint32_t *src, uint32_t *dst, uint64_t n);

/* ---- Blake2s (cuda/src/blake2s.cu:5-170) ---- */
void or_blake2s_compress(uint32_t h[8], const uint32_t m[16], uint32_t t0, uint32_t t1,
                         uint32_t f0, uint32_t f1);
void or_blake2s256(const uint8_t *data, uint64_t len, uint8_t out[32]);
void or_merkle_commit_on_layer(uint32_t log_size, const uint32_t *prev_layer,
                               uint32_t n_cols, const uint32_t *const *cols,
                               uint32_t *out);

/* ---- FRI (cuda/src/fri/fold_line.cu:5-53, fold_circle_into_line.cu:7-69) ---- */
void or_fold_line(const uint32_t *const src[4], uint64_t n, uint32_t line_initial_index,
                  uint32_t line_step_size, const uint32_t alpha[4], uint32_t *const dst[4]);
void or_fold_circle_into_line(const uint32_t *const src[4], uint64_t n,
                              uint32_t half_initial_index, uint32_t half_step_size,
                              const uint32_t alpha[4], uint32_t *const dst[4]);

/* ---- quotients (cuda/src/quotients.cu:31-270) ----
 * sample_points: n_batches x 8 words (x[4], y[4]); per batch sizes[b] entries of
 * (column index, QM31 value) flattened in batch order. */
void or_accumulate_quotients(uint32_t half_initial_index, uint32_t half_step_size,
                             uint32_t log_size, const uint32_t *const *cols,
                             const uint32_t random_coeff[4], uint32_t n_batches,
                             const uint32_t *sample_points, const uint32_t *sizes,
                             const uint32_t *col_indices, const uint32_t *values,
                             uint32_t *const out[4]);

/* ---- accumulate (cuda/src/accumulate.cu:4-26) ---- */
void or_accumulate(uint64_t n, uint32_t *const left[4], const uint32_t *const right[4]);

/* ---- wide Fibonacci constraint quotients (cuda/src/examples/wide_fibonacci.cu:3-66) ---- */
void or_wide_fib_constraint_quotients(uint64_t n_rows, uint32_t n_cols,
                                      const uint32_t *const *trace,
                                      const uint32_t *coeff_powers,
                                      const uint32_t *denom_inv, uint32_t *const out[4]);

/* ---- multilinear ops of the lookups (cuda/src/gkr.cu:4-38, cuda/src/mle.cu:4-30) ----
 * QM31 values AoS (4 words); variable 0 is the most significant index bit. */
void or_gen_eq_evals(const uint32_t v[4], const uint32_t *y, uint32_t n_vars, uint32_t *evals);
void or_fix_first_variable_base(const uint32_t *evals, uint64_t n, const uint32_t a[4], uint32_t *out);
void or_fix_first_variable_secure(const uint32_t *evals, uint64_t n, const uint32_t a[4], uint32_t *out);

/* ---- planted mistakes, for the oracle's pinning tests only ----
 * or_set_mutation(m) makes later calls compute a deliberately wrong variant;
 * 0 (OR_MUT_NONE) restores the restatement.  Nothing but tests/test_oracle.py
 * sets it. */
enum {
    OR_MUT_NONE = 0,
    OR_MUT_QUOT_ALPHA_POW = 1,   /* sample j of a batch weighted alpha^j, not alpha^(j+1) */
    OR_MUT_QUOT_BATCH_POW = 2,   /* batch chain coefficient rc^(|b|+1), not rc^|b| */
    OR_MUT_QUOT_BATCH_ORDER = 3, /* batches accumulated in reverse order */
    OR_MUT_QUOT_REF_OFFSET = 4,  /* line coefficients stored at b*|b|*3 (quotients.cu:74) */
    OR_MUT_WF_POWER_ORDER = 5,   /* constraint i weighted coeff_pow[i], not coeff_pow[n-3-i] */
};
void or_set_mutation(int m);

/* ---- synthetic inputs (counter-based, identical to the device generator) ---- */
uint32_t or_uniform_m31(uint64_t seed, uint64_t index);
void or_fill_uniform_m31(uint32_t *dst, uint64_t n, uint64_t seed, uint64_t offset);

#ifdef __cplusplus
}
#endif
#endif
