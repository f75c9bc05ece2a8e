/*
 * stwo_hip.h -- C ABI of the MI355X (gfx950) Stwo prover backend.
 *
 * Part 1 keeps the names, argument order and argument meaning of the FFI the
 * reference backend's Rust crate binds (stwo_gpu_backend/src/cuda/bindings.rs),
 * so that crate links against libstwo_hip.so with only its #[link] name changed.
 * Each declaration cites the binding it replaces and the CUDA function behind it.
 *
 * Conventions (same as the reference unless noted):
 *   - Device buffers are plain device pointers returned by the allocation
 *     functions below; the caller frees them with cuda_free_memory.
 *   - M31 columns: uint32_t[n] with values in [0, P), P = 2^31 - 1.
 *     QM31 vectors (SecureFieldVec): AoS, 4 x uint32_t per element.
 *     Secure columns (SecureColumnByCoords): four M31 columns (SoA).
 *   - Part-1 functions are synchronous: they return when the device work is
 *     complete, like the reference (which synchronises after every launch).
 *   - Errors: the reference prints and continues.  Here every failure is
 *     printed to stderr once, recorded, and readable with stwo_hip_last_error();
 *     pointer-returning functions return NULL on failure.
 *   - Differences from the reference (all are fixes): every output is canonical
 *     (the reference's neg(0) = P); 64-bit sizes: bit reversal, batch inversion and
 *     the folds take up to 2^31 elements, while columns that the FFT, Merkle,
 *     quotient, OODS and composition kernels read hold at most 2^30 values (their
 *     kernels form byte offsets in 32 bits; larger calls fail with an argument
 *     error before any launch); evaluate_columns
 *     accepts columns of different sizes; accumulate_quotients uses the
 *     cumulative line-coefficient offset; interpolate/evaluate handle circle
 *     domains of log size 1 and 2 (assumed canonic, as stwo constructs them).
 *
 * Part 2 (stwo_hip_*) adds what the drop-in path needs beyond the reference:
 * explicit init, status codes, asynchronous batched variants on the library's
 * stream, whole-tree Merkle commitment with mixed-degree column injection, a
 * twiddle cache, device-side synthetic inputs and timing.
 */
#ifndef STWO_HIP_H
#define STWO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
/* every declaration below is the library's exported interface (it is built with hidden
 * visibility for everything else) */
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

/* bindings.rs:8-14 (CudaSecureField): (a + b i) + (c + d i) u */
typedef struct { uint32_t a, b, c, d; } StwoQM31;
/* bindings.rs:45-49 (CirclePointBaseField) */
typedef struct { uint32_t x, y; } StwoCirclePointM31;
/* bindings.rs:60-64 (CirclePointSecureField) */
typedef struct { StwoQM31 x, y; } StwoCirclePointQM31;
/* cuda/include/utils.cuh:7-9 (Blake2sHash): 32-byte digest as 8 LE words */
typedef struct { uint32_t s[8]; } StwoBlake2sHash;

/* ======================================================================= */
/* Part 1: reference FFI (stwo_gpu_backend/src/cuda/bindings.rs:76-270)     */
/* ======================================================================= */

/* bindings.rs:77-81 -> cuda/src/utils.cu:10-12 */
void copy_uint32_t_vec_from_device_to_host(const uint32_t *device_ptr, uint32_t *host_ptr, uint32_t size);
/* bindings.rs:83 -> utils.cu:14-18 */
uint32_t *copy_uint32_t_vec_from_host_to_device(const uint32_t *host_ptr, uint32_t size);
/* bindings.rs:85-89 -> utils.cu:20-22 (returns dst; the C side of the reference returns void) */
uint32_t *copy_uint32_t_vec_from_device_to_device(const uint32_t *from, uint32_t *dst, uint32_t size);
/* bindings.rs:91 -> utils.cu:24-27 */
uint32_t *cuda_malloc_uint32_t(uint32_t size);
/* bindings.rs:93 -> utils.cu:29-32 */
StwoBlake2sHash *cuda_malloc_blake_2s_hash(size_t size);
/* bindings.rs:95 -> utils.cu:41-45 */
uint32_t *cuda_alloc_zeroes_uint32_t(uint32_t size);
/* bindings.rs:97 -> utils.cu:47-51 (zeroes all 32 bytes per hash; the reference zeroes 4) */
StwoBlake2sHash *cuda_alloc_zeroes_blake_2s_hash(size_t size);
/* bindings.rs:99 -> utils.cu:75-80 */
void cuda_free_memory(void *device_ptr);

/* bindings.rs:101 -> cuda/src/bit_reverse.cu:18-23 (ColumnOps<BaseField>::bit_reverse_column) */
void bit_reverse_base_field(uint32_t *array, size_t size);
/* bindings.rs:103 -> bit_reverse.cu:26-32 (ColumnOps<SecureField>, AoS) */
void bit_reverse_secure_field(uint32_t *array, size_t size);

/* bindings.rs:105 -> cuda/src/batch_inverse.cu:157-166 (FieldOps<BaseField>::batch_inverse) */
void batch_inverse_base_field(const uint32_t *from, uint32_t *dst, size_t size);
/* bindings.rs:107 -> batch_inverse.cu:168-177 (FieldOps<SecureField>, AoS) */
void batch_inverse_secure_field(const uint32_t *from, uint32_t *dst, size_t size);

/* bindings.rs:109-112 -> cuda/src/poly/twiddles.cu:17-27 (PolyOps::new_canonical_ordered) */
uint32_t *sort_values_and_permute_with_bit_reverse_order(const uint32_t *from, size_t size);

/* bindings.rs:114-118 -> twiddles.cu:50-73: twiddle tree of the coset (initial, step, size).
 * The itwiddles are produced by batch_inverse_base_field (poly.rs:180-185). */
uint32_t *precompute_twiddles(StwoCirclePointM31 initial, StwoCirclePointM31 step, size_t total_size);

/* bindings.rs:120-126 -> cuda/src/poly/ifft.cu:41-65 (PolyOps::interpolate), in place.
 * eval_domain_size = domain.half_coset.size(); values_size = 2 * eval_domain_size. */
void interpolate(uint32_t eval_domain_size, uint32_t *values, const uint32_t *inverse_twiddles_tree,
                 uint32_t inverse_twiddle_tree_size, uint32_t values_size);
/* bindings.rs:128-135 -> ifft.cu:123-163 (PolyOps::interpolate_columns); values is a HOST
 * array of values_size device pointers, each column number_of_rows long; in place. */
void interpolate_columns(uint32_t eval_domain_size, uint32_t *const *values,
                         const uint32_t *inverse_twiddles_tree, uint32_t inverse_twiddle_tree_size,
                         uint32_t values_size, uint32_t number_of_rows);
/* bindings.rs:137-143 -> cuda/src/poly/rfft.cu:45-64 (PolyOps::evaluate), in place on the
 * zero-padded coefficients. */
void evaluate(uint32_t eval_domain_size, uint32_t *values, const uint32_t *twiddles_tree,
              uint32_t twiddle_tree_size, uint32_t values_size);
/* bindings.rs:145-152 -> rfft.cu:110-141 (PolyOps::evaluate_polynomials); values is a HOST
 * array of device pointers; eval_domain_sizes / column_sizes are HOST arrays (per column). */
void evaluate_columns(const uint32_t *eval_domain_sizes, uint32_t *const *values,
                      const uint32_t *twiddles_tree, uint32_t twiddle_tree_size,
                      uint32_t number_of_columns, const uint32_t *column_sizes);

/* bindings.rs:154-159 -> cuda/src/poly/eval_at_point.cu:99-155 (PolyOps::eval_at_point) */
StwoQM31 eval_at_point(const uint32_t *coeffs, uint32_t coeffs_size, StwoQM31 point_x, StwoQM31 point_y);

/* bindings.rs:161-168 -> cuda/src/fri/fold_line.cu:36-53 (FriOps::fold_line).
 * eval_values / folded_values are HOST arrays of 4 device pointers (SoA). */
void fold_line(const uint32_t *gpu_domain, size_t twiddle_offset, size_t n,
               uint32_t *const *eval_values, StwoQM31 alpha, uint32_t *const *folded_values);
/* bindings.rs:170-177 -> cuda/src/fri/fold_circle_into_line.cu:48-69 (FriOps::fold_circle_into_line) */
void fold_circle_into_line(const uint32_t *gpu_domain, size_t twiddle_offset, size_t n,
                           uint32_t *const *eval_values, StwoQM31 alpha, uint32_t *const *folded_values);

/* bindings.rs:179 -> cuda/src/accumulate.cu:15-26 (AccumulationOps::accumulate); HOST arrays of 4 */
void accumulate(uint32_t size, uint32_t *const *left_columns, uint32_t *const *right_columns);

/* bindings.rs:181-186 -> cuda/src/blake2s.cu:157-162 (MerkleOps::commit_on_layer, leaves).
 * columns is a DEVICE array of device pointers (see copy_device_pointer_vec_from_host_to_device). */
void commit_on_first_layer(size_t size, size_t amount_of_columns, uint32_t *const *columns,
                           StwoBlake2sHash *result);
/* bindings.rs:188-194 -> blake2s.cu:164-170 (MerkleOps::commit_on_layer, inner layers) */
void commit_on_layer_with_previous(size_t size, size_t amount_of_columns, uint32_t *const *columns,
                                   const StwoBlake2sHash *previous_layer, StwoBlake2sHash *result);

/* bindings.rs:196-211 -> utils.cu:53-68.  Copies infer the direction from the pointers, so the
 * reference wrapper's device-pointer-as-host-pointer calls (blake2s.rs:55-69) also work. */
StwoBlake2sHash *copy_blake_2s_hash_vec_from_host_to_device(const StwoBlake2sHash *from, size_t size);
void copy_blake_2s_hash_vec_from_device_to_host(const StwoBlake2sHash *from, StwoBlake2sHash *to, size_t size);
void copy_blake_2s_hash_vec_from_device_to_device(const StwoBlake2sHash *from, StwoBlake2sHash *dst, size_t size);
/* bindings.rs:213-216 -> utils.cu:70-73: device copy of a host array of device pointers */
uint32_t **copy_device_pointer_vec_from_host_to_device(uint32_t *const *from, size_t size);

/* bindings.rs:218-236 -> cuda/src/quotients.cu:180-270 (QuotientOps::accumulate_quotients).
 * columns: DEVICE array of device pointers.  sample_points: HOST array of sample_size
 * StwoCirclePointQM31.  sample_column_indexes / sample_column_values: HOST, flattened in batch
 * order; sample_column_and_values_sizes: HOST, per batch.  Results: 4 device columns of
 * domain_size.  flattened_line_coeffs_size is accepted and ignored (sizes are derived). */
void accumulate_quotients(uint32_t half_coset_initial_index, uint32_t half_coset_step_size,
                          uint32_t domain_size, uint32_t *const *columns, size_t number_of_columns,
                          StwoQM31 random_coeff, const uint32_t *sample_points,
                          const uint32_t *sample_column_indexes, uint32_t sample_column_indexes_size,
                          const StwoQM31 *sample_column_values, const uint32_t *sample_column_and_values_sizes,
                          uint32_t sample_size, uint32_t *result_column_0, uint32_t *result_column_1,
                          uint32_t *result_column_2, uint32_t *result_column_3,
                          uint32_t flattened_line_coeffs_size);

/* bindings.rs:238-244 -> cuda/src/gkr.cu:20-38 (GkrOps::gen_eq_evals, lookups/gkr.rs:12-27):
 * evals[x] = v * prod_k (y_k if bit (y_size-1-k) of x else 1 - y_k), x < 2^y_size.
 * y: HOST array of y_size QM31; evals: device QM31 (AoS) of evals_size = 2^y_size. */
void gen_eq_evals(StwoQM31 v, const StwoQM31 *y, uint32_t y_size, StwoQM31 *evals, uint32_t evals_size);
/* bindings.rs:246-251 -> cuda/src/mle.cu:14-20 (MleOps<BaseField>::fix_first_variable,
 * lookups/mle.rs:12-31): output_evals[i] = lhs + (rhs - lhs) assignment with lhs = evals[i],
 * rhs = evals[i + evals_size/2]; M31 in, QM31 (AoS) out of evals_size/2 elements. */
void fix_first_variable_base_field(const uint32_t *evals, size_t evals_size, StwoQM31 assignment,
                                   uint32_t *output_evals);
/* bindings.rs:253-258 -> mle.cu:22-28 (MleOps<SecureField>::fix_first_variable, mle.rs:33-57);
 * QM31 (AoS) in and out. */
void fix_first_variable_secure_field(const uint32_t *evals, size_t evals_size, StwoQM31 assignment,
                                     uint32_t *output_evals);

/* bindings.rs:260-270 -> cuda/src/examples/wide_fibonacci.cu:40-65 (WideFib composition).
 * trace_evaluations: HOST array of number_of_columns device pointers; random_coeff_powers:
 * device QM31 (AoS) of number_of_columns - 2; denominator_inverses: device M31. */
void evaluate_wide_fibonacci_constraint_quotients_on_domain(
    uint32_t *quotients_0, uint32_t *quotients_1, uint32_t *quotients_2, uint32_t *quotients_3,
    uint32_t *const *trace_evaluations, const uint32_t *random_coeff_powers,
    const uint32_t *denominator_inverses, uint32_t extended_domain_size, uint32_t number_of_columns);

/* ======================================================================= */
/* Part 2: extensions (no counterpart in the reference)                     */
/* ======================================================================= */

enum {
  STWO_HIP_OK = 0,
  STWO_HIP_ERR_ARG = 1,    /* invalid argument (size, null, not a power of two, ...) */
  STWO_HIP_ERR_HIP = 2,    /* HIP runtime error */
  STWO_HIP_ERR_NOMEM = 3,  /* device allocation failed */
  STWO_HIP_ERR_NODEV = 4,  /* no gfx950 device / library not initialised */
  STWO_HIP_NOT_READY = 5   /* not an error: nothing finished yet (server wait) */
};

/* Initialise on `device` (idempotent).  Every other entry point initialises device 0 lazily. */
int stwo_hip_init(int device);
/* Status and message of the last failure on the calling thread (0 / "" when none). */
int stwo_hip_last_status(void);
const char *stwo_hip_last_error(void);
void stwo_hip_clear_error(void);
/* Wait for all work the library queued on its stream. */
int stwo_hip_synchronize(void);
/* The library's stream (a hipStream_t), for callers that order their own work after it. */
void *stwo_hip_stream(void);
/* Device memory: stream-ordered caching pool used by every entry point. */
int stwo_hip_malloc(void **ptr, size_t bytes);
int stwo_hip_free(void *ptr);
int stwo_hip_memcpy(void *dst, const void *src, size_t bytes); /* synchronous, any direction */
int stwo_hip_release_cached_memory(void);

/* Device-to-device copy by a plain 16-byte-per-lane kernel (async; 16-byte aligned sizes). */
int stwo_hip_device_copy(void *dst, const void *src, size_t bytes);
/* Synthetic input: dst[i] = uniform M31 value number (offset + i) of stream `seed` (async). */
int stwo_hip_fill_uniform_m31(uint32_t *dst, uint64_t n, uint64_t seed, uint64_t offset);

/* Cached twiddle tree for the coset (initial_index, step_size, log_size) of the circle
 * group (indices mod 2^31).  Returns device pointers owned by the library (valid until
 * stwo_hip_release_twiddles); both trees have 2^log_size entries. */
int stwo_hip_twiddles(uint32_t initial_index, uint32_t step_size, uint32_t log_size,
                      const uint32_t **twiddles, const uint32_t **itwiddles);
/* Frees every cached twiddle tree and the cached quotient domain tables (one small table
 * pair per circle domain that accumulate_quotients has run over). */
int stwo_hip_release_twiddles(void);

/* Asynchronous batched circle FFTs on the library stream.  Columns may have different log
 * sizes (one fused plan per size class).  All domains are canonic (CanonicCoset(log).
 * circle_domain()) and served by one twiddle tree whose root coset is a doubling source
 * of every domain (tree_len = 2^root_log).
 *   interpolate: in place; cols[i] has 2^log_sizes[i] values.
 *   evaluate: coeffs[i] (2^coeff_log[i] values) extended with zeros and evaluated on the
 *             canonic domain of log eval_log[i] >= coeff_log[i], written to dst[i]
 *             (dst may alias coeffs only when eval_log == coeff_log). */
int stwo_hip_interpolate_columns_async(uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols,
                                       const uint32_t *itwiddles, uint64_t tree_len);
/* Low-degree extension in one call: evals[i] (2^log_sizes[i] evaluations) are interpolated
 * into coeffs[i] (may alias evals[i]) and evaluated on the canonic domain of log
 * log_sizes[i] + log_blowup into lde[i].  Columns of log 14..21 take a fused path (three
 * passes over HBM instead of four). */
int stwo_hip_lde_columns_async(const uint32_t *const *evals, uint32_t *const *coeffs, uint32_t *const *lde,
                               const uint32_t *log_sizes, uint32_t n_cols, uint32_t log_blowup,
                               const uint32_t *twiddles, const uint32_t *itwiddles, uint64_t tree_len);
/* Out-of-place interpolate: src[i] (evaluations) is read, dst[i] receives the coefficients. */
int stwo_hip_interpolate_columns_oop_async(const uint32_t *const *src, uint32_t *const *dst,
                                           const uint32_t *log_sizes, uint32_t n_cols,
                                           const uint32_t *itwiddles, uint64_t tree_len);
int stwo_hip_evaluate_columns_async(const uint32_t *const *coeffs, const uint32_t *coeff_log,
                                    uint32_t *const *dst, const uint32_t *eval_log, uint32_t n_cols,
                                    const uint32_t *twiddles, uint64_t tree_len);

/* Column::at and decommitment gathers (replace the whole-column D2H copy per element of
 * column.rs:49-52 and the per-query host walks of MerkleProver::decommit).
 * gather_rows: values[i * n_cols + c] = columns[c][rows[i]]; columns: HOST array of device
 * columns of column_size values each; rows, values: HOST.  merkle_auth_paths: layers: HOST
 * array of max_log + 1 device layers (layers[k] holds 2^k hashes, k = max_log the leaves);
 * paths[q * max_log + j] = sibling at layer max_log - j of the node on the path of leaf
 * leaves[q]; leaves, paths: HOST.  Both synchronous; indices are checked on the host. */
int stwo_hip_gather_rows(const uint32_t *const *columns, uint32_t n_cols, uint64_t column_size, const uint32_t *rows,
                         uint32_t n_rows, uint32_t *values);
int stwo_hip_merkle_auth_paths(StwoBlake2sHash *const *layers, uint32_t max_log, const uint32_t *leaves,
                               uint32_t n_leaves, StwoBlake2sHash *paths);

/* Whole Merkle tree (MerkleProver::commit semantics): columns sorted by decreasing size
 * (stable), one layer per log size from the largest down to 0, columns of each size injected
 * into that layer.  layers: HOST array of (max_log + 1) device buffers, layers[k] holding
 * 2^k hashes (the root is layers[0][0]); root_out (host, may be NULL) receives the root. */
int stwo_hip_merkle_commit_async(uint32_t n_cols, const uint32_t *const *cols, const uint32_t *log_sizes,
                                 StwoBlake2sHash *const *layers, uint32_t max_log);
int stwo_hip_merkle_root(const StwoBlake2sHash *root_layer, StwoBlake2sHash *root_out);

/* Asynchronous FRI folds (same arguments as fold_line / fold_circle_into_line). */
int stwo_hip_fold_line_async(const uint32_t *itwiddles, size_t twiddle_offset, size_t n,
                             uint32_t *const *eval_values, StwoQM31 alpha, uint32_t *const *folded_values);
int stwo_hip_fold_circle_into_line_async(const uint32_t *itwiddles, size_t twiddle_offset, size_t n,
                                         uint32_t *const *eval_values, StwoQM31 alpha,
                                         uint32_t *const *folded_values);
/* A whole FRI fold sequence whose coefficients are all known (FriOps::fold_circle_into_line then
 * FriOps::fold_line layer after layer, stwo_gpu_backend/src/fri.rs:15-66): layers[0] accumulates
 * the circle fold of `circle` (2^log_circle values) with alphas[0], exactly as
 * fold_circle_into_line(itwiddles, tree_len - 2^(log_circle-1), ...) does, then
 * layers[j] = fold_line(layers[j-1], alphas[j]) at twiddle offset tree_len - 2^(log_circle-j),
 * j = 1 .. n_layers-1.  layers: HOST array of 4 * n_layers device pointers (four coordinate
 * columns per layer, layer j of 2^(log_circle-1-j) values); alphas: HOST.  Each launch folds
 * 2^13-value ranges up to 13 times in LDS (a log-22 circle to 2^4: two launches). */
int stwo_hip_fri_fold_chain_async(const uint32_t *itwiddles, uint64_t tree_len, uint32_t log_circle,
                                  uint32_t *const *circle, uint32_t n_layers, const StwoQM31 *alphas,
                                  uint32_t *const *layers);
/* Asynchronous DEEP quotients.  columns: HOST array of n_columns device pointers (each of
 * domain_size values); points/batch_sizes/column_indexes/values: HOST, in batch order;
 * out: HOST array of 4 device columns. */
int stwo_hip_accumulate_quotients_async(uint32_t half_coset_initial_index, uint32_t half_coset_step_size,
                                        uint32_t domain_size, const uint32_t *const *columns, uint32_t n_columns,
                                        StwoQM31 random_coeff, uint32_t n_batches, const StwoCirclePointQM31 *points,
                                        const uint32_t *batch_sizes, const uint32_t *column_indexes,
                                        const StwoQM31 *values, uint32_t *const *out);

/* Batched commitment plan: the LDE (interpolate into coeffs[i], which may alias evals[i],
 * then evaluate on the canonic domain of log log_sizes[i] + log_blowup into ldes[i]) of
 * columns of any mix of sizes, then the Merkle tree over all LDE columns with
 * mixed-degree injection into layers[0..max_log] (max_log = largest LDE log size).  The
 * plan keeps device copies of every pointer table; the buffers must outlive it.
 * run(use_graph = 1) captures the whole sequence as a hipGraph on first use and replays
 * it; autotune = 1 times the candidate FFT pass schedules of every size class once and
 * keeps the fastest (process-wide, per log size, blowup and column-count bucket), then
 * times graph replay against stream launches of the plan: where the streams are faster
 * run(1) launches from the streams too (uses_graph reports the choice).  Where coeffs[i]
 * aliases evals[i], every run interpolates that column in place (after a run the buffer
 * holds coefficients, so write the evaluations again before the next run); the
 * autotuner's runs keep a copy of such columns and restore them before create returns.
 * No counterpart in the reference (its commit is one synchronous call per operation). */
typedef struct StwoCommitPlan StwoCommitPlan;
int stwo_hip_commit_plan_create(StwoCommitPlan **plan, uint32_t n_cols, const uint32_t *const *evals,
                                uint32_t *const *coeffs, uint32_t *const *ldes, const uint32_t *log_sizes,
                                uint32_t log_blowup, const uint32_t *twiddles, const uint32_t *itwiddles,
                                uint64_t tree_len, StwoBlake2sHash *const *layers, uint32_t max_log, int autotune);
int stwo_hip_commit_plan_run(StwoCommitPlan *plan, int use_graph);
int stwo_hip_commit_plan_uses_graph(const StwoCommitPlan *plan, int *uses_graph);
int stwo_hip_commit_plan_destroy(StwoCommitPlan *plan);
/* The FFT pass schedule in use for an LDE size class (autotuned or default): layers per
 * contiguous pass (12 or 13) and whether the strided passes are fused. */
int stwo_hip_lde_schedule(uint32_t log_size, uint32_t log_blowup, uint64_t n_cols, uint32_t *contiguous_layers,
                          int *fused);

/* Batched OODS evaluation (PolyOps::eval_at_point, stwo_gpu_backend/src/poly.rs:84-94, for
 * many polynomials at once): values[c] = value at (point_x, point_y) of the circle polynomial
 * with coefficients columns[c] (2^log_size each; HOST array of device columns).  Columns of
 * log >= 12 are evaluated as one matrix-vector product against a factored basis table;
 * 16-byte aligned columns are required there. */
int stwo_hip_eval_at_point_columns(const uint32_t *const *columns, uint32_t n_cols, uint32_t log_size,
                                   StwoQM31 point_x, StwoQM31 point_y, StwoQM31 *values);
/* The same for several points: values[k * n_cols + c] = column c at points[k] (the batched
 * OODS evaluation over (column, point) pairs; points: HOST). */
int stwo_hip_eval_at_points_columns(const uint32_t *const *columns, uint32_t n_cols, uint32_t log_size,
                                    const StwoCirclePointQM31 *points, uint32_t n_points, StwoQM31 *values);
/* No counterpart in the reference: stwo_hip_eval_at_points_columns for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results; no hipMalloc/hipFree or pool allocation: the pointer
 * table, points, values and partial sums live in a fixed 32 MiB device array (one per device,
 * held until the process ends; e.g. 512 columns up to log 23 at one point).  Calls needing
 * more return STWO_HIP_ERR_ARG before any copy or launch.  Calls that overlap on one device
 * share the array and race.  Synchronous, like the allocating form. */
int stwo_hip_eval_at_points_columns_noalloc(const uint32_t *const *columns, uint32_t n_cols, uint32_t log_size,
                                            const StwoCirclePointQM31 *points, uint32_t n_points,
                                            StwoQM31 *values);

/* Proof of work (GrindOps<Blake2sChannel>::grind, which the reference delegates to the CPU
 * SimdBackend, stwo_gpu_backend/src/backend.rs:14-18): the smallest nonce whose
 * BLAKE2s-256(digest || lo32(nonce) || hi32(nonce)) starts (as a little-endian 64-bit
 * value) with pow_bits zero bits.  digest: 8 host words. */
int stwo_hip_grind(const uint32_t *digest, uint32_t pow_bits, uint64_t *nonce);
/* The channel's tree hash of n_words words (16-word leaves F0(block), pairwise levels, an
 * unpaired last node carried up; F0 = BLAKE2s compression from the zero state), computed on
 * the HOST: for a Fiat-Shamir channel kept on the CPU, as stwo keeps its channel.  out: 8 words. */
int stwo_hip_host_tree_hash(const uint32_t *words, uint64_t n_words, uint32_t *out);

/* Wide-Fibonacci prover (the WideFibComponent of stwo_gpu_backend/src/examples/wide_fibonacci.rs
 * proved end to end on the device: trace LDE + commitment, composition, OODS samples, DEEP
 * quotients, FRI, proof of work, decommitments).  The trace is generated on the device from
 * `seed`.  The proof is a word array whose layout the header (first 8 words) describes. */
/* The trace of the wide-Fibonacci statement on the library stream: column i of 2^log_n
 * rows, col[0][r] = U(seed, r), col[1][r] = U(seed, 2^log_n + r) (the stream of
 * stwo_hip_fill_uniform_m31), col[i] = col[i-2]^2 + col[i-1]^2.  columns: HOST array. */
int stwo_hip_wide_fib_trace(uint32_t *const *columns, uint32_t n_cols, uint32_t log_n, uint64_t seed);
/* out[r] = 1 / Z(p_r) on the rows of CanonicCoset(log_ext) (bit-reversed), where
 * Z = pi^(log_n - 1)(x) vanishes on CanonicCoset(log_n) (the denominators of
 * examples/wide_fibonacci.rs:101-104, computed on the device). */
int stwo_hip_vanishing_inverse(uint32_t *out, uint32_t log_ext, uint32_t log_n);

typedef struct {
  uint32_t n_cols;    /* >= 3 */
  uint32_t log_n;     /* trace rows = 2^log_n, 4 .. 26 */
  uint32_t n_queries; /* 1 .. 64 */
  uint32_t pow_bits;  /* proof-of-work bits, <= 32 */
  uint32_t last_log;  /* log size of the last FRI layer, 2 .. 10 */
} StwoWideFibParams;
int stwo_hip_wide_fib_proof_size(const StwoWideFibParams *params, uint64_t *proof_words, uint64_t *device_bytes);
/* One proof on the library stream; returns when the proof is in `proof`. */
int stwo_hip_wide_fib_prove(const StwoWideFibParams *params, uint64_t seed, uint32_t *proof,
                            uint64_t proof_capacity_words);
/* n_jobs (1 .. 16) proofs of one statement shape from seeds[0 .. n_jobs-1], proved together:
 * every phase (trace, LDE, Merkle commitments with their channel steps, composition, OODS
 * evaluation, quotients, FRI folds, proof of work, decommitment gather) is one launch for all
 * of them.  proofs: n_jobs x proof_words words, job k's at k * proof_words, each the proof
 * stwo_hip_wide_fib_prove gives for that seed.  No counterpart in the reference. */
int stwo_hip_wide_fib_prove_batch(const StwoWideFibParams *params, const uint64_t *seeds, uint32_t n_jobs,
                                  uint32_t *proofs, uint64_t proof_capacity_words);

/* Proof server with continuous batching: jobs are admitted in arrival order while fewer than
 * max_inflight run and their device footprints fit hbm_budget_bytes; each runs on its own
 * stream and overlaps with the others.  wait() returns STWO_HIP_NOT_READY after timeout_us
 * without a finished job; otherwise the job's id, status, proof (when it fits) and in
 * times[0..3] its submit / start / end times in seconds since the server was created and
 * the host time spent building and enqueuing it. */
typedef struct StwoProofServer StwoProofServer;
int stwo_hip_server_create(StwoProofServer **server, uint64_t hbm_budget_bytes, uint32_t max_inflight);
/* Also hold back the next job while the trace elements (n_cols x 2^log_n) of the jobs in
 * flight plus its own exceed max_trace_elements (0: no limit, the default); set before
 * submitting.  No counterpart in the reference (it has no server). */
int stwo_hip_server_limit_work(StwoProofServer *server, uint64_t max_trace_elements);
/* Same-phase batching across jobs: admitting a job of at most max_batch_work trace elements
 * also takes up to max_batch - 1 queued jobs of the same statement shape, and they are proved
 * as one batch (as stwo_hip_wide_fib_prove_batch; each result is the job's own proof).
 * max_batch 1 (the default) turns it off; at most 16.  Set before submitting. */
int stwo_hip_server_batch(StwoProofServer *server, uint32_t max_batch, uint64_t max_batch_work);
int stwo_hip_server_submit(StwoProofServer *server, const StwoWideFibParams *params, uint64_t seed,
                           uint64_t *job_id);
int stwo_hip_server_wait(StwoProofServer *server, uint32_t timeout_us, uint64_t *job_id, int *job_status,
                         uint32_t *proof, uint64_t proof_capacity_words, uint64_t *proof_words, double *times);
int stwo_hip_server_destroy(StwoProofServer *server);

/* Timing on the library stream (HIP events). */
int stwo_hip_timer_begin(void);
int stwo_hip_timer_end(float *elapsed_ms);
/* Per-kernel timing: when enabled, every launch is bracketed by events on the library stream
 * and its duration is added to a per-kernel total (for roofline figures). */
int stwo_hip_kernel_timing(int enable);
int stwo_hip_kernel_stats(const char *kernel_name, uint64_t *launches, double *total_ms);
int stwo_hip_kernel_stats_reset(void);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif
#ifdef __cplusplus
}
#endif
#endif /* STWO_HIP_H */
