/*
 * stwo_oracle.c -- CPU restatement of the Stwo backend operations (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.
 *
 * Reference files are cited relative to the reference repository root.  Where
 * the reference CUDA deviates from the trait contract its tests check against
 * (stwo CpuBackend), the contract wins and the deviation is named:
 *   - neg(0) returns P in cuda/src/fields.cu:19-21; here neg(0) = 0.
 *   - the line-coefficient offset in cuda/src/quotients.cu:74 is wrong when
 *     batch sizes differ; here the cumulative offset is used (quotients.cu:150).
 *   - interpolate/evaluate of circle domains of log size 1 and 2 are special
 *     cases in stwo's CpuBackend that the reference CUDA lacks.
 */
#include "stwo_oracle.h"

#include <stdlib.h>
#include <string.h>

#define P OR_P

/* Planted mistakes for the oracle's own pinning tests (tests/test_oracle.py
   sets one, checks that a definition-derived check rejects the result, and
   resets it to 0).  Every other caller leaves it at 0.  See OR_MUT_* in
   stwo_oracle.h. */
static int g_mutation = OR_MUT_NONE;
void or_set_mutation(int m) { g_mutation = m; }

/* ------------------------------------------------------------------------- */
/* M31: cuda/src/fields.cu:3-40                                               */
/* ------------------------------------------------------------------------- */

static inline uint32_t reduce62(uint64_t v) {
    /* v < 2^62; 2^31 = 1 (mod P) */
    uint64_t r = (v & P) + (v >> 31);
    r = (r & P) + (r >> 31);
    return r >= P ? (uint32_t)(r - P) : (uint32_t)r;
}
uint32_t or_m31_add(uint32_t a, uint32_t b) { uint32_t s = a + b; return s >= P ? s - P : s; }
uint32_t or_m31_sub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + P - b; }
uint32_t or_m31_neg(uint32_t a) { return a ? P - a : 0; }
uint32_t or_m31_mul(uint32_t a, uint32_t b) { return reduce62((uint64_t)a * b); }
static uint32_t m31_pow(uint32_t a, uint64_t e) {
    uint32_t r = 1;
    while (e) { if (e & 1) r = or_m31_mul(r, a); a = or_m31_mul(a, a); e >>= 1; }
    return r;
}
/* fields.cu:32-40 uses an addition chain for x^(P-2); the value is the same. */
uint32_t or_m31_inv(uint32_t a) { return m31_pow(a, P - 2); }

/* ------------------------------------------------------------------------- */
/* CM31 = M31[i]/(i^2+1), QM31 = CM31[u]/(u^2 - (2+i)): fields.cu:42-160      */
/* ------------------------------------------------------------------------- */
typedef struct { uint32_t a, b; } cm31;
typedef struct { cm31 x, y; } qm31;

static inline cm31 cm_add(cm31 p, cm31 q) { return (cm31){or_m31_add(p.a, q.a), or_m31_add(p.b, q.b)}; }
static inline cm31 cm_sub(cm31 p, cm31 q) { return (cm31){or_m31_sub(p.a, q.a), or_m31_sub(p.b, q.b)}; }
static inline cm31 cm_neg(cm31 p) { return (cm31){or_m31_neg(p.a), or_m31_neg(p.b)}; }
static inline cm31 cm_mul(cm31 p, cm31 q) {
    /* (a + bi)(c + di) = (ac - bd) + (ad + bc) i   (fields.cu:47-49) */
    return (cm31){or_m31_sub(or_m31_mul(p.a, q.a), or_m31_mul(p.b, q.b)),
                  or_m31_add(or_m31_mul(p.a, q.b), or_m31_mul(p.b, q.a))};
}
static inline cm31 cm_scale(cm31 p, uint32_t s) { return (cm31){or_m31_mul(p.a, s), or_m31_mul(p.b, s)}; }
static cm31 cm_inv(cm31 p) {
    /* (a + bi)^-1 = (a - bi) / (a^2 + b^2)   (fields.cu:80-84) */
    uint32_t f = or_m31_inv(or_m31_add(or_m31_mul(p.a, p.a), or_m31_mul(p.b, p.b)));
    return (cm31){or_m31_mul(p.a, f), or_m31_mul(or_m31_neg(p.b), f)};
}
static const cm31 QM31_R = {2, 1}; /* u^2 = 2 + i  (fields.cuh:19) */

static inline qm31 qm_add(qm31 p, qm31 q) { return (qm31){cm_add(p.x, q.x), cm_add(p.y, q.y)}; }
static inline qm31 qm_sub(qm31 p, qm31 q) { return (qm31){cm_sub(p.x, q.x), cm_sub(p.y, q.y)}; }
static inline qm31 qm_mul(qm31 p, qm31 q) {
    /* (x1 + y1 u)(x2 + y2 u) = (x1 x2 + R y1 y2) + (x1 y2 + y1 x2) u
       (Karatsuba form in fields.cu:92-101; schoolbook here) */
    cm31 xx = cm_mul(p.x, q.x), yy = cm_mul(p.y, q.y);
    cm31 xy = cm_add(cm_mul(p.x, q.y), cm_mul(p.y, q.x));
    return (qm31){cm_add(xx, cm_mul(QM31_R, yy)), xy};
}
static inline qm31 qm_mul_cm(qm31 p, cm31 c) { return (qm31){cm_mul(p.x, c), cm_mul(p.y, c)}; }
static inline qm31 qm_scale(qm31 p, uint32_t s) { return (qm31){cm_scale(p.x, s), cm_scale(p.y, s)}; }
static inline qm31 qm_from_m31(uint32_t v) { return (qm31){{v, 0}, {0, 0}}; }
static qm31 qm_inv(qm31 p) {
    /* (x + y u)^-1 = (x - y u) / (x^2 - R y^2)   (fields.cu:138-144) */
    cm31 d = cm_sub(cm_mul(p.x, p.x), cm_mul(QM31_R, cm_mul(p.y, p.y)));
    cm31 di = cm_inv(d);
    return (qm31){cm_mul(p.x, di), cm_neg(cm_mul(p.y, di))};
}
static qm31 qm_pow(qm31 p, uint64_t e) {
    qm31 r = qm_from_m31(1);
    while (e) { if (e & 1) r = qm_mul(r, p); p = qm_mul(p, p); e >>= 1; }
    return r;
}
static inline qm31 qm_load(const uint32_t *w) { return (qm31){{w[0], w[1]}, {w[2], w[3]}}; }
static inline void qm_store(qm31 v, uint32_t *w) { w[0] = v.x.a; w[1] = v.x.b; w[2] = v.y.a; w[3] = v.y.b; }
static inline qm31 qm_load_soa(const uint32_t *const c[4], uint64_t i) {
    return (qm31){{c[0][i], c[1][i]}, {c[2][i], c[3][i]}};
}
static inline void qm_store_soa(qm31 v, uint32_t *const c[4], uint64_t i) {
    c[0][i] = v.x.a; c[1][i] = v.x.b; c[2][i] = v.y.a; c[3][i] = v.y.b;
}

void or_qm31_add(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) { qm_store(qm_add(qm_load(a), qm_load(b)), out); }
void or_qm31_sub(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) { qm_store(qm_sub(qm_load(a), qm_load(b)), out); }
void or_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) { qm_store(qm_mul(qm_load(a), qm_load(b)), out); }
void or_qm31_inv(const uint32_t a[4], uint32_t out[4]) { qm_store(qm_inv(qm_load(a)), out); }
void or_qm31_pow(const uint32_t a[4], uint64_t e, uint32_t out[4]) { qm_store(qm_pow(qm_load(a), e), out); }
void or_cm31_inv(const uint32_t a[2], uint32_t out[2]) { cm31 r = cm_inv((cm31){a[0], a[1]}); out[0] = r.a; out[1] = r.b; }

/* ------------------------------------------------------------------------- */
/* Circle group: point.cuh:7-65, generator (2, 1268011823) of order 2^31.      */
/* ------------------------------------------------------------------------- */
typedef struct { uint32_t x, y; } pt;
static inline pt pt_mul(pt p, pt q) {
    /* (x1, y1) * (x2, y2) = (x1 x2 - y1 y2, x1 y2 + y1 x2)   (point.cuh:25-30) */
    return (pt){or_m31_sub(or_m31_mul(p.x, q.x), or_m31_mul(p.y, q.y)),
                or_m31_add(or_m31_mul(p.x, q.y), or_m31_mul(p.y, q.x))};
}
static pt pt_at_index(uint32_t index) {
    pt r = {1, 0}, g = {2, 1268011823u};
    index &= 0x7fffffffu;
    while (index) { if (index & 1) r = pt_mul(r, g); g = pt_mul(g, g); index >>= 1; }
    return r;
}
void or_point_at_index(uint32_t index, uint32_t out[2]) { pt p = pt_at_index(index); out[0] = p.x; out[1] = p.y; }
void or_coset_at(uint32_t initial_index, uint32_t step_size, uint32_t i, uint32_t out[2]) {
    or_point_at_index((uint32_t)(((uint64_t)initial_index + (uint64_t)step_size * i) & 0x7fffffffu), out);
}
/* CircleDomain::index_at: quotients.cu:17-29 */
static uint32_t circle_domain_index(uint32_t hi, uint32_t hs, uint32_t half_log, uint64_t i) {
    uint64_t half = 1ull << half_log;
    if (i < half) return (uint32_t)((hi + (uint64_t)hs * i) & 0x7fffffffu);
    uint64_t g = (hi + (uint64_t)hs * (i - half)) & 0x7fffffffu;
    return (uint32_t)((0x80000000ull - g) & 0x7fffffffu);
}
void or_circle_domain_at(uint32_t hi, uint32_t hs, uint32_t half_log, uint64_t i, uint32_t out[2]) {
    or_point_at_index(circle_domain_index(hi, hs, half_log, i), out);
}

/* ------------------------------------------------------------------------- */
/* Bit reversal: cuda/src/bit_reverse.cu:6-33, cuda/src/utils.cu:5-8          */
/* ------------------------------------------------------------------------- */
static inline uint64_t brev(uint64_t i, uint32_t bits) {
    uint64_t r = 0;
    for (uint32_t b = 0; b < bits; b++) { r = (r << 1) | (i & 1); i >>= 1; }
    return r;
}
static uint32_t log2_exact(uint64_t n) { uint32_t l = 0; while ((1ull << l) < n) l++; return l; }
/* each pair (i, brev i) is swapped once, by its smaller index: pairs are disjoint */
void or_bit_reverse_u32(uint32_t *a, uint64_t n) {
    uint32_t bits = log2_exact(n);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        uint64_t i = (uint64_t)ii;
        uint64_t j = brev(i, bits);
        if (j > i) { uint32_t t = a[i]; a[i] = a[j]; a[j] = t; }
    }
}
void or_bit_reverse_qm31(uint32_t *a, uint64_t n) {
    uint32_t bits = log2_exact(n);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        uint64_t i = (uint64_t)ii;
        uint64_t j = brev(i, bits);
        if (j > i) for (int k = 0; k < 4; k++) { uint32_t t = a[4 * i + k]; a[4 * i + k] = a[4 * j + k]; a[4 * j + k] = t; }
    }
}
/* new_canonical_ordered: cuda/src/poly/twiddles.cu:4-27 (evens, then odds reversed, then bit reverse) */
void or_new_canonical_ordered(const uint32_t *in, uint32_t *out, uint64_t n) {
    uint64_t half = n >> 1;
    for (uint64_t i = 0; i < half; i++) out[i] = in[2 * i];
    for (uint64_t i = 0; i < half; i++) out[half + i] = in[n - 1 - 2 * i];
    or_bit_reverse_u32(out, n);
}

/* ------------------------------------------------------------------------- */
/* Twiddle tree: cuda/src/poly/twiddles.cu:29-73; itwiddles = batch inverse   */
/* (stwo_gpu_backend/src/poly.rs:170-192).  Level j holds x of the first half */
/* of the coset doubled j times, in bit-reversed order; a trailing 1 pads.     */
/* ------------------------------------------------------------------------- */
void or_precompute_twiddles(uint32_t initial_index, uint32_t step_size, uint32_t log_size,
                            uint32_t *tw, uint32_t *itw) {
    uint64_t off = 0;
    uint32_t ini = initial_index & 0x7fffffffu, st = step_size & 0x7fffffffu;
    for (uint32_t lvl_log = log_size; lvl_log >= 1; lvl_log--) {
        uint64_t half = 1ull << (lvl_log - 1);
        for (uint64_t i = 0; i < half; i++) {
            uint32_t idx = (uint32_t)((ini + (uint64_t)st * brev(i, lvl_log - 1)) & 0x7fffffffu);
            tw[off + i] = pt_at_index(idx).x;
        }
        off += half;
        ini = (ini * 2u) & 0x7fffffffu;
        st = (st * 2u) & 0x7fffffffu;
    }
    tw[off] = 1;
    if (itw) or_batch_inverse_m31(tw, itw, off + 1);
}

/* ------------------------------------------------------------------------- */
/* Circle FFT.  Butterflies: ifft.cu:16-39 (inverse), rfft.cu:21-41 (forward). */
/* Circle-layer twiddles from the first line layer: poly/utils.cu:3-14.        */
/* Line-layer twiddle slices taken from the end of the tree: ifft.cu:136-152.  */
/* ------------------------------------------------------------------------- */
static inline void ibfly(uint32_t *v0, uint32_t *v1, uint32_t t) {
    uint32_t a = *v0, b = *v1;
    *v0 = or_m31_add(a, b);
    *v1 = or_m31_mul(or_m31_sub(a, b), t);
}
static inline void fbfly(uint32_t *v0, uint32_t *v1, uint32_t t) {
    uint32_t a = *v0, tmp = or_m31_mul(*v1, t);
    *v0 = or_m31_add(a, tmp);
    *v1 = or_m31_sub(a, tmp);
}
/* get_twiddle(index): [y, -y, -x, x] pattern over pairs (x, y) of the first line level */
static inline uint32_t circle_twiddle(const uint32_t *line0, uint64_t h) {
    uint64_t k = h >> 2;
    switch (h & 3) {
        case 0: return line0[2 * k + 1];
        case 1: return or_m31_neg(line0[2 * k + 1]);
        case 2: return or_m31_neg(line0[2 * k]);
        default: return line0[2 * k];
    }
}

void or_interpolate(uint32_t *v, uint32_t half_initial_index, uint32_t half_log,
                    const uint32_t *itw, uint64_t tree_len) {
    uint32_t log_n = half_log + 1;
    uint64_t n = 1ull << log_n;
    if (log_n == 1) {
        /* stwo CpuBackend special case: one butterfly with y^-1, scale by 1/2 */
        pt p = pt_at_index(half_initial_index);
        ibfly(&v[0], &v[1], or_m31_inv(p.y));
        uint32_t ni = or_m31_inv(2);
        v[0] = or_m31_mul(v[0], ni); v[1] = or_m31_mul(v[1], ni);
        return;
    }
    if (log_n == 2) {
        pt p = pt_at_index(half_initial_index);
        uint32_t xi = or_m31_inv(p.x), yi = or_m31_inv(p.y), ni = or_m31_inv(4);
        ibfly(&v[0], &v[1], yi);
        ibfly(&v[2], &v[3], or_m31_neg(yi));
        ibfly(&v[0], &v[2], xi);
        ibfly(&v[1], &v[3], xi);
        for (int i = 0; i < 4; i++) v[i] = or_m31_mul(v[i], ni);
        return;
    }
    /* domain's tree part starts at tree_len - N/2 (ifft.cu:137) */
    const uint32_t *t = itw + (tree_len - (n >> 1));
    for (uint64_t h = 0; h < (n >> 1); h++) ibfly(&v[2 * h], &v[2 * h + 1], circle_twiddle(t, h));
    uint64_t lvl_off = 0, lvl_size = n >> 2;
    for (uint32_t layer = 1; layer < log_n; layer++) {
        uint64_t span = 1ull << layer;
        for (uint64_t h = 0; h < lvl_size; h++)
            for (uint64_t l = 0; l < span; l++) {
                uint64_t i0 = (h << (layer + 1)) + l;
                ibfly(&v[i0], &v[i0 + span], t[lvl_off + h]);
            }
        lvl_off += lvl_size;
        lvl_size >>= 1;
    }
    uint32_t scale = or_m31_inv((uint32_t)(n % P)); /* 1 / 2^log_n  (ifft.cu:157) */
    for (uint64_t i = 0; i < n; i++) v[i] = or_m31_mul(v[i], scale);
}

void or_evaluate(uint32_t *v, uint32_t half_initial_index, uint32_t half_log,
                 const uint32_t *tw, uint64_t tree_len) {
    uint32_t log_n = half_log + 1;
    uint64_t n = 1ull << log_n;
    if (log_n == 1) {
        pt p = pt_at_index(half_initial_index);
        fbfly(&v[0], &v[1], p.y);
        return;
    }
    if (log_n == 2) {
        pt p = pt_at_index(half_initial_index);
        fbfly(&v[0], &v[2], p.x);
        fbfly(&v[1], &v[3], p.x);
        fbfly(&v[0], &v[1], p.y);
        fbfly(&v[2], &v[3], or_m31_neg(p.y));
        return;
    }
    const uint32_t *t = tw + (tree_len - (n >> 1));
    /* line layers from the top (rfft.cu:50-60): layer log_n-1 uses the 1-entry level */
    for (uint32_t layer = log_n - 1; layer >= 1; layer--) {
        uint64_t lvl_size = n >> (layer + 1);
        uint64_t lvl_off = (n >> 1) - 2 * lvl_size; /* levels sized N/4, N/8, ... from 0 */
        uint64_t span = 1ull << layer;
        for (uint64_t h = 0; h < lvl_size; h++)
            for (uint64_t l = 0; l < span; l++) {
                uint64_t i0 = (h << (layer + 1)) + l;
                fbfly(&v[i0], &v[i0 + span], t[lvl_off + h]);
            }
    }
    for (uint64_t h = 0; h < (n >> 1); h++) fbfly(&v[2 * h], &v[2 * h + 1], circle_twiddle(t, h));
}

/* CanonicCoset(log).half_coset initial index = 2^(30 - log) (stwo Coset::half_odds) */
static uint32_t canonic_half_initial(uint32_t log_n) { return 1u << (30 - log_n); }

void or_lde_columns(const uint32_t *const *evals, uint32_t *const *coeffs_out,
                    uint32_t *const *lde_out, uint32_t n_cols, uint32_t log_n,
                    uint32_t log_blowup, const uint32_t *tw, const uint32_t *itw,
                    uint64_t tree_len) {
    uint64_t n = 1ull << log_n, m = n << log_blowup;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < (int64_t)n_cols; c++) {
        uint32_t *coef = coeffs_out[c];
        memcpy(coef, evals[c], n * 4);
        or_interpolate(coef, canonic_half_initial(log_n), log_n - 1, itw, tree_len);
        uint32_t *out = lde_out[c];
        memcpy(out, coef, n * 4); /* extend = zero pad (stwo_gpu_backend/src/poly.rs:96-106) */
        memset(out + n, 0, (m - n) * 4);
        or_evaluate(out, canonic_half_initial(log_n + log_blowup), log_n + log_blowup - 1, tw, tree_len);
    }
}

/* ------------------------------------------------------------------------- */
/* eval_at_point: mappings [pi^(n-2)(x), ..., pi(x), x, y], pi(x) = 2x^2 - 1, */
/* then fold(v, f) = fold(lhs, f[1:]) + fold(rhs, f[1:]) * f[0]                */
/* (cuda/src/poly/eval_at_point.cu:99-110 and its first/second passes).        */
/* ------------------------------------------------------------------------- */
void or_eval_at_point(const uint32_t *coeffs, uint32_t log_n, const uint32_t px[4],
                      const uint32_t py[4], uint32_t out[4]) {
    if (log_n == 0) { qm_store(qm_from_m31(coeffs[0]), out); return; }
    qm31 *f = (qm31 *)malloc(sizeof(qm31) * log_n);
    f[log_n - 1] = qm_load(py);
    qm31 x = qm_load(px);
    for (uint32_t i = 2; i <= log_n; i++) {
        f[log_n - i] = x;
        x = qm_sub(qm_scale(qm_mul(x, x), 2), qm_from_m31(1));
    }
    uint64_t n = 1ull << log_n;
    qm31 *lvl = (qm31 *)malloc(sizeof(qm31) * (n >> 1));
    /* bottom level: pairs split by the last factor (y) */
    for (uint64_t i = 0; i < (n >> 1); i++)
        lvl[i] = qm_add(qm_from_m31(coeffs[2 * i]), qm_scale(f[log_n - 1], coeffs[2 * i + 1]));
    for (int32_t j = (int32_t)log_n - 2; j >= 0; j--) {
        uint64_t cnt = 1ull << j;
        for (uint64_t i = 0; i < cnt; i++) lvl[i] = qm_add(lvl[2 * i], qm_mul(lvl[2 * i + 1], f[j]));
    }
    qm_store(lvl[0], out);
    free(lvl);
    free(f);
}

/* ------------------------------------------------------------------------- */
/* Batch inverse: Montgomery's trick (cuda/src/batch_inverse.cu:24-125).      */
/* Inputs must be non-zero; the output is the exact inverse element-wise.    */
/* ------------------------------------------------------------------------- */
/* Independent chunks of OR_BINV_CHUNK elements run on separate threads: the
   inverse of a non-zero element is unique, so the chunking does not change the
   result. */
#define OR_BINV_CHUNK 65536
void or_batch_inverse_m31(const uint32_t *src, uint32_t *dst, uint64_t n) {
    if (!n) return;
    int64_t n_chunks = (int64_t)((n + OR_BINV_CHUNK - 1) / OR_BINV_CHUNK);
#pragma omp parallel
    {
        uint32_t *pre = (uint32_t *)malloc(OR_BINV_CHUNK * 4);
#pragma omp for schedule(static)
        for (int64_t ch = 0; ch < n_chunks; ch++) {
            uint64_t lo = (uint64_t)ch * OR_BINV_CHUNK, hi = lo + OR_BINV_CHUNK < n ? lo + OR_BINV_CHUNK : n;
            uint32_t acc = 1;
            for (uint64_t i = lo; i < hi; i++) { pre[i - lo] = acc; acc = or_m31_mul(acc, src[i]); }
            uint32_t inv = or_m31_inv(acc);
            for (uint64_t i = hi; i-- > lo;) {
                uint32_t s = src[i];
                dst[i] = or_m31_mul(inv, pre[i - lo]);
                inv = or_m31_mul(inv, s);
            }
        }
        free(pre);
    }
}
void or_batch_inverse_qm31(const uint32_t *src, uint32_t *dst, uint64_t n) {
    if (!n) return;
    int64_t n_chunks = (int64_t)((n + OR_BINV_CHUNK - 1) / OR_BINV_CHUNK);
#pragma omp parallel
    {
        qm31 *pre = (qm31 *)malloc(OR_BINV_CHUNK * sizeof(qm31));
#pragma omp for schedule(static)
        for (int64_t ch = 0; ch < n_chunks; ch++) {
            uint64_t lo = (uint64_t)ch * OR_BINV_CHUNK, hi = lo + OR_BINV_CHUNK < n ? lo + OR_BINV_CHUNK : n;
            qm31 acc = qm_from_m31(1);
            for (uint64_t i = lo; i < hi; i++) { pre[i - lo] = acc; acc = qm_mul(acc, qm_load(src + 4 * i)); }
            qm31 inv = qm_inv(acc);
            for (uint64_t i = hi; i-- > lo;) {
                qm31 s = qm_load(src + 4 * i);
                qm_store(qm_mul(inv, pre[i - lo]), dst + 4 * i);
                inv = qm_mul(inv, s);
            }
        }
        free(pre);
    }
}

/* ------------------------------------------------------------------------- */
/* Blake2s (RFC 7693).  The Merkle hasher uses a zero chaining value and zero */
/* counter/flags (cuda/src/blake2s.cu:56-151); blake2s256 is the standard     */
/* 32-byte digest used by the Fiat-Shamir channel.                            */
/* ------------------------------------------------------------------------- */
static const uint32_t B2S_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                   0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t B2S_SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},
    {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4},
    {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13},
    {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11},
    {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5},
    {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0}};
static inline uint32_t rotr32(uint32_t x, int c) { return (x >> c) | (x << (32 - c)); }
#define B2S_G(a, b, c, d, x, y)            \
    do {                                   \
        a = a + b + x; d = rotr32(d ^ a, 16); \
        c = c + d;     b = rotr32(b ^ c, 12); \
        a = a + b + y; d = rotr32(d ^ a, 8);  \
        c = c + d;     b = rotr32(b ^ c, 7);  \
    } while (0)
void or_blake2s_compress(uint32_t h[8], const uint32_t m[16], uint32_t t0, uint32_t t1,
                         uint32_t f0, uint32_t f1) {
    uint32_t v[16];
    for (int i = 0; i < 8; i++) { v[i] = h[i]; v[8 + i] = B2S_IV[i]; }
    v[12] ^= t0; v[13] ^= t1; v[14] ^= f0; v[15] ^= f1;
    for (int r = 0; r < 10; r++) {
        const uint8_t *s = B2S_SIGMA[r];
        B2S_G(v[0], v[4], v[8], v[12], m[s[0]], m[s[1]]);
        B2S_G(v[1], v[5], v[9], v[13], m[s[2]], m[s[3]]);
        B2S_G(v[2], v[6], v[10], v[14], m[s[4]], m[s[5]]);
        B2S_G(v[3], v[7], v[11], v[15], m[s[6]], m[s[7]]);
        B2S_G(v[0], v[5], v[10], v[15], m[s[8]], m[s[9]]);
        B2S_G(v[1], v[6], v[11], v[12], m[s[10]], m[s[11]]);
        B2S_G(v[2], v[7], v[8], v[13], m[s[12]], m[s[13]]);
        B2S_G(v[3], v[4], v[9], v[14], m[s[14]], m[s[15]]);
    }
    for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[8 + i];
}
void or_blake2s256(const uint8_t *data, uint64_t len, uint8_t out[32]) {
    uint32_t h[8];
    for (int i = 0; i < 8; i++) h[i] = B2S_IV[i];
    h[0] ^= 0x01010000u ^ 32u; /* digest length 32, no key, fanout = depth = 1 */
    uint64_t off = 0;
    uint32_t m[16];
    while (len - off > 64) {
        memcpy(m, data + off, 64);
        off += 64;
        or_blake2s_compress(h, m, (uint32_t)off, (uint32_t)(off >> 32), 0, 0);
    }
    uint8_t block[64] = {0};
    memcpy(block, data + off, len - off);
    memcpy(m, block, 64);
    or_blake2s_compress(h, m, (uint32_t)len, (uint32_t)(len >> 32), 0xffffffffu, 0);
    memcpy(out, h, 32);
}
/* hash of one Merkle node: blake2s.cu:97-151 (16-column chunks, zero-padded
   remainder, no compression when there are no columns and no children). */
static void merkle_hash_row(const uint32_t *prev, uint32_t n_cols, const uint32_t *const *cols,
                            uint64_t row, uint32_t h[8]) {
    memset(h, 0, 32);
    uint32_t m[16];
    if (prev) {
        memcpy(m, prev + 16 * row, 64); /* left child || right child */
        or_blake2s_compress(h, m, 0, 0, 0, 0);
    }
    for (uint32_t c = 0; c < n_cols; c += 16) {
        memset(m, 0, 64);
        for (uint32_t j = 0; j < 16 && c + j < n_cols; j++) m[j] = cols[c + j][row];
        or_blake2s_compress(h, m, 0, 0, 0, 0);
    }
}
void or_merkle_commit_on_layer(uint32_t log_size, const uint32_t *prev, uint32_t n_cols,
                               const uint32_t *const *cols, uint32_t *out) {
    int64_t n = 1ll << log_size;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; r++) merkle_hash_row(prev, n_cols, cols, (uint64_t)r, out + 8 * r);
}

/* ------------------------------------------------------------------------- */
/* FRI folds.  fold_line: out[i] = (f(x) + f(-x)) + alpha (f(x) - f(-x)) / x, */
/* x = domain.at(brev(2i)) (cuda/src/fri/fold_line.cu:15-32).                 */
/* fold_circle_into_line: dst[i] = dst[i] alpha^2 + f0 + alpha f1 with        */
/* f1 = (f(p) - f(-p)) / p.y (cuda/src/fri/fold_circle_into_line.cu:19-44).   */
/* Domain points are computed from the coset directly, not from the tree.    */
/* ------------------------------------------------------------------------- */
void or_fold_line(const uint32_t *const src[4], uint64_t n, uint32_t li, uint32_t ls,
                  const uint32_t alpha_w[4], uint32_t *const dst[4]) {
    uint32_t log_n = log2_exact(n);
    qm31 alpha = qm_load(alpha_w);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)(n >> 1); i++) {
        uint32_t idx = (uint32_t)((li + (uint64_t)ls * brev((uint64_t)i << 1, log_n)) & 0x7fffffffu);
        uint32_t xi = or_m31_inv(pt_at_index(idx).x);
        qm31 a = qm_load_soa(src, 2 * i), b = qm_load_soa(src, 2 * i + 1);
        qm31 f0 = qm_add(a, b), f1 = qm_scale(qm_sub(a, b), xi);
        qm_store_soa(qm_add(f0, qm_mul(alpha, f1)), dst, i);
    }
}
void or_fold_circle_into_line(const uint32_t *const src[4], uint64_t n, uint32_t hi,
                              uint32_t hs, const uint32_t alpha_w[4], uint32_t *const dst[4]) {
    uint32_t log_n = log2_exact(n);
    qm31 alpha = qm_load(alpha_w), alpha2 = qm_mul(alpha, alpha);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)(n >> 1); i++) {
        uint32_t idx = circle_domain_index(hi, hs, log_n - 1, brev((uint64_t)i << 1, log_n));
        uint32_t yi = or_m31_inv(pt_at_index(idx).y);
        qm31 a = qm_load_soa(src, 2 * i), b = qm_load_soa(src, 2 * i + 1);
        qm31 f0 = qm_add(a, b), f1 = qm_scale(qm_sub(a, b), yi);
        qm31 fp = qm_add(f0, qm_mul(alpha, f1));
        qm31 d = qm_load_soa((const uint32_t *const *)dst, i);
        qm_store_soa(qm_add(qm_mul(d, alpha2), fp), dst, i);
    }
}

/* ------------------------------------------------------------------------- */
/* DEEP quotients (cuda/src/quotients.cu:31-178 with the cumulative line-     */
/* coefficient offset of quotients.cu:147-168).                               */
/* ------------------------------------------------------------------------- */
void or_accumulate_quotients(uint32_t hi, uint32_t hs, uint32_t log_size,
                             const uint32_t *const *cols, const uint32_t rc_w[4],
                             uint32_t n_batches, const uint32_t *points, const uint32_t *sizes,
                             const uint32_t *col_idx, const uint32_t *values,
                             uint32_t *const out[4]) {
    qm31 rc = qm_load(rc_w);
    uint32_t total = 0;
    for (uint32_t b = 0; b < n_batches; b++) total += sizes[b];
    /* line coefficients (a, b, c) * alpha^(j+1) per sample; batch coeff rc^|batch|
       (quotients.cu:49-88) */
    qm31 *lc = (qm31 *)malloc(sizeof(qm31) * 3 * (total ? total : 1));
    qm31 *bc = (qm31 *)malloc(sizeof(qm31) * (n_batches ? n_batches : 1));
    uint32_t off = 0;
    for (uint32_t b = 0; b < n_batches; b++) {
        qm31 py = qm_load(points + 8 * b + 4);
        qm31 py_conj = {py.x, cm_neg(py.y)};
        bc[b] = qm_pow(rc, sizes[b] + (g_mutation == OR_MUT_QUOT_BATCH_POW ? 1 : 0));
        qm31 alpha = qm_from_m31(1);
        for (uint32_t j = 0; j < sizes[b]; j++) {
            qm31 v = qm_load(values + 4 * (off + j));
            if (g_mutation != OR_MUT_QUOT_ALPHA_POW) alpha = qm_mul(alpha, rc);
            qm31 v_conj = {v.x, cm_neg(v.y)};
            qm31 a = qm_sub(v_conj, v);
            qm31 c = qm_sub(py_conj, py);
            qm31 bb = qm_sub(qm_mul(v, c), qm_mul(a, py));
            lc[3 * (off + j) + 0] = qm_mul(alpha, a);
            lc[3 * (off + j) + 1] = qm_mul(alpha, bb);
            lc[3 * (off + j) + 2] = qm_mul(alpha, c);
            if (g_mutation == OR_MUT_QUOT_ALPHA_POW) alpha = qm_mul(alpha, rc);
        }
        off += sizes[b];
    }
    if (g_mutation == OR_MUT_QUOT_REF_OFFSET) {
        /* the per-batch offset b * |b| * 3 of quotients.cu:74 (with the
           cumulative read offset), which is wrong when batch sizes differ */
        qm31 *lc2 = (qm31 *)malloc(sizeof(qm31) * 3 * (total ? total : 1));
        memcpy(lc2, lc, sizeof(qm31) * 3 * (total ? total : 1));
        for (uint32_t b = 0, o = 0; b < n_batches; o += sizes[b], b++)
            for (uint32_t k = 0; k < 3 * sizes[b]; k++)
                if (3 * b * sizes[b] + k < 3 * total) lc2[3 * b * sizes[b] + k] = lc[3 * o + k];
        free(lc);
        lc = lc2;
    }
    int64_t n = 1ll << log_size;
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < n; row++) {
        pt dp = pt_at_index(circle_domain_index(hi, hs, log_size - 1, brev((uint64_t)row, log_size)));
        qm31 acc = qm_from_m31(0);
        uint32_t o = 0;
        if (g_mutation == OR_MUT_QUOT_BATCH_ORDER) o = total;
        for (uint32_t bi = 0; bi < n_batches; bi++) {
            uint32_t b = g_mutation == OR_MUT_QUOT_BATCH_ORDER ? n_batches - 1 - bi : bi;
            if (g_mutation == OR_MUT_QUOT_BATCH_ORDER) o -= sizes[b];
            /* denominator (Pr.x - p.x) Pi.y - (Pr.y - p.y) Pi.x in CM31 (quotients.cu:91-109) */
            cm31 prx = {points[8 * b + 0], points[8 * b + 1]}, pix = {points[8 * b + 2], points[8 * b + 3]};
            cm31 pry = {points[8 * b + 4], points[8 * b + 5]}, piy = {points[8 * b + 6], points[8 * b + 7]};
            cm31 d = cm_sub(cm_mul((cm31){or_m31_sub(prx.a, dp.x), prx.b}, piy),
                            cm_mul((cm31){or_m31_sub(pry.a, dp.y), pry.b}, pix));
            cm31 dinv = cm_inv(d);
            qm31 num = qm_from_m31(0);
            for (uint32_t j = 0; j < sizes[b]; j++) {
                qm31 a = lc[3 * (o + j)], bb = lc[3 * (o + j) + 1], c = lc[3 * (o + j) + 2];
                uint32_t cv = cols[col_idx[o + j]][row];
                num = qm_add(num, qm_sub(qm_scale(c, cv), qm_add(qm_scale(a, dp.y), bb)));
            }
            acc = qm_add(qm_mul(acc, bc[b]), qm_mul_cm(num, dinv));
            if (g_mutation != OR_MUT_QUOT_BATCH_ORDER) o += sizes[b];
        }
        qm_store_soa(acc, out, (uint64_t)row);
    }
    free(lc);
    free(bc);
}

/* accumulate.cu:4-13 */
void or_accumulate(uint64_t n, uint32_t *const left[4], const uint32_t *const right[4]) {
    for (int k = 0; k < 4; k++)
        for (uint64_t i = 0; i < n; i++) left[k][i] = or_m31_add(left[k][i], right[k][i]);
}

/* examples/wide_fibonacci.cu:3-38: sum_i (t_i^2 + t_{i+1}^2 - t_{i+2}) * coeff[n-3-i],
   times the M31 vanishing-polynomial inverse of the row. */
void or_wide_fib_constraint_quotients(uint64_t n_rows, uint32_t n_cols,
                                      const uint32_t *const *trace,
                                      const uint32_t *coeff_powers,
                                      const uint32_t *denom_inv, uint32_t *const out[4]) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)n_rows; r++) {
        qm31 num = qm_from_m31(0);
        for (uint32_t i = 0; i + 2 < n_cols; i++) {
            uint32_t a = trace[i][r], b = trace[i + 1][r], c = trace[i + 2][r];
            uint32_t cst = or_m31_sub(or_m31_add(or_m31_mul(a, a), or_m31_mul(b, b)), c);
            uint32_t k = g_mutation == OR_MUT_WF_POWER_ORDER ? i : n_cols - 3 - i;
            num = qm_add(num, qm_scale(qm_load(coeff_powers + 4 * k), cst));
        }
        qm_store_soa(qm_scale(num, denom_inv[r]), out, (uint64_t)r);
    }
}

/* ------------------------------------------------------------------------- */
/* Multilinear ops of the lookups (cuda/src/gkr.cu:4-38, cuda/src/mle.cu:4-30) */
/* ------------------------------------------------------------------------- */
/* evals[x] = v * prod_k (y_k if bit (n-1-k) of x is set else 1 - y_k): the
   reference kernel walks the index from its low bit against y_{n-1} .. y_0
   (gkr.cu:11-15) with factors (1 - y_k, y_k) (gkr.cu:25-28). */
void or_gen_eq_evals(const uint32_t v[4], const uint32_t *y, uint32_t n, uint32_t *evals) {
    for (uint64_t x = 0; x < (1ull << n); x++) {
        qm31 e = qm_load(v);
        uint64_t t = x;
        for (int k = (int)n - 1; k >= 0; k--) {
            qm31 yk = qm_load(y + 4 * k);
            e = qm_mul(e, (t & 1) ? yk : qm_sub(qm_from_m31(1), yk));
            t >>= 1;
        }
        qm_store(e, evals + 4 * x);
    }
}
/* out[i] = lhs + (rhs - lhs) * a, lhs = evals[i], rhs = evals[i + n/2] (mle.cu:4-12);
   M31 evaluations in, QM31 (AoS) out */
void or_fix_first_variable_base(const uint32_t *evals, uint64_t n, const uint32_t a[4], uint32_t *out) {
    for (uint64_t i = 0; i < n / 2; i++) {
        uint32_t l = evals[i], r = evals[i + n / 2];
        qm31 t = qm_scale(qm_load(a), or_m31_sub(r, l));
        t.x.a = or_m31_add(t.x.a, l);
        qm_store(t, out + 4 * i);
    }
}
void or_fix_first_variable_secure(const uint32_t *evals, uint64_t n, const uint32_t a[4], uint32_t *out) {
    for (uint64_t i = 0; i < n / 2; i++) {
        qm31 l = qm_load(evals + 4 * i), r = qm_load(evals + 4 * (i + n / 2));
        qm_store(qm_add(l, qm_mul(qm_sub(r, l), qm_load(a))), out + 4 * i);
    }
}

/* ------------------------------------------------------------------------- */
/* Synthetic inputs: value i of stream `seed` is the top 31 bits of the       */
/* SplitMix64 output mix(seed + (i+1) * golden); the value P is rejected by   */
/* falling back to the low 31 bits (then to 0).                               */
/* ------------------------------------------------------------------------- */
uint32_t or_uniform_m31(uint64_t seed, uint64_t index) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (index + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    uint32_t v = (uint32_t)(z >> 33);
    if (v == P) v = (uint32_t)(z & P);
    if (v == P) v = 0;
    return v;
}
void or_fill_uniform_m31(uint32_t *dst, uint64_t n, uint64_t seed, uint64_t offset) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) dst[i] = or_uniform_m31(seed, offset + (uint64_t)i);
}
