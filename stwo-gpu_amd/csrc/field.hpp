// Device arithmetic over M31 = Z / (2^31 - 1), CM31 = M31[i] / (i^2 + 1) and
// QM31 = CM31[u] / (u^2 - (2 + i)), plus circle-group points.
//
// Semantics follow cuda/src/fields.cu:3-160 and cuda/include/point.cuh:17-65 of
// the reference, with one deliberate difference: every result is canonical
// (in [0, P)); the reference's neg(0) returns P (fields.cu:19-21).
//
// Cost model on gfx950 (measured): v_mad_u64_u32 issues at close to full rate,
// so a product mod P is mad_u64 + and + alignbit + add, and canonicalisation is
// the unsigned-min trick min(r, r - P) (2 instructions).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sh {

constexpr uint32_t P = 0x7fffffffu;

__host__ __device__ __forceinline__ uint32_t canon(uint32_t r) {  // r in [0, 2P] -> [0, P)
  uint32_t t = r - P;
  return r < t ? r : t;
}
// p < 2^62 (product of two values below 2^31): 2^31 = 1 (mod P)
__host__ __device__ __forceinline__ uint32_t reduce62(uint64_t p) {
  return canon(((uint32_t)p & P) + (uint32_t)(p >> 31));
}
// p < 2^63: one extra fold (sums of two products)
__host__ __device__ __forceinline__ uint32_t reduce63(uint64_t p) {
  uint64_t r = (p & P) + (p >> 31);  // < 2^32 + 2^31
  return canon((uint32_t)((r & P) + (r >> 31)));
}
__host__ __device__ __forceinline__ uint32_t add(uint32_t a, uint32_t b) { return canon(a + b); }
__host__ __device__ __forceinline__ uint32_t sub(uint32_t a, uint32_t b) { return canon(a + P - b); }
__host__ __device__ __forceinline__ uint32_t neg(uint32_t a) { return canon(P - a); }
__host__ __device__ __forceinline__ uint32_t mul(uint32_t a, uint32_t b) { return reduce62((uint64_t)a * b); }
__host__ __device__ __forceinline__ uint32_t sqr(uint32_t a) { return mul(a, a); }
// x t for x < P and a doubled factor t2 = 2t (t <= P): 2xt = hi 2^32 + lo with lo even,
// so xt = hi 2^31 + lo / 2 = hi + lo / 2 (mod P); 2xt <= 2 (P - 1) P gives hi <= 2^31 - 3,
// so hi + lo / 2 < 2P and one conditional subtraction makes it canonical.  One
// instruction shorter than mul (a shift replaces the mask and the 64-bit funnel shift).
__host__ __device__ __forceinline__ uint32_t mul_dbl(uint32_t x, uint32_t t2) {
  const uint64_t q = (uint64_t)x * t2;
  return canon((uint32_t)(q >> 32) + ((uint32_t)q >> 1));
}

__host__ __device__ __forceinline__ uint32_t pow2k(uint32_t t, int k) {
  for (int i = 0; i < k; i++) t = sqr(t);
  return t;
}
// x^(P-2) via the chain used in fields.cu:32-40
__host__ __device__ __forceinline__ uint32_t inv(uint32_t t) {
  uint32_t t0 = mul(pow2k(t, 2), t);
  uint32_t t1 = mul(pow2k(t0, 1), t0);
  uint32_t t2 = mul(pow2k(t1, 3), t0);
  uint32_t t3 = mul(pow2k(t2, 1), t0);
  uint32_t t4 = mul(pow2k(t3, 8), t3);
  uint32_t t5 = mul(pow2k(t4, 8), t3);
  return mul(pow2k(t5, 7), t2);
}

struct cm31 {
  uint32_t a, b;
};
struct qm31 {
  cm31 x, y;  // x + y u
};

__host__ __device__ __forceinline__ cm31 cadd(cm31 p, cm31 q) { return {add(p.a, q.a), add(p.b, q.b)}; }
__host__ __device__ __forceinline__ cm31 csub(cm31 p, cm31 q) { return {sub(p.a, q.a), sub(p.b, q.b)}; }
__host__ __device__ __forceinline__ cm31 cneg(cm31 p) { return {neg(p.a), neg(p.b)}; }
__host__ __device__ __forceinline__ cm31 cmul(cm31 p, cm31 q) {
  // (ac - bd) + (ad + bc) i; -bd = b (P - d)
  uint64_t re = (uint64_t)p.a * q.a + (uint64_t)p.b * (P - q.b);
  uint64_t im = (uint64_t)p.a * q.b + (uint64_t)p.b * q.a;
  return {reduce63(re), reduce63(im)};
}
__host__ __device__ __forceinline__ cm31 cscale(cm31 p, uint32_t s) { return {mul(p.a, s), mul(p.b, s)}; }
__host__ __device__ __forceinline__ cm31 cinv(cm31 p) {
  uint32_t f = inv(reduce63((uint64_t)p.a * p.a + (uint64_t)p.b * p.b));
  return {mul(p.a, f), mul(neg(p.b), f)};
}
// (2 + i)(a + b i) = (2a - b) + (a + 2b) i
__host__ __device__ __forceinline__ cm31 cmul_r(cm31 p) {
  return {sub(add(p.a, p.a), p.b), add(p.a, add(p.b, p.b))};
}

__host__ __device__ __forceinline__ qm31 qzero() { return {{0, 0}, {0, 0}}; }
__host__ __device__ __forceinline__ qm31 qone() { return {{1, 0}, {0, 0}}; }
__host__ __device__ __forceinline__ qm31 qadd(qm31 p, qm31 q) { return {cadd(p.x, q.x), cadd(p.y, q.y)}; }
__host__ __device__ __forceinline__ qm31 qsub(qm31 p, qm31 q) { return {csub(p.x, q.x), csub(p.y, q.y)}; }
__host__ __device__ __forceinline__ qm31 qneg(qm31 p) { return {cneg(p.x), cneg(p.y)}; }
__host__ __device__ __forceinline__ qm31 qmul(qm31 p, qm31 q) {
  // Karatsuba over CM31 (fields.cu:92-101): v0 = x1 x2, v1 = y1 y2,
  // x = v0 + R v1, y = (x1 + y1)(x2 + y2) - v0 - v1
  cm31 v0 = cmul(p.x, q.x), v1 = cmul(p.y, q.y);
  cm31 v2 = cmul(cadd(p.x, p.y), cadd(q.x, q.y));
  return {cadd(v0, cmul_r(v1)), csub(v2, cadd(v0, v1))};
}
__host__ __device__ __forceinline__ qm31 qmul_cm(qm31 p, cm31 c) { return {cmul(p.x, c), cmul(p.y, c)}; }
__host__ __device__ __forceinline__ qm31 qscale(qm31 p, uint32_t s) { return {cscale(p.x, s), cscale(p.y, s)}; }
__host__ __device__ __forceinline__ qm31 qsqr(qm31 p) { return qmul(p, p); }
__host__ __device__ __forceinline__ qm31 qinv(qm31 p) {
  // (x + y u)^-1 = (x - y u) / (x^2 - R y^2)   (fields.cu:138-144)
  cm31 d = csub(cmul(p.x, p.x), cmul_r(cmul(p.y, p.y)));
  cm31 di = cinv(d);
  return {cmul(p.x, di), cneg(cmul(p.y, di))};
}
__host__ __device__ __forceinline__ qm31 qpow(qm31 p, uint64_t e) {
  qm31 r = qone();
  while (e) {
    if (e & 1) r = qmul(r, p);
    p = qsqr(p);
    e >>= 1;
  }
  return r;
}
// ---- products by a fixed QM31 factor (FRI folds: one alpha per layer) ------------------
// alpha z = alpha_x z_x + (2 + i) alpha_y z_y + (alpha_x z_y + alpha_y z_x) u is linear in
// z = (z_x.a, z_x.b, z_y.a, z_y.b): a 4 x 4 matrix over M31 with entries in [0, P), so each
// output coordinate is four 64-bit multiply-adds and one reduction (a product by qmul costs
// nine reductions and the Karatsuba additions).  With alpha = (a0 + a1 i) + (a2 + a3 i) u and
// (2 + i)(a2 + a3 i) = b0 + b1 i the rows are
//   [a0, -a1, b0, -b1], [a1, a0, b1, b0], [a2, -a3, a0, -a1], [a3, a2, a1, a0]
// and only their nine distinct entries are kept (uniform values live in scalar registers).
struct qmat {
  uint32_t a0, a1, na1, a2, a3, na3, b0, b1, nb1;
};
__host__ __device__ __forceinline__ qmat qmat_of(qm31 a) {
  const cm31 b = cmul_r(a.y);
  return {a.x.a, a.x.b, neg(a.x.b), a.y.a, a.y.b, neg(a.y.b), b.a, b.b, neg(b.b)};
}
// y < 2^64 reduced mod P: 2^32 = 2 (mod P), so y = 2 hi + lo (mod P), below 3 2^32; one more
// fold of that gives at most 2^31 + 4 <= 2P
__host__ __device__ __forceinline__ uint32_t reduce64(uint64_t y) {
  const uint64_t z = (uint64_t)(uint32_t)(y >> 32) * 2u + (uint32_t)y;
  const uint32_t lo = (uint32_t)z, hi = (uint32_t)(z >> 32);
  return canon((lo & P) + (lo >> 31) + 2 * hi);
}
// a - b for canonical a, b: a - b wraps above 2^31 exactly when a < b, and then a - b + P
// is the smaller of the two
__host__ __device__ __forceinline__ uint32_t sub3(uint32_t a, uint32_t b) {
  const uint32_t x = a - b, y = x + P;
  return x < y ? x : y;
}
__host__ __device__ __forceinline__ uint32_t dot4(uint32_t m0, uint32_t m1, uint32_t m2, uint32_t m3,
                                                  const uint32_t e[4], uint32_t c) {
  // c < 2^32 plus four products of values below P: at most 4 (P - 1)^2 + 2^32 < 2^64
  return reduce64((uint64_t)c + (uint64_t)m0 * e[0] + (uint64_t)m1 * e[1] + (uint64_t)m2 * e[2] +
                  (uint64_t)m3 * e[3]);
}
// alpha e + c coordinatewise (c[k] < 2^32), reduced
__host__ __device__ __forceinline__ qm31 qmat_apply(const qmat& M, const uint32_t e[4], const uint32_t c[4]) {
  return {{dot4(M.a0, M.na1, M.b0, M.nb1, e, c[0]), dot4(M.a1, M.a0, M.b1, M.b0, e, c[1])},
          {dot4(M.a2, M.na3, M.a0, M.na1, e, c[2]), dot4(M.a3, M.a2, M.a1, M.a0, e, c[3])}};
}
// FRI fold of a pair: (a + b) + alpha t (a - b), alpha given by its matrix (t canonical)
__host__ __device__ __forceinline__ qm31 fold_mat(qm31 a, qm31 b, uint32_t t, const qmat& M) {
  const uint32_t t2 = t + t;
  const uint32_t e[4] = {mul_dbl(sub3(a.x.a, b.x.a), t2), mul_dbl(sub3(a.x.b, b.x.b), t2),
                         mul_dbl(sub3(a.y.a, b.y.a), t2), mul_dbl(sub3(a.y.b, b.y.b), t2)};
  const uint32_t c[4] = {a.x.a + b.x.a, a.x.b + b.x.b, a.y.a + b.y.a, a.y.b + b.y.b};
  return qmat_apply(M, e, c);
}
// the circle fold accumulated into acc: acc alpha^2 + (a + b) + alpha t (a - b), M2 the
// matrix of alpha^2 (the acc alpha^2 + a + b part reduced first, so each sum stays below 2^64)
__host__ __device__ __forceinline__ qm31 fold_mat_acc(qm31 a, qm31 b, uint32_t t, const qmat& M, qm31 acc,
                                                     const qmat& M2) {
  const uint32_t ac[4] = {acc.x.a, acc.x.b, acc.y.a, acc.y.b};
  const uint32_t c[4] = {a.x.a + b.x.a, a.x.b + b.x.b, a.y.a + b.y.a, a.y.b + b.y.b};
  const qm31 s = qmat_apply(M2, ac, c);
  const uint32_t t2 = t + t;
  const uint32_t e[4] = {mul_dbl(sub3(a.x.a, b.x.a), t2), mul_dbl(sub3(a.x.b, b.x.b), t2),
                         mul_dbl(sub3(a.y.a, b.y.a), t2), mul_dbl(sub3(a.y.b, b.y.b), t2)};
  const uint32_t c2[4] = {s.x.a, s.x.b, s.y.a, s.y.b};
  return qmat_apply(M, e, c2);
}

__host__ __device__ __forceinline__ bool qis_zero(qm31 p) { return !(p.x.a | p.x.b | p.y.a | p.y.b); }
__host__ __device__ __forceinline__ bool qeq(qm31 p, qm31 q) {
  return p.x.a == q.x.a && p.x.b == q.x.b && p.y.a == q.y.a && p.y.b == q.y.b;
}

// ---- circle group: generator (2, 1268011823) of order 2^31 (point.cuh:17) ----
struct pt {
  uint32_t x, y;
};
__host__ __device__ __forceinline__ pt pmul(pt p, pt q) {
  uint64_t x = (uint64_t)p.x * q.x + (uint64_t)p.y * (P - q.y);
  uint64_t y = (uint64_t)p.x * q.y + (uint64_t)p.y * q.x;
  return {reduce63(x), reduce63(y)};
}
__host__ __device__ __forceinline__ pt point_at_index(uint32_t index) {
  pt r = {1, 0}, g = {2, 1268011823u};
  index &= 0x7fffffffu;
  while (index) {
    if (index & 1) r = pmul(r, g);
    g = pmul(g, g);
    index >>= 1;
  }
  return r;
}
// CircleDomain(half_coset).index_at(i) (cuda/src/quotients.cu:17-29)
__host__ __device__ __forceinline__ uint32_t circle_domain_index(uint32_t hi, uint32_t hs, uint32_t half_log,
                                                        uint32_t i) {
  uint32_t half = 1u << half_log;
  if (i < half) return (uint32_t)((hi + (uint64_t)hs * i) & 0x7fffffffu);
  uint32_t g = (uint32_t)((hi + (uint64_t)hs * (i - half)) & 0x7fffffffu);
  return (0x80000000u - g) & 0x7fffffffu;
}
__host__ __device__ __forceinline__ uint32_t brev(uint32_t i, uint32_t bits) {
  return bits ? (__builtin_bitreverse32(i) >> (32 - bits)) : 0u;
}

// Synthetic uniform M31 stream: value i of (seed, offset) is the top 31 bits of the
// SplitMix64 output for counter seed + (offset + i + 1) * golden gamma, with 2^31 - 1
// mapped to the low 31 bits (then to 0 if still 2^31 - 1).
__host__ __device__ __forceinline__ uint32_t uniform_m31(uint64_t seed, uint64_t i) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  uint32_t v = (uint32_t)(z >> 33);
  if (v == P) v = (uint32_t)(z & P);
  if (v == P) v = 0;
  return v;
}

// SoA secure-column access
__host__ __device__ __forceinline__ qm31 load_soa(const uint32_t* const* c, size_t i) {
  return {{c[0][i], c[1][i]}, {c[2][i], c[3][i]}};
}
__host__ __device__ __forceinline__ void store_soa(uint32_t* const* c, size_t i, qm31 v) {
  c[0][i] = v.x.a;
  c[1][i] = v.x.b;
  c[2][i] = v.y.a;
  c[3][i] = v.y.b;
}
struct qm31w {  // 16-byte AoS element as passed across the C ABI
  uint32_t a, b, c, d;
};
__host__ __device__ __forceinline__ qm31 from_w(qm31w w) { return {{w.a, w.b}, {w.c, w.d}}; }
__host__ __device__ __forceinline__ qm31w to_w(qm31 v) { return {v.x.a, v.x.b, v.y.a, v.y.b}; }

}  // namespace sh
