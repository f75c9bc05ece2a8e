// BLAKE2s on a lane quad: the four G functions of each half-round side by side on four
// lanes (Merkle nodes of narrow layers, the Fiat-Shamir channel's one-block hashes), and
// the channel step that follows a commitment.  Used by the Merkle kernels and the prover.
#pragma once
#include <stdint.h>

#include "blake2s.hpp"
#include "channel.hpp"
#include "field.hpp"
#include "merkle_plan.hpp"

namespace sh {
namespace merkle {

__device__ __forceinline__ uint32_t qrotr(uint32_t x, uint32_t c) { return __builtin_amdgcn_alignbit(x, x, c); }

// ---- one compression on the four lanes of a quad ------------------------------------
// Narrow layers (few nodes) leave most lanes idle and their hash chains are latency
// bound, so there a node is hashed by a quad: lane q keeps state words v[q], v[4+q],
// v[8+q], v[12+q] and computes the column-step G on them; for the diagonal step b, c, d
// are rotated by 1, 2, 3 lanes inside the quad (DPP quad permutes) and back.  The lane
// reads the 40 message words it needs (two per G, ten rounds) from the node's 16-word
// message in LDS before the rounds start.
constexpr uint8_t SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0}};
// message index of word k (0..3: column x, column y, diagonal x, diagonal y) of round r,
// for the four lanes of a quad packed in bytes
constexpr uint32_t quad_index(int r, int k) {
  uint32_t v = 0;
  for (int q = 0; q < 4; q++) v |= (uint32_t)SIGMA[r][(k < 2 ? 0 : 8) + 2 * q + (k & 1)] << (8 * q);
  return v;
}
template <int CTRL>
__device__ __forceinline__ uint32_t qperm(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, false);
}
constexpr int QROT1 = 0x39, QROT2 = 0x4E, QROT3 = 0x93;  // lane q reads lane q+1, q+2, q+3 (mod 4)

#define QG(a, b, c, d, x, y) \
  a = a + b + (x);           \
  d = qrotr(d ^ a, 16);       \
  c = c + d;                 \
  b = qrotr(b ^ c, 12);       \
  a = a + b + (y);           \
  d = qrotr(d ^ a, 8);        \
  c = c + d;                 \
  b = qrotr(b ^ c, 7);

// (h_lo, h_hi) = (h[q], h[4+q]) of the node's state; msg: the node's 16 message words in LDS.
// dx is XORed into v[12 + q] (BLAKE2s counter and last-block flag; zero for tree nodes).
__device__ __forceinline__ void quad_compress(uint32_t& h_lo, uint32_t& h_hi, const uint32_t* msg, uint32_t q,
                                              uint32_t dx = 0) {
  const uint32_t sh = 8 * q;
  uint32_t m[40];
#pragma unroll
  for (int r = 0; r < 10; r++)
#pragma unroll
    for (int k = 0; k < 4; k++) m[4 * r + k] = msg[(quad_index(r, k) >> sh) & 15];
  constexpr uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                              0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
  uint32_t a = h_lo, b = h_hi;
  uint32_t c = q == 0 ? IV[0] : q == 1 ? IV[1] : q == 2 ? IV[2] : IV[3];
  uint32_t d = (q == 0 ? IV[4] : q == 1 ? IV[5] : q == 2 ? IV[6] : IV[7]) ^ dx;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    QG(a, b, c, d, m[4 * r], m[4 * r + 1])
    b = qperm<QROT1>(b);
    c = qperm<QROT2>(c);
    d = qperm<QROT3>(d);
    QG(a, b, c, d, m[4 * r + 2], m[4 * r + 3])
    b = qperm<QROT3>(b);
    c = qperm<QROT2>(c);
    d = qperm<QROT1>(d);
  }
  h_lo ^= a ^ c;
  h_hi ^= b ^ d;
}
#undef QG

// One BLAKE2s-256 block of the 16 words at msg (LDS) on a lane quad, for a message of
// t_bytes bytes (blake2s256_block): returns (h[q], h[4+q]) of the digest.
__device__ __forceinline__ void quad_blake2s256_block(const uint32_t* msg, uint32_t q, uint32_t t_bytes,
                                                      uint32_t& h_lo, uint32_t& h_hi) {
  constexpr uint32_t IVS[8] = {b2s::IV0, b2s::IV1, b2s::IV2, b2s::IV3, b2s::IV4, b2s::IV5, b2s::IV6, b2s::IV7};
  h_lo = (q == 0 ? IVS[0] ^ 0x01010020u : q == 1 ? IVS[1] : q == 2 ? IVS[2] : IVS[3]);
  h_hi = (q == 0 ? IVS[4] : q == 1 ? IVS[5] : q == 2 ? IVS[6] : IVS[7]);
  quad_compress(h_lo, h_hi, msg, q, q == 0 ? t_bytes : q == 2 ? 0xFFFFFFFFu : 0u);
}

// Channel::mix8 on lanes 0..3 of a wave (one quad): digest <- B2S(digest || w); returns
// digest words q and 4 + q.  mb: 16 words of LDS owned by the quad.
__device__ __forceinline__ void quad_mix8(const Channel* ch, const uint32_t* w, uint32_t* mb, uint32_t q,
                                          uint32_t& g_lo, uint32_t& g_hi) {
  mb[q] = ch->digest[q];
  mb[4 + q] = ch->digest[4 + q];
  mb[8 + q] = w[q];
  mb[12 + q] = w[4 + q];
  __builtin_amdgcn_wave_barrier();
  quad_blake2s256_block(mb, q, 64, g_lo, g_hi);
}
// Channel::draw_qm31 on a quad, from the digest words (g_lo, g_hi) and draw counter n
// (advanced past every draw it takes): draws B2S(digest || n || 0) until the first four
// words are all below 2P.
__device__ __forceinline__ qm31 quad_draw_qm31(uint32_t g_lo, uint32_t g_hi, uint32_t& n, uint32_t* mb,
                                               uint32_t q) {
  for (;;) {
    __builtin_amdgcn_wave_barrier();  // the quad's reads of the previous message are done
    mb[q] = g_lo;
    mb[4 + q] = g_hi;
    mb[8 + q] = q == 0 ? n : 0u;
    mb[12 + q] = 0u;
    __builtin_amdgcn_wave_barrier();
    uint32_t w_lo, w_hi;
    quad_blake2s256_block(mb, q, 40, w_lo, w_hi);
    n++;
    __builtin_amdgcn_wave_barrier();
    mb[q] = w_lo;  // words 0..3 of the draw
    __builtin_amdgcn_wave_barrier();
    const uint32_t w0 = mb[0], w1 = mb[1], w2 = mb[2], w3 = mb[3];
    if (w0 < 2 * P && w1 < 2 * P && w2 < 2 * P && w3 < 2 * P)  // uniform over the quad
      return qm31{{w0 >= P ? w0 - P : w0, w1 >= P ? w1 - P : w1}, {w2 >= P ? w2 - P : w2, w3 >= P ? w3 - P : w3}};
  }
}

// The channel step after a commitment, on one quad: Channel::mix8 of the root, then
// Channel::draw_qm31 into *rd.dst, or Channel::draw_point into rd.dst[0], rd.dst[1]
// (rd.point); the compressions run four G functions side by side.
static __device__ __noinline__ void channel_step_quad(RootDraw rd, const uint32_t* root, uint32_t* mb, uint32_t q) {
  uint32_t g_lo, g_hi;
  quad_mix8(rd.ch, root, mb, q, g_lo, g_hi);
  uint32_t n = 0;
  qm31 x = qzero(), y = qzero(), t = qzero();
  for (;;) {
    t = quad_draw_qm31(g_lo, g_hi, n, mb, q);
    if (!rd.point) break;
    const qm31 t2 = qmul(t, t);
    const qm31 den = qadd(qone(), t2);
    if (qis_zero(den)) continue;
    const qm31 di = qinv(den);
    x = qmul(qsub(qone(), t2), di);
    y = qmul(qadd(t, t), di);
    break;
  }
  rd.ch->digest[q] = g_lo;
  rd.ch->digest[4 + q] = g_hi;
  if (q == 0) {
    rd.ch->n_draws = n;
    if (rd.point) {
      rd.dst[0] = x;
      rd.dst[1] = y;
    } else {
      *rd.dst = t;
    }
  }
}

}  // namespace merkle
}  // namespace sh
