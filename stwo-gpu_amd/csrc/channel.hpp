// Fiat-Shamir channel over BLAKE2s-256, usable on the host and on the device (the
// prover runs it in single-lane kernels so a proof is enqueued without host round
// trips).  It plays the role of stwo's Blake2sChannel (the channel type that
// stwo_gpu_backend/src/backend.rs:14-20 binds to the Blake2s Merkle channel); the
// transcript byte layout below is this library's own:
//   mix(words):   digest <- B2S(digest || words), n_draws <- 0
//   mix_felts(v): mix(tree_root(words of v)) (tree hash of blake2s.hpp, QM31 as 4 words)
//   draw_words(): B2S(digest || n_draws || 0), then n_draws += 1   (8 words)
//   draw_qm31():  first 4 words of a draw, redrawn unless all are < 2P, each reduced mod P
//   draw_point(): t = draw_qm31(), (x, y) = ((1 - t^2) / (1 + t^2), 2t / (1 + t^2))
//   queries:      successive draw words masked to the domain size
//   proof of work: nonce n is valid for b bits when the 64-bit little-endian value of the
//                 first two words of B2S(digest || lo(n) || hi(n)) has b trailing zeros;
//                 the prover takes the smallest such n and mixes (lo, hi).
#pragma once
#include "blake2s.hpp"
#include "field.hpp"

namespace sh {

struct Channel {
  uint32_t digest[8];
  uint32_t n_draws;

  __host__ __device__ void reset() {
    for (int i = 0; i < 8; i++) digest[i] = 0;
    n_draws = 0;
  }
  __host__ __device__ void mix(const uint32_t* w, uint32_t n) {
    if (n <= 8) {  // digest || w is one block: message and state in registers
      uint32_t m[16];
#pragma unroll
      for (int i = 0; i < 16; i++) m[i] = i < 8 ? digest[i] : ((uint32_t)(i - 8) < n ? w[i - 8] : 0u);
      b2s::blake2s256_padded_block(m, 8 + n, digest);
      n_draws = 0;
      return;
    }
    b2s::Hasher hs;
    hs.words(digest, 8);
    hs.words(w, n);
    hs.finish(digest);
    n_draws = 0;
  }
  // mix(w, 8) with the message in registers (one block: digest || w)
  __host__ __device__ void mix8(const uint32_t* w) {
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      m[i] = digest[i];
      m[8 + i] = w[i];
    }
    b2s::blake2s256_block<16>(m, digest);
    n_draws = 0;
  }
  // Field-element vectors are mixed through their tree hash (blake2s.hpp): the prover
  // computes the root in parallel (prover.hip) and mixes it with mix(root, 8).
  __host__ __device__ void draw_words(uint32_t* out) {
    uint32_t m[10];
#pragma unroll
    for (int i = 0; i < 8; i++) m[i] = digest[i];
    m[8] = n_draws++;
    m[9] = 0;
    uint32_t h[8];
    b2s::blake2s256_block<10>(m, h);
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = h[i];
  }
  __host__ __device__ qm31 draw_qm31() {
    for (;;) {
      uint32_t w[8];
      draw_words(w);
      if (w[0] < 2 * P && w[1] < 2 * P && w[2] < 2 * P && w[3] < 2 * P) {
        uint32_t r[4];
        for (int k = 0; k < 4; k++) r[k] = w[k] >= P ? w[k] - P : w[k];
        return qm31{{r[0], r[1]}, {r[2], r[3]}};
      }
    }
  }
  __host__ __device__ void draw_point(qm31& x, qm31& y) {
    for (;;) {
      const qm31 t = draw_qm31();
      const qm31 t2 = qmul(t, t);
      const qm31 den = qadd(qone(), t2);
      if (qis_zero(den)) continue;
      const qm31 di = qinv(den);
      x = qmul(qsub(qone(), t2), di);
      y = qmul(qadd(t, t), di);
      return;
    }
  }
  __host__ __device__ void draw_queries(uint32_t* out, uint32_t n, uint32_t log_size) {
    const uint32_t mask = log_size >= 32 ? 0xffffffffu : (1u << log_size) - 1;
    uint32_t k = 0;
    while (k < n) {
      uint32_t w[8];
      draw_words(w);
      for (int j = 0; j < 8 && k < n; j++) out[k++] = w[j] & mask;
    }
  }
};

// proof-of-work check shared by the grinder and the host side
__host__ __device__ inline uint32_t pow_trailing_zeros(const uint32_t* digest, uint64_t nonce) {
  uint32_t m[10];
#pragma unroll
  for (int i = 0; i < 8; i++) m[i] = digest[i];
  m[8] = (uint32_t)nonce;
  m[9] = (uint32_t)(nonce >> 32);
  uint32_t h[8];
  b2s::blake2s256_block<10>(m, h);
  uint32_t tz = 0;
  for (int k = 0; k < 2; k++) {
    if (h[k] == 0) {
      tz += 32;
      continue;
    }
    uint32_t v = h[k];
    while (!(v & 1)) {
      v >>= 1;
      tz++;
    }
    break;
  }
  return tz;
}

}  // namespace sh
