// BLAKE2s (RFC 7693) for the Fiat-Shamir channel and proof-of-work: the standard
// 32-byte digest (parameter block: digest length 32, no key, fanout = depth = 1), and
// the bare compression with zero counter and flags that the Merkle hasher uses
// (blake2s_merkle.hip has its own copy specialised for its kernels).
//
// The compression is written out round by round with literal message indices, so on
// the device the state and the message stay in registers even in a single-lane kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sh {
namespace b2s {

__host__ __device__ __forceinline__ uint32_t rotr32(uint32_t x, uint32_t c) { return (x >> c) | (x << (32 - c)); }

constexpr uint32_t IV0 = 0x6A09E667u, IV1 = 0xBB67AE85u, IV2 = 0x3C6EF372u, IV3 = 0xA54FF53Au;
constexpr uint32_t IV4 = 0x510E527Fu, IV5 = 0x9B05688Cu, IV6 = 0x1F83D9ABu, IV7 = 0x5BE0CD19u;

#define SH_B2S_G(a, b, c, d, x, y) \
  a = a + b + (x);                 \
  d = rotr32(d ^ a, 16);           \
  c = c + d;                       \
  b = rotr32(b ^ c, 12);           \
  a = a + b + (y);                 \
  d = rotr32(d ^ a, 8);            \
  c = c + d;                       \
  b = rotr32(b ^ c, 7);

// message schedule of each round (RFC 7693, section 2.7)
#define SH_B2S_ROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
  SH_B2S_G(v0, v4, v8, v12, m[s0], m[s1])                                                   \
  SH_B2S_G(v1, v5, v9, v13, m[s2], m[s3])                                                   \
  SH_B2S_G(v2, v6, v10, v14, m[s4], m[s5])                                                  \
  SH_B2S_G(v3, v7, v11, v15, m[s6], m[s7])                                                  \
  SH_B2S_G(v0, v5, v10, v15, m[s8], m[s9])                                                  \
  SH_B2S_G(v1, v6, v11, v12, m[s10], m[s11])                                                \
  SH_B2S_G(v2, v7, v8, v13, m[s12], m[s13])                                                 \
  SH_B2S_G(v3, v4, v9, v14, m[s14], m[s15])

// h <- F(h, m, t, last)
__host__ __device__ __forceinline__ void compress(uint32_t (&h)[8], const uint32_t (&m)[16], uint64_t t,
                                                  bool last) {
  uint32_t v0 = h[0], v1 = h[1], v2 = h[2], v3 = h[3], v4 = h[4], v5 = h[5], v6 = h[6], v7 = h[7];
  uint32_t v8 = IV0, v9 = IV1, v10 = IV2, v11 = IV3;
  uint32_t v12 = IV4 ^ (uint32_t)t, v13 = IV5 ^ (uint32_t)(t >> 32), v14 = last ? ~IV6 : IV6, v15 = IV7;
  SH_B2S_ROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
  SH_B2S_ROUND(14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3)
  SH_B2S_ROUND(11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4)
  SH_B2S_ROUND(7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8)
  SH_B2S_ROUND(9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13)
  SH_B2S_ROUND(2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9)
  SH_B2S_ROUND(12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11)
  SH_B2S_ROUND(13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10)
  SH_B2S_ROUND(6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5)
  SH_B2S_ROUND(10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0)
  h[0] ^= v0 ^ v8;
  h[1] ^= v1 ^ v9;
  h[2] ^= v2 ^ v10;
  h[3] ^= v3 ^ v11;
  h[4] ^= v4 ^ v12;
  h[5] ^= v5 ^ v13;
  h[6] ^= v6 ^ v14;
  h[7] ^= v7 ^ v15;
}
#undef SH_B2S_ROUND
#undef SH_B2S_G

// Streaming BLAKE2s-256 over little-endian 32-bit words.  A full block is compressed
// only once more input arrives, since the last block carries the final flag.
struct Hasher {
  uint32_t h[8];
  uint32_t buf[16];
  uint32_t nbuf = 0;
  uint64_t bytes = 0;
  __host__ __device__ Hasher() {
    h[0] = IV0 ^ 0x01010020u;  // digest length 32, fanout 1, depth 1
    h[1] = IV1;
    h[2] = IV2;
    h[3] = IV3;
    h[4] = IV4;
    h[5] = IV5;
    h[6] = IV6;
    h[7] = IV7;
  }
  __host__ __device__ void word(uint32_t w) {
    if (nbuf == 16) {
      compress(h, buf, bytes, false);
      nbuf = 0;
    }
    buf[nbuf++] = w;
    bytes += 4;
  }
  __host__ __device__ void words(const uint32_t* w, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) word(w[i]);
  }
  __host__ __device__ void finish(uint32_t* out) {
    for (uint32_t j = nbuf; j < 16; j++) buf[j] = 0;
    compress(h, buf, bytes, true);
    for (int i = 0; i < 8; i++) out[i] = h[i];
  }
};

// BLAKE2s-256 of a message of N <= 16 words (one block, so one compression with the
// final flag and counter 4N); the indices are compile-time constants, so on the device the
// message and state stay in registers (the streaming Hasher indexes its buffer at run time).
template <int N>
__host__ __device__ __forceinline__ void blake2s256_block(const uint32_t* words, uint32_t (&out)[8]) {
  static_assert(N >= 1 && N <= 16, "one block");
  uint32_t h[8] = {IV0 ^ 0x01010020u, IV1, IV2, IV3, IV4, IV5, IV6, IV7};
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 16; i++) m[i] = i < N ? words[i] : 0u;
  compress(h, m, 4ull * N, true);
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = h[i];
}

// BLAKE2s-256 of a message of n_words <= 16 words given zero-padded to one block.
__host__ __device__ __forceinline__ void blake2s256_padded_block(const uint32_t (&m)[16], uint32_t n_words,
                                                                uint32_t (&out)[8]) {
  uint32_t h[8] = {IV0 ^ 0x01010020u, IV1, IV2, IV3, IV4, IV5, IV6, IV7};
  compress(h, m, 4ull * n_words, true);
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = h[i];
}

// BLAKE2s-256 of a message given as little-endian 32-bit words (n_words * 4 bytes).
__host__ __device__ inline void blake2s256_words(const uint32_t* words, uint32_t n_words, uint32_t* out) {
  Hasher hs;
  hs.words(words, n_words);
  hs.finish(out);
}

// Tree hash of a word sequence (the channel mixes field-element vectors through it, so
// that long vectors hash in parallel): blocks of 16 words (the last one zero-padded)
// are leaves h_i = F0(block_i); a level pairs nodes as F0(h_2i || h_2i+1) and carries an
// unpaired last node up unchanged; the single remaining node is the root.  F0 is the
// compression from a zero state with zero counter and flags (the Merkle node function).
__host__ __device__ inline void f0(const uint32_t (&m)[16], uint32_t (&out)[8]) {
  uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  compress(h, m, 0, false);
  for (int i = 0; i < 8; i++) out[i] = h[i];
}

}  // namespace b2s
}  // namespace sh
