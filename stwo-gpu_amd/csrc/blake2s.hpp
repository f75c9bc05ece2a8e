// BLAKE2s (RFC 7693) for the Fiat-Shamir channel and proof-of-work: the standard
// 32-byte digest (parameter block: digest length 32, no key, fanout = depth = 1).
// The Merkle hasher uses its own specialised compression (zero IV state, zero counter
// and flags) in blake2s_merkle.hip; this one carries the byte counter and the final
// flag, so blake2s256() equals any standard BLAKE2s-256 implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sh {
namespace b2s {

__host__ __device__ inline uint32_t rotr32(uint32_t x, uint32_t c) { return (x >> c) | (x << (32 - c)); }

__host__ __device__ constexpr uint32_t iv(int i) {
  return i == 0   ? 0x6A09E667u
         : i == 1 ? 0xBB67AE85u
         : i == 2 ? 0x3C6EF372u
         : i == 3 ? 0xA54FF53Au
         : i == 4 ? 0x510E527Fu
         : i == 5 ? 0x9B05688Cu
         : i == 6 ? 0x1F83D9ABu
                  : 0x5BE0CD19u;
}

// message word schedule of round r (RFC 7693, section 2.7)
__host__ __device__ inline uint32_t sigma(int r, int j) {
  constexpr uint8_t S[10][16] = {
      {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
      {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
      {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
      {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
      {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0}};
  return S[r][j];
}

__host__ __device__ inline void g(uint32_t* v, int a, int b, int c, int d, uint32_t x, uint32_t y) {
  v[a] = v[a] + v[b] + x;
  v[d] = rotr32(v[d] ^ v[a], 16);
  v[c] = v[c] + v[d];
  v[b] = rotr32(v[b] ^ v[c], 12);
  v[a] = v[a] + v[b] + y;
  v[d] = rotr32(v[d] ^ v[a], 8);
  v[c] = v[c] + v[d];
  v[b] = rotr32(v[b] ^ v[c], 7);
}

// h <- F(h, m, t, last)
__host__ __device__ inline void compress(uint32_t* h, const uint32_t* m, uint64_t t, bool last) {
  uint32_t v[16];
  for (int i = 0; i < 8; i++) {
    v[i] = h[i];
    v[8 + i] = iv(i);
  }
  v[12] ^= (uint32_t)t;
  v[13] ^= (uint32_t)(t >> 32);
  if (last) v[14] = ~v[14];
  for (int r = 0; r < 10; r++) {
    g(v, 0, 4, 8, 12, m[sigma(r, 0)], m[sigma(r, 1)]);
    g(v, 1, 5, 9, 13, m[sigma(r, 2)], m[sigma(r, 3)]);
    g(v, 2, 6, 10, 14, m[sigma(r, 4)], m[sigma(r, 5)]);
    g(v, 3, 7, 11, 15, m[sigma(r, 6)], m[sigma(r, 7)]);
    g(v, 0, 5, 10, 15, m[sigma(r, 8)], m[sigma(r, 9)]);
    g(v, 1, 6, 11, 12, m[sigma(r, 10)], m[sigma(r, 11)]);
    g(v, 2, 7, 8, 13, m[sigma(r, 12)], m[sigma(r, 13)]);
    g(v, 3, 4, 9, 14, m[sigma(r, 14)], m[sigma(r, 15)]);
  }
  for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[8 + i];
}

// Streaming BLAKE2s-256 over little-endian 32-bit words.  A full block is compressed
// only once more input arrives, since the last block carries the final flag.
struct Hasher {
  uint32_t h[8];
  uint32_t buf[16];
  uint32_t nbuf = 0;
  uint64_t bytes = 0;
  __host__ __device__ Hasher() {
    for (int i = 0; i < 8; i++) h[i] = iv(i);
    h[0] ^= 0x01010020u;  // digest length 32, fanout 1, depth 1
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

// BLAKE2s-256 of a message given as little-endian 32-bit words (n_words * 4 bytes).
__host__ __device__ inline void blake2s256_words(const uint32_t* words, uint32_t n_words, uint32_t* out) {
  uint32_t h[8];
  for (int i = 0; i < 8; i++) h[i] = iv(i);
  h[0] ^= 0x01010020u;  // digest length 32, fanout 1, depth 1
  const uint64_t n_bytes = 4ull * n_words;
  uint32_t done = 0;
  do {
    uint32_t m[16];
    const uint32_t take = n_words - done > 16 ? 16 : n_words - done;
    for (uint32_t j = 0; j < 16; j++) m[j] = j < take ? words[done + j] : 0u;
    done += take;
    const bool last = done == n_words;
    compress(h, m, last ? n_bytes : 4ull * done, last);
  } while (done < n_words);
  for (int i = 0; i < 8; i++) out[i] = h[i];
}

}  // namespace b2s
}  // namespace sh
