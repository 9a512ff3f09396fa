// Host restatement of the reference's randomness (rand 0.8.5 StdRng =
// rand_chacha 0.3.1 ChaCha12, rand_core 0.6.4 seeding; Cargo.lock:374-397):
//   * StdRng::seed_from_u64: PCG32 output expanded into the 32-byte key;
//   * ChaCha with 6 double rounds, 64-bit block counter in words 12-13,
//     stream 0 in words 14-15;
//   * BlockRng buffer of 64 words = 4 consecutive blocks; next_u64 reads two
//     words and straddles a refill at index 63;
//   * Standard u8/u16 = truncated next_u32, u32 = next_u32, u64/usize =
//     next_u64; gen_range(0..=i) over usize by widening-multiply rejection
//     with zone = (range << lz(range)) - 1.
// Call sites: galois-ring/src/sampling.rs:48-68 (challenges),
// pcs-ring/src/raa.rs:44-64, 156-169 (RAA permutations and diagonals).
#pragma once
#include <cstdint>

namespace nenya {

struct StdRng {
  uint32_t key[8];
  uint64_t counter = 0;
  uint32_t buf[64];
  int idx = 64;

  static uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
  static void block(const uint32_t key[8], uint64_t ctr, uint32_t out[16]) {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2], key[3],
                      key[4], key[5], key[6], key[7], (uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0u};
    uint32_t x[16];
    for (int i = 0; i < 16; i++) x[i] = s[i];
    auto qr = [&](int a, int b, int c, int d) {
      x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 16);
      x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 12);
      x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 8);
      x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 7);
    };
    for (int r = 0; r < 6; r++) {
      qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
      qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
  }

  static StdRng seed_from_u64(uint64_t state) {
    StdRng r;
    const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
    for (int c = 0; c < 8; c++) {
      state = state * MUL + INC;
      uint64_t s = state;
      uint32_t xs = (uint32_t)(((s >> 18) ^ s) >> 27);
      uint32_t rot = (uint32_t)(s >> 59);
      r.key[c] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    return r;
  }
  void refill() {
    for (int b = 0; b < 4; b++) block(key, counter + (uint64_t)b, &buf[16 * b]);
    counter += 4;
  }
  uint32_t next_u32() {
    if (idx >= 64) { refill(); idx = 0; }
    return buf[idx++];
  }
  uint64_t next_u64() {
    if (idx < 63) {
      uint64_t v = (uint64_t)buf[idx] | ((uint64_t)buf[idx + 1] << 32);
      idx += 2;
      return v;
    }
    if (idx >= 64) {
      refill();
      idx = 2;
      return (uint64_t)buf[0] | ((uint64_t)buf[1] << 32);
    }
    uint64_t lo = buf[63];
    refill();
    idx = 1;
    return ((uint64_t)buf[0] << 32) | lo;
  }
  uint64_t gen_range_incl(uint64_t low, uint64_t high) {
    uint64_t range = high - low + 1;
    if (range == 0) return next_u64();
    uint64_t zone = (range << __builtin_clzll(range)) - 1;
    for (;;) {
      uint64_t v = next_u64();
      unsigned __int128 m = (unsigned __int128)v * range;
      if ((uint64_t)m <= zone) return low + (uint64_t)(m >> 64);
    }
  }
  // L::random (limb.rs:103-105) for width k.
  uint64_t limb(int k) { return k == 64 ? next_u64() : (uint64_t)next_u32() & ((1ull << k) - 1); }
};

}  // namespace nenya
