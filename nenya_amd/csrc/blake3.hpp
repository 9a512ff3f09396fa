// BLAKE3 (published specification; the reference uses the blake3 1.8.4 crate,
// Cargo.lock:24-25): a compression function shared by host and device, a
// host incremental hasher with XOF (Fiat-Shamir transcript, query indices),
// and device helpers that hash one message whose 32-bit words come from an
// arbitrary loader (Merkle column leaves, pcs-ring/src/merkle.rs:20-44).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

namespace nenya {
namespace b3 {

enum : uint32_t { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };

__host__ __device__ inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

#define NENYA_B3_G(a, b, c, d, x, y)               \
  s[a] = s[a] + s[b] + (x); s[d] = rotr(s[d] ^ s[a], 16); \
  s[c] = s[c] + s[d]; s[b] = rotr(s[b] ^ s[c], 12);       \
  s[a] = s[a] + s[b] + (y); s[d] = rotr(s[d] ^ s[a], 8);  \
  s[c] = s[c] + s[d]; s[b] = rotr(s[b] ^ s[c], 7);

// One 7-round compression; writes 16 output words (first 8 = chaining value).
__host__ __device__ inline void compress(const uint32_t cv[8], const uint32_t m_in[16], uint64_t counter,
                                         uint32_t blen, uint32_t flags, uint32_t out[16]) {
  uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                    0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                    (uint32_t)counter, (uint32_t)(counter >> 32), blen, flags};
  uint32_t m[16];
  for (int i = 0; i < 16; i++) m[i] = m_in[i];
#pragma unroll
  for (int r = 0; r < 7; r++) {
    NENYA_B3_G(0, 4, 8, 12, m[0], m[1]);
    NENYA_B3_G(1, 5, 9, 13, m[2], m[3]);
    NENYA_B3_G(2, 6, 10, 14, m[4], m[5]);
    NENYA_B3_G(3, 7, 11, 15, m[6], m[7]);
    NENYA_B3_G(0, 5, 10, 15, m[8], m[9]);
    NENYA_B3_G(1, 6, 11, 12, m[10], m[11]);
    NENYA_B3_G(2, 7, 8, 13, m[12], m[13]);
    NENYA_B3_G(3, 4, 9, 14, m[14], m[15]);
    // message permutation [2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8]
    uint32_t t[16] = {m[2], m[6], m[3], m[10], m[7], m[0], m[4], m[13],
                      m[1], m[11], m[12], m[5], m[9], m[14], m[15], m[8]};
    for (int i = 0; i < 16; i++) m[i] = t[i];
  }
  for (int i = 0; i < 8; i++) {
    out[i] = s[i] ^ s[i + 8];
    out[i + 8] = s[i + 8] ^ cv[i];
  }
}
#undef NENYA_B3_G

__host__ __device__ inline void iv(uint32_t cv[8]) {
  cv[0] = 0x6A09E667u; cv[1] = 0xBB67AE85u; cv[2] = 0x3C6EF372u; cv[3] = 0xA54FF53Au;
  cv[4] = 0x510E527Fu; cv[5] = 0x9B05688Cu; cv[6] = 0x1F83D9ABu; cv[7] = 0x5BE0CD19u;
}

// Hash a message of `len` bytes (len a multiple of 4) whose word w is
// produced by `word(w)`; writes the 32-byte root hash.  Follows the
// incremental hasher's chunk/CV-stack order, so the tree shape is BLAKE3's.
template <class F>
__host__ __device__ inline void hash_words(uint64_t len, F word, uint32_t out[8]) {
  uint32_t stack[16][8];
  int sp = 0;
  uint64_t nchunks = len == 0 ? 1 : (len + 1023) / 1024;
  uint32_t key[8];
  iv(key);
  for (uint64_t c = 0; c < nchunks; c++) {
    uint64_t cbeg = c * 1024;
    uint64_t clen = len - cbeg < 1024 ? len - cbeg : 1024;
    uint32_t nblk = clen == 0 ? 1 : (uint32_t)((clen + 63) / 64);
    uint32_t cv[8];
    iv(cv);
    uint32_t m[16], o[16];
    for (uint32_t b = 0; b + 1 < nblk; b++) {
      uint64_t wb = (cbeg + 64 * (uint64_t)b) / 4;
      for (int i = 0; i < 16; i++) m[i] = word(wb + i);
      compress(cv, m, c, 64, b == 0 ? CHUNK_START : 0, o);
      for (int i = 0; i < 8; i++) cv[i] = o[i];
    }
    uint32_t lb = nblk - 1;
    uint32_t blen = (uint32_t)(clen - 64 * (uint64_t)lb);
    uint64_t wb = (cbeg + 64 * (uint64_t)lb) / 4;
    for (int i = 0; i < 16; i++) m[i] = (4u * i < blen) ? word(wb + i) : 0u;
    uint32_t fl = (lb == 0 ? CHUNK_START : 0) | CHUNK_END;
    if (c + 1 < nchunks) {
      compress(cv, m, c, blen, fl, o);
      uint32_t ncv[8];
      for (int i = 0; i < 8; i++) ncv[i] = o[i];
      uint64_t total = c + 1;
      while ((total & 1) == 0) {
        sp--;
        uint32_t pm[16];
        for (int i = 0; i < 8; i++) { pm[i] = stack[sp][i]; pm[8 + i] = ncv[i]; }
        compress(key, pm, 0, 64, PARENT, o);
        for (int i = 0; i < 8; i++) ncv[i] = o[i];
        total >>= 1;
      }
      for (int i = 0; i < 8; i++) stack[sp][i] = ncv[i];
      sp++;
    } else {
      // last chunk: fold the stack with its output node, ROOT on the final node
      uint32_t ocv[8], ob[16];
      uint64_t octr = c;
      uint32_t oblen = blen, ofl = fl;
      for (int i = 0; i < 8; i++) ocv[i] = cv[i];
      for (int i = 0; i < 16; i++) ob[i] = m[i];
      while (sp > 0) {
        compress(ocv, ob, octr, oblen, ofl, o);
        sp--;
        for (int i = 0; i < 8; i++) { ob[i] = stack[sp][i]; ob[8 + i] = o[i]; }
        for (int i = 0; i < 8; i++) ocv[i] = key[i];
        octr = 0; oblen = 64; ofl = PARENT;
      }
      compress(ocv, ob, octr, oblen, ofl | ROOT, o);
      for (int i = 0; i < 8; i++) out[i] = o[i];
    }
  }
}

// Parent node hash of two 32-byte children (a 64-byte single-block message).
__host__ __device__ inline void hash64(const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
  uint32_t cv[8], m[16], o[16];
  iv(cv);
  for (int i = 0; i < 8; i++) { m[i] = l[i]; m[8 + i] = r[i]; }
  compress(cv, m, 0, 64, CHUNK_START | CHUNK_END | ROOT, o);
  for (int i = 0; i < 8; i++) out[i] = o[i];
}

// ---------------- incremental hasher with XOF ----------------
// The hasher's state as plain data, identical on host and device, so a
// transcript can move between them as bytes.
struct State {
  uint32_t key[8];
  uint32_t cv[8];
  uint64_t chunk;
  uint8_t block[64];
  uint32_t blen;
  uint32_t blocks;
  uint32_t stack[56][8];
  int sp;
  int pad_;
};

struct Hasher : State {
  Hasher() { reset(); }
  void reset() {
    iv(key);
    iv(cv);
    chunk = 0;
    std::memset(block, 0, 64);
    blen = 0;
    blocks = 0;
    sp = 0;
    pad_ = 0;
  }
  static void words(const uint8_t* b, uint32_t* w) {
    for (int i = 0; i < 16; i++)
      w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
             ((uint32_t)b[4 * i + 3] << 24);
  }
  uint64_t chunk_len() const { return 64ull * blocks + blen; }
  void update(const void* data, size_t len) {
    const uint8_t* in = (const uint8_t*)data;
    while (len) {
      if (chunk_len() == 1024) {
        uint32_t m[16], o[16];
        words(block, m);
        compress(cv, m, chunk, blen, (blocks == 0 ? CHUNK_START : 0) | CHUNK_END, o);
        uint32_t ncv[8];
        for (int i = 0; i < 8; i++) ncv[i] = o[i];
        uint64_t total = chunk + 1;
        while ((total & 1) == 0) {
          sp--;
          uint32_t pm[16];
          for (int i = 0; i < 8; i++) { pm[i] = stack[sp][i]; pm[8 + i] = ncv[i]; }
          compress(key, pm, 0, 64, PARENT, o);
          for (int i = 0; i < 8; i++) ncv[i] = o[i];
          total >>= 1;
        }
        std::memcpy(stack[sp++], ncv, 32);
        chunk++;
        iv(cv);
        std::memset(block, 0, 64);
        blen = 0;
        blocks = 0;
      }
      // fill the chunk, compressing full blocks only when more input follows
      while (len && chunk_len() < 1024) {
        if (blen == 64) {
          uint32_t m[16], o[16];
          words(block, m);
          compress(cv, m, chunk, 64, blocks == 0 ? CHUNK_START : 0, o);
          std::memcpy(cv, o, 32);
          blocks++;
          std::memset(block, 0, 64);
          blen = 0;
        }
        size_t take = 64 - blen;
        if (take > len) take = len;
        std::memcpy(block + blen, in, take);
        blen += (uint32_t)take;
        in += take;
        len -= take;
      }
    }
  }
  // XOF output starting at byte `seek`.
  void finalize(uint64_t seek, uint8_t* out, size_t olen) const {
    uint32_t ocv[8], ob[16], o[16];
    std::memcpy(ocv, cv, 32);
    words(block, ob);
    uint64_t octr = chunk;
    uint32_t oblen = blen, ofl = (blocks == 0 ? CHUNK_START : 0) | CHUNK_END;
    for (int s = sp - 1; s >= 0; s--) {
      compress(ocv, ob, octr, oblen, ofl, o);
      for (int i = 0; i < 8; i++) { ob[i] = stack[s][i]; ob[8 + i] = o[i]; }
      std::memcpy(ocv, key, 32);
      octr = 0;
      oblen = 64;
      ofl = PARENT;
    }
    uint64_t blk = seek / 64;
    size_t off = seek % 64;
    while (olen) {
      compress(ocv, ob, blk, oblen, ofl | ROOT, o);
      uint8_t bytes[64];
      for (int i = 0; i < 16; i++)
        for (int j = 0; j < 4; j++) bytes[4 * i + j] = (uint8_t)(o[i] >> (8 * j));
      size_t take = 64 - off;
      if (take > olen) take = olen;
      std::memcpy(out, bytes + off, take);
      out += take;
      olen -= take;
      off = 0;
      blk++;
    }
  }
};

}  // namespace b3
}  // namespace nenya
