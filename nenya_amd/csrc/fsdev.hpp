// Fiat-Shamir on the device: the transcript's BLAKE3 state absorbs a
// sumcheck round polynomial and squeezes the round challenge's sampler seed,
// and ChaCha12 expands it into the natural sample, all in one wave of the
// kernel that then lifts the challenge.  Same bytes as the host transcript
// (snark-ring/src/transcript.rs:23-105: label "L" || u32 len || label,
// element "E" || limbs, challenge = 8 XOF bytes of a clone, absorbed as
// "S" || bytes, then StdRng::seed_from_u64 and 128 L::random draws,
// galois-ring/src/sampling.rs:48-68).
//
// Compression on a quad of lanes: lane c holds column c of the 4x4 state
// (s[c], s[4+c], s[8+c], s[12+c]) and applies the column function G_c; the
// diagonal step rotates rows 1..3 of the quad by 1, 2, 3 lanes (DPP
// quad_perm), applies G again and rotates back.  Every quad of the wave
// computes the same compression, so the whole wave runs in lock step and
// the DPP reads always see active lanes.  ChaCha uses the same layout, one
// 16-word block per quad.
#pragma once
#include "blake3.hpp"
#include "kernels.hpp"

namespace nenya {
namespace fsd {

constexpr uint8_t SCHED[7][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},  {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8},
    {3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1},  {10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6},
    {12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4},  {9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7},
    {11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13}};

// The message word lane c reads for slot j of round r (j: column x, y,
// diagonal x, y): byte c of the packed constant.
__device__ __forceinline__ uint32_t sched_idx(int r, int j, int c) {
  uint32_t packed = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int pos = (j < 2 ? 2 * q + j : 8 + 2 * q + (j - 2));
    packed |= (uint32_t)SCHED[r][pos] << (8 * q);
  }
  return (packed >> (8 * c)) & 0xffu;
}

template <int CTRL> __device__ __forceinline__ uint32_t qperm(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
}
// quad_perm controls: lane c reads lane c+1, c+2, c+3 (mod 4)
constexpr int Q1 = 0x39, Q2 = 0x4E, Q3 = 0x93;

// Orders this wave's LDS writes before the reads that follow (LDS operations
// of one wave complete in order; this keeps the compiler from caching).
__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

__device__ __forceinline__ void g_mix(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d, uint32_t x, uint32_t y) {
  a = a + b + x;
  d = b3::rotr(d ^ a, 16);
  c = c + d;
  b = b3::rotr(b ^ c, 12);
  a = a + b + y;
  d = b3::rotr(d ^ a, 8);
  c = c + d;
  b = b3::rotr(b ^ c, 7);
}

// One BLAKE3 compression by the whole wave (every quad the same): cv (8
// words) and m (16 words) in LDS; lanes 0..3 write the 16 output words to
// out (LDS).  Inputs are read before any output is written, so out may
// alias cv or m.
__device__ __forceinline__ void compress(const uint32_t* cv, const uint32_t* m, uint64_t counter, uint32_t blen,
                                         uint32_t flags, uint32_t* out) {
  const int lane = threadIdx.x & 63, c = lane & 3;
  const uint32_t IVc[4] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au};
  uint32_t a = cv[c], b = cv[4 + c];
  uint32_t cc = c == 0 ? IVc[0] : c == 1 ? IVc[1] : c == 2 ? IVc[2] : IVc[3];
  uint32_t d = c == 0 ? (uint32_t)counter : c == 1 ? (uint32_t)(counter >> 32) : c == 2 ? blen : flags;
  const uint32_t cv0 = a, cv1 = b;
  uint32_t mx[28];
#pragma unroll
  for (int r = 0; r < 7; r++)
#pragma unroll
    for (int j = 0; j < 4; j++) mx[4 * r + j] = m[sched_idx(r, j, c)];
  wsync();
#pragma unroll
  for (int r = 0; r < 7; r++) {
    g_mix(a, b, cc, d, mx[4 * r], mx[4 * r + 1]);
    b = qperm<Q1>(b);
    cc = qperm<Q2>(cc);
    d = qperm<Q3>(d);
    g_mix(a, b, cc, d, mx[4 * r + 2], mx[4 * r + 3]);
    b = qperm<Q3>(b);
    cc = qperm<Q2>(cc);
    d = qperm<Q1>(d);
  }
  if (lane < 4) {
    out[c] = a ^ cc;
    out[4 + c] = b ^ d;
    out[8 + c] = cc ^ cv0;
    out[12 + c] = d ^ cv1;
  }
  wsync();
}

// The same compression with this quad's chaining value in registers: lane c
// holds cv[c] in ca and cv[4+c] in cb, and receives the new chaining value
// (output words c and 4+c) there.  m: 16 words in LDS.  Only the quads that
// call it take part (each quad is entirely active or not).
__device__ __forceinline__ void compress_cv(uint32_t& ca, uint32_t& cb, const uint32_t* m, uint64_t counter,
                                            uint32_t blen, uint32_t flags) {
  const int c = threadIdx.x & 3;
  uint32_t a = ca, b = cb;
  uint32_t cc = c == 0 ? 0x6A09E667u : c == 1 ? 0xBB67AE85u : c == 2 ? 0x3C6EF372u : 0xA54FF53Au;
  uint32_t d = c == 0 ? (uint32_t)counter : c == 1 ? (uint32_t)(counter >> 32) : c == 2 ? blen : flags;
  uint32_t mx[28];
#pragma unroll
  for (int r = 0; r < 7; r++)
#pragma unroll
    for (int j = 0; j < 4; j++) mx[4 * r + j] = m[sched_idx(r, j, c)];
#pragma unroll
  for (int r = 0; r < 7; r++) {
    g_mix(a, b, cc, d, mx[4 * r], mx[4 * r + 1]);
    b = qperm<Q1>(b);
    cc = qperm<Q2>(cc);
    d = qperm<Q3>(d);
    g_mix(a, b, cc, d, mx[4 * r + 2], mx[4 * r + 3]);
    b = qperm<Q3>(b);
    cc = qperm<Q2>(cc);
    d = qperm<Q1>(d);
  }
  ca = a ^ cc;
  cb = b ^ d;
}

// Scratch of the wave-level hasher (LDS).
struct Work {
  uint32_t out[16];
  uint32_t pm[16];
};

// Appends len bytes of msg (LDS) to the hasher state S (LDS); the whole
// wave calls this (the host Hasher::update, blake3.hpp, step for step).
__device__ __forceinline__ void update(b3::State* S, const uint8_t* msg, uint32_t len, Work* w) {
  const int lane = threadIdx.x & 63;
  uint32_t blen = S->blen, blocks = S->blocks;
  uint64_t chunk = S->chunk;
  int sp = S->sp;
  uint32_t pos = 0;
  while (pos < len) {
    if (64u * blocks + blen == 1024u) {
      // the chunk is full and input follows: finish it, push its chaining
      // value, merging completed subtrees
      compress(S->cv, reinterpret_cast<const uint32_t*>(S->block), chunk, blen,
               (blocks == 0 ? b3::CHUNK_START : 0u) | b3::CHUNK_END, w->out);
      uint64_t total = chunk + 1;
      while ((total & 1) == 0) {
        sp--;
        if (lane < 8) {
          w->pm[lane] = S->stack[sp][lane];
          w->pm[8 + lane] = w->out[lane];
        }
        wsync();
        compress(S->key, w->pm, 0, 64, b3::PARENT, w->out);
        total >>= 1;
      }
      if (lane < 8) {
        S->stack[sp][lane] = w->out[lane];
        S->cv[lane] = S->key[lane];
      }
      if (lane < 16) reinterpret_cast<uint32_t*>(S->block)[lane] = 0;
      wsync();
      sp++;
      chunk++;
      blen = 0;
      blocks = 0;
    }
    while (pos < len && 64u * blocks + blen < 1024u) {
      if (blen == 64) {
        compress(S->cv, reinterpret_cast<const uint32_t*>(S->block), chunk, 64, blocks == 0 ? b3::CHUNK_START : 0u,
                 w->out);
        if (lane < 8) S->cv[lane] = w->out[lane];
        if (lane < 16) reinterpret_cast<uint32_t*>(S->block)[lane] = 0;
        wsync();
        blocks++;
        blen = 0;
      }
      const uint32_t take = min(64u - blen, len - pos);
      if ((uint32_t)lane < take) S->block[blen + lane] = msg[pos + lane];
      wsync();
      blen += take;
      pos += take;
    }
  }
  if (lane == 0) {
    S->blen = blen;
    S->blocks = blocks;
    S->chunk = chunk;
    S->sp = sp;
  }
  wsync();
}

// Appends len bytes of msg (LDS) to S like update(), with the chunks the
// input completes compressed in parallel: the current block's bytes and msg
// are laid out from the current block's start in buf (LDS, 64-byte blocks,
// zero padded); every block but the last is compressed, quad q taking the
// blocks of the q-th chunk from the current one (its chaining value in
// registers), and the last block stays buffered, as the lazy host hasher
// leaves it.  Then the finished chunks' chaining values enter the CV stack in
// order.  At most 16 chunks (len < 15 KiB).
__device__ __forceinline__ void absorb(b3::State* S, const uint8_t* msg, uint32_t len, Work* w, uint8_t* buf,
                                       uint32_t (*cvs)[8]) {
  const int lane = threadIdx.x & 63, c = lane & 3, q = lane >> 2;
  if (len == 0) return;
  const uint32_t blen0 = S->blen, blocks0 = S->blocks;
  const uint64_t chunk0 = S->chunk;
  const uint32_t tot = blen0 + len, nbk = (tot + 63) / 64, ncomp = nbk - 1;
  for (uint32_t i = lane; i < 64 * nbk; i += 64) buf[i] = i < blen0 ? S->block[i] : i < tot ? msg[i - blen0] : (uint8_t)0;
  wsync();
  const uint32_t last = (blocks0 + ncomp) / 16;  // chunk (from the current one) of the buffered block
  if ((uint32_t)q <= last) {
    uint32_t ca = q == 0 ? S->cv[c] : S->key[c], cb = q == 0 ? S->cv[4 + c] : S->key[4 + c];
    // blocks j of chunk q: blocks0 + j in [16q, 16q + 16), j < ncomp
    const uint32_t jb = q == 0 ? 0u : 16u * q - blocks0;
    const uint32_t je = min(16u * (q + 1) - blocks0, ncomp);
    for (uint32_t j = jb; j < je; j++) {
      const uint32_t wb = (blocks0 + j) & 15u;
      compress_cv(ca, cb, reinterpret_cast<const uint32_t*>(buf + 64 * j), chunk0 + q, 64,
                  (wb == 0 ? b3::CHUNK_START : 0u) | (wb == 15 ? b3::CHUNK_END : 0u));
    }
    if ((lane & ~3) == 4 * q) {  // one quad per chunk writes
      if ((uint32_t)q < last) {
        cvs[q][c] = ca;
        cvs[q][4 + c] = cb;
      } else {
        S->cv[c] = ca;
        S->cv[4 + c] = cb;
      }
    }
  }
  wsync();
  int sp = S->sp;
  for (uint32_t k = 0; k < last; k++) {
    if (lane < 8) w->out[lane] = cvs[k][lane];
    wsync();
    uint64_t total = chunk0 + k + 1;
    while ((total & 1) == 0) {
      sp--;
      if (lane < 8) {
        w->pm[lane] = S->stack[sp][lane];
        w->pm[8 + lane] = w->out[lane];
      }
      wsync();
      compress(S->key, w->pm, 0, 64, b3::PARENT, w->out);
      total >>= 1;
    }
    if (lane < 8) S->stack[sp][lane] = w->out[lane];
    wsync();
    sp++;
  }
  if (lane < 16) reinterpret_cast<uint32_t*>(S->block)[lane] = reinterpret_cast<const uint32_t*>(buf + 64 * ncomp)[lane];
  if (lane == 0) {
    S->blen = tot - 64 * ncomp;
    S->blocks = blocks0 + ncomp - 16 * last;
    S->chunk = chunk0 + last;
    S->sp = sp;
  }
  wsync();
}

// The first 8 XOF bytes (seek 0) of S's root, little-endian (the host
// Hasher::finalize).
__device__ __forceinline__ uint64_t squeeze8(const b3::State* S, Work* w, uint32_t* ocv /*LDS, 8*/) {
  const int lane = threadIdx.x & 63;
  if (lane < 8) ocv[lane] = S->cv[lane];
  if (lane < 16) w->pm[lane] = reinterpret_cast<const uint32_t*>(S->block)[lane];
  wsync();
  uint64_t octr = S->chunk;
  uint32_t oblen = S->blen, ofl = (S->blocks == 0 ? b3::CHUNK_START : 0u) | b3::CHUNK_END;
  for (int s = S->sp - 1; s >= 0; s--) {
    compress(ocv, w->pm, octr, oblen, ofl, w->out);
    if (lane < 8) {
      w->pm[lane] = S->stack[s][lane];
      w->pm[8 + lane] = w->out[lane];
      ocv[lane] = S->key[lane];
    }
    wsync();
    octr = 0;
    oblen = 64;
    ofl = b3::PARENT;
  }
  compress(ocv, w->pm, 0, oblen, ofl | b3::ROOT, w->out);
  return (uint64_t)w->out[0] | ((uint64_t)w->out[1] << 32);
}

// StdRng::seed_from_u64 (PCG32 expansion into the ChaCha key, rng.hpp).
__device__ __forceinline__ void pcg_key(uint64_t state, uint32_t key[8]) {
  const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    state = state * MUL + INC;
    const uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
    const uint32_t rot = (uint32_t)(state >> 59);
    key[c] = (xs >> rot) | (xs << ((32 - rot) & 31));
  }
}

__device__ __forceinline__ uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
__device__ __forceinline__ void qr_mix(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
  a += b; d ^= a; d = rotl(d, 16);
  c += d; b ^= c; b = rotl(b, 12);
  a += b; d ^= a; d = rotl(d, 8);
  c += d; b ^= c; b = rotl(b, 7);
}

// The first 16 * nblocks words of the ChaCha12 keystream (counter from 0,
// stream 0) into ks (LDS): quad q computes block q (nblocks <= 16).
__device__ __forceinline__ void chacha12(const uint32_t key[8], int nblocks, uint32_t* ks) {
  const int lane = threadIdx.x & 63, c = lane & 3, q = lane >> 2;
  const uint32_t K[4] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
  const uint32_t s0 = c == 0 ? K[0] : c == 1 ? K[1] : c == 2 ? K[2] : K[3];
  const uint32_t s1 = c == 0 ? key[0] : c == 1 ? key[1] : c == 2 ? key[2] : key[3];
  const uint32_t s2 = c == 0 ? key[4] : c == 1 ? key[5] : c == 2 ? key[6] : key[7];
  const uint32_t s3 = c == 0 ? (uint32_t)q : 0u;  // 64-bit block counter q, stream 0
  uint32_t a = s0, b = s1, cc = s2, d = s3;
#pragma unroll
  for (int r = 0; r < 6; r++) {
    qr_mix(a, b, cc, d);
    b = qperm<Q1>(b);
    cc = qperm<Q2>(cc);
    d = qperm<Q3>(d);
    qr_mix(a, b, cc, d);
    b = qperm<Q3>(b);
    cc = qperm<Q2>(cc);
    d = qperm<Q1>(d);
  }
  if (q < nblocks) {
    ks[16 * q + c] = a + s0;
    ks[16 * q + 4 + c] = b + s1;
    ks[16 * q + 8 + c] = cc + s2;
    ks[16 * q + 12 + c] = d + s3;
  }
  wsync();
}

// LDS of one device Fiat-Shamir round.
template <class T> struct RoundLds {
  static constexpr int MSG = 40 + 3 * (1 + E * sizeof(LT<T>));  // round label, count, q, challenge label
  b3::State st;
  Work w;
  uint32_t ocv[8];
  uint32_t key[8];
  uint32_t ks[256];
  uint32_t cvs[16][8];
  uint8_t msg[MSG];
  __attribute__((aligned(16))) uint8_t buf[(MSG + 64 + 63) / 64 * 64];
};

// Label bytes "L" || u32 len || label at dst; returns the length.
__device__ __forceinline__ uint32_t put_label(uint8_t* dst, const char* label, uint32_t n, int lane) {
  for (uint32_t i = lane; i < 5 + n; i += 64) {
    uint8_t b;
    if (i == 0) b = 'L';
    else if (i < 5) b = (uint8_t)(n >> (8 * (i - 1)));
    else b = (uint8_t)label[i - 5];
    dst[i] = b;
  }
  return 5 + n;
}

// One round on the device transcript `gst` (global): absorb the round
// polynomial q (3 elements, global) under "sc/round", squeeze the
// "sc/challenge" sampler and write its 128 natural limbs to nat (LDS).  Run
// by wave 0 of the block; the caller synchronizes the block afterwards.
template <class T>
__device__ __forceinline__ void round_challenge(b3::State* gst, const LT<T>* q, RoundLds<T>& L, LT<T>* nat) {
  using Lm = LT<T>;
  const int lane = threadIdx.x & 63;
  constexpr int NBY = sizeof(Lm), EB = E * NBY;
  {
    const uint32_t* g = reinterpret_cast<const uint32_t*>(gst);
    uint32_t* s = reinterpret_cast<uint32_t*>(&L.st);
    for (int i = lane; i < (int)(sizeof(b3::State) / 4); i += 64) s[i] = g[i];
  }
  // "L" u32(8) "sc/round", u32(3), then 3 x ("E" || limbs LE)
  uint32_t n = put_label(L.msg, "sc/round", 8, lane);
  if (lane < 4) L.msg[n + lane] = (uint8_t)(3u >> (8 * lane));
  n += 4;
  const uint8_t* qb = reinterpret_cast<const uint8_t*>(q);
  for (int i = lane; i < 3 * (1 + EB); i += 64) {
    const int e = i / (1 + EB), o = i % (1 + EB);
    L.msg[n + i] = o == 0 ? (uint8_t)'E' : qb[e * EB + o - 1];
  }
  n += 3 * (1 + EB);
  n += put_label(L.msg + n, "sc/challenge", 12, lane);
  wsync();
  absorb(&L.st, L.msg, n, &L.w, L.buf, L.cvs);
  const uint64_t seed = squeeze8(&L.st, &L.w, L.ocv);
  if (lane < 9) L.msg[lane] = lane == 0 ? (uint8_t)'S' : (uint8_t)(seed >> (8 * (lane - 1)));
  wsync();
  update(&L.st, L.msg, 9, &L.w);
  {
    uint32_t* g = reinterpret_cast<uint32_t*>(gst);
    const uint32_t* s = reinterpret_cast<const uint32_t*>(&L.st);
    for (int i = lane; i < (int)(sizeof(b3::State) / 4); i += 64) g[i] = s[i];
  }
  uint32_t key[8];
  pcg_key(seed, key);
  constexpr int NW = T::K == 64 ? 2 * E : E;  // keystream words for 128 draws
  chacha12(key, NW / 16, L.ks);
  for (int i = lane; i < E; i += 64) {
    if constexpr (T::K == 64) nat[i] = (Lm)((uint64_t)L.ks[2 * i] | ((uint64_t)L.ks[2 * i + 1] << 32));
    else nat[i] = (Lm)L.ks[i];
  }
}

}  // namespace fsd
}  // namespace nenya

#include "mfma.hpp"

namespace nenya {

// Device transcript of a running sumcheck: the BLAKE3 state and the round
// polynomial that the lift kernel absorbs before drawing the challenge.
template <class T> struct FsArg {
  b3::State* st = nullptr;
  const LT<T>* q = nullptr;
};

// Round challenge drawn on the device and lifted (k = 64: no fragments).
template <class T>
__global__ void __launch_bounds__(256) k_lift_fs(FsArg<T> f, LT<T>* x, const LT<T>* frob_t, int squarings) {
  __shared__ __attribute__((aligned(16))) CT<T> sx[E];
  __shared__ __attribute__((aligned(16))) LT<T> nat[E];
  __shared__ fsd::RoundLds<T> fl;
  if (threadIdx.x < 64) fsd::round_challenge<T>(f.st, f.q, fl, nat);
  __syncthreads();
  teich_lift_block<T>(nat, x, frob_t, squarings, sx);
}

// Round challenge drawn on the device, lifted, and its M_r fragments emitted
// (k <= 32).
template <class T>
__global__ void __launch_bounds__(256) k_lift_mulmat_fs(FsArg<T> f, LT<T>* x, const LT<T>* frob_t, int squarings,
                                                        uint32_t* frag) {
  using C = CT<T>;
  __shared__ __attribute__((aligned(16))) C v[32][E];
  __shared__ __attribute__((aligned(16))) LT<T> nat[E];
  __shared__ fsd::RoundLds<T> fl;
  if (threadIdx.x < 64) fsd::round_challenge<T>(f.st, f.q, fl, nat);
  __syncthreads();
  teich_lift_block<T>(nat, x, frob_t, squarings, v[0]);
  __syncthreads();
  mulmat_chain<C>(v);
  mulmat_frag<T>(v, frag);
}

}  // namespace nenya
