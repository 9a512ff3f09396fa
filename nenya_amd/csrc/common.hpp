// Nenya MI355X prover: shared width traits, host-side GR(2^k,128) scalar
// arithmetic, error handling.
//
// A base-ring limb of Z/2^k is stored in its native width L (u8/u16/u32/u64)
// so that device memory bytes are exactly the little-endian serialization the
// reference hashes and emits (low k/8 bytes per limb, e.g.
// snark-ring/src/transcript.rs:48-56).  Arithmetic runs in the wider compute
// type C (u32 for k <= 32, u64 for k = 64): wrapping C arithmetic truncated to
// L is exact mod 2^k.  A GR(2^k,128) element is 128 limbs in the reference
// layout [beta power i][alpha power c] -> 4*i+c (galois-ring/src/gr128.rs:23-27).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nenya.h"

namespace nenya {

template <int K> struct W;
template <> struct W<8> { using L = uint8_t; using C = uint32_t; static constexpr int K = 8; };
template <> struct W<16> { using L = uint16_t; using C = uint32_t; static constexpr int K = 16; };
template <> struct W<32> { using L = uint32_t; using C = uint32_t; static constexpr int K = 32; };
template <> struct W<64> { using L = uint64_t; using C = uint64_t; static constexpr int K = 64; };

constexpr int E = 128;  // limbs per GR(2^k,128) element

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(x)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      throw ::nenya::Error(e_ == hipErrorOutOfMemory ? NENYA_E_OOM : NENYA_E_DEVICE,      \
                           std::string(#x) + ": " + hipGetErrorString(e_));               \
  } while (0)

// ---------------------------------------------------------------------------
// Host-side element arithmetic (scalars of the protocol: challenges, claims,
// round messages).  Values are exact, so any correct algorithm gives the same
// bytes as the reference's schoolbook (gr128.rs:215-296).
// ---------------------------------------------------------------------------
template <class T> struct Gr {
  typename T::L v[E];
  static Gr zero() { Gr g; std::memset(g.v, 0, sizeof g.v); return g; }
  static Gr one() { Gr g = zero(); g.v[0] = 1; return g; }
  bool operator==(const Gr& o) const { return std::memcmp(v, o.v, sizeof v) == 0; }
};

template <class T> inline Gr<T> operator+(const Gr<T>& a, const Gr<T>& b) {
  Gr<T> r;
  for (int i = 0; i < E; i++) r.v[i] = (typename T::L)(a.v[i] + b.v[i]);
  return r;
}
template <class T> inline Gr<T> operator-(const Gr<T>& a, const Gr<T>& b) {
  Gr<T> r;
  for (int i = 0; i < E; i++) r.v[i] = (typename T::L)(a.v[i] - b.v[i]);
  return r;
}

// R1 = (Z/2^k)[X]/(X^4+X+1) product from the 7 raw alpha-convolution terms.
template <class C> inline void g4_reduce(const C p[7], C o[4]) {
  o[0] = p[0] - p[4];
  o[1] = p[1] - p[4] - p[5];
  o[2] = p[2] - p[5] - p[6];
  o[3] = p[3] - p[6];
}
template <class C> inline void g4_raw(const C a[4], const C b[4], C p[7]) {
  p[0] = a[0] * b[0];
  p[1] = a[0] * b[1] + a[1] * b[0];
  p[2] = a[0] * b[2] + a[1] * b[1] + a[2] * b[0];
  p[3] = a[0] * b[3] + a[1] * b[2] + a[2] * b[1] + a[3] * b[0];
  p[4] = a[1] * b[3] + a[2] * b[2] + a[3] * b[1];
  p[5] = a[2] * b[3] + a[3] * b[2];
  p[6] = a[3] * b[3];
}

// Full product: raw beta convolution of 63 R1 coefficients, then
// Y^32 = -Y^3 - Y - alpha applied from the top (positions 62 down to 32).
template <class T> inline Gr<T> operator*(const Gr<T>& a, const Gr<T>& b) {
  using C = typename T::C;
  C raw[63][4];
  std::memset(raw, 0, sizeof raw);
  for (int i = 0; i < 32; i++) {
    C ai[4] = {a.v[4 * i], a.v[4 * i + 1], a.v[4 * i + 2], a.v[4 * i + 3]};
    bool nz = ai[0] | ai[1] | ai[2] | ai[3];
    if (!nz) continue;
    for (int j = 0; j < 32; j++) {
      C bj[4] = {b.v[4 * j], b.v[4 * j + 1], b.v[4 * j + 2], b.v[4 * j + 3]};
      C p[7], o[4];
      g4_raw(ai, bj, p);
      g4_reduce(p, o);
      for (int c = 0; c < 4; c++) raw[i + j][c] += o[c];
    }
  }
  for (int s = 62; s >= 32; s--) {
    C h[4] = {raw[s][0], raw[s][1], raw[s][2], raw[s][3]};
    int i = s - 32;
    for (int c = 0; c < 4; c++) { raw[i + 3][c] -= h[c]; raw[i + 1][c] -= h[c]; }
    // alpha*h = (-h3, h0-h3, h1, h2)
    raw[i][0] += h[3];
    raw[i][1] -= h[0] - h[3];
    raw[i][2] -= h[1];
    raw[i][3] -= h[2];
  }
  Gr<T> r;
  for (int i = 0; i < 32; i++)
    for (int c = 0; c < 4; c++) r.v[4 * i + c] = (typename T::L)raw[i][c];
  return r;
}

// Serialize limbs (low k/8 bytes each, little endian): the in-memory bytes of L.
template <class T> inline void put_gr(std::vector<uint8_t>& out, const Gr<T>& g) {
  const uint8_t* p = reinterpret_cast<const uint8_t*>(g.v);
  out.insert(out.end(), p, p + sizeof g.v);
}
inline void put_u64(std::vector<uint8_t>& out, uint64_t x) {
  for (int i = 0; i < 8; i++) out.push_back((uint8_t)(x >> (8 * i)));
}
inline void put_u32(std::vector<uint8_t>& out, uint32_t x) {
  for (int i = 0; i < 4; i++) out.push_back((uint8_t)(x >> (8 * i)));
}

}  // namespace nenya
