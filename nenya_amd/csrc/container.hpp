// Versioned proof container: a 72-byte header in front of the serialized
// proof (the reference's serialize() bytes, e.g. snark-ring/src/general.rs:
// 120-150, simd.rs:150-174, mimd.rs:128-158), so stored or transmitted
// proofs say which model, width, shape and PCS seeds they belong to and
// carry a BLAKE3 digest of the bytes.  The reference has no container (it
// serializes the proof structs only).  Layout, little-endian:
//    0  4  magic "NNYA"
//    4  2  version (1)
//    6  1  model: 0 SIMD, 1 General, 2 MIMD
//    7  1  k (8, 16, 32, 64)
//    8  1  log_n (SIMD, MIMD) or log_blocks (General)
//    9  1  log_m (SIMD, MIMD) or 0 (General)
//   10  6  zero
//   16  8  seed_w
//   24  8  seed_q (SIMD) or seed_e (General, MIMD)
//   32  8  payload length
//   40 32  BLAKE3 of the payload
//   72  -  payload
#pragma once
#include "snark.hpp"

namespace nenya {
namespace container {

constexpr size_t HEADER = 72;
constexpr uint16_t VERSION = 1;

struct Info {
  int model = 0, k = 0, shape0 = 0, shape1 = 0;
  uint64_t seed_w = 0, seed_e = 0, len = 0;
};

inline void put64(uint8_t* d, uint64_t x) {
  for (int i = 0; i < 8; i++) d[i] = (uint8_t)(x >> (8 * i));
}
inline uint64_t get64(const uint8_t* d) {
  uint64_t x = 0;
  for (int i = 0; i < 8; i++) x |= (uint64_t)d[i] << (8 * i);
  return x;
}
inline void digest(const uint8_t* p, size_t n, uint8_t out[32]) {
  b3::Hasher h;
  h.update(p, n);
  h.finalize(0, out, 32);
}
inline bool is_container(const uint8_t* c, size_t len) {
  return c && len >= 4 && c[0] == 'N' && c[1] == 'N' && c[2] == 'Y' && c[3] == 'A';
}

// Writes header + payload into out (HEADER + n bytes).
inline void wrap(const ParamsBase& p, const uint8_t* payload, size_t n, uint8_t* out) {
  std::memset(out, 0, HEADER);
  std::memcpy(out, "NNYA", 4);
  out[4] = (uint8_t)VERSION;
  out[5] = (uint8_t)(VERSION >> 8);
  out[6] = (uint8_t)p.kind;
  out[7] = (uint8_t)p.k;
  out[8] = (uint8_t)p.shape[0];
  out[9] = (uint8_t)p.shape[1];
  put64(out + 16, p.seeds[0]);
  put64(out + 24, p.seeds[1]);
  put64(out + 32, n);
  digest(payload, n, out + 40);
  std::memcpy(out + HEADER, payload, n);
}

// A well-formed container whose digest matches: fills info and the payload
// range and returns true.
inline bool parse(const uint8_t* c, size_t len, Info& info, const uint8_t** payload) {
  if (!is_container(c, len) || len < HEADER) return false;
  if ((uint16_t)(c[4] | (c[5] << 8)) != VERSION) return false;
  for (int i = 10; i < 16; i++)
    if (c[i]) return false;
  info.model = c[6];
  info.k = c[7];
  info.shape0 = c[8];
  info.shape1 = c[9];
  info.seed_w = get64(c + 16);
  info.seed_e = get64(c + 24);
  info.len = get64(c + 32);
  if (info.len != len - HEADER) return false;
  uint8_t d[32];
  digest(c + HEADER, info.len, d);
  if (std::memcmp(d, c + 40, 32) != 0) return false;
  *payload = c + HEADER;
  return true;
}

// For a verifier of params p: raw proofs pass unchanged; a container is
// replaced by its payload when it is intact and made for params of p's
// model, width, shape and seeds.  False means "does not parse" (verdict 8).
inline bool unwrap_for(const ParamsBase& p, const uint8_t*& proof, size_t& len) {
  if (!is_container(proof, len)) return true;
  Info in;
  const uint8_t* pl = nullptr;
  if (!parse(proof, len, in, &pl)) return false;
  if (in.model != p.kind || in.k != p.k || in.shape0 != p.shape[0] || in.shape1 != p.shape[1] ||
      in.seed_w != p.seeds[0] || in.seed_e != p.seeds[1])
    return false;
  proof = pl;
  len = (size_t)in.len;
  return true;
}

}  // namespace container
}  // namespace nenya
