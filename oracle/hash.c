/*
 * Oracle: BLAKE3, rand 0.8.5 StdRng (ChaCha12), the Teichmuller sampler and
 * the Fiat-Shamir transcript.  TEST INFRASTRUCTURE ONLY.
 *
 * Third-party algorithms restated (absent from the reference tree):
 *   blake3 1.8.4 (Cargo.lock:24-25): the published BLAKE3 spec, incremental
 *     hasher + XOF.  Pinned by the official vectors and libclang's
 *     llvm_blake3_* (tests/test_oracle_hash.py).
 *   rand 0.8.5 / rand_chacha 0.3.1 / rand_core 0.6.4 (Cargo.lock:374-397):
 *     StdRng = ChaCha12 (64-bit block counter in words 12-13, stream 0),
 *     BlockRng with a 64-word (4 block) buffer, next_u64 straddle rule,
 *     SeedableRng::seed_from_u64 via PCG32 expansion, Standard u8/u16 as
 *     truncated next_u32, usize gen_range(0..=i) by widening-multiply
 *     rejection with zone = (range << lz(range)) - 1.  The ChaCha block is
 *     pinned against libsodium's chacha20 at 20 rounds; the 12-round count,
 *     PCG seeding, buffering and gen_range are "parity unpinned".
 * Call sites followed: galois-ring/src/sampling.rs:48-68,
 *   snark-ring/src/transcript.rs:27-105, pcs-ring/src/raa.rs:156-169.
 */
#include "oracle.h"

/* ======================= BLAKE3 ======================= */
static const u32 B3_IV[8] = {0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
                             0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19};
static const int B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };

static inline u32 rotr(u32 x, int n) { return (x >> n) | (x << (32 - n)); }
static inline void b3g(u32 *s, int a, int b, int c, int d, u32 mx, u32 my) {
  s[a] = s[a] + s[b] + mx; s[d] = rotr(s[d] ^ s[a], 16);
  s[c] = s[c] + s[d];      s[b] = rotr(s[b] ^ s[c], 12);
  s[a] = s[a] + s[b] + my; s[d] = rotr(s[d] ^ s[a], 8);
  s[c] = s[c] + s[d];      s[b] = rotr(s[b] ^ s[c], 7);
}
static void b3_compress(const u32 cv[8], const u32 m_in[16], u64 counter, u32 blen, u32 flags, u32 out[16]) {
  u32 s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
               B3_IV[0], B3_IV[1], B3_IV[2], B3_IV[3], (u32)counter, (u32)(counter >> 32), blen, flags};
  u32 m[16], t[16];
  memcpy(m, m_in, sizeof m);
  for (int r = 0; r < 7; r++) {
    b3g(s, 0, 4, 8, 12, m[0], m[1]);  b3g(s, 1, 5, 9, 13, m[2], m[3]);
    b3g(s, 2, 6, 10, 14, m[4], m[5]); b3g(s, 3, 7, 11, 15, m[6], m[7]);
    b3g(s, 0, 5, 10, 15, m[8], m[9]); b3g(s, 1, 6, 11, 12, m[10], m[11]);
    b3g(s, 2, 7, 8, 13, m[12], m[13]); b3g(s, 3, 4, 9, 14, m[14], m[15]);
    if (r < 6) { for (int i = 0; i < 16; i++) t[i] = m[B3_PERM[i]]; memcpy(m, t, sizeof m); }
  }
  for (int i = 0; i < 8; i++) { out[i] = s[i] ^ s[i + 8]; out[i + 8] = s[i + 8] ^ cv[i]; }
}
static void words_le(const u8 *b, u32 *w, int n) {
  for (int i = 0; i < n; i++) w[i] = (u32)b[4 * i] | ((u32)b[4 * i + 1] << 8) | ((u32)b[4 * i + 2] << 16) | ((u32)b[4 * i + 3] << 24);
}

void b3_init(b3_t *h) {
  memset(h, 0, sizeof *h);
  memcpy(h->key, B3_IV, 32);
  memcpy(h->cv, B3_IV, 32);
}
static size_t cs_len(const b3_t *h) { return 64 * (size_t)h->blocks_compressed + h->block_len; }
static void cs_update(b3_t *h, const u8 *in, size_t len) {
  while (len) {
    if (h->block_len == 64) {
      u32 w[16], o[16];
      words_le(h->block, w, 16);
      b3_compress(h->cv, w, h->chunk_counter, 64, h->flags | (h->blocks_compressed == 0 ? CHUNK_START : 0), o);
      memcpy(h->cv, o, 32);
      h->blocks_compressed++;
      memset(h->block, 0, 64);
      h->block_len = 0;
    }
    size_t take = 64 - h->block_len;
    if (take > len) take = len;
    memcpy(h->block + h->block_len, in, take);
    h->block_len += (u8)take;
    in += take;
    len -= take;
  }
}
typedef struct { u32 cv[8]; u32 block[16]; u64 counter; u32 blen; u32 flags; } b3out_t;
static b3out_t cs_output(const b3_t *h) {
  b3out_t o;
  memcpy(o.cv, h->cv, 32);
  words_le(h->block, o.block, 16);
  o.counter = h->chunk_counter;
  o.blen = h->block_len;
  o.flags = h->flags | (h->blocks_compressed == 0 ? CHUNK_START : 0) | CHUNK_END;
  return o;
}
static void out_cv(const b3out_t *o, u32 cv[8]) { u32 t[16]; b3_compress(o->cv, o->block, o->counter, o->blen, o->flags, t); memcpy(cv, t, 32); }
static b3out_t parent_out(const u32 l[8], const u32 r[8], const u32 key[8], u32 flags) {
  b3out_t o;
  memcpy(o.cv, key, 32);
  memcpy(o.block, l, 32);
  memcpy(o.block + 8, r, 32);
  o.counter = 0; o.blen = 64; o.flags = flags | PARENT;
  return o;
}
void b3_update(b3_t *h, const void *data, size_t len) {
  const u8 *in = data;
  while (len) {
    if (cs_len(h) == 1024) {
      b3out_t o = cs_output(h);
      u32 cv[8];
      out_cv(&o, cv);
      u64 total = h->chunk_counter + 1;
      while ((total & 1) == 0) {
        h->cv_stack_len--;
        b3out_t p = parent_out(h->cv_stack[h->cv_stack_len], cv, h->key, h->flags);
        out_cv(&p, cv);
        total >>= 1;
      }
      memcpy(h->cv_stack[h->cv_stack_len++], cv, 32);
      u64 nc = h->chunk_counter + 1;
      memcpy(h->cv, h->key, 32);
      h->chunk_counter = nc;
      memset(h->block, 0, 64);
      h->block_len = 0;
      h->blocks_compressed = 0;
    }
    size_t want = 1024 - cs_len(h);
    size_t take = want < len ? want : len;
    cs_update(h, in, take);
    in += take;
    len -= take;
  }
}
void b3_finalize_seek(const b3_t *h, u64 seek, u8 *out, size_t out_len) {
  b3out_t o = cs_output(h);
  size_t rem = h->cv_stack_len;
  while (rem > 0) {
    rem--;
    u32 cv[8];
    out_cv(&o, cv);
    o = parent_out(h->cv_stack[rem], cv, h->key, h->flags);
  }
  u64 blk = seek / 64;
  size_t off = (size_t)(seek % 64);
  while (out_len) {
    u32 w[16];
    b3_compress(o.cv, o.block, blk, o.blen, o.flags | ROOT, w);
    u8 bytes[64];
    for (int i = 0; i < 16; i++) for (int j = 0; j < 4; j++) bytes[4 * i + j] = (u8)(w[i] >> (8 * j));
    size_t take = 64 - off;
    if (take > out_len) take = out_len;
    memcpy(out, bytes + off, take);
    out += take;
    out_len -= take;
    off = 0;
    blk++;
  }
}
void b3_hash(const void *data, size_t len, u8 out[32]) {
  b3_t h;
  b3_init(&h);
  b3_update(&h, data, len);
  b3_finalize_seek(&h, 0, out, 32);
}

/* ======================= ChaCha / rand ======================= */
static inline u32 rotl(u32 x, int n) { return (x << n) | (x >> (32 - n)); }
#define QR(a, b, c, d) a += b; d ^= a; d = rotl(d, 16); c += d; b ^= c; b = rotl(b, 12); \
                       a += b; d ^= a; d = rotl(d, 8);  c += d; b ^= c; b = rotl(b, 7);
void chacha_block(u32 out[16], const u32 key[8], u64 counter, u64 stream, int double_rounds) {
  u32 s[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574,
               key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
               (u32)counter, (u32)(counter >> 32), (u32)stream, (u32)(stream >> 32)};
  u32 x[16];
  memcpy(x, s, sizeof x);
  for (int r = 0; r < double_rounds; r++) {
    QR(x[0], x[4], x[8], x[12]); QR(x[1], x[5], x[9], x[13]);
    QR(x[2], x[6], x[10], x[14]); QR(x[3], x[7], x[11], x[15]);
    QR(x[0], x[5], x[10], x[15]); QR(x[1], x[6], x[11], x[12]);
    QR(x[2], x[7], x[8], x[13]); QR(x[3], x[4], x[9], x[14]);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}
void rng_from_seed(rng_t *r, const u8 seed[32]) {
  words_le(seed, r->key, 8);
  r->counter = 0;
  r->idx = 64;
  r->rounds = 6;
}
static void rng_refill(rng_t *r) {
  for (int b = 0; b < 4; b++) chacha_block(&r->buf[16 * b], r->key, r->counter + (u64)b, 0, r->rounds);
  r->counter += 4;
}
/* rand_core 0.6 SeedableRng::seed_from_u64: PCG32 output per 4-byte chunk */
void rng_seed_from_u64(rng_t *r, u64 state) {
  const u64 MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
  u8 seed[32];
  for (int c = 0; c < 8; c++) {
    state = state * MUL + INC;
    u64 s = state;
    u32 xs = (u32)(((s >> 18) ^ s) >> 27);
    u32 rot = (u32)(s >> 59);
    u32 x = (xs >> rot) | (xs << ((32 - rot) & 31));
    for (int j = 0; j < 4; j++) seed[4 * c + j] = (u8)(x >> (8 * j));
  }
  rng_from_seed(r, seed);
}
u32 rng_next_u32(rng_t *r) {
  if (r->idx >= 64) { rng_refill(r); r->idx = 0; }
  return r->buf[r->idx++];
}
u64 rng_next_u64(rng_t *r) {
  if (r->idx < 63) {
    u64 v = (u64)r->buf[r->idx] | ((u64)r->buf[r->idx + 1] << 32);
    r->idx += 2;
    return v;
  } else if (r->idx >= 64) {
    rng_refill(r);
    r->idx = 2;
    return (u64)r->buf[0] | ((u64)r->buf[1] << 32);
  } else {
    u64 x = r->buf[63];
    rng_refill(r);
    r->idx = 1;
    return ((u64)r->buf[0] << 32) | x;
  }
}
/* rand 0.8.5 UniformInt<usize>::sample_single_inclusive */
u64 rng_gen_range_incl(rng_t *r, u64 low, u64 high) {
  u64 range = high - low + 1;
  if (range == 0) return rng_next_u64(r);
  u64 zone = (range << __builtin_clzll(range)) - 1;
  for (;;) {
    u64 v = rng_next_u64(r);
    unsigned __int128 m = (unsigned __int128)v * range;
    u64 hi = (u64)(m >> 64), lo = (u64)m;
    if (lo <= zone) return low + hi;
  }
}
u64 rng_limb(rng_t *r) {
  if (g_k == 64) return rng_next_u64(r);
  return (u64)rng_next_u32(r) & g_mask;
}
void r2_sample(rng_t *r, gr *out) {
  for (int i = 0; i < E; i++) out->v[i] = rng_limb(r);
  teich_lift(out);
}

/* ======================= transcript ======================= */
static void h_u32(b3_t *h, u32 x) { u8 t[4] = {(u8)x, (u8)(x >> 8), (u8)(x >> 16), (u8)(x >> 24)}; b3_update(h, t, 4); }
void tr_new(tr_t *t, const char *label) {
  b3_init(&t->h);
  b3_update(&t->h, "snark-ring/v0", 13);
  h_u32(&t->h, (u32)strlen(label));
  b3_update(&t->h, label, strlen(label));
}
void tr_label(tr_t *t, const char *label) {
  b3_update(&t->h, "L", 1);
  h_u32(&t->h, (u32)strlen(label));
  b3_update(&t->h, label, strlen(label));
}
void tr_root(tr_t *t, const u8 root[32]) { b3_update(&t->h, "R", 1); b3_update(&t->h, root, 32); }
static void tr_gr(tr_t *t, const gr *x) {
  u8 tmp[E * 8];
  size_t o = 0;
  for (int i = 0; i < E; i++) for (int j = 0; j < g_nb; j++) tmp[o++] = (u8)(x->v[i] >> (8 * j));
  b3_update(&t->h, "E", 1);
  b3_update(&t->h, tmp, o);
}
void tr_gr_slice(tr_t *t, const char *label, const gr *xs, size_t n) {
  tr_label(t, label);
  h_u32(&t->h, (u32)n);
  for (size_t i = 0; i < n; i++) tr_gr(t, &xs[i]);
}
void tr_sampler(tr_t *t, const char *label, rng_t *out) {
  tr_label(t, label);
  u8 bytes[8];
  b3_finalize_seek(&t->h, 0, bytes, 8);
  u64 seed = 0;
  for (int i = 0; i < 8; i++) seed |= (u64)bytes[i] << (8 * i);
  b3_update(&t->h, "S", 1);
  b3_update(&t->h, bytes, 8);
  rng_seed_from_u64(out, seed);
}
void tr_challenge(tr_t *t, const char *label, gr *out) {
  rng_t r;
  tr_sampler(t, label, &r);
  r2_sample(&r, out);
}
void tr_target(tr_t *t, const char *label, size_t n, gr *out) {
  rng_t r;
  tr_sampler(t, label, &r);
  for (size_t i = 0; i < n; i++) r2_sample(&r, &out[i]);
}

/* piop-ring/src/fs.rs + transcript.rs:95-105 */
void ch_observe(ch_t *c, const gr *msg, size_t n) {
  if (c->lockstep) return;
  tr_gr_slice(c->tr, "sc/round", msg, n);
}
void ch_sample(ch_t *c, gr *out) {
  if (c->lockstep) { r2_sample(&c->rng, out); return; }
  tr_challenge(c->tr, "sc/challenge", out);
}
