/*
 * Oracle: RAA code, Merkle tree and the two tensor PCSs (packed slot-form
 * pcs_ring::Pcs and generic snark_ring::R2Pcs).  TEST INFRASTRUCTURE ONLY.
 *
 * Follows: pcs-ring/src/raa.rs:26-169, merkle.rs:20-127,
 *          pcs.rs:45-293 (setup/shape), 489-1168 (commit/open/verify);
 *          snark-ring/src/r2_pcs.rs:41-619.
 * Pinned by: t=1060 KAT (pcs.rs:1298), RAA linearity and encode/pack and
 * encode/rotation commutation (raa.rs:189-301), open value == direct MLE
 * evaluation (pcs-ring lib.rs:55-90), rotated == real commitment
 * (pcs.rs:1193-1228), stacked == weighted sum (pcs.rs:1234-1277),
 * proof_size_bytes == serialized length (pcs.rs:1341-1356), verifier
 * acceptance / tamper rejection.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>

/* ---------------- RAA (raa.rs) ---------------- */
void raa_setup(raa_t *p, size_t k, size_t q, u64 seed) {
  size_t n = k * q;
  rng_t r;
  rng_seed_from_u64(&r, seed);
  p->k = k; p->q = q; p->n = n;
  p->perm1 = xmalloc(n * sizeof(u32));
  p->perm2 = xmalloc(n * sizeof(u32));
  p->diag1 = xmalloc(n * sizeof(u64));
  p->diag2 = xmalloc(n * sizeof(u64));
  u32 *perms[2] = {p->perm1, p->perm2};
  for (int w = 0; w < 2; w++) {
    u32 *pp = perms[w];
    for (size_t i = 0; i < n; i++) pp[i] = (u32)i;
    for (size_t i = n - 1; i >= 1; i--) {
      size_t j = (size_t)rng_gen_range_incl(&r, 0, i);
      u32 t = pp[i]; pp[i] = pp[j]; pp[j] = t;
    }
  }
  for (size_t i = 0; i < n; i++) p->diag1[i] = mk(rng_next_u64(&r) | 1);
  for (size_t i = 0; i < n; i++) p->diag2[i] = mk(rng_next_u64(&r) | 1);
}
void raa_free(raa_t *p) { free(p->perm1); free(p->perm2); free(p->diag1); free(p->diag2); memset(p, 0, sizeof *p); }

/* generic: `w` limbs per entry (32 packed, 128 for R2) */
static void raa_encode_w(const raa_t *p, const u64 *msg, u64 *out, int w) {
  size_t n = p->n, q = p->q;
  u64 *t = xmalloc(n * w * sizeof(u64));
  for (size_t i = 0; i < n; i++) {
    const u64 *src = &msg[(p->perm1[i] / q) * w]; /* y[src] = msg[src/q] */
    for (int s = 0; s < w; s++) t[i * w + s] = mk(src[s] * p->diag1[i]);
  }
  for (size_t i = 1; i < n; i++)
    for (int s = 0; s < w; s++) t[i * w + s] = mk(t[i * w + s] + t[(i - 1) * w + s]);
  for (size_t i = 0; i < n; i++) {
    const u64 *src = &t[(size_t)p->perm2[i] * w];
    for (int s = 0; s < w; s++) out[i * w + s] = mk(src[s] * p->diag2[i]);
  }
  for (size_t i = 1; i < n; i++)
    for (int s = 0; s < w; s++) out[i * w + s] = mk(out[i * w + s] + out[(i - 1) * w + s]);
  free(t);
}
void raa_encode(const raa_t *p, const gr *msg, gr *out) { raa_encode_w(p, (const u64 *)msg, (u64 *)out, E); }
void raa_encode_packed(const raa_t *p, const u64 *msg, u64 *out) { raa_encode_w(p, msg, out, 32); }

/* ---------------- Merkle (merkle.rs) ---------------- */
void merkle_build(merkle_t *m, u8 (*leaves)[32], size_t nl) {
  size_t n = 1;
  while (n < nl) n <<= 1;
  m->n = n;
  m->nodes = xcalloc(2 * n - 1, 32);
  memcpy(m->nodes, leaves, nl * 32);
  size_t start = 0, len = n;
  while (len > 1) {
    size_t next = start + len;
#pragma omp parallel for if (len >= 512)
    for (size_t i = 0; i < len / 2; i++) {
      u8 pair[64];
      memcpy(pair, m->nodes[start + 2 * i], 32);
      memcpy(pair + 32, m->nodes[start + 2 * i + 1], 32);
      b3_hash(pair, 64, m->nodes[next + i]);
    }
    start = next;
    len /= 2;
  }
}
void merkle_open(const merkle_t *m, size_t j, u8 (*path)[32]) {
  size_t idx = j, start = 0, len = m->n, k = 0;
  while (len > 1) {
    memcpy(path[k++], m->nodes[start + (idx ^ 1)], 32);
    idx /= 2;
    start += len;
    len /= 2;
  }
}
int merkle_verify(const u8 leaf[32], size_t idx, const u8 (*path)[32], size_t plen, const u8 root[32]) {
  u8 cur[32], pair[64];
  memcpy(cur, leaf, 32);
  for (size_t i = 0; i < plen; i++) {
    if ((idx & 1) == 0) { memcpy(pair, cur, 32); memcpy(pair + 32, path[i], 32); }
    else { memcpy(pair, path[i], 32); memcpy(pair + 32, cur, 32); }
    b3_hash(pair, 64, cur);
    idx /= 2;
  }
  return memcmp(cur, root, 32) == 0;
}
void merkle_free(merkle_t *m) { free(m->nodes); m->nodes = NULL; }

/* ---------------- parameters (pcs.rs:198-293, r2_pcs.rs:57-62) ---------------- */
size_t recommended_query_size(double delta, size_t lambda) {
  double lam = (double)lambda;
  double bt = 1.0 - delta / 3.0, bo = 1.0 - 2.0 * delta / 3.0;
  size_t t = (size_t)ceil(lam / (-log2(bt)));
  double target = exp2(-lam);
  /* Rust f64::powi: repeated multiplication (LLVM powi) */
  for (;;) {
    double pt = 1.0, po = 1.0;
    for (size_t i = 0; i < t; i++) { pt *= bt; po *= bo; }
    if (pt + 2.0 * po > target) t++; else break;
  }
  return t;
}
size_t proof_size_bytes_for(size_t k1, size_t k2, size_t n, size_t t, size_t mb, size_t cb) {
  size_t np = 1;
  while (np < n) np <<= 1;
  size_t path_len = (size_t)__builtin_ctzll(np);
  return 32 + k2 * mb + t * (8 + k1 * cb) + t * (8 + path_len * 32) + t * 4;
}
static void shape_for(size_t nv, size_t q, size_t t, size_t mb, size_t cb, int floor_t, size_t *k1o, size_t *k2o, int *found) {
  size_t best = 0;
  *found = 0;
  for (size_t lk1 = 1; lk1 < nv; lk1++) {
    size_t k1 = (size_t)1 << lk1, k2 = (size_t)1 << (nv - lk1);
    if (floor_t && q * k2 < t) continue;
    size_t sz = proof_size_bytes_for(k1, k2, q * k2, t, mb, cb);
    if (!*found || sz < best) { best = sz; *k1o = k1; *k2o = k2; *found = 1; }
  }
}
void pcs_new(pcs_t *p, size_t k1, size_t k2, size_t q, size_t ell, u64 seed, int packed) {
  p->k1 = k1; p->k2 = k2; p->ell = ell; p->packed = packed;
  raa_setup(&p->raa, k2, q, seed);
}
void pcs_free(pcs_t *p) { raa_free(&p->raa); }
void pcs_setup(pcs_t *p, size_t nv, size_t q, double delta, size_t lambda, u64 seed) {
  size_t t = recommended_query_size(delta, lambda), k1 = 0, k2 = 0;
  int f;
  shape_for(nv, q, t, 128 * g_nb, 32 * g_nb, 0, &k1, &k2, &f);
  pcs_new(p, k1, k2, q, t, seed, 1);
}
void pcs_setup_batched(pcs_t *p, size_t nv, size_t q, double delta, size_t lambda, u64 seed, size_t nopen) {
  size_t t = recommended_query_size(delta, lambda), k1 = 0, k2 = 0;
  int f;
  shape_for(nv, q, t, nopen * 128 * g_nb, 32 * g_nb, 1, &k1, &k2, &f);
  if (!f) shape_for(nv, q, t, 128 * g_nb, 32 * g_nb, 0, &k1, &k2, &f);
  pcs_new(p, k1, k2, q, t, seed, 1);
}
void pcs_setup_stacked(pcs_t *p, size_t nv, size_t q, double delta, size_t lambda, u64 seed, size_t stack) {
  size_t t = recommended_query_size(delta, lambda), k1 = 0, k2 = 0;
  int f;
  shape_for(nv, q, t, 128 * g_nb, stack * 32 * g_nb, 1, &k1, &k2, &f);
  if (!f) shape_for(nv, q, t, 128 * g_nb, stack * 32 * g_nb, 0, &k1, &k2, &f);
  pcs_new(p, k1, k2, q, t, seed, 1);
}
void r2pcs_setup(pcs_t *p, size_t nv, size_t q, double delta, size_t lambda, u64 seed) {
  size_t t = recommended_query_size(delta, lambda), k1 = 0, k2 = 0;
  int f;
  shape_for(nv, q, t, 128 * g_nb, 128 * g_nb, 0, &k1, &k2, &f);
  pcs_new(p, k1, k2, q, t, seed, 0);
}

/* ---------------- commit ---------------- */
static void hash_columns(const pcs_t *p, comm_t *c) {
  size_t n = p->raa.n, w = p->packed ? 32 : E;
  u8 (*leaves)[32] = xmalloc(n * 32);
#pragma omp parallel
  {
    u8 *tmp = xmalloc(c->rows * w * 8);
#pragma omp for
    for (size_t j = 0; j < n; j++) {
      size_t o = 0;
      for (size_t i = 0; i < c->rows; i++) {
        const u64 *ent = &c->encoded[(i * n + j) * w];
        for (size_t s = 0; s < w; s++) for (int b = 0; b < g_nb; b++) tmp[o++] = (u8)(ent[s] >> (8 * b));
      }
      b3_hash(tmp, o, leaves[j]);
    }
    free(tmp);
  }
  merkle_build(&c->tree, leaves, n);
  memcpy(c->root, c->tree.nodes[2 * c->tree.n - 2], 32);
  free(leaves);
}
void pcs_commit_stacked(const pcs_t *p, const u64 *const *msgs, size_t s, comm_t *c) {
  size_t n = p->raa.n, k1 = p->k1, k2 = p->k2, rows = s * k1;
  memset(c, 0, sizeof *c);
  c->rows = rows; c->s = s;
  c->encoded = xmalloc(rows * n * 32 * sizeof(u64));
  c->slots = xmalloc(rows * k2 * 32 * sizeof(u64));
  c->message = xmalloc(rows * k2 * sizeof(gr));
#pragma omp parallel for
  for (size_t i = 0; i < rows; i++) {
    const u64 *row = &msgs[i / k1][(i % k1) * k2 * 32];
    memcpy(&c->slots[i * k2 * 32], row, k2 * 32 * sizeof(u64));
    raa_encode_packed(&p->raa, row, &c->encoded[i * n * 32]);
  }
  hash_columns(p, c);
#pragma omp parallel for
  for (size_t i = 0; i < rows * k2; i++) pack_r2(&c->message[i], &c->slots[i * 32]);
}
void pcs_commit(const pcs_t *p, const u64 *slots, comm_t *c) { pcs_commit_stacked(p, &slots, 1, c); }
void r2pcs_commit_stacked(const pcs_t *p, const gr *const *msgs, size_t s, comm_t *c) {
  size_t n = p->raa.n, k1 = p->k1, k2 = p->k2, rows = s * k1;
  memset(c, 0, sizeof *c);
  c->rows = rows; c->s = s;
  c->encoded = xmalloc(rows * n * E * sizeof(u64));
  c->message = xmalloc(rows * k2 * sizeof(gr));
#pragma omp parallel for
  for (size_t i = 0; i < rows; i++) {
    const gr *row = &msgs[i / k1][(i % k1) * k2];
    memcpy(&c->message[i * k2], row, k2 * sizeof(gr));
    raa_encode(&p->raa, row, (gr *)&c->encoded[i * n * E]);
  }
  hash_columns(p, c);
}
void r2pcs_commit(const pcs_t *p, const gr *evals, comm_t *c) { r2pcs_commit_stacked(p, &evals, 1, c); }
void comm_free(comm_t *c) { free(c->encoded); free(c->message); free(c->slots); merkle_free(&c->tree); memset(c, 0, sizeof *c); }

/* ---------------- open ---------------- */
static void split_eq(const pcs_t *p, const gr *z, gr **e0, gr **e1) {
  size_t lk1 = (size_t)__builtin_ctzll(p->k1), lk2 = (size_t)__builtin_ctzll(p->k2);
  *e0 = xmalloc(p->k1 * sizeof(gr));
  *e1 = xmalloc(p->k2 * sizeof(gr));
  build_eq_table(z, lk1, *e0);
  build_eq_table(z + lk1, lk2, *e1);
}
static gr *stacked_e0(const gr *e0, size_t k1, const gr *w, size_t s) {
  gr *o = xmalloc(s * k1 * sizeof(gr));
  for (size_t t = 0; t < s; t++)
    for (size_t i = 0; i < k1; i++) {
      if (w) gr_mul(&o[t * k1 + i], &w[t], &e0[i]);
      else o[t * k1 + i] = e0[i];
    }
  return o;
}
static void derive_queries(int packed, const u8 root[32], const gr *mbar, size_t cnt, size_t n, size_t t, u32 *out) {
  b3_t h;
  b3_init(&h);
  const char *dom = packed ? "pcs-ring/fs/queries" : "snark-ring/r2-pcs/fs/queries";
  b3_update(&h, dom, strlen(dom));
  b3_update(&h, root, 32);
  u8 tmp[E * 8];
  for (size_t i = 0; i < cnt; i++) {
    size_t o = 0;
    for (int l = 0; l < E; l++) for (int b = 0; b < g_nb; b++) tmp[o++] = (u8)(mbar[i].v[l] >> (8 * b));
    b3_update(&h, tmp, o);
  }
  size_t np = 1;
  while (np < n) np <<= 1;
  u32 mask = (u32)(np - 1);
  u64 pos = 0;
  size_t got = 0;
  while (got < t) {
    u8 b[4];
    b3_finalize_seek(&h, pos, b, 4);
    pos += 4;
    u32 v = ((u32)b[0] | ((u32)b[1] << 8) | ((u32)b[2] << 16) | ((u32)b[3] << 24)) & mask;
    if (v < (u32)n) out[got++] = v;
  }
}
static void open_columns(const pcs_t *p, const comm_t *c, open_t *o) {
  size_t n = p->raa.n, w = p->packed ? 32 : E;
  o->col_len = c->rows;
  o->col_elem = (int)w;
  o->cols = xmalloc(o->t * c->rows * w * sizeof(u64));
  o->plen = (size_t)__builtin_ctzll(c->tree.n);
  o->paths = xmalloc((o->t * o->plen + 1) * 32);
#pragma omp parallel for
  for (size_t q = 0; q < o->t; q++) {
    size_t j = o->idx[q];
    for (size_t i = 0; i < c->rows; i++)
      memcpy(&o->cols[(q * c->rows + i) * w], &c->encoded[(i * n + j) * w], w * sizeof(u64));
    merkle_open(&c->tree, j, &o->paths[q * o->plen]);
  }
}
/* Pcs::open / open_stacked (pcs.rs:547-591, 734-779) and R2Pcs::open /
 * open_stacked (r2_pcs.rs:297-338, 512-559).  w == NULL: single message. */
void pcs_open_weighted(const pcs_t *p, const comm_t *c, const gr *z, const gr *w, gr *y, open_t *o) {
  gr *e0, *e1;
  split_eq(p, z, &e0, &e1);
  size_t k2 = p->k2, rows = c->rows;
  gr *e0s = stacked_e0(e0, p->k1, w, c->s);
  memset(o, 0, sizeof *o);
  o->nmbar = 1; o->k2 = k2;
  o->mbar = xmalloc(k2 * sizeof(gr));
  gr t;
#pragma omp parallel for
  for (size_t j = 0; j < k2; j++) {
    gr acc, tj; gr_zero(&acc);
    for (size_t i = 0; i < rows; i++) { gr_mul(&tj, &e0s[i], &c->message[i * k2 + j]); gr_add(&acc, &acc, &tj); }
    o->mbar[j] = acc;
  }
  gr_zero(y);
  for (size_t j = 0; j < k2; j++) { gr_mul(&t, &o->mbar[j], &e1[j]); gr_add(y, y, &t); }
  o->t = p->ell;
  o->idx = xmalloc(o->t * sizeof(u32));
  derive_queries(p->packed, c->root, o->mbar, k2, p->raa.n, o->t, o->idx);
  open_columns(p, c, o);
  free(e0); free(e1); free(e0s);
}
/* pcs.rs:973-1040 */
void pcs_rotated_open_batch(const pcs_t *p, const comm_t *c, const gr *z, size_t nl, gr *ys, open_t *o) {
  gr *e0, *e1;
  split_eq(p, z, &e0, &e1);
  size_t k1 = p->k1, k2 = p->k2;
  memset(o, 0, sizeof *o);
  o->nmbar = nl; o->k2 = k2;
  o->mbar = xmalloc(nl * k2 * sizeof(gr));
  gr t;
#pragma omp parallel for collapse(2)
  for (size_t l = 0; l < nl; l++)
    for (size_t j = 0; j < k2; j++) {
      gr acc, pk, tj; gr_zero(&acc);
      for (size_t i = 0; i < k1; i++) {
        const u64 *sl = &c->slots[(i * k2 + j) * 32];
        u64 rot[32];
        for (int s = 0; s < 32; s++) rot[s] = sl[(s + l) % 32];
        pack_r2(&pk, rot);
        gr_mul(&tj, &e0[i], &pk);
        gr_add(&acc, &acc, &tj);
      }
      o->mbar[l * k2 + j] = acc;
    }
  for (size_t l = 0; l < nl; l++) {
    gr_zero(&ys[l]);
    for (size_t j = 0; j < k2; j++) { gr_mul(&t, &o->mbar[l * k2 + j], &e1[j]); gr_add(&ys[l], &ys[l], &t); }
  }
  o->t = p->ell;
  o->idx = xmalloc(o->t * sizeof(u32));
  derive_queries(1, c->root, o->mbar, nl * k2, p->raa.n, o->t, o->idx);
  open_columns(p, c, o);
  free(e0); free(e1);
}
void open_free(open_t *o) { free(o->mbar); free(o->idx); free(o->cols); free(o->paths); memset(o, 0, sizeof *o); }

/* OpeningProof / R2OpeningProof / RotatedOpeningBatch ::serialize
 * (pcs.rs:357-406, 426-471; r2_pcs.rs:109-160) */
void open_serialize(buf_t *b, const open_t *o, int batch) {
  if (batch) {
    buf_u64(b, o->nmbar);
    for (size_t l = 0; l < o->nmbar; l++) {
      buf_u64(b, o->k2);
      for (size_t j = 0; j < o->k2; j++) buf_gr(b, &o->mbar[l * o->k2 + j]);
    }
  } else {
    buf_u64(b, o->k2);
    for (size_t j = 0; j < o->k2; j++) buf_gr(b, &o->mbar[j]);
  }
  buf_u64(b, o->t);
  for (size_t q = 0; q < o->t; q++) buf_u32(b, o->idx[q]);
  buf_u64(b, o->t);
  for (size_t q = 0; q < o->t; q++) {
    buf_u64(b, o->col_len);
    for (size_t i = 0; i < o->col_len * (size_t)o->col_elem; i++) buf_limb(b, o->cols[q * o->col_len * o->col_elem + i]);
  }
  buf_u64(b, o->t);
  for (size_t q = 0; q < o->t; q++) {
    buf_u64(b, o->plen);
    for (size_t i = 0; i < o->plen; i++) buf_put(b, o->paths[q * o->plen + i], 32);
  }
}

/* ---------------- verify ---------------- */
static void hash_col(const open_t *o, size_t q, u8 out[32]) {
  size_t w = o->col_elem, L = o->col_len;
  u8 *tmp = xmalloc(L * w * 8 + 1);
  size_t off = 0;
  for (size_t i = 0; i < L * w; i++) for (int b = 0; b < g_nb; b++) tmp[off++] = (u8)(o->cols[q * L * w + i] >> (8 * b));
  b3_hash(tmp, off, out);
  free(tmp);
}
/* acc_j = sum_i e0s[i] * lift(col[i]) with slot rotation `rot` (packed only) */
static void col_combine(const pcs_t *p, const open_t *o, size_t q, const gr *e0s, size_t rows, size_t rot, gr *acc) {
  const consts_t *C = consts();
  gr t;
  gr_zero(acc);
  if (p->packed) {
    for (int k = 0; k < 32; k++) {
      gr inner; gr_zero(&inner);
      size_t src = (k + rot) % 32;
      for (size_t i = 0; i < rows; i++) { gr_mul_gr1(&t, &e0s[i], o->cols[(q * rows + i) * 32 + src]); gr_add(&inner, &inner, &t); }
      gr_mul(&t, &C->basis[k], &inner);
      gr_add(acc, acc, &t);
    }
  } else {
    for (size_t i = 0; i < rows; i++) { gr_mul(&t, &e0s[i], (const gr *)&o->cols[(q * rows + i) * E]); gr_add(acc, acc, &t); }
  }
}
int pcs_verify_weighted(const pcs_t *p, const u8 root[32], const gr *z, const gr *w, size_t s, const gr *val, const open_t *o) {
  size_t n = p->raa.n, k2 = p->k2, rows = s * p->k1;
  if (o->k2 != k2 || o->nmbar != 1 || o->t != p->ell || o->col_len != rows) return 0;
  gr *e0, *e1;
  split_eq(p, z, &e0, &e1);
  gr *e0s = stacked_e0(e0, p->k1, w, s);
  int ok = 1;
  gr y, t;
  gr_zero(&y);
  for (size_t j = 0; j < k2; j++) { gr_mul(&t, &o->mbar[j], &e1[j]); gr_add(&y, &y, &t); }
  if (!gr_eq(&y, val)) ok = 0;
  u32 *exp = xmalloc(p->ell * sizeof(u32));
  if (ok) {
    derive_queries(p->packed, root, o->mbar, k2, n, p->ell, exp);
    if (memcmp(exp, o->idx, p->ell * sizeof(u32))) ok = 0;
  }
  gr *ubar = NULL;
  if (ok) { ubar = xmalloc(n * sizeof(gr)); raa_encode(&p->raa, o->mbar, ubar); }
  for (size_t q = 0; ok && q < p->ell; q++) {
    size_t j = o->idx[q];
    if (j >= n) { ok = 0; break; }
    gr acc;
    col_combine(p, o, q, e0s, rows, 0, &acc);
    if (!gr_eq(&acc, &ubar[j])) { ok = 0; break; }
    u8 leaf[32];
    hash_col(o, q, leaf);
    if (!merkle_verify(leaf, j, (const u8(*)[32])&o->paths[q * o->plen], o->plen, root)) ok = 0;
  }
  free(ubar); free(exp); free(e0); free(e1); free(e0s);
  return ok;
}
int pcs_rotated_verify_batch(const pcs_t *p, const u8 root[32], const gr *z, const gr *ys, size_t nl, const open_t *o) {
  size_t n = p->raa.n, k2 = p->k2, k1 = p->k1;
  if (o->nmbar != nl || o->k2 != k2 || o->t != p->ell || o->col_len != k1) return 0;
  gr *e0, *e1;
  split_eq(p, z, &e0, &e1);
  int ok = 1;
  gr t;
  for (size_t l = 0; ok && l < nl; l++) {
    gr y; gr_zero(&y);
    for (size_t j = 0; j < k2; j++) { gr_mul(&t, &o->mbar[l * k2 + j], &e1[j]); gr_add(&y, &y, &t); }
    if (!gr_eq(&y, &ys[l])) ok = 0;
  }
  u32 *exp = xmalloc(p->ell * sizeof(u32));
  if (ok) {
    derive_queries(1, root, o->mbar, nl * k2, n, p->ell, exp);
    if (memcmp(exp, o->idx, p->ell * sizeof(u32))) ok = 0;
  }
  for (size_t q = 0; ok && q < p->ell; q++) {
    if (o->idx[q] >= n) { ok = 0; break; }
    u8 leaf[32];
    hash_col(o, q, leaf);
    if (!merkle_verify(leaf, o->idx[q], (const u8(*)[32])&o->paths[q * o->plen], o->plen, root)) ok = 0;
  }
  gr *ubar = xmalloc(n * sizeof(gr));
  for (size_t l = 0; ok && l < nl; l++) {
    raa_encode(&p->raa, &o->mbar[l * k2], ubar);
    for (size_t q = 0; ok && q < p->ell; q++) {
      gr acc;
      col_combine(p, o, q, e0, k1, l, &acc);
      if (!gr_eq(&acc, &ubar[o->idx[q]])) ok = 0;
    }
  }
  free(ubar); free(exp); free(e0); free(e1);
  return ok;
}
/* r2_pcs.rs:182-188 */
int columns_in_kernel(const open_t *o) {
  for (size_t q = 0; q < o->t; q++)
    for (size_t i = 0; i < o->col_len; i++)
      if (!in_kernel((const gr *)&o->cols[(q * o->col_len + i) * E])) return 0;
  return 1;
}
