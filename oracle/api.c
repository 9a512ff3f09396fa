/*
 * Oracle C entry points (loaded with ctypes by tests/ and bench.py's
 * cpu_baseline leg only).  TEST INFRASTRUCTURE ONLY.
 * All limb buffers are uint64_t, one limb per element, reduced mod 2^k.
 */
#include "oracle.h"
#include <stdlib.h>

#define API __attribute__((visibility("default")))

API void orc_free(void *p) { free(p); }
/* threads the prover loops run on (OpenMP; OMP_NUM_THREADS sets it) */
API int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
API void orc_set_width(int k) { or_set_width(k); }

/* ---- ring primitives ---- */
API void orc_gr_mul(int k, const u64 *a, const u64 *b, u64 *o) { or_set_width(k); gr_mul((gr *)o, (const gr *)a, (const gr *)b); }
API void orc_gr_mul_gr1(int k, const u64 *a, u64 s, u64 *o) { or_set_width(k); gr_mul_gr1((gr *)o, (const gr *)a, s); }
API void orc_gr_mul_gr4(int k, const u64 *a, const u64 *s, u64 *o) { or_set_width(k); gr_mul_gr4((gr *)o, (const gr *)a, (const g4 *)s); }
API void orc_g4_mul(int k, const u64 *a, const u64 *b, u64 *o) { or_set_width(k); g4_mul((g4 *)o, (const g4 *)a, (const g4 *)b); }
API u64 orc_inv_2adic(int k, u64 x) { or_set_width(k); return inv_2adic(x); }
API void orc_lift(int k, u64 *x) { or_set_width(k); teich_lift((gr *)x); }
API void orc_t1(int k, u64 *out /*64*/) { or_set_width(k); compute_t1((g4 *)out); }
API void orc_pack(int k, const u64 *slots, size_t cnt, u64 *out) {
  or_set_width(k);
  for (size_t i = 0; i < cnt; i++) pack_r2((gr *)&out[i * E], &slots[i * 32]);
}
API void orc_psi(int k, const u64 *x, size_t cnt, u64 *out) {
  or_set_width(k);
  for (size_t i = 0; i < cnt; i++) psi_r2(&out[i * 32], (const gr *)&x[i * E]);
}
API int orc_in_kernel(int k, const u64 *x) { or_set_width(k); return in_kernel((const gr *)x); }
API void orc_quotients(int k, const u64 *a, const u64 *b, u64 *q1, u64 *q2) {
  or_set_width(k);
  mul_with_quotients_only((gr *)q1, (gr *)q2, a, b);
}
/* which: 0 l_beta[16], 1 v_beta, 2 v_alpha(4), 3 basis[32], 4 q_beta[16][16], 5 inv15(1) */
API void orc_consts(int k, int which, u64 *out) {
  or_set_width(k);
  const consts_t *C = consts();
  switch (which) {
    case 0: memcpy(out, C->l_beta, sizeof C->l_beta); break;
    case 1: memcpy(out, &C->v_beta, sizeof C->v_beta); break;
    case 2: memcpy(out, &C->v_alpha, sizeof C->v_alpha); break;
    case 3: memcpy(out, C->basis, sizeof C->basis); break;
    case 4: memcpy(out, C->q_beta, sizeof C->q_beta); break;
    case 5: out[0] = C->inv15; break;
  }
}
API void orc_eq_table(int k, const u64 *r, size_t n, u64 *out) { or_set_width(k); build_eq_table((const gr *)r, n, (gr *)out); }
API void orc_fold(int k, u64 *t, size_t len, const u64 *r) { or_set_width(k); mle_fold((gr *)t, len, (const gr *)r); }

/* ---- hashing / rng ---- */
API void orc_blake3(const u8 *d, size_t len, u8 *out, size_t olen) {
  b3_t h; b3_init(&h); b3_update(&h, d, len); b3_finalize_seek(&h, 0, out, olen);
}
API void orc_chacha_block(const u32 *key, u64 counter, u64 stream, int drounds, u32 *out) { chacha_block(out, key, counter, stream, drounds); }
/* kind 0: next_u32 (as u64), 1: next_u64, 2: gen_range(0..=arg) repeated, 3: limbs of width k, 4: interleaved u32,u64,u64,... */
API void orc_rng_stream(int k, u64 seed, int kind, u64 arg, size_t cnt, u64 *out) {
  or_set_width(k);
  rng_t r;
  rng_seed_from_u64(&r, seed);
  for (size_t i = 0; i < cnt; i++) {
    switch (kind) {
      case 0: out[i] = rng_next_u32(&r); break;
      case 1: out[i] = rng_next_u64(&r); break;
      case 2: out[i] = rng_gen_range_incl(&r, 0, arg); break;
      case 3: out[i] = rng_limb(&r); break;
      case 4: out[i] = (i % 3 == 0) ? rng_next_u32(&r) : rng_next_u64(&r); break;
    }
  }
}
API void orc_r2_sample(int k, u64 seed, size_t cnt, u64 *out) {
  or_set_width(k);
  rng_t r;
  rng_seed_from_u64(&r, seed);
  for (size_t i = 0; i < cnt; i++) r2_sample(&r, (gr *)&out[i * E]);
}
/* transcript smoke: new(label) then challenge_target(tlabel, cnt) */
API void orc_transcript_target(int k, const char *label, const u8 *root, const char *tlabel, size_t cnt, u64 *out) {
  or_set_width(k);
  tr_t t;
  tr_new(&t, label);
  if (root) tr_root(&t, root);
  tr_target(&t, tlabel, cnt, (gr *)out);
}

/* ---- RAA / PCS ---- */
API size_t orc_query_size(double delta, size_t lambda) { return recommended_query_size(delta, lambda); }
API void orc_raa_setup(int k, size_t kk, size_t q, u64 seed, u32 *perm1, u32 *perm2, u64 *d1, u64 *d2) {
  or_set_width(k);
  raa_t p;
  raa_setup(&p, kk, q, seed);
  memcpy(perm1, p.perm1, p.n * 4); memcpy(perm2, p.perm2, p.n * 4);
  memcpy(d1, p.diag1, p.n * 8); memcpy(d2, p.diag2, p.n * 8);
  raa_free(&p);
}
API void orc_raa_encode(int k, size_t kk, size_t q, u64 seed, const u64 *msg, u64 *out, int packed) {
  or_set_width(k);
  raa_t p;
  raa_setup(&p, kk, q, seed);
  if (packed) raa_encode_packed(&p, msg, out); else raa_encode(&p, (const gr *)msg, (gr *)out);
  raa_free(&p);
}
/* kind 0 Pcs::setup, 1 setup_batched(nopen=arg), 2 setup_stacked(stack=arg), 3 R2Pcs::setup */
API void orc_pcs_shape(int k, size_t nv, int kind, size_t arg, size_t *k1, size_t *k2, size_t *ell) {
  or_set_width(k);
  pcs_t p;
  switch (kind) {
    case 0: pcs_setup(&p, nv, 4, 0.19, 100, 1); break;
    case 1: pcs_setup_batched(&p, nv, 4, 0.19, 100, 1, arg); break;
    case 2: pcs_setup_stacked(&p, nv, 4, 0.19, 100, 1, arg); break;
    default: r2pcs_setup(&p, nv, 4, 0.19, 100, 1); break;
  }
  *k1 = p.k1; *k2 = p.k2; *ell = p.ell;
  pcs_free(&p);
}
API size_t orc_proof_size_bytes_for(size_t k1, size_t k2, size_t n, size_t t, size_t mb, size_t cb) {
  return proof_size_bytes_for(k1, k2, n, t, mb, cb);
}
/* Commit + open + verify round trip on an explicit (insecure) shape.
 * mode 0: packed single, 1: packed stacked(s), 2: r2 single, 3: r2 stacked(s), 4: rotated batch (nl layers)
 * msgs: s*k1*k2 entries (32 limbs packed / 128 limbs r2).  Returns the opened value(s) in `vals`,
 * verify result, and (tamper != 0) flips one column limb before verifying. */
API int orc_pcs_roundtrip(int k, size_t k1, size_t k2, size_t ell, u64 seed, int mode, size_t s, const u64 *msgs,
                          const u64 *z, const u64 *w, int tamper, u64 *vals, u8 *root_out) {
  or_set_width(k);
  pcs_t p;
  int packed = (mode == 0 || mode == 1 || mode == 4);
  pcs_new(&p, k1, k2, 4, ell, seed, packed);
  comm_t c;
  size_t per = k1 * k2 * (packed ? 32 : E);
  size_t ns = (mode == 4) ? 1 : s; /* rotated batch: one message, s layers */
  if (packed) {
    const u64 **ms = xmalloc(ns * sizeof(u64 *));
    for (size_t i = 0; i < ns; i++) ms[i] = msgs + i * per;
    pcs_commit_stacked(&p, ms, ns, &c);
    free(ms);
  } else {
    const gr **ms = xmalloc(s * sizeof(gr *));
    for (size_t i = 0; i < s; i++) ms[i] = (const gr *)(msgs + i * per);
    r2pcs_commit_stacked(&p, ms, s, &c);
    free(ms);
  }
  memcpy(root_out, c.root, 32);
  open_t o;
  int ok;
  if (mode == 4) {
    pcs_rotated_open_batch(&p, &c, (const gr *)z, s, (gr *)vals, &o);
    if (tamper) o.cols[0] = mk(o.cols[0] + 1);
    ok = pcs_rotated_verify_batch(&p, c.root, (const gr *)z, (const gr *)vals, s, &o);
  } else {
    const gr *wt = (mode == 1 || mode == 3) ? (const gr *)w : NULL;
    pcs_open_weighted(&p, &c, (const gr *)z, wt, (gr *)vals, &o);
    if (tamper) o.cols[0] = mk(o.cols[0] + 1);
    ok = pcs_verify_weighted(&p, c.root, (const gr *)z, wt, (mode == 1 || mode == 3) ? s : 1, (const gr *)vals, &o);
  }
  open_free(&o);
  comm_free(&c);
  pcs_free(&p);
  return ok;
}

/* ---- sumchecks with the lock-step sampler (piop-ring tests) ----
 * which: 0 mult (Gruen), 1 mult sub, 2 linear simd, 3 linear general.
 * tables: packed consecutively (a,b[,c,e] / f,g / 32 rows then 32 rhos) of 2^nv entries each.
 * tamper: add 1 to round 0 message coefficient 0 before verifying.
 * Returns verifier status (0 ok, 1+round round failure, -1 final failure); claimed sum out. */
API int orc_sumcheck(int k, int which, size_t nv, const u64 *tabs, const u64 *rt, u64 seed, int tamper, u64 *claim_out) {
  or_set_width(k);
  size_t N = (size_t)1 << nv;
  ch_t pc = {0}, vc = {0};
  pc.lockstep = vc.lockstep = 1;
  rng_seed_from_u64(&pc.rng, seed);
  rng_seed_from_u64(&vc.rng, seed);
  gr *chal = xmalloc((nv + 1) * sizeof(gr)), *vchal = xmalloc((nv + 1) * sizeof(gr));
  sc_t pf;
  int res;
  const gr *T = (const gr *)tabs;
  if (which == 0 || which == 1) {
    gr *eq = xmalloc(N * sizeof(gr));
    build_eq_table((const gr *)rt, nv, eq);
    gr claim, t;
    gr_zero(&claim);
    for (size_t h = 0; h < N; h++) {
      gr_mul(&t, &T[h], &T[N + h]);
      if (which == 1) { gr_sub(&t, &t, &T[2 * N + h]); gr_sub(&t, &t, &T[3 * N + h]); }
      gr_mul(&t, &eq[h], &t);
      gr_add(&claim, &claim, &t);
    }
    memcpy(claim_out, &claim, sizeof claim);
    if (which == 0) prove_mult_check(T, T + N, nv, (const gr *)rt, &claim, &pc, &pf, chal);
    else prove_mult_check_sub(T, T + N, T + 2 * N, T + 3 * N, nv, (const gr *)rt, &pc, &pf, chal);
    if (tamper) pf.rounds[0].v[0] = mk(pf.rounds[0].v[0] + 1);
    res = which == 0 ? verify_mult_check(&pf, (const gr *)rt, nv, &vc, vchal) : verify_mult_check_sub(&pf, (const gr *)rt, nv, &vc, vchal);
    free(eq);
  } else if (which == 2) {
    gr *f = xmalloc(N * sizeof(gr)), *g = xmalloc(N * sizeof(gr));
    memcpy(f, T, N * sizeof(gr)); memcpy(g, T + N, N * sizeof(gr));
    gr claim, t;
    gr_zero(&claim);
    for (size_t h = 0; h < N; h++) { gr_mul(&t, &f[h], &g[h]); gr_add(&claim, &claim, &t); }
    memcpy(claim_out, &claim, sizeof claim);
    prove_linear_simd(f, g, nv, &claim, &pc, &pf, chal);
    if (tamper) pf.rounds[0].v[0] = mk(pf.rounds[0].v[0] + 1);
    res = verify_linear_simd(&pf, &vc, vchal);
    free(f); free(g);
  } else {
    gr *rows[32], *rhos[32];
    gr claim, t;
    gr_zero(&claim);
    for (int l = 0; l < 32; l++) {
      rows[l] = xmalloc(N * sizeof(gr)); rhos[l] = xmalloc(N * sizeof(gr));
      memcpy(rows[l], T + l * N, N * sizeof(gr));
      memcpy(rhos[l], T + (32 + l) * N, N * sizeof(gr));
      for (size_t h = 0; h < N; h++) { gr_mul(&t, &rows[l][h], &rhos[l][h]); gr_add(&claim, &claim, &t); }
    }
    memcpy(claim_out, &claim, sizeof claim);
    prove_linear_general(rows, rhos, nv, &claim, &pc, &pf, chal);
    if (tamper) pf.rounds[0].v[0] = mk(pf.rounds[0].v[0] + 1);
    res = verify_linear_general(&pf, &vc, vchal);
    for (int l = 0; l < 32; l++) { free(rows[l]); free(rhos[l]); }
  }
  sc_free(&pf);
  free(chal); free(vchal);
  return res;
}

/* ---- SNARKs ---- */
static coo_t mkcoo(size_t rows, size_t cols, const u32 *r, const u32 *c, const u64 *v, size_t nnz) {
  coo_t m = {rows, cols, nnz, (u32 *)r, (u32 *)c, (u64 *)v};
  return m;
}
API int orc_simd_prove(int k, int log_n, int log_m,
                       const u32 *ar, const u32 *ac, const u64 *av, size_t an,
                       const u32 *br, const u32 *bc, const u64 *bv, size_t bn,
                       const u32 *cr, const u32 *cc, const u64 *cv, size_t cn,
                       const u64 *wit, u64 seed_w, u64 seed_q, u8 **out, size_t *len) {
  or_set_width(k);
  size_t n = (size_t)1 << log_n, m = (size_t)1 << log_m;
  coo_t A = mkcoo(n, m, ar, ac, av, an), B = mkcoo(n, m, br, bc, bv, bn), C = mkcoo(n, m, cr, cc, cv, cn);
  buf_t b = {0};
  int rc = simd_prove(log_n, log_m, &A, &B, &C, wit, seed_w, seed_q, &b);
  *out = b.p;
  *len = b.n;
  return rc;
}
API int orc_simd_verify(int k, int log_n, int log_m,
                        const u32 *ar, const u32 *ac, const u64 *av, size_t an,
                        const u32 *br, const u32 *bc, const u64 *bv, size_t bn,
                        const u32 *cr, const u32 *cc, const u64 *cv, size_t cn,
                        u64 seed_w, u64 seed_q, const u8 *proof, size_t len) {
  or_set_width(k);
  size_t n = (size_t)1 << log_n, m = (size_t)1 << log_m;
  coo_t A = mkcoo(n, m, ar, ac, av, an), B = mkcoo(n, m, br, bc, bv, bn), C = mkcoo(n, m, cr, cc, cv, cn);
  return simd_verify(log_n, log_m, &A, &B, &C, seed_w, seed_q, proof, len);
}
API int orc_general_prove(int k, int log_b,
                          const u32 *ar, const u32 *ac, const u64 *av, size_t an,
                          const u32 *br, const u32 *bc, const u64 *bv, size_t bn,
                          const u32 *cr, const u32 *cc, const u64 *cv, size_t cn,
                          const u64 *w, u64 seed_w, u64 seed_e, u8 **out, size_t *len) {
  or_set_width(k);
  size_t n = (size_t)32 << log_b;
  coo_t A = mkcoo(n, n, ar, ac, av, an), B = mkcoo(n, n, br, bc, bv, bn), C = mkcoo(n, n, cr, cc, cv, cn);
  const coo_t *m3[3] = {&A, &B, &C};
  gdiag_t g;
  gdiag_build(&g, m3, n);
  buf_t b = {0};
  int rc = general_prove(log_b, &g, w, seed_w, seed_e, &b);
  gdiag_free(&g);
  *out = b.p;
  *len = b.n;
  return rc;
}
API int orc_general_verify(int k, int log_b,
                           const u32 *ar, const u32 *ac, const u64 *av, size_t an,
                           const u32 *br, const u32 *bc, const u64 *bv, size_t bn,
                           const u32 *cr, const u32 *cc, const u64 *cv, size_t cn,
                           u64 seed_w, u64 seed_e, const u8 *proof, size_t len) {
  or_set_width(k);
  size_t n = (size_t)32 << log_b;
  coo_t A = mkcoo(n, n, ar, ac, av, an), B = mkcoo(n, n, br, bc, bv, bn), C = mkcoo(n, n, cr, cc, cv, cn);
  const coo_t *m3[3] = {&A, &B, &C};
  gdiag_t g;
  gdiag_build(&g, m3, n);
  int rc = general_verify(log_b, &g, seed_w, seed_e, proof, len);
  gdiag_free(&g);
  return rc;
}

/* MIMD: each matrix set is the 32 instance matrices concatenated, instance k
 * holding entries off[k] .. off[k+1]-1. */
static void mimd_sets(size_t n, size_t m, const size_t *off[3], const u32 *r[3], const u32 *c[3], const u64 *v[3],
                      coo_t store[3][32], const coo_t *ptrs[3][32]) {
  for (int w = 0; w < 3; w++)
    for (int k = 0; k < 32; k++) {
      store[w][k] = mkcoo(n, m, r[w] + off[w][k], c[w] + off[w][k], v[w] + off[w][k], off[w][k + 1] - off[w][k]);
      ptrs[w][k] = &store[w][k];
    }
}
API int orc_mimd_prove(int k, int log_n, int log_m,
                       const size_t *aoff, const u32 *ar, const u32 *ac, const u64 *av,
                       const size_t *boff, const u32 *br, const u32 *bc, const u64 *bv,
                       const size_t *coff, const u32 *cr, const u32 *cc, const u64 *cv,
                       const u64 *wit, u64 seed_w, u64 seed_e, u8 **out, size_t *len) {
  or_set_width(k);
  size_t n = (size_t)1 << log_n, m = (size_t)1 << log_m;
  const size_t *off[3] = {aoff, boff, coff};
  const u32 *r[3] = {ar, br, cr}, *c[3] = {ac, bc, cc};
  const u64 *v[3] = {av, bv, cv};
  coo_t store[3][32];
  const coo_t *ptrs[3][32];
  mimd_sets(n, m, off, r, c, v, store, ptrs);
  const coo_t *const *sets[3] = {ptrs[0], ptrs[1], ptrs[2]};
  mmats_t mm;
  mimd_pack(&mm, n, m, sets);
  buf_t b = {0};
  int rc = mimd_prove(log_n, log_m, &mm, wit, seed_w, seed_e, &b);
  mimd_free(&mm);
  *out = b.p;
  *len = b.n;
  return rc;
}
API int orc_mimd_verify(int k, int log_n, int log_m,
                        const size_t *aoff, const u32 *ar, const u32 *ac, const u64 *av,
                        const size_t *boff, const u32 *br, const u32 *bc, const u64 *bv,
                        const size_t *coff, const u32 *cr, const u32 *cc, const u64 *cv,
                        u64 seed_w, u64 seed_e, const u8 *proof, size_t len) {
  or_set_width(k);
  size_t n = (size_t)1 << log_n, m = (size_t)1 << log_m;
  const size_t *off[3] = {aoff, boff, coff};
  const u32 *r[3] = {ar, br, cr}, *c[3] = {ac, bc, cc};
  const u64 *v[3] = {av, bv, cv};
  coo_t store[3][32];
  const coo_t *ptrs[3][32];
  mimd_sets(n, m, off, r, c, v, store, ptrs);
  const coo_t *const *sets[3] = {ptrs[0], ptrs[1], ptrs[2]};
  mmats_t mm;
  mimd_pack(&mm, n, m, sets);
  int rc = mimd_verify(log_n, log_m, &mm, seed_w, seed_e, proof, len);
  mimd_free(&mm);
  return rc;
}
