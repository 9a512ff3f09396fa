/*
 * Oracle: end-to-end SIMD, MIMD and General provers/verifiers, proof
 * (de)serialization, the MIMD union-support packing and the
 * general-circuit diagonal packing.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Follows: reduction/src/reduce.rs:33-104 (simd_to_packed),
 *          reduce.rs:372-438, 618-651 (diagonal layers), 670-744
 *          (general_reduce_witness), matrix.rs:73-154;
 *          reduce.rs:119-141, 249-273, 287-345 (MIMD packing and
 *          witness reduction);
 *          snark-ring/src/mimd.rs:196-219, 242-397 (prove), 422-551 (verify);
 *          snark-ring/src/simd.rs:220-373 (prove), 391-522 (verify),
 *          528-580 (row builders), 742-748 (finals);
 *          snark-ring/src/general.rs:179-271 (helpers), 354-359,
 *          374-537 (prove), 565-698 (verify); serialization per
 *          simd.rs:150-174, general.rs:120-150, mult_check.rs:127-172,
 *          398-424, linear_check_simd.rs:87-131,
 *          linear_check_general.rs:142-190.
 * Pinned by: packed is_satisfied after reduction (instance.rs:64-79),
 * the quotient identity phi(a)phi(b) = phi(a.b) + q1 v_a + q2 v_b
 * (implementation.rs:165-203), honest proofs accepted and tampered proofs
 * rejected with the reference's error variant (end_to_end_*.rs).
 */
#include "oracle.h"
#include <stdlib.h>

#define SECURE_DELTA 0.19
#define SECURE_LAMBDA 100
#define SECURE_Q 4

static u64 *matvec_r0(const coo_t *m, const u64 *w) {
  u64 *o = xcalloc(m->rows, sizeof(u64));
  for (size_t k = 0; k < m->nnz; k++) o[m->r[k]] = mk(o[m->r[k]] + m->v[k] * w[m->c[k]]);
  return o;
}
static gr *matvec_r2(const coo_t *m, const gr *w) { /* matrix.rs:93-105 */
  gr *o = xcalloc(m->rows, sizeof(gr));
  gr t;
  for (size_t k = 0; k < m->nnz; k++) { gr_mul_gr1(&t, &w[m->c[k]], m->v[k]); gr_add(&o[m->r[k]], &o[m->r[k]], &t); }
  return o;
}
static gr *accumulate_row(const coo_t *m, const gr *eq, size_t cols) { /* simd.rs:570-580 */
  gr *o = xcalloc(cols, sizeof(gr));
  gr t;
  for (size_t k = 0; k < m->nnz; k++) { gr_mul_gr1(&t, &eq[m->r[k]], m->v[k]); gr_add(&o[m->c[k]], &o[m->c[k]], &t); }
  return o;
}
static void inner(const gr *a, const gr *b, size_t n, gr *o) {
  gr s;
  gr_zero(&s);
#pragma omp parallel for if (n >= 256) reduction(gr_plus : s)
  for (size_t i = 0; i < n; i++) { gr t; gr_mul(&t, &a[i], &b[i]); gr_add(&s, &s, &t); }
  *o = s;
}

static void ser_mult(buf_t *b, const sc_t *s, int with_claim) {
  if (with_claim) buf_gr(b, &s->claimed);
  buf_u64(b, s->n);
  for (size_t i = 0; i < 3 * s->n; i++) buf_gr(b, &s->rounds[i]);
  for (size_t i = 0; i < s->nfin; i++) buf_gr(b, &s->fin[i]);
}
static void ser_lin_general(buf_t *b, const sc_t *s) {
  buf_gr(b, &s->claimed);
  buf_u64(b, s->n);
  for (size_t i = 0; i < 3 * s->n; i++) buf_gr(b, &s->rounds[i]);
  for (int l = 0; l < 32; l++) buf_gr(b, &s->fin[l]);
  for (int l = 0; l < 32; l++) buf_gr(b, &s->fin2[l]);
}

/* ========================= SIMD ========================= */
static void simd_params(int log_n, int log_m, u64 seed_w, u64 seed_q, pcs_t *pw, pcs_t *pq) {
  pcs_setup(pw, log_m, SECURE_Q, SECURE_DELTA, SECURE_LAMBDA, seed_w);
  r2pcs_setup(pq, log_n, SECURE_Q, SECURE_DELTA, SECURE_LAMBDA, seed_q);
}

int simd_prove(int log_n, int log_m, const coo_t *A, const coo_t *B, const coo_t *C, const u64 *wit, u64 seed_w,
               u64 seed_q, buf_t *out) {
  size_t n = (size_t)1 << log_n, m = (size_t)1 << log_m;
  const consts_t *K = consts();
  pcs_t pw, pq;
  simd_params(log_n, log_m, seed_w, seed_q, &pw, &pq);
  /* 1. simd_to_packed */
  gr *w_phi = xmalloc(m * sizeof(gr));
  u64 *w_slots = xmalloc(m * 32 * sizeof(u64));
#pragma omp parallel for
  for (size_t j = 0; j < m; j++) {
    for (int k = 0; k < 32; k++) w_slots[j * 32 + k] = wit[(size_t)k * m + j];
    pack_r2(&w_phi[j], &w_slots[j * 32]);
  }
  u64 *aw0[32], *bw0[32];
#pragma omp parallel for
  for (int k = 0; k < 32; k++) { aw0[k] = matvec_r0(A, wit + (size_t)k * m); bw0[k] = matvec_r0(B, wit + (size_t)k * m); }
  gr *e = xmalloc(n * sizeof(gr));
#pragma omp parallel for
  for (size_t i = 0; i < n; i++) {
    u64 x[32], y[32];
    for (int k = 0; k < 32; k++) { x[k] = aw0[k][i]; y[k] = bw0[k][i]; }
    gr q1, q2, t1, t2;
    mul_with_quotients_only(&q1, &q2, x, y);
    gr_mul_gr4(&t1, &q1, &K->v_alpha);
    gr_mul(&t2, &q2, &K->v_beta);
    gr_add(&e[i], &t1, &t2);
  }
  for (int k = 0; k < 32; k++) { free(aw0[k]); free(bw0[k]); }
  /* 2. commits */
  comm_t cw, ce;
  pcs_commit(&pw, w_slots, &cw);
  r2pcs_commit(&pq, e, &ce);
  /* 3. transcript */
  tr_t tr;
  tr_new(&tr, "simd");
  tr_root(&tr, cw.root);
  tr_root(&tr, ce.root);
  gr *rt = xmalloc(log_n * sizeof(gr));
  tr_target(&tr, "mult_check_target", log_n, rt);
  /* 4. matvecs */
  gr *aw = matvec_r2(A, w_phi), *bw = matvec_r2(B, w_phi), *cwv = matvec_r2(C, w_phi);
  /* 5. claim */
  gr *eq_r = xmalloc(n * sizeof(gr));
  build_eq_table(rt, log_n, eq_r);
  gr cw_at_r, e_at_r, claim_mult;
  inner(eq_r, cwv, n, &cw_at_r);
  inner(eq_r, e, n, &e_at_r);
  gr_add(&claim_mult, &cw_at_r, &e_at_r);
  tr_gr_slice(&tr, "claim_mult", &claim_mult, 1);
  /* 6. mult check */
  tr_label(&tr, "mult_check_sumcheck");
  ch_t ch = {0};
  ch.tr = &tr;
  sc_t mp, lp;
  gr *s1 = xmalloc(log_n * sizeof(gr));
  prove_mult_check(aw, bw, log_n, rt, &claim_mult, &ch, &mp, s1);
  tr_gr_slice(&tr, "mult/finals", mp.fin, 2);
  /* 7. gamma */
  gr gam, gam2, t;
  tr_challenge(&tr, "linear_check_gamma", &gam);
  gr_mul(&gam2, &gam, &gam);
  /* 8. combined row */
  gr *eq_s1 = xmalloc(n * sizeof(gr));
  build_eq_table(s1, log_n, eq_s1);
  gr *ra = accumulate_row(A, eq_s1, m), *rb = accumulate_row(B, eq_s1, m), *rc = accumulate_row(C, eq_r, m);
  for (size_t h = 0; h < m; h++) {
    gr_mul(&t, &gam, &rb[h]); gr_add(&ra[h], &ra[h], &t);
    gr_mul(&t, &gam2, &rc[h]); gr_add(&ra[h], &ra[h], &t);
  }
  /* 9. claim_lin */
  gr claim_lin;
  gr_mul(&t, &gam, &mp.fin[1]); gr_add(&claim_lin, &mp.fin[0], &t);
  gr_mul(&t, &gam2, &cw_at_r); gr_add(&claim_lin, &claim_lin, &t);
  tr_gr_slice(&tr, "claim_lin", &claim_lin, 1);
  /* 10. linear check (w_phi folded in place) */
  tr_label(&tr, "linear_check_sumcheck");
  gr *s2 = xmalloc(log_m * sizeof(gr));
  prove_linear_simd(ra, w_phi, log_m, &claim_lin, &ch, &lp, s2);
  tr_gr_slice(&tr, "lin/finals", lp.fin, 2);
  /* 11. opens */
  open_t ow, oe;
  gr yw, ye;
  pcs_open_weighted(&pw, &cw, s2, NULL, &yw, &ow);
  pcs_open_weighted(&pq, &ce, rt, NULL, &ye, &oe);
  /* serialize */
  buf_put(out, cw.root, 32);
  buf_put(out, ce.root, 32);
  ser_mult(out, &mp, 1);
  ser_mult(out, &lp, 1);
  open_serialize(out, &ow, 0);
  open_serialize(out, &oe, 0);
  int ok = gr_eq(&yw, &lp.fin[1]) && gr_eq(&ye, &e_at_r);
  open_free(&ow); open_free(&oe); sc_free(&mp); sc_free(&lp);
  free(s1); free(s2); free(eq_s1); free(ra); free(rb); free(rc); free(eq_r);
  free(aw); free(bw); free(cwv); free(rt); free(e); free(w_phi); free(w_slots);
  comm_free(&cw); comm_free(&ce); pcs_free(&pw); pcs_free(&pq);
  return ok ? 0 : -1;
}

/* ---------------- proof reader ---------------- */
typedef struct { const u8 *p; size_t n, o; int bad; } rd_t;
static const u8 *rd_take(rd_t *r, size_t n) {
  if (r->bad || r->o + n > r->n) { r->bad = 1; return NULL; }
  const u8 *p = r->p + r->o;
  r->o += n;
  return p;
}
static u64 rd_u64(rd_t *r) { const u8 *p = rd_take(r, 8); u64 x = 0; if (p) for (int i = 0; i < 8; i++) x |= (u64)p[i] << (8 * i); return x; }
static u32 rd_u32(rd_t *r) { const u8 *p = rd_take(r, 4); u32 x = 0; if (p) for (int i = 0; i < 4; i++) x |= (u32)p[i] << (8 * i); return x; }
static u64 rd_limb(rd_t *r) { const u8 *p = rd_take(r, g_nb); u64 x = 0; if (p) for (int i = 0; i < g_nb; i++) x |= (u64)p[i] << (8 * i); return x; }
static void rd_gr(rd_t *r, gr *x) { for (int i = 0; i < E; i++) x->v[i] = rd_limb(r); }
static int rd_sane(rd_t *r, u64 cnt, size_t unit) { if (r->bad || cnt > (r->n - r->o) / (unit ? unit : 1) + 1) { r->bad = 1; return 0; } return 1; }
static void rd_mult(rd_t *r, sc_t *s, int with_claim, size_t nfin) {
  memset(s, 0, sizeof *s);
  if (with_claim) rd_gr(r, &s->claimed);
  u64 n = rd_u64(r);
  if (!rd_sane(r, n, 3 * E * g_nb)) return;
  s->n = n;
  s->rounds = xmalloc((3 * n + 1) * sizeof(gr));
  for (size_t i = 0; i < 3 * n; i++) rd_gr(r, &s->rounds[i]);
  s->nfin = nfin;
  for (size_t i = 0; i < nfin; i++) rd_gr(r, &s->fin[i]);
}
static void rd_lin_general(rd_t *r, sc_t *s) {
  memset(s, 0, sizeof *s);
  rd_gr(r, &s->claimed);
  u64 n = rd_u64(r);
  if (!rd_sane(r, n, 3 * E * g_nb)) return;
  s->n = n;
  s->rounds = xmalloc((3 * n + 1) * sizeof(gr));
  for (size_t i = 0; i < 3 * n; i++) rd_gr(r, &s->rounds[i]);
  s->nfin = 32;
  for (int l = 0; l < 32; l++) rd_gr(r, &s->fin[l]);
  for (int l = 0; l < 32; l++) rd_gr(r, &s->fin2[l]);
}
static void rd_open(rd_t *r, open_t *o, int batch, int col_elem) {
  memset(o, 0, sizeof *o);
  o->col_elem = col_elem;
  if (batch) {
    u64 nl = rd_u64(r);
    if (!rd_sane(r, nl, 8)) return;
    o->nmbar = nl;
    for (size_t l = 0; l < nl && !r->bad; l++) {
      u64 k2 = rd_u64(r);
      if (!rd_sane(r, k2, E * g_nb)) return;
      if (l == 0) { o->k2 = k2; o->mbar = xmalloc((nl * k2 + 1) * sizeof(gr)); }
      else if (k2 != o->k2) { r->bad = 1; return; }
      for (size_t j = 0; j < k2; j++) rd_gr(r, &o->mbar[l * k2 + j]);
    }
  } else {
    o->nmbar = 1;
    u64 k2 = rd_u64(r);
    if (!rd_sane(r, k2, E * g_nb)) return;
    o->k2 = k2;
    o->mbar = xmalloc((k2 + 1) * sizeof(gr));
    for (size_t j = 0; j < k2; j++) rd_gr(r, &o->mbar[j]);
  }
  u64 t = rd_u64(r);
  if (!rd_sane(r, t, 4)) return;
  o->t = t;
  o->idx = xmalloc((t + 1) * sizeof(u32));
  for (size_t q = 0; q < t; q++) o->idx[q] = rd_u32(r);
  if (rd_u64(r) != t) { r->bad = 1; return; }
  for (size_t q = 0; q < t && !r->bad; q++) {
    u64 cl = rd_u64(r);
    if (!rd_sane(r, cl, (size_t)col_elem * g_nb)) return;
    if (q == 0) { o->col_len = cl; o->cols = xmalloc((t * cl * col_elem + 1) * sizeof(u64)); }
    else if (cl != o->col_len) { r->bad = 1; return; }
    for (size_t i = 0; i < cl * (size_t)col_elem; i++) o->cols[q * cl * col_elem + i] = rd_limb(r);
  }
  if (t == 0) o->cols = xmalloc(8);
  if (rd_u64(r) != t) { r->bad = 1; return; }
  for (size_t q = 0; q < t && !r->bad; q++) {
    u64 pl = rd_u64(r);
    if (!rd_sane(r, pl, 32)) return;
    if (q == 0) { o->plen = pl; o->paths = xmalloc((t * pl + 1) * 32); }
    else if (pl != o->plen) { r->bad = 1; return; }
    for (size_t i = 0; i < pl; i++) { const u8 *h = rd_take(r, 32); if (h) memcpy(o->paths[q * pl + i], h, 32); }
  }
  if (t == 0) o->paths = xmalloc(32);
}

int simd_verify(int log_n, int log_m, const coo_t *A, const coo_t *B, const coo_t *C, u64 seed_w, u64 seed_q,
                const u8 *proof, size_t len) {
  size_t n = (size_t)1 << log_n, m = (size_t)1 << log_m;
  rd_t r = {proof, len, 0, 0};
  u8 root_w[32], root_e[32];
  const u8 *p = rd_take(&r, 32); if (p) memcpy(root_w, p, 32);
  p = rd_take(&r, 32); if (p) memcpy(root_e, p, 32);
  sc_t mp, lp;
  open_t ow, oe;
  rd_mult(&r, &mp, 1, 2);
  rd_mult(&r, &lp, 1, 2);
  rd_open(&r, &ow, 0, 32);
  rd_open(&r, &oe, 0, E);
  if (r.bad || r.o != r.n) { sc_free(&mp); sc_free(&lp); open_free(&ow); open_free(&oe); return VE_PARSE; }
  pcs_t pw, pq;
  simd_params(log_n, log_m, seed_w, seed_q, &pw, &pq);
  int res = VE_OK;
  tr_t tr;
  tr_new(&tr, "simd");
  tr_root(&tr, root_w);
  tr_root(&tr, root_e);
  gr *rt = xmalloc(log_n * sizeof(gr)), *s1 = xmalloc((log_n + 1) * sizeof(gr)), *s2 = xmalloc((lp.n + 1) * sizeof(gr));
  gr *eq_s1 = NULL, *eq_r = NULL, *eq_s2 = NULL, *ra = NULL, *rb = NULL, *rc = NULL;
  tr_target(&tr, "mult_check_target", log_n, rt);
  size_t lk1 = (size_t)__builtin_ctzll(pq.k1);
  gr *e1 = xmalloc(pq.k2 * sizeof(gr));
  build_eq_table(rt + lk1, (size_t)log_n - lk1, e1);
  gr e_at_r, t;
  if (oe.k2 != pq.k2) { res = VE_PCSE; goto done; }
  inner(oe.mbar, e1, pq.k2, &e_at_r);
  if (!pcs_verify_weighted(&pq, root_e, rt, NULL, 1, &e_at_r, &oe)) { res = VE_PCSE; goto done; }
  if (!columns_in_kernel(&oe)) { res = VE_KERNEL; goto done; }
  gr claim_mult = mp.claimed, cw_at_r;
  gr_sub(&cw_at_r, &claim_mult, &e_at_r);
  tr_gr_slice(&tr, "claim_mult", &claim_mult, 1);
  tr_label(&tr, "mult_check_sumcheck");
  ch_t ch = {0};
  ch.tr = &tr;
  if (mp.n != (size_t)log_n || verify_mult_check(&mp, rt, log_n, &ch, s1) != 0) { res = VE_MULT; goto done; }
  tr_gr_slice(&tr, "mult/finals", mp.fin, 2);
  gr gam, gam2, claim_lin;
  tr_challenge(&tr, "linear_check_gamma", &gam);
  gr_mul(&gam2, &gam, &gam);
  gr_mul(&t, &gam, &mp.fin[1]); gr_add(&claim_lin, &mp.fin[0], &t);
  gr_mul(&t, &gam2, &cw_at_r); gr_add(&claim_lin, &claim_lin, &t);
  if (!gr_eq(&claim_lin, &lp.claimed)) { res = VE_LINCLAIM; goto done; }
  tr_gr_slice(&tr, "claim_lin", &claim_lin, 1);
  tr_label(&tr, "linear_check_sumcheck");
  if (lp.n == 0 || verify_linear_simd(&lp, &ch, s2) != 0) { res = VE_LIN; goto done; }
  if (lp.n != (size_t)log_m) { res = VE_LIN; goto done; }
  eq_s1 = xmalloc(n * sizeof(gr)); eq_r = xmalloc(n * sizeof(gr)); eq_s2 = xmalloc(m * sizeof(gr));
  build_eq_table(s1, log_n, eq_s1);
  build_eq_table(rt, log_n, eq_r);
  build_eq_table(s2, log_m, eq_s2);
  ra = accumulate_row(A, eq_s1, m); rb = accumulate_row(B, eq_s1, m); rc = accumulate_row(C, eq_r, m);
  for (size_t h = 0; h < m; h++) {
    gr_mul(&t, &gam, &rb[h]); gr_add(&ra[h], &ra[h], &t);
    gr_mul(&t, &gam2, &rc[h]); gr_add(&ra[h], &ra[h], &t);
  }
  gr at_s2;
  inner(eq_s2, ra, m, &at_s2);
  if (!gr_eq(&at_s2, &lp.fin[0])) { res = VE_AROW; goto done; }
  tr_gr_slice(&tr, "lin/finals", lp.fin, 2);
  if (!pcs_verify_weighted(&pw, root_w, s2, NULL, 1, &lp.fin[1], &ow)) { res = VE_PCSW; goto done; }
done:
  free(rt); free(s1); free(s2); free(e1); free(eq_s1); free(eq_r); free(eq_s2); free(ra); free(rb); free(rc);
  sc_free(&mp); sc_free(&lp); open_free(&ow); open_free(&oe); pcs_free(&pw); pcs_free(&pq);
  return res;
}

/* ========================= General ========================= */
typedef struct { u32 l, p, q, i; u64 v; } item_t;
static int item_cmp(const void *x, const void *y) {
  const item_t *a = x, *b = y;
  if (a->l != b->l) return a->l < b->l ? -1 : 1;
  if (a->p != b->p) return a->p < b->p ? -1 : 1;
  if (a->q != b->q) return a->q < b->q ? -1 : 1;
  return 0;
}
void gdiag_build(gdiag_t *g, const coo_t *m3[3], size_t n) {
  (void)n;
  memset(g, 0, sizeof *g);
  for (int w = 0; w < 3; w++) {
    const coo_t *m = m3[w];
    item_t *it = xmalloc((m->nnz + 1) * sizeof(item_t));
    for (size_t k = 0; k < m->nnz; k++) {
      u32 r = m->r[k], c = m->c[k];
      it[k].p = r / 32; it[k].i = r % 32; it[k].q = c / 32;
      it[k].l = ((c % 32) + 32 - (r % 32)) % 32;
      it[k].v = m->v[k];
    }
    qsort(it, m->nnz, sizeof(item_t), item_cmp);
    size_t cnt[32] = {0};
    for (size_t k = 0; k < m->nnz; k++)
      if (k == 0 || item_cmp(&it[k], &it[k - 1]) != 0) cnt[it[k].l]++;
    for (int l = 0; l < 32; l++) {
      g->ncell[w][l] = 0;
      g->p[w][l] = xmalloc((cnt[l] + 1) * sizeof(u32));
      g->q[w][l] = xmalloc((cnt[l] + 1) * sizeof(u32));
      g->d[w][l] = xcalloc((cnt[l] + 1) * 32, sizeof(u64));
      g->dhat[w][l] = xmalloc((cnt[l] + 1) * sizeof(gr));
    }
    for (size_t k = 0; k < m->nnz; k++) {
      u32 l = it[k].l;
      if (k == 0 || item_cmp(&it[k], &it[k - 1]) != 0) {
        size_t c = g->ncell[w][l]++;
        g->p[w][l][c] = it[k].p;
        g->q[w][l][c] = it[k].q;
      }
      size_t c = g->ncell[w][l] - 1;
      g->d[w][l][c * 32 + it[k].i] = mk(g->d[w][l][c * 32 + it[k].i] + it[k].v);
    }
    for (int l = 0; l < 32; l++)
      for (size_t c = 0; c < g->ncell[w][l]; c++) pack_r2(&g->dhat[w][l][c], &g->d[w][l][c * 32]);
    free(it);
  }
}
void gdiag_free(gdiag_t *g) {
  for (int w = 0; w < 3; w++)
    for (int l = 0; l < 32; l++) { free(g->p[w][l]); free(g->q[w][l]); free(g->d[w][l]); free(g->dhat[w][l]); }
  memset(g, 0, sizeof *g);
}

static void general_params(int log_b, u64 seed_w, u64 seed_e, pcs_t *pw, pcs_t *pabc, pcs_t *pe) {
  pcs_setup_batched(pw, log_b, SECURE_Q, SECURE_DELTA, SECURE_LAMBDA, seed_w, 32);
  pcs_setup_stacked(pabc, log_b, SECURE_Q, SECURE_DELTA, SECURE_LAMBDA, seed_w ^ 0xABC, 3);
  r2pcs_setup(pe, log_b, SECURE_Q, SECURE_DELTA, SECURE_LAMBDA, seed_e);
}
static gr *layered_matvec(const gdiag_t *g, int w, gr *const *rho, size_t nb) { /* general.rs:211-230 */
  gr *o = xcalloc(nb, sizeof(gr));
  for (int l = 0; l < 32; l++) {
    size_t *run, nr = key_runs(g->p[w][l], g->ncell[w][l], &run);
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t j = 0; j < nr; j++)
      for (size_t c = run[j]; c < run[j + 1]; c++) {
        gr t;
        gr_mul(&t, &g->dhat[w][l][c], &rho[l][g->q[w][l][c]]);
        gr_add(&o[g->p[w][l][c]], &o[g->p[w][l][c]], &t);
      }
    free(run);
  }
  return o;
}
/* general.rs:233-271 */
static void build_layer_rows(const gdiag_t *g, const gr *eq, const gr *gam, const gr *gam2, size_t nb, gr **rows) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int l = 0; l < 32; l++) {
    gr t;
    gr *acc[3];
    for (int w = 0; w < 3; w++) {
      acc[w] = xcalloc(nb, sizeof(gr));
      for (size_t c = 0; c < g->ncell[w][l]; c++) {
        gr_mul(&t, &eq[g->p[w][l][c]], &g->dhat[w][l][c]);
        gr_add(&acc[w][g->q[w][l][c]], &acc[w][g->q[w][l][c]], &t);
      }
    }
    for (size_t q = 0; q < nb; q++) {
      gr x = acc[0][q];
      gr_mul(&t, gam, &acc[1][q]); gr_add(&x, &x, &t);
      gr_mul(&t, gam2, &acc[2][q]); gr_add(&x, &x, &t);
      rows[l][q] = x;
    }
    for (int w = 0; w < 3; w++) free(acc[w]);
  }
}

int general_prove(int log_b, const gdiag_t *g, const u64 *w, u64 seed_w, u64 seed_e, buf_t *out) {
  size_t nb = (size_t)1 << log_b;
  const consts_t *K = consts();
  pcs_t pw, pabc, pe;
  general_params(log_b, seed_w, seed_e, &pw, &pabc, &pe);
  /* 1. general_reduce_witness (reduce.rs:670-744) */
  gr *rho[32];
  for (int l = 0; l < 32; l++) rho[l] = xmalloc(nb * sizeof(gr));
#pragma omp parallel for collapse(2)
  for (int l = 0; l < 32; l++) {
    for (size_t q = 0; q < nb; q++) {
      u64 rot[32];
      for (int i = 0; i < 32; i++) rot[i] = w[32 * q + (i + l) % 32];
      pack_r2(&rho[l][q], rot);
    }
  }
  gr *q1[2], *q2[2];
  for (int mtx = 0; mtx < 2; mtx++) {
    q1[mtx] = xcalloc(nb, sizeof(gr));
    q2[mtx] = xcalloc(nb, sizeof(gr));
    for (int l = 0; l < 32; l++) {
      /* cells are (p, q) sorted: a run of equal p is one thread's */
      size_t *run, nr = key_runs(g->p[mtx][l], g->ncell[mtx][l], &run);
#pragma omp parallel for schedule(dynamic, 16)
      for (size_t j = 0; j < nr; j++)
        for (size_t c = run[j]; c < run[j + 1]; c++) {
          u64 rv[32];
          size_t q = g->q[mtx][l][c], p = g->p[mtx][l][c];
          for (int i = 0; i < 32; i++) rv[i] = w[32 * q + (i + l) % 32];
          gr d1, d2;
          mul_with_quotients_only(&d1, &d2, &g->d[mtx][l][c * 32], rv);
          gr_add(&q1[mtx][p], &q1[mtx][p], &d1);
          gr_add(&q2[mtx][p], &q2[mtx][p], &d2);
        }
      free(run);
    }
  }
  /* 2. errors and packed matvec outputs (general.rs:389-428) */
  gr va; gr_from_g4(&va, &K->v_alpha);
  gr *ea = xmalloc(nb * sizeof(gr)), *eb = xmalloc(nb * sizeof(gr)), *ec = xmalloc(nb * sizeof(gr));
  gr *e_[2] = {ea, eb};
  gr t;
#pragma omp parallel for collapse(2)
  for (int mtx = 0; mtx < 2; mtx++)
    for (size_t i = 0; i < nb; i++) {
      gr t1, u1;
      gr_mul(&t1, &q1[mtx][i], &va); gr_mul(&u1, &q2[mtx][i], &K->v_beta); gr_add(&e_[mtx][i], &t1, &u1);
    }
  for (int mtx = 0; mtx < 2; mtx++) { free(q1[mtx]); free(q2[mtx]); }
  gr *asum = layered_matvec(g, 0, rho, nb), *bsum = layered_matvec(g, 1, rho, nb), *csum = layered_matvec(g, 2, rho, nb);
  gr *ah = xmalloc(nb * sizeof(gr)), *bh = xmalloc(nb * sizeof(gr)), *chh = xmalloc(nb * sizeof(gr)), *em = xmalloc(nb * sizeof(gr));
  u64 *as = xmalloc(nb * 32 * sizeof(u64)), *bs = xmalloc(nb * 32 * sizeof(u64)), *cs = xmalloc(nb * 32 * sizeof(u64));
#pragma omp parallel for
  for (size_t i = 0; i < nb; i++) {
    gr t;
    gr_sub(&ah[i], &asum[i], &ea[i]);
    gr_sub(&bh[i], &bsum[i], &eb[i]);
    u64 sl[32];
    psi_r2(sl, &csum[i]);
    pack_r2(&chh[i], sl);
    gr_sub(&ec[i], &csum[i], &chh[i]);
    gr_mul(&t, &ah[i], &bh[i]);
    gr_sub(&em[i], &t, &chh[i]);
    psi_r2(&as[i * 32], &ah[i]);
    psi_r2(&bs[i * 32], &bh[i]);
    psi_r2(&cs[i * 32], &chh[i]);
  }
  free(asum); free(bsum); free(csum);
  /* 3. commits + transcript */
  comm_t cw, cabc, cem, ces;
  pcs_commit(&pw, w, &cw);
  const u64 *abc[3] = {as, bs, cs};
  pcs_commit_stacked(&pabc, abc, 3, &cabc);
  r2pcs_commit(&pe, em, &cem);
  const gr *est[3] = {ea, eb, ec};
  r2pcs_commit_stacked(&pe, est, 3, &ces);
  tr_t tr;
  tr_new(&tr, "general");
  tr_root(&tr, cw.root); tr_root(&tr, cabc.root); tr_root(&tr, cem.root); tr_root(&tr, ces.root);
  gr *rt = xmalloc(log_b * sizeof(gr));
  tr_target(&tr, "mult_check_target", log_b, rt);
  /* 4. mult check (sub) */
  tr_label(&tr, "mult_check_sumcheck");
  ch_t ch = {0};
  ch.tr = &tr;
  sc_t mp, lp;
  gr *s1 = xmalloc(log_b * sizeof(gr)), *s2 = xmalloc(log_b * sizeof(gr));
  prove_mult_check_sub(ah, bh, chh, em, log_b, rt, &ch, &mp, s1);
  tr_gr_slice(&tr, "mult/finals", mp.fin, 4);
  gr gam, gam2;
  tr_challenge(&tr, "linear_check_gamma", &gam);
  gr_mul(&gam2, &gam, &gam);
  /* 5. opens at s1 */
  gr wts[3];
  gr_one(&wts[0]); wts[1] = gam; wts[2] = gam2;
  open_t oes, oabc, oem, orho;
  gr estack_val, abc_val, em_val;
  pcs_open_weighted(&pe, &ces, s1, wts, &estack_val, &oes);
  pcs_open_weighted(&pabc, &cabc, s1, wts, &abc_val, &oabc);
  pcs_open_weighted(&pe, &cem, s1, NULL, &em_val, &oem);
  /* 6. layer rows + linear check */
  gr *eq_s1 = xmalloc(nb * sizeof(gr));
  build_eq_table(s1, log_b, eq_s1);
  gr *rows[32];
  for (int l = 0; l < 32; l++) rows[l] = xmalloc(nb * sizeof(gr));
  build_layer_rows(g, eq_s1, &gam, &gam2, nb, rows);
  gr claim_lin;
  gr_mul(&t, &gam, &mp.fin[1]); gr_add(&claim_lin, &mp.fin[0], &t);
  gr_mul(&t, &gam2, &mp.fin[2]); gr_add(&claim_lin, &claim_lin, &t);
  gr_add(&claim_lin, &claim_lin, &estack_val);
  tr_label(&tr, "linear_check_sumcheck");
  prove_linear_general(rows, rho, log_b, &claim_lin, &ch, &lp, s2);
  /* 7. rotated opens */
  gr ys[32];
  pcs_rotated_open_batch(&pw, &cw, s2, 32, ys, &orho);
  buf_put(out, cw.root, 32); buf_put(out, cabc.root, 32); buf_put(out, cem.root, 32); buf_put(out, ces.root, 32);
  ser_mult(out, &mp, 0);
  ser_lin_general(out, &lp);
  open_serialize(out, &oem, 0);
  open_serialize(out, &oabc, 0);
  open_serialize(out, &oes, 0);
  open_serialize(out, &orho, 1);
  int ok = gr_eq(&em_val, &mp.fin[3]);
  for (int l = 0; l < 32; l++) ok = ok && gr_eq(&ys[l], &lp.fin2[l]);
  for (int l = 0; l < 32; l++) { free(rows[l]); free(rho[l]); }
  free(eq_s1); free(s1); free(s2); free(rt);
  open_free(&oes); open_free(&oabc); open_free(&oem); open_free(&orho); sc_free(&mp); sc_free(&lp);
  comm_free(&cw); comm_free(&cabc); comm_free(&cem); comm_free(&ces);
  free(ea); free(eb); free(ec); free(ah); free(bh); free(chh); free(em); free(as); free(bs); free(cs);
  pcs_free(&pw); pcs_free(&pabc); pcs_free(&pe);
  return ok ? 0 : -1;
}

int general_verify(int log_b, const gdiag_t *g, u64 seed_w, u64 seed_e, const u8 *proof, size_t len) {
  size_t nb = (size_t)1 << log_b;
  rd_t r = {proof, len, 0, 0};
  u8 roots[4][32];
  for (int i = 0; i < 4; i++) { const u8 *p = rd_take(&r, 32); if (p) memcpy(roots[i], p, 32); }
  sc_t mp, lp;
  open_t oem, oabc, oes, orho;
  rd_mult(&r, &mp, 0, 4);
  rd_lin_general(&r, &lp);
  rd_open(&r, &oem, 0, E);
  rd_open(&r, &oabc, 0, 32);
  rd_open(&r, &oes, 0, E);
  rd_open(&r, &orho, 1, 32);
  if (r.bad || r.o != r.n) {
    sc_free(&mp); sc_free(&lp); open_free(&oem); open_free(&oabc); open_free(&oes); open_free(&orho);
    return VE_PARSE;
  }
  pcs_t pw, pabc, pe;
  general_params(log_b, seed_w, seed_e, &pw, &pabc, &pe);
  int res = VE_OK;
  gr *rt = xmalloc(log_b * sizeof(gr)), *s1 = xmalloc((log_b + 1) * sizeof(gr)), *s2 = xmalloc((lp.n + 1) * sizeof(gr));
  gr *e1e = xmalloc(pe.k2 * sizeof(gr)), *e1abc = xmalloc(pabc.k2 * sizeof(gr));
  gr *eq_s1 = NULL, *eq_s2 = NULL, *rows[32] = {0};
  tr_t tr;
  tr_new(&tr, "general");
  for (int i = 0; i < 4; i++) tr_root(&tr, roots[i]);
  tr_target(&tr, "mult_check_target", log_b, rt);
  tr_label(&tr, "mult_check_sumcheck");
  ch_t ch = {0};
  ch.tr = &tr;
  if (mp.n != (size_t)log_b || verify_mult_check_sub(&mp, rt, log_b, &ch, s1) != 0) { res = VE_MULT; goto done; }
  tr_gr_slice(&tr, "mult/finals", mp.fin, 4);
  gr gam, gam2, t;
  tr_challenge(&tr, "linear_check_gamma", &gam);
  gr_mul(&gam2, &gam, &gam);
  tr_label(&tr, "linear_check_sumcheck");
  if (lp.n == 0 || verify_linear_general(&lp, &ch, s2) != 0) { res = VE_LIN; goto done; }
  if (lp.n != (size_t)log_b) { res = VE_LIN; goto done; }
  size_t lk1e = (size_t)__builtin_ctzll(pe.k1), lk1a = (size_t)__builtin_ctzll(pabc.k1);
  build_eq_table(s1 + lk1e, (size_t)log_b - lk1e, e1e);
  build_eq_table(s1 + lk1a, (size_t)log_b - lk1a, e1abc);
  if (oes.k2 != pe.k2 || oem.k2 != pe.k2 || oabc.k2 != pabc.k2) { res = VE_LINCLAIM; goto done; }
  gr estack_val, em_s1, abc_val, abc_exp;
  inner(oes.mbar, e1e, pe.k2, &estack_val);
  inner(oem.mbar, e1e, pe.k2, &em_s1);
  inner(oabc.mbar, e1abc, pabc.k2, &abc_val);
  gr_mul(&t, &gam, &mp.fin[1]); gr_add(&abc_exp, &mp.fin[0], &t);
  gr_mul(&t, &gam2, &mp.fin[2]); gr_add(&abc_exp, &abc_exp, &t);
  if (!gr_eq(&abc_val, &abc_exp) || !gr_eq(&em_s1, &mp.fin[3])) { res = VE_LINCLAIM; goto done; }
  gr exp_claim;
  gr_add(&exp_claim, &abc_exp, &estack_val);
  if (!gr_eq(&lp.claimed, &exp_claim)) { res = VE_LINCLAIM; goto done; }
  eq_s1 = xmalloc(nb * sizeof(gr)); eq_s2 = xmalloc(nb * sizeof(gr));
  build_eq_table(s1, log_b, eq_s1);
  build_eq_table(s2, log_b, eq_s2);
  for (int l = 0; l < 32; l++) rows[l] = xmalloc(nb * sizeof(gr));
  build_layer_rows(g, eq_s1, &gam, &gam2, nb, rows);
  for (int l = 0; l < 32; l++) {
    gr v; inner(eq_s2, rows[l], nb, &v);
    if (!gr_eq(&v, &lp.fin[l])) { res = VE_AROW; goto done; }
  }
  gr wts[3];
  gr_one(&wts[0]); wts[1] = gam; wts[2] = gam2;
  if (!pcs_verify_weighted(&pe, roots[2], s1, NULL, 1, &em_s1, &oem)) { res = VE_PCSE; goto done; }
  if (!pcs_verify_weighted(&pabc, roots[1], s1, wts, 3, &abc_val, &oabc)) { res = VE_PCSW; goto done; }
  if (!pcs_verify_weighted(&pe, roots[3], s1, wts, 3, &estack_val, &oes)) { res = VE_PCSE; goto done; }
  if (orho.nmbar != 32) { res = VE_PCSW; goto done; }
  if (!pcs_rotated_verify_batch(&pw, roots[0], s2, lp.fin2, 32, &orho)) { res = VE_PCSW; goto done; }
  if (!columns_in_kernel(&oem) || !columns_in_kernel(&oes)) { res = VE_KERNEL; goto done; }
done:
  for (int l = 0; l < 32; l++) free(rows[l]);
  free(rt); free(s1); free(s2); free(e1e); free(e1abc); free(eq_s1); free(eq_s2);
  sc_free(&mp); sc_free(&lp); open_free(&oem); open_free(&oabc); open_free(&oes); open_free(&orho);
  pcs_free(&pw); pcs_free(&pabc); pcs_free(&pe);
  return res;
}

/* ========================= MIMD ========================= */
typedef struct { u32 r, c; int k; u64 v; } mitem_t;
static int mitem_cmp(const void *x, const void *y) {
  const mitem_t *a = x, *b = y;
  if (a->r != b->r) return a->r < b->r ? -1 : 1;
  if (a->c != b->c) return a->c < b->c ? -1 : 1;
  return 0;
}
/* pack_matrix_set (reduce.rs:119-141): union support of the 32 instance
 * matrices, entry (r,c) = pack_r2 of the 32 instance values (summed per
 * instance when a triple repeats), cells in (r,c) order. */
void mimd_pack(mmats_t *mm, size_t rows, size_t cols, const coo_t *const *sets[3]) {
  memset(mm, 0, sizeof *mm);
  mm->rows = rows;
  mm->cols = cols;
  for (int w = 0; w < 3; w++) {
    size_t tot = 0;
    for (int k = 0; k < 32; k++) tot += sets[w][k]->nnz;
    mitem_t *it = xmalloc((tot + 1) * sizeof(mitem_t));
    size_t o = 0;
    for (int k = 0; k < 32; k++)
      for (size_t j = 0; j < sets[w][k]->nnz; j++, o++) {
        it[o].r = sets[w][k]->r[j]; it[o].c = sets[w][k]->c[j]; it[o].k = k; it[o].v = sets[w][k]->v[j];
      }
    qsort(it, tot, sizeof(mitem_t), mitem_cmp);
    size_t nc = 0;
    for (size_t j = 0; j < tot; j++)
      if (j == 0 || mitem_cmp(&it[j], &it[j - 1]) != 0) nc++;
    mm->ncell[w] = nc;
    mm->r[w] = xmalloc((nc + 1) * sizeof(u32));
    mm->c[w] = xmalloc((nc + 1) * sizeof(u32));
    mm->d[w] = xcalloc((nc + 1) * 32, sizeof(u64));
    mm->dhat[w] = xmalloc((nc + 1) * sizeof(gr));
    size_t cur = (size_t)-1;
    for (size_t j = 0; j < tot; j++) {
      if (j == 0 || mitem_cmp(&it[j], &it[j - 1]) != 0) { cur++; mm->r[w][cur] = it[j].r; mm->c[w][cur] = it[j].c; }
      mm->d[w][cur * 32 + it[j].k] = mk(mm->d[w][cur * 32 + it[j].k] + it[j].v);
    }
    for (size_t j = 0; j < nc; j++) pack_r2(&mm->dhat[w][j], &mm->d[w][j * 32]);
    free(it);
  }
}
void mimd_free(mmats_t *mm) {
  for (int w = 0; w < 3; w++) { free(mm->r[w]); free(mm->c[w]); free(mm->d[w]); free(mm->dhat[w]); }
  memset(mm, 0, sizeof *mm);
}
static void mimd_params(int log_n, int log_m, u64 seed_w, u64 seed_e, pcs_t *pw, pcs_t *pabc, pcs_t *pe) {
  pcs_setup(pw, log_m, SECURE_Q, SECURE_DELTA, SECURE_LAMBDA, seed_w);
  pcs_setup_stacked(pabc, log_n, SECURE_Q, SECURE_DELTA, SECURE_LAMBDA, seed_w ^ 0xABC, 3);
  r2pcs_setup(pe, log_n, SECURE_Q, SECURE_DELTA, SECURE_LAMBDA, seed_e);
}
/* matvec_gr128_dense_csr over the union-support cells (matrix.rs) */
static gr *mimd_matvec(const mmats_t *mm, int w, const gr *x) {
  gr *o = xcalloc(mm->rows, sizeof(gr));
  gr t;
  for (size_t j = 0; j < mm->ncell[w]; j++) {
    gr_mul(&t, &mm->dhat[w][j], &x[mm->c[w][j]]);
    gr_add(&o[mm->r[w][j]], &o[mm->r[w][j]], &t);
  }
  return o;
}
/* build_combined_row (mimd.rs:196-219) with eq_s1 for all three matrices,
 * as both the prover (mimd.rs:364) and the verifier (mimd.rs:492-499) call it */
static gr *mimd_row(const mmats_t *mm, const gr *eq, const gr *gam, const gr *gam2) {
  gr *acc[3], t;
  for (int w = 0; w < 3; w++) {
    acc[w] = xcalloc(mm->cols, sizeof(gr));
    for (size_t j = 0; j < mm->ncell[w]; j++) {
      gr_mul(&t, &eq[mm->r[w][j]], &mm->dhat[w][j]);
      gr_add(&acc[w][mm->c[w][j]], &acc[w][mm->c[w][j]], &t);
    }
  }
  for (size_t h = 0; h < mm->cols; h++) {
    gr_mul(&t, gam, &acc[1][h]); gr_add(&acc[0][h], &acc[0][h], &t);
    gr_mul(&t, gam2, &acc[2][h]); gr_add(&acc[0][h], &acc[0][h], &t);
  }
  free(acc[1]); free(acc[2]);
  return acc[0];
}

/* prove_mimd_with_timing (mimd.rs:242-397), mimd_reduce_witness
 * (reduce.rs:287-345) with the A,B quotients from mul_with_quotients_only */
int mimd_prove(int log_n, int log_m, const mmats_t *mm, const u64 *wit, u64 seed_w, u64 seed_e, buf_t *out) {
  size_t n = (size_t)1 << log_n, m = (size_t)1 << log_m;
  const consts_t *K = consts();
  pcs_t pw, pabc, pe;
  mimd_params(log_n, log_m, seed_w, seed_e, &pw, &pabc, &pe);
  /* 1. witness reduction */
  gr *w_phi = xmalloc(m * sizeof(gr));
  u64 *w_slots = xmalloc(m * 32 * sizeof(u64));
  for (size_t j = 0; j < m; j++) {
    for (int k = 0; k < 32; k++) w_slots[j * 32 + k] = wit[(size_t)k * m + j];
    pack_r2(&w_phi[j], &w_slots[j * 32]);
  }
  gr va; gr_from_g4(&va, &K->v_alpha);
  gr *e_[3], t, u;
  for (int w = 0; w < 2; w++) {
    gr *q1 = xcalloc(n, sizeof(gr)), *q2 = xcalloc(n, sizeof(gr));
    for (size_t j = 0; j < mm->ncell[w]; j++) {
      gr d1, d2;
      mul_with_quotients_only(&d1, &d2, &mm->d[w][j * 32], &w_slots[(size_t)mm->c[w][j] * 32]);
      gr_add(&q1[mm->r[w][j]], &q1[mm->r[w][j]], &d1);
      gr_add(&q2[mm->r[w][j]], &q2[mm->r[w][j]], &d2);
    }
    e_[w] = xmalloc(n * sizeof(gr));
    for (size_t i = 0; i < n; i++) {
      gr_mul(&t, &q1[i], &va); gr_mul(&u, &q2[i], &K->v_beta); gr_add(&e_[w][i], &t, &u);
    }
    free(q1); free(q2);
  }
  /* 2. matvecs, hats and errors (mimd.rs:258-295) */
  gr *asum = mimd_matvec(mm, 0, w_phi), *bsum = mimd_matvec(mm, 1, w_phi), *csum = mimd_matvec(mm, 2, w_phi);
  gr *ah = xmalloc(n * sizeof(gr)), *bh = xmalloc(n * sizeof(gr)), *chh = xmalloc(n * sizeof(gr));
  gr *em = xmalloc(n * sizeof(gr));
  e_[2] = xmalloc(n * sizeof(gr));
  u64 *as = xmalloc(n * 32 * sizeof(u64)), *bs = xmalloc(n * 32 * sizeof(u64)), *cs = xmalloc(n * 32 * sizeof(u64));
  for (size_t i = 0; i < n; i++) {
    gr_sub(&ah[i], &asum[i], &e_[0][i]);
    gr_sub(&bh[i], &bsum[i], &e_[1][i]);
    u64 sl[32];
    psi_r2(sl, &csum[i]);
    pack_r2(&chh[i], sl);
    gr_sub(&e_[2][i], &csum[i], &chh[i]);
    gr_mul(&t, &ah[i], &bh[i]);
    gr_sub(&em[i], &t, &chh[i]);
    psi_r2(&as[i * 32], &ah[i]);
    psi_r2(&bs[i * 32], &bh[i]);
    psi_r2(&cs[i * 32], &chh[i]);
  }
  free(asum); free(bsum); free(csum);
  /* 3. commits + transcript */
  comm_t cw, cabc, cem, ces;
  pcs_commit(&pw, w_slots, &cw);
  const u64 *abc[3] = {as, bs, cs};
  pcs_commit_stacked(&pabc, abc, 3, &cabc);
  r2pcs_commit(&pe, em, &cem);
  const gr *est[3] = {e_[0], e_[1], e_[2]};
  r2pcs_commit_stacked(&pe, est, 3, &ces);
  tr_t tr;
  tr_new(&tr, "mimd");
  tr_root(&tr, cw.root); tr_root(&tr, cabc.root); tr_root(&tr, cem.root); tr_root(&tr, ces.root);
  gr *rt = xmalloc(log_n * sizeof(gr));
  tr_target(&tr, "mult_check_target", log_n, rt);
  /* 4. mult check (sub form) */
  tr_label(&tr, "mult_check_sumcheck");
  ch_t ch = {0};
  ch.tr = &tr;
  sc_t mp, lp;
  gr *s1 = xmalloc(log_n * sizeof(gr)), *s2 = xmalloc(log_m * sizeof(gr));
  prove_mult_check_sub(ah, bh, chh, em, log_n, rt, &ch, &mp, s1);
  tr_gr_slice(&tr, "mult/finals", mp.fin, 4);
  gr gam, gam2;
  tr_challenge(&tr, "linear_check_gamma", &gam);
  gr_mul(&gam2, &gam, &gam);
  /* 5. opens at s1 */
  gr wts[3];
  gr_one(&wts[0]); wts[1] = gam; wts[2] = gam2;
  open_t oes, oabc, oem, ow;
  gr estack_val, abc_val, em_val, yw;
  pcs_open_weighted(&pe, &ces, s1, wts, &estack_val, &oes);
  pcs_open_weighted(&pabc, &cabc, s1, wts, &abc_val, &oabc);
  pcs_open_weighted(&pe, &cem, s1, NULL, &em_val, &oem);
  /* 6. combined row + linear check (m_row first, w_phi second) */
  gr *eq_s1 = xmalloc(n * sizeof(gr));
  build_eq_table(s1, log_n, eq_s1);
  gr *row = mimd_row(mm, eq_s1, &gam, &gam2);
  gr claim_lin;
  gr_mul(&t, &gam, &mp.fin[1]); gr_add(&claim_lin, &mp.fin[0], &t);
  gr_mul(&t, &gam2, &mp.fin[2]); gr_add(&claim_lin, &claim_lin, &t);
  gr abc_exp = claim_lin;
  gr_add(&claim_lin, &claim_lin, &estack_val);
  tr_label(&tr, "linear_check_sumcheck");
  prove_linear_simd(row, w_phi, log_m, &claim_lin, &ch, &lp, s2);
  /* 7. open w at s2 */
  pcs_open_weighted(&pw, &cw, s2, NULL, &yw, &ow);
  buf_put(out, cw.root, 32); buf_put(out, cabc.root, 32); buf_put(out, cem.root, 32); buf_put(out, ces.root, 32);
  ser_mult(out, &mp, 0);
  ser_mult(out, &lp, 1);
  open_serialize(out, &ow, 0);
  open_serialize(out, &oem, 0);
  open_serialize(out, &oabc, 0);
  open_serialize(out, &oes, 0);
  int ok = gr_eq(&em_val, &mp.fin[3]) && gr_eq(&abc_val, &abc_exp) && gr_eq(&yw, &lp.fin[1]);
  free(row); free(eq_s1); free(s1); free(s2); free(rt);
  open_free(&oes); open_free(&oabc); open_free(&oem); open_free(&ow); sc_free(&mp); sc_free(&lp);
  comm_free(&cw); comm_free(&cabc); comm_free(&cem); comm_free(&ces);
  for (int w = 0; w < 3; w++) free(e_[w]);
  free(ah); free(bh); free(chh); free(em); free(as); free(bs); free(cs); free(w_phi); free(w_slots);
  pcs_free(&pw); pcs_free(&pabc); pcs_free(&pe);
  return ok ? 0 : -1;
}

/* verify_mimd_with_timing (mimd.rs:422-551) */
int mimd_verify(int log_n, int log_m, const mmats_t *mm, u64 seed_w, u64 seed_e, const u8 *proof, size_t len) {
  size_t n = (size_t)1 << log_n, m = (size_t)1 << log_m;
  rd_t r = {proof, len, 0, 0};
  u8 roots[4][32];
  for (int i = 0; i < 4; i++) { const u8 *p = rd_take(&r, 32); if (p) memcpy(roots[i], p, 32); }
  sc_t mp, lp;
  open_t ow, oem, oabc, oes;
  rd_mult(&r, &mp, 0, 4);
  rd_mult(&r, &lp, 1, 2);
  rd_open(&r, &ow, 0, 32);
  rd_open(&r, &oem, 0, E);
  rd_open(&r, &oabc, 0, 32);
  rd_open(&r, &oes, 0, E);
  if (r.bad || r.o != r.n) {
    sc_free(&mp); sc_free(&lp); open_free(&ow); open_free(&oem); open_free(&oabc); open_free(&oes);
    return VE_PARSE;
  }
  pcs_t pw, pabc, pe;
  mimd_params(log_n, log_m, seed_w, seed_e, &pw, &pabc, &pe);
  int res = VE_OK;
  gr *rt = xmalloc(log_n * sizeof(gr)), *s1 = xmalloc((log_n + 1) * sizeof(gr)), *s2 = xmalloc((lp.n + 1) * sizeof(gr));
  gr *e1e = xmalloc(pe.k2 * sizeof(gr)), *e1abc = xmalloc(pabc.k2 * sizeof(gr));
  gr *eq_s1 = NULL, *eq_s2 = NULL, *row = NULL;
  tr_t tr;
  tr_new(&tr, "mimd");
  for (int i = 0; i < 4; i++) tr_root(&tr, roots[i]);
  tr_target(&tr, "mult_check_target", log_n, rt);
  tr_label(&tr, "mult_check_sumcheck");
  ch_t ch = {0};
  ch.tr = &tr;
  if (mp.n != (size_t)log_n || verify_mult_check_sub(&mp, rt, log_n, &ch, s1) != 0) { res = VE_MULT; goto done; }
  tr_gr_slice(&tr, "mult/finals", mp.fin, 4);
  gr gam, gam2, t;
  tr_challenge(&tr, "linear_check_gamma", &gam);
  gr_mul(&gam2, &gam, &gam);
  tr_label(&tr, "linear_check_sumcheck");
  if (lp.n == 0 || verify_linear_simd(&lp, &ch, s2) != 0) { res = VE_LIN; goto done; }
  if (lp.n != (size_t)log_m) { res = VE_LIN; goto done; }
  size_t lk1e = (size_t)__builtin_ctzll(pe.k1), lk1a = (size_t)__builtin_ctzll(pabc.k1);
  build_eq_table(s1 + lk1e, (size_t)log_n - lk1e, e1e);
  build_eq_table(s1 + lk1a, (size_t)log_n - lk1a, e1abc);
  if (oes.k2 != pe.k2 || oem.k2 != pe.k2 || oabc.k2 != pabc.k2) { res = VE_LINCLAIM; goto done; }
  gr estack_val, em_s1, abc_val, abc_exp;
  inner(oes.mbar, e1e, pe.k2, &estack_val);
  inner(oem.mbar, e1e, pe.k2, &em_s1);
  inner(oabc.mbar, e1abc, pabc.k2, &abc_val);
  gr_mul(&t, &gam, &mp.fin[1]); gr_add(&abc_exp, &mp.fin[0], &t);
  gr_mul(&t, &gam2, &mp.fin[2]); gr_add(&abc_exp, &abc_exp, &t);
  if (!gr_eq(&abc_val, &abc_exp) || !gr_eq(&em_s1, &mp.fin[3])) { res = VE_LINCLAIM; goto done; }
  gr exp_claim;
  gr_add(&exp_claim, &abc_exp, &estack_val);
  if (!gr_eq(&lp.claimed, &exp_claim)) { res = VE_LINCLAIM; goto done; }
  eq_s1 = xmalloc(n * sizeof(gr)); eq_s2 = xmalloc(m * sizeof(gr));
  build_eq_table(s1, log_n, eq_s1);
  build_eq_table(s2, log_m, eq_s2);
  row = mimd_row(mm, eq_s1, &gam, &gam2);
  gr at_s2;
  inner(eq_s2, row, m, &at_s2);
  if (!gr_eq(&at_s2, &lp.fin[0])) { res = VE_AROW; goto done; }
  gr wts[3];
  gr_one(&wts[0]); wts[1] = gam; wts[2] = gam2;
  if (!pcs_verify_weighted(&pw, roots[0], s2, NULL, 1, &lp.fin[1], &ow)) { res = VE_PCSW; goto done; }
  if (!pcs_verify_weighted(&pe, roots[2], s1, NULL, 1, &em_s1, &oem)) { res = VE_PCSE; goto done; }
  if (!pcs_verify_weighted(&pabc, roots[1], s1, wts, 3, &abc_val, &oabc)) { res = VE_PCSW; goto done; }
  if (!pcs_verify_weighted(&pe, roots[3], s1, wts, 3, &estack_val, &oes)) { res = VE_PCSE; goto done; }
  if (!columns_in_kernel(&oem) || !columns_in_kernel(&oes)) { res = VE_KERNEL; goto done; }
done:
  free(rt); free(s1); free(s2); free(e1e); free(e1abc); free(eq_s1); free(eq_s2); free(row);
  sc_free(&mp); sc_free(&lp); open_free(&ow); open_free(&oem); open_free(&oabc); open_free(&oes);
  pcs_free(&pw); pcs_free(&pabc); pcs_free(&pe);
  return res;
}
