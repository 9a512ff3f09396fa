/*
 * Oracle: eq tables, MLE folding and the three sumchecks (Gruen mult check,
 * its "sub" variant, SIMD/MIMD linear check, 32-layer general linear check)
 * with their verifiers.  TEST INFRASTRUCTURE ONLY.
 *
 * Follows: piop-ring/src/eq.rs:35-58, mle.rs:42-54,
 *          mult_check.rs:180-313 (prove), 327-377 (verify), 426-549 (sub),
 *          linear_check_simd.rs:196-328, linear_check_general.rs:214-402.
 * Pinned by: eq partition of unity + 2-variable closed form (eq.rs:95-129),
 * sumcheck completeness (claimed sum == direct sum, verifier accepts) and
 * soundness (tampered round -> RoundCheckFailed{0}); see
 * tests/test_oracle_piop.py.
 */
#include "oracle.h"
#include <stdlib.h>

void build_eq_table(const gr *r, size_t n, gr *out) {
  gr_one(&out[0]);
  size_t len = 1;
  gr *tmp = xmalloc(((size_t)1 << n) * sizeof(gr));
  for (size_t l = 0; l < n; l++) {
#pragma omp parallel for if (len >= 256)
    for (size_t k = 0; k < len; k++) {
      gr hi; gr_mul(&hi, &out[k], &r[l]);
      gr_sub(&tmp[2 * k], &out[k], &hi);
      tmp[2 * k + 1] = hi;
    }
    len *= 2;
    memcpy(out, tmp, len * sizeof(gr));
  }
  free(tmp);
}

void mle_fold(gr *t, size_t len, const gr *r) {
  size_t m = len / 2;
#pragma omp parallel for if (m >= 256)
  for (size_t i = 0; i < m; i++) {
    gr d, p;
    gr_sub(&d, &t[i + m], &t[i]);
    gr_mul(&p, r, &d);
    gr_add(&t[i], &t[i], &p);
  }
}

static void eq_one(gr *o, const gr *t, const gr *r) { /* t*r + (1-t)(1-r) */
  gr one, a, b, x;
  gr_one(&one);
  gr_mul(&x, t, r);
  gr_sub(&a, &one, t);
  gr_sub(&b, &one, r);
  gr_mul(&a, &a, &b);
  gr_add(o, &x, &a);
}
static void marginalize_and_scale(gr *eg, size_t len, const gr *scale) {
  size_t m = len / 2;
#pragma omp parallel for if (m >= 256)
  for (size_t i = 0; i < m; i++) {
    gr s; gr_add(&s, &eg[i], &eg[i + m]);
    gr_mul(&eg[i], scale, &s);
  }
}
void sc_free(sc_t *s) { free(s->rounds); s->rounds = NULL; }

static gr *clone_tab(const gr *a, size_t n) { gr *c = xmalloc(n * sizeof(gr)); memcpy(c, a, n * sizeof(gr)); return c; }

/* mult_check.rs:180-241 + compute_round_poly_gruen 250-296 */
void prove_mult_check(const gr *a0, const gr *b0, size_t nv, const gr *rt, const gr *claim, ch_t *ch, sc_t *pf, gr *chal) {
  size_t N = (size_t)1 << nv;
  gr *a = clone_tab(a0, N), *b = clone_tab(b0, N);
  size_t eglen = N / 2;
  gr *eg = xmalloc((eglen ? eglen : 1) * sizeof(gr));
  if (nv > 1) build_eq_table(rt + 1, nv - 1, eg); else gr_one(&eg[0]);
  memset(pf, 0, sizeof *pf);
  pf->claimed = *claim;
  pf->n = nv;
  pf->rounds = xmalloc(3 * nv * sizeof(gr));
  size_t len = N;
  for (size_t r = 0; r < nv; r++) {
    size_t m = len / 2;
    gr s0, s1, si;
    gr_zero(&s0); gr_zero(&s1); gr_zero(&si);
#pragma omp parallel for if (m >= 64) reduction(gr_plus : s0, s1, si)
    for (size_t i = 0; i < m; i++) {
      gr da, db, t, u;
      gr_sub(&da, &a[i + m], &a[i]);
      gr_sub(&db, &b[i + m], &b[i]);
      gr_mul(&t, &a[i], &b[i]); gr_mul(&u, &eg[i], &t); gr_add(&s0, &s0, &u);
      gr_mul(&t, &a[i + m], &b[i + m]); gr_mul(&u, &eg[i], &t); gr_add(&s1, &s1, &u);
      gr_mul(&t, &da, &db); gr_mul(&u, &eg[i], &t); gr_add(&si, &si, &u);
    }
    gr *q = &pf->rounds[3 * r];
    q[0] = s0; q[2] = si;
    gr_sub(&q[1], &s1, &s0); gr_sub(&q[1], &q[1], &si);
    ch_observe(ch, q, 3);
    ch_sample(ch, &chal[r]);
    mle_fold(a, len, &chal[r]);
    mle_fold(b, len, &chal[r]);
    len = m;
    if (r + 1 < nv) { gr sc; eq_one(&sc, &rt[r], &chal[r]); marginalize_and_scale(eg, m, &sc); }
  }
  pf->nfin = 2;
  pf->fin[0] = a[0]; pf->fin[1] = b[0];
  free(a); free(b); free(eg);
}

static int verify_gruen(const sc_t *pf, const gr *rt, size_t nv, ch_t *ch, gr *chal, const gr *start, gr *c_out) {
  if (pf->n != nv) return -1;
  gr c = *start, t, u;
  for (size_t r = 0; r < nv; r++) {
    const gr *q = &pf->rounds[3 * r];
    gr lhs; gr_add(&t, &q[1], &q[2]); gr_mul(&t, &rt[r], &t); gr_add(&lhs, &q[0], &t);
    if (!gr_eq(&lhs, &c)) return (int)r;
    ch_observe(ch, q, 3);
    ch_sample(ch, &chal[r]);
    gr qh; gr_mul(&t, &chal[r], &q[2]); gr_add(&t, &q[1], &t); gr_mul(&t, &chal[r], &t); gr_add(&qh, &q[0], &t);
    gr e1; eq_one(&e1, &rt[r], &chal[r]);
    gr_mul(&c, &e1, &qh);
  }
  (void)u;
  *c_out = c;
  return -2;
}
static void eq_final(const gr *rt, const gr *chal, size_t nv, gr *o) {
  gr_one(o);
  for (size_t i = 0; i < nv; i++) { gr e; eq_one(&e, &rt[i], &chal[i]); gr_mul(o, o, &e); }
}
/* returns 0 ok, 1 + round on round failure, -1 on final failure */
int verify_mult_check(const sc_t *pf, const gr *rt, size_t nv, ch_t *ch, gr *chal) {
  gr c;
  int r = verify_gruen(pf, rt, nv, ch, chal, &pf->claimed, &c);
  if (r >= 0) return 1 + r;
  if (r == -1) return -1;
  gr ef, t;
  eq_final(rt, chal, nv, &ef);
  gr_mul(&t, &pf->fin[0], &pf->fin[1]);
  gr_mul(&t, &ef, &t);
  return gr_eq(&c, &t) ? 0 : -1;
}

/* mult_check.rs:426-514 */
void prove_mult_check_sub(const gr *a0, const gr *b0, const gr *c0, const gr *e0, size_t nv, const gr *rt, ch_t *ch, sc_t *pf, gr *chal) {
  size_t N = (size_t)1 << nv;
  gr *a = clone_tab(a0, N), *b = clone_tab(b0, N), *c = clone_tab(c0, N), *e = clone_tab(e0, N);
  size_t eglen = N / 2;
  gr *eg = xmalloc((eglen ? eglen : 1) * sizeof(gr));
  if (nv > 1) build_eq_table(rt + 1, nv - 1, eg); else gr_one(&eg[0]);
  memset(pf, 0, sizeof *pf);
  pf->n = nv;
  pf->rounds = xmalloc(3 * nv * sizeof(gr));
  size_t len = N;
  for (size_t r = 0; r < nv; r++) {
    size_t m = len / 2;
    gr s0, s1, si;
    gr_zero(&s0); gr_zero(&s1); gr_zero(&si);
#pragma omp parallel for if (m >= 64) reduction(gr_plus : s0, s1, si)
    for (size_t i = 0; i < m; i++) {
      gr da, db, t, u;
      gr_sub(&da, &a[i + m], &a[i]);
      gr_sub(&db, &b[i + m], &b[i]);
      gr_mul(&t, &a[i], &b[i]); gr_sub(&t, &t, &c[i]); gr_sub(&t, &t, &e[i]); gr_mul(&u, &eg[i], &t); gr_add(&s0, &s0, &u);
      gr_mul(&t, &a[i + m], &b[i + m]); gr_sub(&t, &t, &c[i + m]); gr_sub(&t, &t, &e[i + m]); gr_mul(&u, &eg[i], &t); gr_add(&s1, &s1, &u);
      gr_mul(&t, &da, &db); gr_mul(&u, &eg[i], &t); gr_add(&si, &si, &u);
    }
    gr *q = &pf->rounds[3 * r];
    q[0] = s0; q[2] = si;
    gr_sub(&q[1], &s1, &s0); gr_sub(&q[1], &q[1], &si);
    ch_observe(ch, q, 3);
    ch_sample(ch, &chal[r]);
    mle_fold(a, len, &chal[r]); mle_fold(b, len, &chal[r]);
    mle_fold(c, len, &chal[r]); mle_fold(e, len, &chal[r]);
    len = m;
    if (r + 1 < nv) { gr sc; eq_one(&sc, &rt[r], &chal[r]); marginalize_and_scale(eg, m, &sc); }
  }
  pf->nfin = 4;
  pf->fin[0] = a[0]; pf->fin[1] = b[0]; pf->fin[2] = c[0]; pf->fin[3] = e[0];
  free(a); free(b); free(c); free(e); free(eg);
}
int verify_mult_check_sub(const sc_t *pf, const gr *rt, size_t nv, ch_t *ch, gr *chal) {
  gr zero, c;
  gr_zero(&zero);
  int r = verify_gruen(pf, rt, nv, ch, chal, &zero, &c);
  if (r >= 0) return 1 + r;
  if (r == -1) return -1;
  gr ef, t;
  eq_final(rt, chal, nv, &ef);
  gr_mul(&t, &pf->fin[0], &pf->fin[1]);
  gr_sub(&t, &t, &pf->fin[2]);
  gr_sub(&t, &t, &pf->fin[3]);
  gr_mul(&t, &ef, &t);
  return gr_eq(&c, &t) ? 0 : -1;
}

/* linear_check_simd.rs:196-286 (tables are folded in place) */
void prove_linear_simd(gr *f, gr *g, size_t nv, const gr *claim, ch_t *ch, sc_t *pf, gr *chal) {
  memset(pf, 0, sizeof *pf);
  pf->claimed = *claim;
  pf->n = nv;
  pf->rounds = xmalloc(3 * nv * sizeof(gr));
  gr c = *claim, t;
  size_t len = (size_t)1 << nv;
  for (size_t r = 0; r < nv; r++) {
    size_t m = len / 2;
    gr s0, si;
    gr_zero(&s0); gr_zero(&si);
#pragma omp parallel for if (m >= 64) reduction(gr_plus : s0, si)
    for (size_t i = 0; i < m; i++) {
      gr df, dg, t;
      gr_mul(&t, &f[i], &g[i]); gr_add(&s0, &s0, &t);
      gr_sub(&df, &f[i + m], &f[i]); gr_sub(&dg, &g[i + m], &g[i]);
      gr_mul(&t, &df, &dg); gr_add(&si, &si, &t);
    }
    gr *q = &pf->rounds[3 * r];
    q[0] = s0; q[2] = si;
    gr two; gr_add(&two, &s0, &s0);
    gr_sub(&q[1], &c, &two); gr_sub(&q[1], &q[1], &si);
    ch_observe(ch, q, 3);
    ch_sample(ch, &chal[r]);
    gr_mul(&t, &chal[r], &q[2]); gr_add(&t, &q[1], &t); gr_mul(&t, &chal[r], &t); gr_add(&c, &q[0], &t);
    mle_fold(f, len, &chal[r]);
    mle_fold(g, len, &chal[r]);
    len = m;
  }
  pf->nfin = 2;
  pf->fin[0] = f[0]; pf->fin[1] = g[0];
}
static int verify_lin_rounds(const sc_t *pf, ch_t *ch, gr *chal, gr *c_out) {
  gr c = pf->claimed, t;
  for (size_t r = 0; r < pf->n; r++) {
    const gr *q = &pf->rounds[3 * r];
    gr lhs; gr_add(&lhs, &q[0], &q[0]); gr_add(&lhs, &lhs, &q[1]); gr_add(&lhs, &lhs, &q[2]);
    if (!gr_eq(&lhs, &c)) return (int)r;
    ch_observe(ch, q, 3);
    ch_sample(ch, &chal[r]);
    gr_mul(&t, &chal[r], &q[2]); gr_add(&t, &q[1], &t); gr_mul(&t, &chal[r], &t); gr_add(&c, &q[0], &t);
  }
  *c_out = c;
  return -2;
}
int verify_linear_simd(const sc_t *pf, ch_t *ch, gr *chal) {
  gr c, t;
  int r = verify_lin_rounds(pf, ch, chal, &c);
  if (r >= 0) return 1 + r;
  gr_mul(&t, &pf->fin[0], &pf->fin[1]);
  return gr_eq(&c, &t) ? 0 : -1;
}

/* linear_check_general.rs:214-350: rows[l], rhos[l] (32 layers) folded in place */
void prove_linear_general(gr *const *rows, gr *const *rhos, size_t nv, const gr *claim, ch_t *ch, sc_t *pf, gr *chal) {
  memset(pf, 0, sizeof *pf);
  pf->claimed = *claim;
  pf->n = nv;
  pf->rounds = xmalloc(3 * nv * sizeof(gr));
  gr c = *claim, t;
  size_t len = (size_t)1 << nv;
  for (size_t r = 0; r < nv; r++) {
    size_t m = len / 2;
    gr s0, si;
    gr_zero(&s0); gr_zero(&si);
#pragma omp parallel for collapse(2) if (m >= 4) reduction(gr_plus : s0, si)
    for (int l = 0; l < 32; l++)
      for (size_t i = 0; i < m; i++) {
        gr dr, dp, t;
        gr_mul(&t, &rows[l][i], &rhos[l][i]); gr_add(&s0, &s0, &t);
        gr_sub(&dr, &rows[l][i + m], &rows[l][i]); gr_sub(&dp, &rhos[l][i + m], &rhos[l][i]);
        gr_mul(&t, &dr, &dp); gr_add(&si, &si, &t);
      }
    gr *q = &pf->rounds[3 * r];
    q[0] = s0; q[2] = si;
    gr two; gr_add(&two, &s0, &s0);
    gr_sub(&q[1], &c, &two); gr_sub(&q[1], &q[1], &si);
    ch_observe(ch, q, 3);
    ch_sample(ch, &chal[r]);
    gr_mul(&t, &chal[r], &q[2]); gr_add(&t, &q[1], &t); gr_mul(&t, &chal[r], &t); gr_add(&c, &q[0], &t);
#pragma omp parallel for if (m >= 4)
    for (int l = 0; l < 64; l++) mle_fold(l < 32 ? rows[l] : rhos[l - 32], len, &chal[r]);
    len = m;
  }
  pf->nfin = 32;
  for (int l = 0; l < 32; l++) { pf->fin[l] = rows[l][0]; pf->fin2[l] = rhos[l][0]; }
}
int verify_linear_general(const sc_t *pf, ch_t *ch, gr *chal) {
  gr c, t, s;
  int r = verify_lin_rounds(pf, ch, chal, &c);
  if (r >= 0) return 1 + r;
  gr_zero(&s);
  for (int l = 0; l < 32; l++) { gr_mul(&t, &pf->fin[l], &pf->fin2[l]); gr_add(&s, &s, &t); }
  return gr_eq(&c, &s) ? 0 : -1;
}
