/*
 * Oracle: GR(2^k,4) / GR(2^k,128) arithmetic, Teichmuller lift, packing
 * constants and the RMFE phi/psi maps.  TEST INFRASTRUCTURE ONLY.
 *
 * Follows: galois-ring/src/{limb.rs,gr4.rs,gr128.rs,teichmuller.rs},
 *          reduction/src/{constants.rs,embedding.rs}.
 * Pinned by: T1 table KAT (teichmuller.rs:21-38), 15^-1 KATs
 * (limb.rs:139-143), ring laws, X^4=-X-1 and f2(beta)=0 closed forms,
 * psi(phi(a))=a, quotient identity -- see tests/test_oracle_ring.py.
 */
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>

int g_k = 32;
u64 g_mask = 0xffffffffull;
int g_nb = 4;
static consts_t *g_consts[4];

void or_set_width(int k) {
  g_k = k;
  g_mask = (k == 64) ? ~0ull : ((1ull << k) - 1);
  g_nb = k / 8;
}

void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}
void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}
size_t key_runs(const u32 *key, size_t n, size_t **out) {
  size_t *st = xmalloc((n + 1) * sizeof(size_t)), nr = 0;
  for (size_t i = 0; i < n; i++)
    if (i == 0 || key[i] != key[i - 1]) st[nr++] = i;
  st[nr] = n;
  *out = st;
  return nr;
}

/* limb.rs:91-101 -- 6 Newton doublings */
u64 inv_2adic(u64 a) {
  u64 x = 1;
  for (int i = 0; i < 6; i++) x = mk(x * mk(2 - mk(a * x)));
  return x;
}
int ceil_log2_u(unsigned n) {
  if (n <= 1) return 0;
  return 32 - __builtin_clz(n - 1);
}

/* gr4.rs:316-341 */
void g4_mul(g4 *o, const g4 *a, const g4 *b) {
  u64 c[7] = {0};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) c[i + j] += a->v[i] * b->v[j];
  g4 r;
  r.v[0] = mk(c[0] - c[4]);
  r.v[1] = mk(c[1] - c[4] - c[5]);
  r.v[2] = mk(c[2] - c[5] - c[6]);
  r.v[3] = mk(c[3] - c[6]);
  *o = r;
}
static void g4_add(g4 *o, const g4 *a, const g4 *b) { for (int i = 0; i < 4; i++) o->v[i] = mk(a->v[i] + b->v[i]); }
static void g4_pow(g4 *o, const g4 *x, u64 n) { /* gr4.rs:259-270 */
  g4 base = *x, acc = {{1, 0, 0, 0}};
  while (n) {
    if (n & 1) g4_mul(&acc, &acc, &base);
    g4_mul(&base, &base, &base);
    n >>= 1;
  }
  *o = acc;
}

void gr_zero(gr *o) { memset(o, 0, sizeof *o); }
void gr_one(gr *o) { gr_zero(o); o->v[0] = 1; }
void gr_add(gr *o, const gr *a, const gr *b) { for (int i = 0; i < E; i++) o->v[i] = mk(a->v[i] + b->v[i]); }
void gr_sub(gr *o, const gr *a, const gr *b) { for (int i = 0; i < E; i++) o->v[i] = mk(a->v[i] - b->v[i]); }
void gr_neg(gr *o, const gr *a) { for (int i = 0; i < E; i++) o->v[i] = mk(0 - a->v[i]); }
int gr_eq(const gr *a, const gr *b) { return memcmp(a, b, sizeof *a) == 0; }
int gr_is_zero(const gr *a) { for (int i = 0; i < E; i++) if (a->v[i]) return 0; return 1; }
void gr_from_g4(gr *o, const g4 *x) { gr_zero(o); for (int c = 0; c < 4; c++) o->v[c] = x->v[c]; }

/* gr128.rs:80-93 */
void gr_mul_gr1(gr *o, const gr *a, u64 s) { for (int i = 0; i < E; i++) o->v[i] = mk(a->v[i] * s); }

/* gr128.rs:105-149 */
void gr_mul_gr4(gr *o, const gr *a, const g4 *s) {
  gr r;
  for (int i = 0; i < 32; i++) {
    g4 v = {{a->v[4 * i], a->v[4 * i + 1], a->v[4 * i + 2], a->v[4 * i + 3]}}, p;
    g4_mul(&p, s, &v);
    for (int c = 0; c < 4; c++) r.v[4 * i + c] = p.v[c];
  }
  *o = r;
}

/* gr128.rs:215-296 -- schoolbook over the 32 beta slots with the alpha
 * product reduced per term, then the cascading Y^32 = -Y^3 - Y - alpha
 * reduction from i=30 down to 0. */
static int top_slot(const gr *a) {
  for (int i = 31; i >= 0; i--)
    if (a->v[4 * i] | a->v[4 * i + 1] | a->v[4 * i + 2] | a->v[4 * i + 3]) return i;
  return -1;
}
/* Zero high slots contribute nothing to the schoolbook, so skipping them is
 * exact; it only speeds up the constant-coefficient polynomial algebra. */
void gr_mul(gr *o, const gr *a, const gr *b) {
  u64 cp[4][63];
  memset(cp, 0, sizeof cp);
  int na = top_slot(a), nbb = top_slot(b);
  for (int i = 0; i <= na; i++) {
    const u64 *s = &a->v[4 * i];
    for (int j = 0; j <= nbb; j++) {
      const u64 *v = &b->v[4 * j];
      u64 t0 = s[0] * v[0];
      u64 t1 = s[0] * v[1] + s[1] * v[0];
      u64 t2 = s[0] * v[2] + s[1] * v[1] + s[2] * v[0];
      u64 t3 = s[0] * v[3] + s[1] * v[2] + s[2] * v[1] + s[3] * v[0];
      u64 t4 = s[1] * v[3] + s[2] * v[2] + s[3] * v[1];
      u64 t5 = s[2] * v[3] + s[3] * v[2];
      u64 t6 = s[3] * v[3];
      int k = i + j;
      cp[0][k] += t0 - t4;
      cp[1][k] += t1 - t4 - t5;
      cp[2][k] += t2 - t5 - t6;
      cp[3][k] += t3 - t6;
    }
  }
  for (int i = 30; i >= 0; i--) {
    u64 h0 = cp[0][32 + i], h1 = cp[1][32 + i], h2 = cp[2][32 + i], h3 = cp[3][32 + i];
    cp[0][3 + i] -= h0; cp[1][3 + i] -= h1; cp[2][3 + i] -= h2; cp[3][3 + i] -= h3;
    cp[0][1 + i] -= h0; cp[1][1 + i] -= h1; cp[2][1 + i] -= h2; cp[3][1 + i] -= h3;
    cp[0][i] -= (0 - h3); cp[1][i] -= (h0 - h3); cp[2][i] -= h1; cp[3][i] -= h2;
  }
  gr r;
  for (int k = 0; k < 32; k++)
    for (int c = 0; c < 4; c++) r.v[4 * k + c] = mk(cp[c][k]);
  *o = r;
}

/* teichmuller.rs:104-114: (ceil_log2(k)+1) x 128 squarings */
void teich_lift(gr *x) {
  int outer = ceil_log2_u((unsigned)g_k) + 1;
  for (int o = 0; o < outer; o++)
    for (int s = 0; s < 128; s++) gr_mul(x, x, x);
}

/* teichmuller.rs:54-76 */
void compute_t1(g4 t1[16]) {
  memset(&t1[0], 0, sizeof(g4));
  g4 alpha = {{0, 1, 0, 0}}, pw = {{1, 0, 0, 0}};
  for (int i = 1; i < 16; i++) {
    g4 x = pw;
    for (int it = 0; it < g_k; it++) g4_pow(&x, &x, 16);
    t1[i] = x;
    g4_mul(&pw, &pw, &alpha);
  }
}

/* ---------------- polynomials over R2 (constants.rs:43-159) ----------------
 * Coefficients here are all constants embedded from R1 (only slot 0 can be
 * non-zero) until evaluation at beta; products use the exact Gr128 multiply. */
typedef struct { gr *c; int n; } poly;
static void p_norm(poly *p) { while (p->n > 0 && gr_is_zero(&p->c[p->n - 1])) p->n--; }
static poly p_new(int n) { poly p; p.c = xcalloc(n > 0 ? n : 1, sizeof(gr)); p.n = n; return p; }
static void p_free(poly *p) { free(p->c); p->c = NULL; p->n = 0; }
static poly p_mul(const poly *a, const poly *b) {
  if (!a->n || !b->n) return p_new(0);
  poly r = p_new(a->n + b->n - 1);
  gr t;
  for (int i = 0; i < a->n; i++)
    for (int j = 0; j < b->n; j++) { gr_mul(&t, &a->c[i], &b->c[j]); gr_add(&r.c[i + j], &r.c[i + j], &t); }
  p_norm(&r);
  return r;
}
static poly p_sub(const poly *a, const poly *b) {
  int n = a->n > b->n ? a->n : b->n;
  poly r = p_new(n);
  for (int i = 0; i < a->n; i++) gr_add(&r.c[i], &r.c[i], &a->c[i]);
  for (int i = 0; i < b->n; i++) gr_sub(&r.c[i], &r.c[i], &b->c[i]);
  p_norm(&r);
  return r;
}
static poly p_add(const poly *a, const poly *b) {
  int n = a->n > b->n ? a->n : b->n;
  poly r = p_new(n);
  for (int i = 0; i < a->n; i++) gr_add(&r.c[i], &r.c[i], &a->c[i]);
  for (int i = 0; i < b->n; i++) gr_add(&r.c[i], &r.c[i], &b->c[i]);
  p_norm(&r);
  return r;
}
/* constants.rs:97-124 (monic divisor) */
static void p_div(const poly *a, const poly *b, poly *q, poly *rem) {
  if (a->n < b->n) { *q = p_new(0); *rem = p_new(a->n); memcpy(rem->c, a->c, a->n * sizeof(gr)); return; }
  poly r = p_new(a->n);
  memcpy(r.c, a->c, a->n * sizeof(gr));
  poly qq = p_new(a->n - b->n + 1);
  gr t;
  while (r.n >= b->n) {
    int dr = r.n - 1, db = b->n - 1, qi = dr - db;
    gr f = r.c[r.n - 1];
    gr_add(&qq.c[qi], &qq.c[qi], &f);
    for (int i = 0; i < b->n; i++) { gr_mul(&t, &f, &b->c[i]); gr_sub(&r.c[qi + i], &r.c[qi + i], &t); }
    p_norm(&r);
  }
  p_norm(&qq);
  *q = qq;
  *rem = r;
}
static void eval_at_beta(gr *o, const poly *p) { /* constants.rs:152-159 */
  gr b, acc;
  gr_zero(&b); b.v[4] = 1;
  gr_zero(&acc);
  for (int i = p->n - 1; i >= 0; i--) { gr_mul(&acc, &acc, &b); gr_add(&acc, &acc, &p->c[i]); }
  *o = acc;
}

static void consts_compute(consts_t *C) {
  C->inv15 = inv_2adic(15);
  gr inv15r; gr_zero(&inv15r); inv15r.v[0] = C->inv15;
  g4 one = {{1, 0, 0, 0}}, a = {{0, 1, 0, 0}}, aa;
  C->l_alpha[0] = (g4){{1, mk(0 - 1), 0, 0}};
  C->l_alpha[1] = a;
  g4_mul(&aa, &a, &a);
  C->v_alpha = (g4){{mk(aa.v[0] - a.v[0]), mk(aa.v[1] - a.v[1]), mk(aa.v[2] - a.v[2]), mk(aa.v[3] - a.v[3])}};
  g4 mone = {{mk(0 - 1), 0, 0, 0}};
  C->q_alpha[0][0] = one; C->q_alpha[0][1] = mone; C->q_alpha[1][0] = mone; C->q_alpha[1][1] = one;
  compute_t1(C->t1);
  gr e[16];
  for (int i = 0; i < 16; i++) gr_from_g4(&e[i], &C->t1[i]);
  /* build_l_beta (constants.rs:163-187) */
  poly lp[16];
  lp[0] = p_new(16); gr_one(&lp[0].c[0]); gr_one(&lp[0].c[15]); gr_neg(&lp[0].c[15], &lp[0].c[15]); p_norm(&lp[0]);
  for (int i = 1; i < 16; i++) {
    poly p = p_new(1); gr_one(&p.c[0]);
    for (int j = 0; j < 16; j++) {
      if (j == i) continue;
      poly f = p_new(2); gr_neg(&f.c[0], &e[j]); gr_one(&f.c[1]);
      poly np = p_mul(&p, &f);
      p_free(&p); p_free(&f);
      p = np;
    }
    for (int c = 0; c < p.n; c++) gr_mul(&p.c[c], &p.c[c], &inv15r);
    p_norm(&p);
    lp[i] = p;
  }
  poly vb = p_new(17); gr_one(&vb.c[1]); gr_neg(&vb.c[1], &vb.c[1]); gr_one(&vb.c[16]); p_norm(&vb);
  for (int i = 0; i < 16; i++) eval_at_beta(&C->l_beta[i], &lp[i]);
  eval_at_beta(&C->v_beta, &vb);
  /* sanity: partition of unity as polynomials */
  poly sum = p_new(0);
  for (int i = 0; i < 16; i++) { poly s2 = p_add(&sum, &lp[i]); p_free(&sum); sum = s2; }
  if (!(sum.n == 1 && sum.c[0].v[0] == 1)) { fprintf(stderr, "oracle: beta partition of unity failed\n"); abort(); }
  p_free(&sum);
  for (int i = 0; i < 16; i++)
    for (int j = 0; j < 16; j++) {
      poly prod = p_mul(&lp[i], &lp[j]);
      poly td = (i == j) ? p_sub(&prod, &lp[i]) : prod;
      poly q, rem;
      p_div(&td, &vb, &q, &rem);
      if (rem.n != 0) { fprintf(stderr, "oracle: beta quotient remainder nonzero\n"); abort(); }
      eval_at_beta(&C->q_beta[i][j], &q);
      p_free(&q); p_free(&rem);
      if (i == j) p_free(&td);
      p_free(&prod);
    }
  for (int i = 0; i < 16; i++)
    for (int j = 0; j < 2; j++)
      for (int k = 0; k < 2; k++) {
        gr qa; gr_from_g4(&qa, &C->q_alpha[j][k]);
        gr_mul(&C->q_alpha_l_beta[i][j][k], &qa, &C->l_beta[i]);
      }
  for (int i = 0; i < 16; i++) p_free(&lp[i]);
  p_free(&vb);
  g_consts[0] = g_consts[0]; /* no-op: keep cache array referenced */
  for (int n = 0; n < 32; n++) {
    u64 s[32] = {0};
    s[n] = 1;
    /* pack_r2 needs C filled already (l_beta, l_alpha) */
    gr acc; gr_zero(&acc);
    for (int i = 0; i < 16; i++) {
      g4 row;
      for (int c = 0; c < 4; c++) row.v[c] = mk(s[2 * i] * C->l_alpha[0].v[c] + s[2 * i + 1] * C->l_alpha[1].v[c]);
      gr t; gr_mul_gr4(&t, &C->l_beta[i], &row);
      gr_add(&acc, &acc, &t);
    }
    C->basis[n] = acc;
  }
}

const consts_t *consts(void) {
  int w = g_k == 8 ? 0 : g_k == 16 ? 1 : g_k == 32 ? 2 : 3;
  consts_t *c = __atomic_load_n(&g_consts[w], __ATOMIC_ACQUIRE);
  if (c) return c;
  /* first use of a width, possibly from several threads: compute once */
#pragma omp critical(oracle_consts)
  {
    c = __atomic_load_n(&g_consts[w], __ATOMIC_ACQUIRE);
    if (!c) {
      c = xcalloc(1, sizeof(consts_t));
      consts_compute(c);
      __atomic_store_n(&g_consts[w], c, __ATOMIC_RELEASE);
    }
  }
  return c;
}

/* embedding.rs:30-65: phi_2 on the flat 32-vector, 16x2 view m[i][j]=a[2i+j] */
void pack_r2(gr *o, const u64 a[32]) {
  const consts_t *C = consts();
  gr acc; gr_zero(&acc);
  for (int i = 0; i < 16; i++) {
    g4 x0 = {{a[2 * i], 0, 0, 0}}, x1 = {{a[2 * i + 1], 0, 0, 0}}, p0, p1, row;
    g4_mul(&p0, &x0, &C->l_alpha[0]);
    g4_mul(&p1, &x1, &C->l_alpha[1]);
    g4_add(&row, &p0, &p1);
    gr t; gr_mul_gr4(&t, &C->l_beta[i], &row);
    gr_add(&acc, &acc, &t);
  }
  *o = acc;
}

/* embedding.rs:247-261, flattened as general.rs:192-200 (out[2i]=g(0), out[2i+1]=g(1)) */
void psi_r2(u64 out[32], const gr *x) {
  const consts_t *C = consts();
  for (int i = 0; i < 16; i++) {
    g4 g = {{0, 0, 0, 0}};
    for (int k = 31; k >= 0; k--) {
      g4_mul(&g, &g, &C->t1[i]);
      g4 xk = {{x->v[4 * k], x->v[4 * k + 1], x->v[4 * k + 2], x->v[4 * k + 3]}};
      g4_add(&g, &g, &xk);
    }
    out[2 * i] = g.v[0];
    out[2 * i + 1] = mk(g.v[0] + g.v[1] + g.v[2] + g.v[3]);
  }
}

int in_kernel(const gr *x) { /* embedding.rs:274-289 */
  u64 s[32];
  psi_r2(s, x);
  for (int i = 0; i < 32; i++) if (s[i]) return 0;
  return 1;
}

/* embedding.rs:172-212 */
void mul_with_quotients_only(gr *q1, gr *q2, const u64 a[32], const u64 b[32]) {
  const consts_t *C = consts();
  gr acc, t;
  gr_zero(&acc);
  for (int i = 0; i < 16; i++)
    for (int j = 0; j < 2; j++)
      for (int k = 0; k < 2; k++) {
        u64 coef = mk(a[2 * i + j] * b[2 * i + k]);
        gr_mul_gr1(&t, &C->q_alpha_l_beta[i][j][k], coef);
        gr_add(&acc, &acc, &t);
      }
  *q1 = acc;
  g4 a1[16], b1[16];
  for (int i = 0; i < 16; i++) {
    for (int c = 0; c < 4; c++) {
      a1[i].v[c] = mk(a[2 * i] * C->l_alpha[0].v[c] + a[2 * i + 1] * C->l_alpha[1].v[c]);
      b1[i].v[c] = mk(b[2 * i] * C->l_alpha[0].v[c] + b[2 * i + 1] * C->l_alpha[1].v[c]);
    }
  }
  gr_zero(&acc);
  for (int i = 0; i < 16; i++) {
    g4 pd; g4_mul(&pd, &a1[i], &b1[i]);
    gr_mul_gr4(&t, &C->q_beta[i][i], &pd);
    gr_add(&acc, &acc, &t);
    for (int k = i + 1; k < 16; k++) {
      g4 x, y, po;
      g4_mul(&x, &a1[i], &b1[k]);
      g4_mul(&y, &a1[k], &b1[i]);
      g4_add(&po, &x, &y);
      gr_mul_gr4(&t, &C->q_beta[i][k], &po);
      gr_add(&acc, &acc, &t);
    }
  }
  *q2 = acc;
}

/* ---------------- byte buffer ---------------- */
void buf_put(buf_t *b, const void *d, size_t n) {
  if (b->n + n > b->cap) {
    size_t nc = b->cap ? b->cap * 2 : 4096;
    while (nc < b->n + n) nc *= 2;
    b->p = realloc(b->p, nc);
    if (!b->p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    b->cap = nc;
  }
  memcpy(b->p + b->n, d, n);
  b->n += n;
}
void buf_u64(buf_t *b, u64 x) { u8 t[8]; for (int i = 0; i < 8; i++) t[i] = (u8)(x >> (8 * i)); buf_put(b, t, 8); }
void buf_u32(buf_t *b, u32 x) { u8 t[4]; for (int i = 0; i < 4; i++) t[i] = (u8)(x >> (8 * i)); buf_put(b, t, 4); }
void buf_limb(buf_t *b, u64 x) { u8 t[8]; for (int i = 0; i < g_nb; i++) t[i] = (u8)(x >> (8 * i)); buf_put(b, t, g_nb); }
void buf_gr(buf_t *b, const gr *x) { for (int i = 0; i < E; i++) buf_limb(b, x->v[i]); }
void buf_free(buf_t *b) { free(b->p); b->p = NULL; b->n = b->cap = 0; }
