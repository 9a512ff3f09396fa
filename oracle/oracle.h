/*
 * Nenya CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A bit-exact C restatement of the reference prover/verifier
 * (tnyuzg/nenya, the crates under crates/nenya) used solely as the parity
 * checker for the
 * HIP product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in nenya_amd/ links or calls it.
 *
 * Representation: every base-ring limb is a uint64_t holding a value of
 * Z/2^k (k in {8,16,32,64}), always kept reduced (masked).  Wrapping u64
 * arithmetic followed by a mask is exact mod 2^k because 2^k | 2^64.
 * A GR(2^k,4) element is 4 limbs, a GR(2^k,128) element is 128 limbs laid
 * out [slot i (beta power)][alpha power c] -> index 4*i+c, exactly the
 * reference layout (galois-ring/src/gr128.rs:23-27).
 *
 * The active width is process-global (or_set_width); the oracle is not
 * re-entrant across widths, which is fine for a test checker.
 *
 * Parity status per function is stated in the header of each source file:
 * BLAKE3 pinned by published vectors + libclang's
 * llvm_blake3; ChaCha block pinned by libsodium (20 rounds); rand 0.8.5
 * stream semantics (seed_from_u64 PCG32 expansion, 4-block buffer,
 * next_u64 straddle, gen_range widening multiply) are restated from the
 * published crates and are "parity unpinned" (no golden vector exists in
 * the reference).
 */
#ifndef NENYA_ORACLE_H
#define NENYA_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#include <string.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint8_t u8;

/* ---------------- width ---------------- */
extern int g_k;      /* bits */
extern u64 g_mask;   /* 2^k - 1 */
extern int g_nb;     /* bytes per limb = k/8 */
void or_set_width(int k);
static inline u64 mk(u64 x) { return x & g_mask; }

/* ---------------- ring (galois-ring) ---------------- */
#define E 128 /* limbs per Gr128 */
typedef struct { u64 v[E]; } gr;
typedef struct { u64 v[4]; } g4;

void g4_mul(g4 *o, const g4 *a, const g4 *b);
void gr_zero(gr *o);
void gr_one(gr *o);
void gr_add(gr *o, const gr *a, const gr *b);
void gr_sub(gr *o, const gr *a, const gr *b);
void gr_neg(gr *o, const gr *a);
void gr_mul(gr *o, const gr *a, const gr *b);
void gr_mul_gr1(gr *o, const gr *a, u64 s);
void gr_mul_gr4(gr *o, const gr *a, const g4 *s);
int gr_eq(const gr *a, const gr *b);
int gr_is_zero(const gr *a);
void gr_from_g4(gr *o, const g4 *x);
u64 inv_2adic(u64 x);
int ceil_log2_u(unsigned n);
void teich_lift(gr *x);                /* teichmuller.rs:104-114 */
void compute_t1(g4 t1[16]);            /* teichmuller.rs:65-76 */

/* packing constants (reduction/src/constants.rs:198-302) */
typedef struct {
  u64 inv15;
  g4 l_alpha[2];
  g4 v_alpha;
  g4 q_alpha[2][2];
  gr l_beta[16];
  gr v_beta;
  gr q_beta[16][16];
  gr q_alpha_l_beta[16][2][2];
  g4 t1[16];
  gr basis[32]; /* pack_r2(e_n): pcs.rs:49-55 */
} consts_t;
const consts_t *consts(void);

void pack_r2(gr *o, const u64 a[32]);                  /* embedding.rs:63-65 */
void psi_r2(u64 out[32], const gr *x);                 /* embedding.rs:247-261, flattened */
int in_kernel(const gr *x);                            /* embedding.rs:274-289 */
void mul_with_quotients_only(gr *q1, gr *q2, const u64 a[32], const u64 b[32]); /* embedding.rs:172-212 */

/* ---------------- byte buffer ---------------- */
typedef struct { u8 *p; size_t n, cap; } buf_t;
void buf_put(buf_t *b, const void *d, size_t n);
void buf_u64(buf_t *b, u64 x);
void buf_u32(buf_t *b, u32 x);
void buf_limb(buf_t *b, u64 x);
void buf_gr(buf_t *b, const gr *x);
void buf_free(buf_t *b);

/* ---------------- BLAKE3 ---------------- */
typedef struct {
  u32 key[8];
  u32 cv[8];
  u64 chunk_counter;
  u8 block[64];
  u8 block_len;
  u8 blocks_compressed;
  u32 flags;
  u32 cv_stack[54][8];
  u8 cv_stack_len;
} b3_t;
void b3_init(b3_t *h);
void b3_update(b3_t *h, const void *data, size_t len);
void b3_finalize_seek(const b3_t *h, u64 seek, u8 *out, size_t out_len);
void b3_hash(const void *data, size_t len, u8 out[32]);

/* ---------------- rand 0.8.5 StdRng (ChaCha12) ---------------- */
typedef struct {
  u32 key[8];
  u64 counter; /* next block index */
  u32 buf[64];
  int idx;
  int rounds; /* double rounds; 6 for ChaCha12 */
} rng_t;
void chacha_block(u32 out[16], const u32 key[8], u64 counter, u64 stream, int double_rounds);
void rng_from_seed(rng_t *r, const u8 seed[32]);
void rng_seed_from_u64(rng_t *r, u64 state);
u32 rng_next_u32(rng_t *r);
u64 rng_next_u64(rng_t *r);
u64 rng_gen_range_incl(rng_t *r, u64 low, u64 high);
u64 rng_limb(rng_t *r); /* L::random = rng.gen::<L>() (limb.rs:103-105) */
void r2_sample(rng_t *r, gr *out); /* sampling.rs:57-68 */

/* ---------------- transcript (snark-ring/src/transcript.rs) ---------------- */
typedef struct { b3_t h; } tr_t;
void tr_new(tr_t *t, const char *label);
void tr_label(tr_t *t, const char *label);
void tr_root(tr_t *t, const u8 root[32]);
void tr_gr_slice(tr_t *t, const char *label, const gr *xs, size_t n);
void tr_sampler(tr_t *t, const char *label, rng_t *out);
void tr_challenge(tr_t *t, const char *label, gr *out);
void tr_target(tr_t *t, const char *label, size_t n, gr *out);

/* ---------------- RAA + Merkle (pcs-ring) ---------------- */
typedef struct {
  size_t k, q, n;
  u32 *perm1, *perm2;
  u64 *diag1, *diag2;
} raa_t;
void raa_setup(raa_t *p, size_t k, size_t q, u64 seed);
void raa_free(raa_t *p);
void raa_encode(const raa_t *p, const gr *msg, gr *out);            /* raa.rs:67-98 */
void raa_encode_packed(const raa_t *p, const u64 *msg, u64 *out);   /* raa.rs:106-153, 32 limbs/entry */

typedef struct { size_t n; u8 (*nodes)[32]; } merkle_t;
void merkle_build(merkle_t *m, u8 (*leaves)[32], size_t nleaves);
void merkle_open(const merkle_t *m, size_t j, u8 (*path)[32]); /* log2(n) hashes */
int merkle_verify(const u8 leaf[32], size_t idx, const u8 (*path)[32], size_t plen, const u8 root[32]);
void merkle_free(merkle_t *m);

/* ---------------- PCS ---------------- */
typedef struct {
  raa_t raa;
  size_t k1, k2, ell;
  int packed; /* 1: pcs_ring::Pcs (32-limb columns); 0: R2Pcs (128-limb columns) */
} pcs_t;
size_t recommended_query_size(double delta, size_t lambda);
size_t proof_size_bytes_for(size_t k1, size_t k2, size_t n, size_t t, size_t mbar_b, size_t col_b);
void pcs_setup(pcs_t *p, size_t nv, size_t q, double delta, size_t lambda, u64 seed);
void pcs_setup_batched(pcs_t *p, size_t nv, size_t q, double delta, size_t lambda, u64 seed, size_t nopen);
void pcs_setup_stacked(pcs_t *p, size_t nv, size_t q, double delta, size_t lambda, u64 seed, size_t stack);
void r2pcs_setup(pcs_t *p, size_t nv, size_t q, double delta, size_t lambda, u64 seed);
void pcs_new(pcs_t *p, size_t k1, size_t k2, size_t q, size_t ell, u64 seed, int packed);
void pcs_free(pcs_t *p);

typedef struct {
  u8 root[32];
  size_t rows; /* s*k1 */
  size_t s;
  u64 *encoded;   /* rows*n entries (32 or 128 limbs) */
  gr *message;    /* rows*k2 (lifted) */
  u64 *slots;     /* rows*k2*32 slot form (packed PCS only) */
  merkle_t tree;
} comm_t;
void comm_free(comm_t *c);

typedef struct {
  size_t nmbar;   /* number of mbar vectors (1, or 32 for rotated batch) */
  size_t k2;
  gr *mbar;       /* nmbar*k2 */
  size_t t;
  u32 *idx;
  size_t col_len; /* entries per column */
  int col_elem;   /* 32 or 128 limbs per entry */
  u64 *cols;      /* t*col_len*col_elem */
  size_t plen;
  u8 (*paths)[32]; /* t*plen */
} open_t;
void open_free(open_t *o);
void open_serialize(buf_t *b, const open_t *o, int batch);

void pcs_commit(const pcs_t *p, const u64 *slots, comm_t *c);                     /* pcs.rs:489-543 */
void pcs_commit_stacked(const pcs_t *p, const u64 *const *msgs, size_t s, comm_t *c); /* pcs.rs:670-730 */
void r2pcs_commit(const pcs_t *p, const gr *evals, comm_t *c);                    /* r2_pcs.rs:259-294 */
void r2pcs_commit_stacked(const pcs_t *p, const gr *const *msgs, size_t s, comm_t *c); /* r2_pcs.rs:452-494 */
void pcs_open_weighted(const pcs_t *p, const comm_t *c, const gr *z, const gr *w, gr *y, open_t *o);
void pcs_rotated_open_batch(const pcs_t *p, const comm_t *c, const gr *z, size_t nl, gr *ys, open_t *o);
int pcs_verify_weighted(const pcs_t *p, const u8 root[32], const gr *z, const gr *w, size_t s,
                        const gr *val, const open_t *o);
int pcs_rotated_verify_batch(const pcs_t *p, const u8 root[32], const gr *z, const gr *ys, size_t nl,
                             const open_t *o);
int columns_in_kernel(const open_t *o);

/* ---------------- PIOP ---------------- */
void build_eq_table(const gr *r, size_t n, gr *out); /* eq.rs:35-58, out has 2^n entries */
void mle_fold(gr *t, size_t len, const gr *r);        /* mle.rs:42-54, in place; len halves */

typedef struct {
  gr claimed;           /* plain mult check only */
  size_t n;
  gr *rounds;           /* n*3 */
  gr fin[32];           /* finals */
  size_t nfin;
  gr fin2[32];          /* second finals array (general lin: final_rho) */
} sc_t;
void sc_free(sc_t *s);

/* challenger: transcript or lock-step sampler */
typedef struct { int lockstep; tr_t *tr; rng_t rng; } ch_t;
void ch_observe(ch_t *c, const gr *msg, size_t n);
void ch_sample(ch_t *c, gr *out);

void prove_mult_check(const gr *a, const gr *b, size_t nv, const gr *rt, const gr *claim, ch_t *ch, sc_t *pf, gr *chal);
int verify_mult_check(const sc_t *pf, const gr *rt, size_t nv, ch_t *ch, gr *chal);
void prove_mult_check_sub(const gr *a, const gr *b, const gr *c, const gr *e, size_t nv, const gr *rt, ch_t *ch, sc_t *pf, gr *chal);
int verify_mult_check_sub(const sc_t *pf, const gr *rt, size_t nv, ch_t *ch, gr *chal);
void prove_linear_simd(gr *arow, gr *w, size_t nv, const gr *claim, ch_t *ch, sc_t *pf, gr *chal);
int verify_linear_simd(const sc_t *pf, ch_t *ch, gr *chal);
void prove_linear_general(gr *const *rows, gr *const *rhos, size_t nv, const gr *claim, ch_t *ch, sc_t *pf, gr *chal);
int verify_linear_general(const sc_t *pf, ch_t *ch, gr *chal);

/* ---------------- sparse matrices ---------------- */
typedef struct { size_t rows, cols, nnz; u32 *r, *c; u64 *v; } coo_t;

/* ---------------- SNARKs ---------------- */
enum { VE_OK = 0, VE_MULT = 1, VE_LIN = 2, VE_LINCLAIM = 3, VE_AROW = 4, VE_PCSW = 5, VE_PCSE = 6, VE_KERNEL = 7, VE_PARSE = 8 };
int simd_prove(int log_n, int log_m, const coo_t *a, const coo_t *b, const coo_t *c, const u64 *wit /*32*m*/,
               u64 seed_w, u64 seed_q, buf_t *out);
int simd_verify(int log_n, int log_m, const coo_t *a, const coo_t *b, const coo_t *c, u64 seed_w, u64 seed_q,
                const u8 *proof, size_t len);

typedef struct {
  size_t ncell[3][32];
  u32 *p[3][32], *q[3][32]; /* cell coordinates, (l,p,q) sorted */
  u64 *d[3][32];            /* 32 slots per cell */
  gr *dhat[3][32];          /* pack_r2(d) */
} gdiag_t;
void gdiag_build(gdiag_t *g, const coo_t *m3[3], size_t n); /* reduce.rs:372-438, 618-651 */
void gdiag_free(gdiag_t *g);
int general_prove(int log_b, const gdiag_t *diag, const u64 *w, u64 seed_w, u64 seed_e, buf_t *out);
int general_verify(int log_b, const gdiag_t *diag, u64 seed_w, u64 seed_e, const u8 *proof, size_t len);

typedef struct {
  size_t rows, cols;
  size_t ncell[3];
  u32 *r[3], *c[3]; /* union-support cells, (r,c) sorted */
  u64 *d[3];        /* 32 instance values per cell */
  gr *dhat[3];      /* pack_r2(d) */
} mmats_t;
/* pack_mimd_matrices (reduce.rs:249-273): sets[w][k] = instance k of matrix w */
void mimd_pack(mmats_t *mm, size_t rows, size_t cols, const coo_t *const *sets[3]);
void mimd_free(mmats_t *mm);
int mimd_prove(int log_n, int log_m, const mmats_t *mm, const u64 *wit /*32*m*/, u64 seed_w, u64 seed_e, buf_t *out);
int mimd_verify(int log_n, int log_m, const mmats_t *mm, u64 seed_w, u64 seed_e, const u8 *proof, size_t len);

/* ---------------- misc ---------------- */
void *xmalloc(size_t n);
void *xcalloc(size_t n, size_t s);

/* ---------------- threads ----------------
 * The prover loops run on OpenMP threads (the CPU-baseline timing runs the
 * reference algorithm over all host cores).  Every sum is exact arithmetic
 * mod 2^k, so partial sums combined in any order give the same bytes as one
 * thread; scatters are split so that no two threads write one output. */
#ifdef _OPENMP
#include <omp.h>
#endif
#pragma omp declare reduction(gr_plus : gr : gr_add(&omp_out, &omp_out, &omp_in)) initializer(gr_zero(&omp_priv))
/* starts of the runs of equal keys in key[0..n): writes them to *out
 * (allocated, n_runs + 1 entries, the last = n) and returns n_runs */
size_t key_runs(const u32 *key, size_t n, size_t **out);
#endif
