// SIMD and General SNARK provers on the device (the drop-in prove_simd /
// prove_general of snark-ring/src/{simd,general}.rs).
#pragma once
#include <algorithm>
#include <atomic>
#include <numeric>
#include <thread>
#include <tuple>

#include "prover.hpp"

namespace nenya {

// Per-proof execution state: one HIP stream, its stream-ordered device
// pool, the pinned buffer the proof is assembled in, and the profiling
// switches.  Params own one (setup and direct calls); each server worker
// owns its own, so proofs on different contexts run concurrently.
struct CtxBase {
  Pool own_pool;
  Pool& pool;  // own_pool, or a pool its owner shares between contexts
  hipStream_t st = nullptr;
  Pinned proofbuf;  // the last proof's bytes (valid until the next proof)
  KProf prof;
  bool timing = false;
  std::vector<std::pair<std::string, double>> last_timing;
  bool own_st = true;
  CtxBase() : pool(own_pool) { HIP_CHECK(hipStreamCreate(&st)); }
  // a context on a stream and a pool its owner keeps (and destroys after it)
  CtxBase(hipStream_t shared, Pool& shared_pool) : pool(shared_pool), st(shared), own_st(false) {}
  CtxBase(const CtxBase&) = delete;
  CtxBase& operator=(const CtxBase&) = delete;
  virtual ~CtxBase() {
    if (st) {
      (void)hipStreamSynchronize(st);
      if (own_st) (void)hipStreamDestroy(st);
    }
  }
};
template <class T> struct Ctx : CtxBase {
  std::unique_ptr<Engine<T>> eng;
  Ctx() { eng.reset(new Engine<T>(&pool, st)); }
  Ctx(hipStream_t shared, Pool& shared_pool) : CtxBase(shared, shared_pool) { eng.reset(new Engine<T>(&pool, st)); }
  ~Ctx() override { eng.reset(); }
};

// Debug switch: NENYA_SKIP_SELFCHECK=1 returns proofs whose openings do not
// match the sumcheck finals instead of failing (to localize a divergence).
inline bool skip_selfcheck() {
  static const bool on = [] {
    const char* e = std::getenv("NENYA_SKIP_SELFCHECK");
    return e && e[0] == '1';
  }();
  return on;
}

struct ParamsBase {
  int k = 32;
  int kind = 0;  // 0 simd, 1 general, 2 mimd
  // the shape and seeds the params were made with (log_n, log_m for SIMD and
  // MIMD; log_blocks, 0 for General), recorded in proof containers
  int shape[2] = {0, 0};
  uint64_t seeds[2] = {0, 0};
  // device bytes of one proof: an estimate until a proof has run, then the
  // largest measured pool growth of a proof
  std::atomic<size_t> footprint{0};
  void note_footprint(size_t bytes) {
    size_t cur = footprint.load();
    while (bytes > cur && !footprint.compare_exchange_weak(cur, bytes)) {
    }
  }
  // bytes of the last proof (0 before the first): sizes the proof buffer a
  // server hands to the next job with these params
  std::atomic<size_t> proof_len{0};
  // wall-clock milliseconds of the setup's stages, in order (each stage's
  // device work finished before its mark)
  std::vector<std::pair<std::string, double>> setup_timing;
  std::chrono::steady_clock::time_point setup_last = std::chrono::steady_clock::now();
  void setup_mark(const char* name, hipStream_t st) {
    HIP_CHECK(hipStreamSynchronize(st));
    const auto now = std::chrono::steady_clock::now();
    setup_timing.push_back({name, std::chrono::duration<double, std::milli>(now - setup_last).count()});
    setup_last = now;
  }
  virtual CtxBase& ctx0() = 0;
  virtual ~ParamsBase() = default;
};
using ProofView = std::pair<const uint8_t*, size_t>;

struct HostCoo {
  size_t rows, cols, nnz;
  const uint32_t* r;
  const uint32_t* c;
  const void* v;  // nnz limbs of k bits (native width)
};

// A matrix grouped by major index (rows or columns) on the device, built by
// DevSetup::csr.
template <class T> struct Csr {
  using L = typename T::L;
  DBuf<uint32_t> ptr, idx;
  DBuf<L> val;
  size_t nnz = 0;
  // majors with at least HEAVY entries (summed by k_spmv_heavy)
  static constexpr uint32_t HEAVY = 1024;
  DBuf<uint32_t> heavy;
  size_t nheavy = 0;
  uint32_t maxheavy = 0;
};

// out[o] (+)= sum over o's entries of val * X[idx] (mul_by_gr1 matvecs,
// matrix.rs:93-105, simd.rs:570-580): a 32-lane group per output, and the
// outputs with at least Csr::HEAVY entries split over blocks of ~128
// entries each (one group walking 32k entries took 10 ms).
template <class T, class Eng>
void spmv_gr1(Eng& e, const Csr<T>& m, const typename T::L* X, typename T::L* out, uint64_t nout, int accumulate) {
  using L = typename T::L;
  e.launch("k_spmv_gr1", 0, 0, (k_spmv_gr1<T>), dim3(grid_for(nout, GPB)), dim3(GB), (const uint32_t*)m.ptr.p,
           (const uint32_t*)m.idx.p, (const L*)m.val.p, X, out, nout, accumulate,
           m.nheavy ? Csr<T>::HEAVY : 0xFFFFFFFFu);
  if (!m.nheavy) return;
  const uint32_t nch = std::min<uint32_t>(1024, (m.maxheavy + 127) / 128);
  DBuf<L> part(e.pool, m.nheavy * nch * E);
  e.launch("k_spmv_heavy", 0, 0, (k_spmv_heavy<T>), dim3(nch, (unsigned)m.nheavy), dim3(GB), (const uint32_t*)m.ptr.p,
           (const uint32_t*)m.idx.p, (const L*)m.val.p, X, (const uint32_t*)m.heavy.p, part.p);
  e.launch("k_spmv_heavy_sum", 0, 0, (k_spmv_heavy_sum<T>), dim3((unsigned)m.nheavy), dim3(E), (const L*)part.p, nch,
           (const uint32_t*)m.heavy.p, out);
}

inline void check_coo(const HostCoo& m, size_t rows, size_t cols) {
  if (m.rows != rows || m.cols != cols) throw Error(NENYA_E_SHAPE, "matrix shape mismatch");
  for (size_t k = 0; k < m.nnz; k++)
    if (m.r[k] >= rows || m.c[k] >= cols) throw Error(NENYA_E_SHAPE, "matrix entry out of range");
}

}  // namespace nenya

#include "setup.hpp"  // device matrix packing (uses HostCoo and Csr above)

namespace nenya {

// ============================================================================
// SIMD (simd.rs:220-373)
// ============================================================================
template <class T> struct SimdParams : ParamsBase {
  using Width = T;
  using L = typename T::L;
  using G = Gr<T>;
  int log_n = 0, log_m = 0;
  Ctx<T> cx0;
  PcsDev<T> pw, pq;
  CtxBase& ctx0() override { return cx0; }

  SimdParams(int ln, int lm, uint64_t seed_w, uint64_t seed_q) {
    k = T::K;
    kind = 0;
    shape[0] = ln;
    shape[1] = lm;
    seeds[0] = seed_w;
    seeds[1] = seed_q;
    log_n = ln;
    log_m = lm;
    setup_mark("context", cx0.st);  // the params' stream, pool and constants
    pw.P = PcsParams::setup(lm, seed_w, T::K);
    pq.P = PcsParams::setup_r2(ln, seed_q, T::K);
    setup_mark("pcs_params", cx0.st);
    pw.upload(*cx0.eng);
    pq.upload(*cx0.eng);
    setup_mark("pcs_upload", cx0.st);
    // setup data is read by other streams (server workers): finish it here
    HIP_CHECK(hipStreamSynchronize(cx0.st));
    footprint = (size_t)((32.0 * 2 + 2 * E) * sizeof(L) * (1ull << lm) + (10.0 * E + 96) * sizeof(L) * (1ull << ln));
  }
  ~SimdParams() override {
    pw = PcsDev<T>();
    pq = PcsDev<T>();
  }

  ProofView prove(Ctx<T>& cx, const HostCoo& A, const HostCoo& B, const HostCoo& Cm, const L* wit /*32 x m*/) {
    Engine<T>& e = *cx.eng;
    e.finish_downloads();  // a failed earlier proof may have left copies in flight
    Pool& pool = cx.pool;
    const size_t pool_base = pool.mark();
    hipStream_t st = cx.st;
    KProf& prof = cx.prof;
    e.prof = &prof;
    size_t n = size_t(1) << log_n, m = size_t(1) << log_m;
    check_coo(A, n, m);
    check_coo(B, n, m);
    check_coo(Cm, n, m);
    Stages stg;
    stg.on = cx.timing;
    stg.st = st;
    stg.start();
    // the six CSRs (by row and by column) grouped on the device from one
    // upload of each matrix's entries
    Csr<T> ar, br, cr, ac, bc, cc;
    {
      DevSetup<T> ds(e);
      DevCoo<T> coo(e, {&A, &B, &Cm});
      const HostCoo* hm[3] = {&A, &B, &Cm};
      Csr<T>* byrow[3] = {&ar, &br, &cr};
      Csr<T>* bycol[3] = {&ac, &bc, &cc};
      for (int w = 0; w < 3; w++) {
        const uint32_t* r = coo.r.p + coo.off[w];
        const uint32_t* c = coo.c.p + coo.off[w];
        const L* v = coo.v.p + coo.off[w];
        ds.csr(*byrow[w], r, c, v, hm[w]->nnz, n);
        ds.csr(*bycol[w], c, r, v, hm[w]->nnz, m);
      }
    }
    // witnesses [32][m] and their slot transpose [m][32] (made on the device)
    DBuf<L> dw(&pool, 32 * m), slots(&pool, m * 32);
    e.h2d(dw.p, wit, 32 * m * sizeof(L));
    e.launch("k_slots", 64.0 * m * sizeof(L), 0, (k_slots<T>), dim3(grid_for(m, 64, 1u << 16)), dim3(256),
             (const L*)dw.p, (uint64_t)m, slots.p);
    stg.mark("upload");
    // 1. simd_to_packed (reduce.rs:33-104): w_phi = phi(slots); x_i = (A w_k)_i, y_i = (B w_k)_i
    //    and e_i = phi(x_i) phi(y_i) - phi(x_i . y_i)  (= q1 v_alpha + q2 v_beta exactly,
    //    embedding.rs:147-167).
    DBuf<L> wphi(&pool, m * E);
    e.pack(slots.p, wphi.p, m, 1);
    DBuf<L> xs(&pool, n * 32), ys(&pool, n * 32), xy(&pool, n * 32);
    unsigned g0 = grid_for(n * 32, 256, 1u << 20);
    e.launch("k_r0_matvec", 0, 0, (k_r0_matvec<T>), dim3(g0), dim3(256), (const uint32_t*)ar.ptr.p,
             (const uint32_t*)ar.idx.p, (const L*)ar.val.p, (const L*)dw.p, (uint64_t)m, xs.p, (uint64_t)n, 32, (uint64_t)32,
             (uint64_t)1);
    e.launch("k_r0_matvec", 0, 0, (k_r0_matvec<T>), dim3(g0), dim3(256), (const uint32_t*)br.ptr.p,
             (const uint32_t*)br.idx.p, (const L*)br.val.p, (const L*)dw.p, (uint64_t)m, ys.p, (uint64_t)n, 32, (uint64_t)32,
             (uint64_t)1);
    e.launch("k_hadamard", 0, 0, (k_hadamard<T>), dim3(g0), dim3(256), (const L*)xs.p, (const L*)ys.p, xy.p,
             (uint64_t)n * 32);
    DBuf<L> px(&pool, n * E), py(&pool, n * E), ev(&pool, n * E);
    e.pack(xs.p, px.p, n, 1);
    e.pack(ys.p, py.p, n, 1);
    e.pack(xy.p, ev.p, n, 1);
    e.template mul_sub<16>(px.p, py.p, ev.p, ev.p, n);
    stg.mark("reduce");
    // 2. commits
    // the commitments are independent: one stream each
    e.fork(2);
    e.on(1);
    CommitDev<T> ce = e.commit(pq, ev.p, pq.P.k1, ev.p);
    e.on(0);
    CommitDev<T> cw = e.commit(pw, slots.p, pw.P.k1, wphi.p);
    e.join();
    e.resolve_roots({&cw, &ce});
    stg.mark("commit");
    // 3. transcript
    Transcript tr("simd");
    tr.root(cw.root);
    tr.root(ce.root);
    StdRng rs = tr.sampler("mult_check_target");
    std::vector<G> rt = e.sample(rs, log_n);
    // proof layout (simd.rs:150-174): roots, claims and sumchecks, then the
    // openings of w at s2 and of e at r
    const size_t EBs = E * sizeof(L);
    size_t s_mult = EBs + 8 + 3 * (size_t)log_n * EBs + 2 * EBs, s_lin = EBs + 8 + 3 * (size_t)log_m * EBs + 2 * EBs;
    size_t s_ow = opening_bytes(pw.P, pw.P.k1, sizeof(L), 0), s_oe = opening_bytes(pq.P, pq.P.k1, sizeof(L), 0);
    size_t total = 64 + s_mult + s_lin + s_ow + s_oe;
    uint8_t* buf = cx.proofbuf.get(total);
    Sink k_oe(buf + total - s_oe, s_oe);
    // 4. a_phi, b_phi, c_phi (matrix.rs:93-105)
    DBuf<L> aw(&pool, n * E), bw(&pool, n * E), cwv(&pool, n * E);
    spmv_gr1(e, ar, (const L*)wphi.p, aw.p, (uint64_t)n, 0);
    spmv_gr1(e, br, (const L*)wphi.p, bw.p, (uint64_t)n, 0);
    spmv_gr1(e, cr, (const L*)wphi.p, cwv.p, (uint64_t)n, 0);
    // 5. claim_mult = c_phi(r) + e(r)
    DBuf<L> eq_r = e.eq_table(rt.data(), log_n);
    // the opening of e at r needs only r: it runs on the background stream,
    // beside the rest of the proof (enqueued once the matvecs are)
    e.bg_begin();
    DBuf<L> ye_d = e.open_dev(pq, ce, rt.data(), nullptr, 1, k_oe);
    e.bg_end();
    G cw_at_r = e.template dot<32>(eq_r.p, cwv.p, n);
    G e_at_r = e.template dot<32>(eq_r.p, ev.p, n);
    G claim_mult = cw_at_r + e_at_r;
    tr.elems("claim_mult", &claim_mult, 1);
    stg.mark("matvec");
    // 6. mult check (Gruen)
    tr.label("mult_check_sumcheck");
    const L* mt[2] = {aw.p, bw.p};
    auto mp = e.mult_check(mt, 2, log_n, rt, tr);
    G fin_ab[2] = {mp.finals[0], mp.finals[1]};
    tr.elems("mult/finals", fin_ab, 2);
    stg.mark("mult_sumcheck");
    // 7. gamma
    G gam = e.challenge(tr, "linear_check_gamma");
    G gam2 = gam * gam;
    // 8. M_gamma row (simd.rs:539-580) = A_row(s1) + gamma B_row(s1) + gamma^2 C_row(r),
    //    accumulated as sum over columns of (eq_s1 | gamma eq_s1 | gamma^2 eq_r)[row] * v.
    DBuf<L> eq_s1 = e.eq_table(mp.chal.data(), log_n);
    DBuf<L> geq(&pool, n * E), g2eq(&pool, n * E), mrow(&pool, m * E);
    DBuf<L> dg = e.put(&gam, 1), dg2 = e.put(&gam2, 1);
    e.scale(dg.p, eq_s1.p, geq.p, n);
    e.scale(dg2.p, eq_r.p, g2eq.p, n);
    spmv_gr1(e, ac, (const L*)eq_s1.p, mrow.p, (uint64_t)m, 0);
    spmv_gr1(e, bc, (const L*)geq.p, mrow.p, (uint64_t)m, 1);
    spmv_gr1(e, cc, (const L*)g2eq.p, mrow.p, (uint64_t)m, 1);
    // 9. claim_lin
    G claim_lin = mp.finals[0] + gam * mp.finals[1] + gam2 * cw_at_r;
    tr.elems("claim_lin", &claim_lin, 1);
    stg.mark("linear_rows");
    // 10. linear check on (w_phi, M_gamma row); w_phi is low-degree
    tr.label("linear_check_sumcheck");
    auto lp = e.template linear<16>(wphi.p, 0, mrow.p, 0, 1, log_m, claim_lin, tr);
    G final_w = lp.finals[0], final_row = lp.finals[1];
    G fin_lin[2] = {final_row, final_w};
    tr.elems("lin/finals", fin_lin, 2);
    stg.mark("linear_sumcheck");
    // 11. serialize (simd.rs:150-174) with the opening of w
    Sink out(buf, total - s_oe);
    out.bytes(cw.root, 32);
    out.bytes(ce.root, 32);
    out.gr(claim_mult);
    out.u64(mp.chal.size());
    for (auto& g : mp.rounds) out.gr(g);
    out.gr(mp.finals[0]);
    out.gr(mp.finals[1]);
    out.gr(claim_lin);
    out.u64(lp.chal.size());
    for (auto& g : lp.rounds) out.gr(g);
    out.gr(final_row);
    out.gr(final_w);
    G yw = e.open(pw, cw, lp.chal.data(), nullptr, 1, out);
    e.bg_join();
    G ye = e.get(ye_d.p);
    stg.mark("open");
    if ((!(yw == final_w) || !(ye == e_at_r)) && !skip_selfcheck())
      throw Error(NENYA_E_DEVICE, "internal opening self-check failed");
    if (!out.full() || !k_oe.full()) throw Error(NENYA_E_DEVICE, "proof size mismatch");
    e.finish_downloads();
    cx.last_timing = stg.collect();
    if (prof.on) prof.collect();
    note_footprint(pool.peak - pool_base);
    return {buf, total};
  }
};

// ============================================================================
// MIMD (mimd.rs:242-397)
// ============================================================================
// One MIMD matrix set: the 32 instance matrices concatenated, instance s at
// entries off[s] .. off[s+1]-1 (values: limbs of the native width).
struct MimdSet {
  const size_t* off;
  const uint32_t* r;
  const uint32_t* c;
  const void* v;
};

template <class T> struct MimdParams : ParamsBase {
  using L = typename T::L;
  using G = Gr<T>;
  int log_n = 0, log_m = 0;
  size_t n = 0, m = 0;
  Ctx<T> cx0;
  PcsDev<T> pw, pabc, pe;
  // union-support cells of A, B, C (pack_matrix_set, reduce.rs:119-141),
  // concatenated: the 32 instance values per cell
  size_t ncell[3] = {0, 0, 0};
  DBuf<uint64_t> cptr;  // 3 x (n+1) row pointers (global cell indices)
  DBuf<uint32_t> ccol;
  DBuf<L> cval;
  // M_gamma-row entries by column, one per nonzero instance value,
  // code = y << 7 | w << 5 | s (k_rows_gather)
  DBuf<uint64_t> eptr;
  DBuf<uint32_t> ecode;
  DBuf<L> eval;
  CtxBase& ctx0() override { return cx0; }

  // MimdSnarkParams::secure (mimd.rs:84-107) with pack_mimd_matrices
  // (reduce.rs:249-273) done once here (setup, not prover time).
  MimdParams(int ln, int lm, const MimdSet* sets, uint64_t seed_w, uint64_t seed_e) {
    k = T::K;
    kind = 2;
    shape[0] = ln;
    shape[1] = lm;
    seeds[0] = seed_w;
    seeds[1] = seed_e;
    log_n = ln;
    log_m = lm;
    n = size_t(1) << ln;
    m = size_t(1) << lm;
    Pool& pool = cx0.pool;
    Engine<T>* eng = cx0.eng.get();
    setup_mark("context", cx0.st);  // the params' stream, pool and constants
    pw.P = PcsParams::setup(lm, seed_w, T::K);
    pabc.P = PcsParams::setup_stacked(ln, seed_w ^ 0xABC, 3, T::K);
    pe.P = PcsParams::setup_r2(ln, seed_e, T::K);
    setup_mark("pcs_params", cx0.st);
    pw.upload(*eng);
    pabc.upload(*eng);
    pe.upload(*eng);
    setup_mark("pcs_upload", cx0.st);
    // pack_mimd_matrices on the device: per matrix the entries of the 32
    // instances sorted (stably) by (row, col); runs of one key are the
    // union-support cells, each with its 32 instance values summed
    std::vector<HostCoo> hs(3);
    std::vector<uint64_t> hoff(3 * 33);
    for (int w = 0; w < 3; w++) {
      const MimdSet& S = sets[w];
      const size_t tot = S.off[32];
      for (int inst = 0; inst < 32; inst++) {
        if (S.off[inst + 1] < S.off[inst] || S.off[inst + 1] > tot || S.off[0] != 0)
          throw Error(NENYA_E_ARG, "bad instance offsets");
      }
      for (size_t j = 0; j < tot; j++)
        if (S.r[j] >= n || S.c[j] >= m) throw Error(NENYA_E_SHAPE, "matrix entry out of range");
      hs[w] = HostCoo{n, m, tot, S.r, S.c, S.v};
      for (int i = 0; i <= 32; i++) hoff[w * 33 + i] = S.off[i];
    }
    DevSetup<T> ds(*eng);
    DevCoo<T> coo(*eng, {&hs[0], &hs[1], &hs[2]});
    const uint64_t ne = coo.off[3];
    if (ne >= (uint64_t(1) << 32)) throw Error(NENYA_E_SHAPE, "too many matrix entries");
    DBuf<uint8_t> inst(&pool, std::max<uint64_t>(ne, 1));
    {
      DBuf<uint64_t> doff(&pool, hoff.size());
      eng->h2d(doff.p, hoff.data(), hoff.size() * 8);
      for (int w = 0; w < 3; w++)
        eng->launch("k_inst_of", 0, 0, k_inst_of, dim3(ds.grid(hs[w].nnz)), dim3(256), (const uint64_t*)(doff.p + 33 * w),
                    (uint64_t)hs[w].nnz, inst.p + coo.off[w]);
      eng->sync();
    }
    DBuf<uint64_t> ckey(&pool, std::max<uint64_t>(ne, 1));
    DBuf<L> slots(&pool, std::max<uint64_t>(ne, 1) * 32);
    HIP_CHECK(hipMemsetAsync(slots.p, 0, std::max<uint64_t>(ne, 1) * 32 * sizeof(L), cx0.st));
    uint64_t cbase[4] = {0, 0, 0, 0};
    for (int w = 0; w < 3; w++) {
      const uint64_t nw = hs[w].nnz;
      cbase[w + 1] = cbase[w];
      if (!nw) continue;
      DBuf<uint64_t> key(&pool, nw), skey(&pool, nw);
      DBuf<uint32_t> perm(&pool, nw), head(&pool, nw);
      eng->launch("k_rc_keys", 0, 0, k_rc_keys, dim3(ds.grid(nw)), dim3(256), (const uint32_t*)(coo.r.p + coo.off[w]),
                  (const uint32_t*)(coo.c.p + coo.off[w]), nw, (uint64_t)m, key.p);
      {
        DBuf<uint32_t> id = ds.iota(nw, (uint32_t)coo.off[w]);
        ds.template sort_pairs<uint64_t>(key.p, skey.p, id.p, perm.p, nw, bits_for((uint64_t)n * m - 1));
      }
      eng->launch("k_run_heads", 0, 0, (k_run_heads<uint64_t>), dim3(ds.grid(nw)), dim3(256), (const uint64_t*)skey.p, nw,
                  head.p);
      uint32_t* id = reinterpret_cast<uint32_t*>(key.p);  // run numbers (1-based), over the unsorted keys
      ds.scan(head.p, id, nw, true);
      const uint64_t nc = ds.last_of(id, nw);
      eng->launch("k_cell_sum", 0, 0, (k_cell_sum<L>), dim3(ds.grid(nw)), dim3(256), (const uint64_t*)skey.p,
                  (const uint32_t*)perm.p, id, nw, (const uint8_t*)inst.p, (const L*)coo.v.p, ckey.p + cbase[w],
                  slots.p + cbase[w] * 32);
      ncell[w] = nc;
      cbase[w + 1] = cbase[w] + nc;
    }
    const size_t total = cbase[3];
    if (total >= (size_t(1) << 27)) throw Error(NENYA_E_SHAPE, "too many matrix cells");
    ccol = DBuf<uint32_t>(&pool, std::max<size_t>(total, 1));
    cval = DBuf<L>(&pool, std::max<size_t>(total, 1) * 32);
    DBuf<uint32_t> crow(&pool, std::max<size_t>(total, 1));
    if (total) {
      eng->d2d(cval.p, slots.p, total * 32 * sizeof(L));
      eng->launch("k_cell_rc", 0, 0, k_cell_rc, dim3(ds.grid(total)), dim3(256), (const uint64_t*)ckey.p, (uint64_t)total,
                  (uint64_t)m, crow.p, ccol.p);
    }
    slots.release();
    // row pointers of each matrix's cells (global cell indices)
    cptr = DBuf<uint64_t>(&pool, 3 * (n + 1));
    for (int w = 0; w < 3; w++)
      eng->launch("k_group_ptr", 0, 0, (k_group_ptr<uint32_t, uint64_t>), dim3(ds.grid(n + 1)), dim3(256),
                  (const uint32_t*)(crow.p + cbase[w]), (uint64_t)ncell[w], (uint64_t)n, (uint64_t)cbase[w],
                  cptr.p + w * (n + 1));
    {
      if (n > (size_t(1) << 25)) throw Error(NENYA_E_SHAPE, "too many rows for the row-entry code");
      // M_gamma-row entries: the nonzero instance values by column, in
      // (cell, slot) order within a column
      const uint64_t ns = (uint64_t)total * 32;
      DBuf<uint32_t> flag(&pool, std::max<uint64_t>(ns, 1)), pos(&pool, std::max<uint64_t>(ns, 1));
      if (ns) {
        eng->launch("k_nz_flags", 0, 0, (k_nz_flags<L>), dim3(ds.grid(ns)), dim3(256), (const L*)cval.p, ns, flag.p);
        ds.scan(flag.p, pos.p, ns, false);
      }
      const uint64_t nz = ns ? (uint64_t)ds.last_of(pos.p, ns) + ds.last_of(flag.p, ns) : 0;
      eptr = DBuf<uint64_t>(&pool, m + 1);
      ecode = DBuf<uint32_t>(&pool, std::max<uint64_t>(nz, 1));
      eval = DBuf<L>(&pool, std::max<uint64_t>(nz, 1));
      DBuf<uint32_t> key(&pool, std::max<uint64_t>(nz, 1)), at(&pool, std::max<uint64_t>(nz, 1));
      DBuf<uint32_t> skey(&pool, std::max<uint64_t>(nz, 1)), perm(&pool, std::max<uint64_t>(nz, 1));
      if (nz) {
        eng->launch("k_nz_place", 0, 0, k_nz_place, dim3(ds.grid(ns)), dim3(256), (const uint32_t*)flag.p,
                    (const uint32_t*)pos.p, (const uint32_t*)ccol.p, ns, key.p, at.p);
        DBuf<uint32_t> id = ds.iota(nz);
        ds.template sort_pairs<uint32_t>(key.p, skey.p, id.p, perm.p, nz, bits_for(m - 1));
        eng->launch("k_nz_codes", 0, 0, (k_nz_codes<L>), dim3(ds.grid(nz)), dim3(256), (const uint32_t*)perm.p,
                    (const uint32_t*)at.p, (const uint32_t*)crow.p, (const L*)cval.p, nz, (uint64_t)cbase[1],
                    (uint64_t)cbase[2], ecode.p, eval.p);
      }
      eng->launch("k_group_ptr", 0, 0, (k_group_ptr<uint32_t, uint64_t>), dim3(ds.grid(m + 1)), dim3(256),
                  (const uint32_t*)skey.p, nz, (uint64_t)m, (uint64_t)0, eptr.p);
      eng->sync();
      setup_mark("matrices", cx0.st);
      footprint = (size_t)((24.0 * E + 3 * 32) * sizeof(L) * n + (8.0 * E + 64) * sizeof(L) * m +
                           (48.0 * E + 1024) * sizeof(L) * n);
    }
  }
  ~MimdParams() override {
    pw = PcsDev<T>();
    pabc = PcsDev<T>();
    pe = PcsDev<T>();
    cptr.release(); ccol.release(); cval.release();
    eptr.release(); ecode.release(); eval.release();
  }

  // sum_c Xhat[r, c] w_phi[c] for the three matrices into out (3 x n)
  void matvecs(Engine<T>& e, const L* slots, const L* wphi, L* out) {
    if constexpr (kMx<T>) {
      // the three matrices are independent: one stream each, with a C-matrix
      // slice per stream (one slice when the work stays on one stream)
      e.fork(3);
      const size_t ns = (size_t)e.forked + 1;
      DBuf<L> cm(e.pool, ns * n * 1024);
      for (int w = 0; w < 3; w++) {
        e.on(w);
        L* cmw = cm.p + (size_t)(w % ns) * n * 1024;
        e.launch("k_cmat_mimd", (double)n * 1024 * sizeof(L), (double)ncell[w] * 1024, (k_cmat_mimd<T>),
                 dim3(grid_for(n, 8, 4096)), dim3(256), (const uint64_t*)(cptr.p + w * (n + 1)), (const uint32_t*)ccol.p,
                 (const L*)cval.p, slots, cmw, (uint64_t)n);
        e.cgemm(cmw, out + w * n * E, n);
      }
      e.join();
    } else {
      // C_r from the cells' instance values and the witness slots, then the
      // basis-product GEMM with 64-bit limbs (as the layered matvec)
      (void)wphi;
      DBuf<L> cm(e.pool, n * 1024);
      for (int w = 0; w < 3; w++) {
        e.launch("k_cmat_mimd", (double)n * 1024 * sizeof(L), (double)ncell[w] * 1024, (k_cmat_mimd<T>),
                 dim3(grid_for(n, 8, 4096)), dim3(256), (const uint64_t*)(cptr.p + w * (n + 1)), (const uint32_t*)ccol.p,
                 (const L*)cval.p, slots, cm.p, (uint64_t)n);
        e.cgemm64(cm.p, out + w * n * E, (uint64_t)n);
      }
    }
  }
  // build_combined_row (mimd.rs:196-219): M[col] = sum over the three
  // matrices of eqs[w][y] Xhat[y, col], eqs = {eq, gamma eq, gamma^2 eq} (3 x n)
  void m_row(Engine<T>& e, const L* eqs, L* mrow) {
    // G[w][y][i] = eqs[w][y] l_beta_i, then entries v * G[w][y][s/2] u_{s%2}
    DBuf<L> G(e.pool, 3 * n * 16 * E);
    e.mulfix_lb(eqs, 3 * n, G.p);
    e.launch("k_rows_gather", 0, 0, (k_rows_gather<T>), dim3(grid_for(m, 8, 1u << 16)), dim3(256),
             (const uint64_t*)eptr.p, (const uint32_t*)ecode.p, (const L*)eval.p, (const L*)G.p, (uint64_t)n, mrow,
             (uint64_t)m);
  }

  ProofView prove(Ctx<T>& cx, const L* wit /* 32 x m */) {
    Engine<T>& e = *cx.eng;
    e.finish_downloads();  // a failed earlier proof may have left copies in flight
    Pool& pool = cx.pool;
    const size_t pool_base = pool.mark();
    KProf& prof = cx.prof;
    e.prof = &prof;
    Stages stg;
    stg.on = cx.timing;
    stg.st = cx.st;
    stg.start();
    // witnesses [32][m] and their slot transpose [m][32] (the slot-form
    // commitment message), made on the device
    DBuf<L> dw(&pool, 32 * m), slots(&pool, m * 32);
    e.h2d(dw.p, wit, 32 * m * sizeof(L));
    e.launch("k_slots", 64.0 * m * sizeof(L), 0, (k_slots<T>), dim3(grid_for(m, 64, 1u << 16)), dim3(256),
             (const L*)dw.p, (uint64_t)m, slots.p);
    stg.mark("upload");
    // 1. w_phi = pack_r2(witness column j) (reduce.rs:294-312)
    DBuf<L> wphi(&pool, m * E);
    e.pack(slots.p, wphi.p, m, 1);
    // 2. a_hat, b_hat, c_hat.  With Ahat w_phi = phi(A_s w_s) + q1 v_alpha +
    //    q2 v_beta (the quotient identity summed over a row, reduce.rs:172-192),
    //    a_hat = Ahat w_phi - e_A = pack_r2((A_s w_s)_s) exactly, and for C,
    //    pack_r2(psi(Chat w_phi)) = pack_r2((C_s w_s)_s) (mimd.rs:276-287).
    DBuf<L> abc_slots(&pool, 3 * n * 32);
    for (int w = 0; w < 3; w++)
      e.launch("k_mimd_slots", 0, 0, (k_mimd_slots<T>), dim3(grid_for(n * 32, 256, 1u << 16)), dim3(256),
               (const uint64_t*)(cptr.p + w * (n + 1)), (const uint32_t*)ccol.p, (const L*)cval.p,
               (const L*)slots.p, abc_slots.p + w * n * 32, (uint64_t)n);
    DBuf<L> hats(&pool, 4 * n * E);  // a_hat, b_hat, c_hat, e_mult
    e.pack(abc_slots.p, hats.p, 3 * n, 1);
    L* emult = hats.p + 3 * n * E;
    e.template mul_sub<16>(hats.p, hats.p + n * E, hats.p + 2 * n * E, emult, n);
    // 3. Ahat w_phi, Bhat w_phi, Chat w_phi (matvec_gr128_dense_csr) and the
    //    errors e_X = Xhat w_phi - x_hat
    DBuf<L> estack(&pool, 3 * n * E);
    matvecs(e, slots.p, wphi.p, estack.p);
    e.sub(estack.p, hats.p, estack.p, 3 * n * E);
    stg.mark("reduce_matvec");
    // 4. commits (mimd.rs:302-316)
    // the four commitments are independent: one stream each
    e.fork(4);
    e.on(1);
    CommitDev<T> ces = e.commit(pe, estack.p, 3 * pe.P.k1, estack.p);
    e.on(2);
    CommitDev<T> cabc = e.commit(pabc, abc_slots.p, 3 * pabc.P.k1, hats.p);
    e.on(3);
    CommitDev<T> cem = e.commit(pe, emult, pe.P.k1, emult);
    e.on(0);
    CommitDev<T> cw = e.commit(pw, slots.p, pw.P.k1, wphi.p);
    e.join();
    e.resolve_roots({&cw, &cabc, &cem, &ces});
    stg.mark("commit");
    Transcript tr("mimd");
    tr.root(cw.root);
    tr.root(cabc.root);
    tr.root(cem.root);
    tr.root(ces.root);
    StdRng rs = tr.sampler("mult_check_target");
    std::vector<G> rt = e.sample(rs, log_n);
    // 5. mult check, sub form, on (a_hat, b_hat, c_hat, e_mult)
    tr.label("mult_check_sumcheck");
    const L* mt[4] = {hats.p, hats.p + n * E, hats.p + 2 * n * E, emult};
    auto mp = e.mult_check(mt, 4, log_n, rt, tr);
    tr.elems("mult/finals", mp.finals.data(), 4);
    G gam = e.challenge(tr, "linear_check_gamma");
    G gam2 = gam * gam;
    stg.mark("mult_sumcheck");
    // proof layout (mimd.rs:128-152): roots, mult, lin, open_w, open_emult,
    // open_abc, open_estack -- sections sized up front
    const size_t EBs = E * sizeof(L);
    size_t s_mult = 8 + 3 * (size_t)log_n * EBs + 4 * EBs, s_lin = EBs + 8 + 3 * (size_t)log_m * EBs + 2 * EBs;
    size_t s_w = opening_bytes(pw.P, pw.P.k1, sizeof(L), 0), s_em = opening_bytes(pe.P, pe.P.k1, sizeof(L), 0);
    size_t s_abc = opening_bytes(pabc.P, 3 * pabc.P.k1, sizeof(L), 0), s_es = opening_bytes(pe.P, 3 * pe.P.k1, sizeof(L), 0);
    size_t total = 128 + s_mult + s_lin + s_w + s_em + s_abc + s_es;
    uint8_t* buf = cx.proofbuf.get(total);
    uint8_t* o_w = buf + 128 + s_mult + s_lin;
    uint8_t* o_em = o_w + s_w;
    uint8_t* o_abc = o_em + s_em;
    uint8_t* o_es = o_abc + s_abc;
    // 6. openings at s1 (mimd.rs:340-352)
    G wts[3] = {G::one(), gam, gam2};
    Sink k_es(o_es, s_es), k_abc(o_abc, s_abc), k_em(o_em, s_em);
    // estack's value enters the linear claim: it is formed first on the
    // prover stream (the rest of its opening in the background); the other
    // two openings feed only the final self-check and run on the background
    // stream, enqueued after the M_gamma row (critical path first)
    DBuf<L> es_y = e.open_dev(pe, ces, mp.chal.data(), wts, 3, k_es, true);
    stg.mark("open_s1");
    // 7. M_gamma row (mimd.rs:196-219, eq_s1 for all three matrices)
    DBuf<L> eqs(&pool, 3 * n * E), mrow(&pool, m * E);
    {
      DBuf<L> eq1 = e.eq_table(mp.chal.data(), log_n);
      e.d2d(eqs.p, eq1.p, n * E * sizeof(L));
      DBuf<L> dg = e.put(&gam, 1), dg2 = e.put(&gam2, 1);
      e.scale(dg.p, eq1.p, eqs.p + n * E, n);
      e.scale(dg2.p, eq1.p, eqs.p + 2 * n * E, n);
    }
    m_row(e, eqs.p, mrow.p);
    eqs.release();
    stg.mark("linear_rows");
    e.bg_begin();
    DBuf<L> abc_y = e.open_dev(pabc, cabc, mp.chal.data(), wts, 3, k_abc);
    DBuf<L> em_y = e.open_dev(pe, cem, mp.chal.data(), nullptr, 1, k_em);
    e.bg_end();
    const G estack_val = e.get(es_y.p);
    // 8. linear check on (M_gamma row, w_phi); w_phi is low-degree
    G claim_lin = mp.finals[0] + gam * mp.finals[1] + gam2 * mp.finals[2] + estack_val;
    tr.label("linear_check_sumcheck");
    auto lp = e.template linear<16>(wphi.p, 0, mrow.p, 0, 1, log_m, claim_lin, tr);
    G final_w = lp.finals[0], final_row = lp.finals[1];
    stg.mark("linear_sumcheck");
    // 9. open w at s2
    Sink k_w(o_w, s_w);
    DBuf<L> yw_d = e.open_dev(pw, cw, lp.chal.data(), nullptr, 1, k_w);
    stg.mark("open_w");
    // the opening values of the self-check, one synchronization
    e.bg_join();
    std::vector<G> late = e.pick_many({{abc_y.p, 1, 1}, {em_y.p, 1, 1}, {yw_d.p, 1, 1}});
    const G abc_val = late[0], em_val = late[1], yw = late[2];
    Sink hd(buf, 128 + s_mult + s_lin);
    for (auto* r : {cw.root, cabc.root, cem.root, ces.root}) hd.bytes(r, 32);
    hd.u64(mp.chal.size());
    for (auto& g : mp.rounds) hd.gr(g);
    for (int i = 0; i < 4; i++) hd.gr(mp.finals[i]);
    hd.gr(claim_lin);
    hd.u64(lp.chal.size());
    for (auto& g : lp.rounds) hd.gr(g);
    hd.gr(final_row);
    hd.gr(final_w);
    if (!hd.full() || !k_es.full() || !k_abc.full() || !k_em.full() || !k_w.full())
      throw Error(NENYA_E_DEVICE, "proof size mismatch");
    const bool ok = em_val == mp.finals[3] && abc_val == mp.finals[0] + gam * mp.finals[1] + gam2 * mp.finals[2] &&
                    yw == final_w;
    if (!ok && !skip_selfcheck()) throw Error(NENYA_E_DEVICE, "internal opening self-check failed");
    e.finish_downloads();
    cx.last_timing = stg.collect();
    if (prof.on) prof.collect();
    note_footprint(pool.peak - pool_base);
    return {buf, total};
  }
};

// ============================================================================
// General (general.rs:374-537)
// ============================================================================
template <class T> struct GeneralParams : ParamsBase {
  using L = typename T::L;
  using G = Gr<T>;
  int log_b = 0;
  size_t nb = 0, N = 0;
  Ctx<T> cx0;
  PcsDev<T> pw, pabc, pe;
  Csr<T> mr[3];  // A, B, C by row (base-ring matvecs)
  DBuf<uint64_t> eptr;  // layer-row entries by output
  DBuf<uint32_t> ecode;
  DBuf<L> eval;

  // GeneralSnarkParams::secure (general.rs:69-95) with pack_general_matrices
  // (reduce.rs:372-438, 618-651) done once here (setup, not prover time).
  CtxBase& ctx0() override { return cx0; }
  GeneralParams(int lb, const HostCoo* m3, uint64_t seed_w, uint64_t seed_e) {
    k = T::K;
    kind = 1;
    shape[0] = lb;
    seeds[0] = seed_w;
    seeds[1] = seed_e;
    log_b = lb;
    nb = size_t(1) << lb;
    N = 32 * nb;
    Pool& pool = cx0.pool;
    hipStream_t st = cx0.st;
    Engine<T>* eng = cx0.eng.get();
    footprint = (size_t)(170.0 * E * sizeof(L) * nb) + (size_t)nb * 1024 * sizeof(L);
    setup_mark("context", st);  // the params' stream, pool and constants
    pw.P = PcsParams::setup_batched(lb, seed_w, 32, T::K);
    pabc.P = PcsParams::setup_stacked(lb, seed_w ^ 0xABC, 3, T::K);
    pe.P = PcsParams::setup_r2(lb, seed_e, T::K);
    setup_mark("pcs_params", st);
    pw.upload(*eng);
    pabc.upload(*eng);
    pe.upload(*eng);
    setup_mark("pcs_upload", st);
    for (int w = 0; w < 3; w++) check_coo(m3[w], N, N);
    // pack_general_matrices on the device (stable sorts of the uploaded
    // entries): the row CSRs, then the layer-row entries (the diagonal cells
    // of reduce.rs:387-397 are never formed: the layer rows gather the
    // entries against the basis products instead)
    DevSetup<T> ds(*eng);
    DevCoo<T> coo(*eng, {&m3[0], &m3[1], &m3[2]});
    const uint64_t ne = coo.off[3];
    if (ne >= (uint64_t(1) << 32)) throw Error(NENYA_E_SHAPE, "too many matrix entries");
    for (int w = 0; w < 3; w++)
      ds.csr(mr[w], coo.r.p + coo.off[w], coo.c.p + coo.off[w], coo.v.p + coo.off[w], m3[w].nnz, N);
    {
      // layer-row entries (row 32p+s of matrix w, value v) by output
      // (l = (c - r) mod 32, q = c / 32), code p << 7 | w << 5 | s
      if (nb > (size_t(1) << 25)) throw Error(NENYA_E_SHAPE, "too many blocks for the layer-row entry code");
      const uint64_t nn = std::max<uint64_t>(ne, 1);
      eptr = DBuf<uint64_t>(&pool, 32 * nb + 1);
      ecode = DBuf<uint32_t>(&pool, nn);
      eval = DBuf<L>(&pool, nn);
      DBuf<uint32_t> key(&pool, nn), skey(&pool, nn), perm(&pool, nn);
      if (ne) {
        eng->launch("k_layer_keys", 0, 0, k_layer_keys, dim3(ds.grid(ne)), dim3(256), (const uint32_t*)coo.r.p,
                    (const uint32_t*)coo.c.p, ne, (uint64_t)nb, key.p);
        DBuf<uint32_t> id = ds.iota(ne);
        ds.template sort_pairs<uint32_t>(key.p, skey.p, id.p, perm.p, ne, bits_for(32 * (uint64_t)nb - 1));
        eng->launch("k_layer_place", 0, 0, (k_layer_place<L>), dim3(ds.grid(ne)), dim3(256), (const uint32_t*)perm.p,
                    (const uint32_t*)coo.r.p, (const L*)coo.v.p, ne, coo.off[1], coo.off[2], ecode.p, eval.p);
      }
      eng->launch("k_group_ptr", 0, 0, (k_group_ptr<uint32_t, uint64_t>), dim3(ds.grid(32 * nb + 1)), dim3(256),
                  (const uint32_t*)skey.p, ne, (uint64_t)(32 * nb), (uint64_t)0, eptr.p);
      eng->sync();
    }
    setup_mark("matrices", st);
    // the proof buffer and the pinned staging areas exist before the first
    // proof (a no-alloc call may be the first)
    cx0.proofbuf.get(proof_size());
    eng->reserve_staging();
    eng->sync();
    setup_mark("staging", st);
  }
  ~GeneralParams() override {
    pw = PcsDev<T>();
    pabc = PcsDev<T>();
    pe = PcsDev<T>();
    for (auto& m : mr) m = Csr<T>();
    eptr.release(); ecode.release(); eval.release();
  }

  // build_layer_rows (general.rs:233-271): rows[l][q] (32 x nb) from
  // eqs = {eq_s1, gamma eq_s1, gamma^2 eq_s1} (3 x nb, formed here): each
  // w's third of eqs and of the products G[w][p][i] = eqs[w][p] l_beta_i on
  // its own stream, so the scalings overlap the other thirds' products; then
  // an entry (row 32p+s of matrix w, value v) adds v G[w][p][s/2] u_{s%2} to
  // the row of its output (l = (c - r) mod 32, q = c / 32).
  void layer_rows_from(Engine<T>& e, const L* eq1, const L* dg, const L* dg2, L* eqs, L* rows) {
    DBuf<L> G(e.pool, 3 * nb * 16 * E);
    e.fork(3);
    for (int w = 0; w < 3; w++) {
      e.on(w);
      L* ew = eqs + (size_t)w * nb * E;
      if (w == 0) e.d2d(ew, eq1, nb * E * sizeof(L));
      else e.scale(w == 1 ? dg : dg2, eq1, ew, nb);
      e.mulfix_lb(ew, nb, G.p + (size_t)w * nb * 16 * E);
    }
    e.join();
    const double ne = (double)(mr[0].nnz + mr[1].nnz + mr[2].nnz);
    e.launch("k_rows_gather", ne * (Engine<T>::EB + 8) + 32.0 * nb * Engine<T>::EB, ne * E, (k_rows_gather<T>),
             dim3(grid_for(32 * nb, 8, 1u << 16)), dim3(256), (const uint64_t*)eptr.p, (const uint32_t*)ecode.p,
             (const L*)eval.p, (const L*)G.p, (uint64_t)nb, rows, (uint64_t)(32 * nb));
  }

  // Section sizes of the serialized proof (general.rs:120-150): mult check,
  // linear check, and the openings of e_mult, abc, the stacked errors and the
  // rotated rho batch; returns the total with the four roots.
  size_t sections(size_t& s_mult, size_t& s_lin, size_t& s_em, size_t& s_abc, size_t& s_es, size_t& s_rho) const {
    const size_t EBs = E * sizeof(L);
    s_mult = 8 + 3 * (size_t)log_b * EBs + 4 * EBs;
    s_lin = EBs + 8 + 3 * (size_t)log_b * EBs + 64 * EBs;
    s_em = opening_bytes(pe.P, pe.P.k1, sizeof(L), 0);
    s_abc = opening_bytes(pabc.P, 3 * pabc.P.k1, sizeof(L), 0);
    s_es = opening_bytes(pe.P, 3 * pe.P.k1, sizeof(L), 0);
    s_rho = opening_bytes(pw.P, pw.P.k1, sizeof(L), 32);
    return 128 + s_mult + s_lin + s_em + s_abc + s_es + s_rho;
  }
  size_t proof_size() const {
    size_t a, b, c, d, e, f;
    return sections(a, b, c, d, e, f);
  }

  ProofView prove_device(Ctx<T>& cx, const L* w /* device, N limbs */) {
    Engine<T>& e = *cx.eng;
    e.finish_downloads();  // a failed earlier proof may have left copies in flight
    Pool& pool = cx.pool;
    const size_t pool_base = pool.mark();
    KProf& prof = cx.prof;
    e.prof = &prof;
    Stages stg;
    stg.on = cx.timing;
    stg.st = cx.st;
    stg.start();
    // 1. rho_l[q] = pack_r2(rot_l(w_q)) for all 32 l (reduce.rs:690-707)
    DBuf<L> rho(&pool, 32 * nb * E);
    e.pack(w, rho.p, nb, 32);
    // 2. base-ring matvecs: slots of a_hat, b_hat are A w, B w blocks
    //    (psi(a_hat) = A w since a_hat = sum D rho - e_A = phi(A w) exactly)
    DBuf<L> abc_slots(&pool, 3 * N);
    unsigned g0 = grid_for(N, 256, 1u << 20);
    for (int m = 0; m < 2; m++)
      e.launch("k_r0_matvec", 0, 0, (k_r0_matvec<T>), dim3(g0), dim3(256), (const uint32_t*)mr[m].ptr.p,
               (const uint32_t*)mr[m].idx.p, (const L*)mr[m].val.p, w, (uint64_t)0, abc_slots.p + m * N, (uint64_t)N, 1,
               (uint64_t)1, (uint64_t)0);
    // 3. layered matvecs sum_l D_l rho_l (general.rs:211-230) for A, B, C
    DBuf<L> sums(&pool, 3 * nb * E);
    if constexpr (kMx<T>) {
      // sum_{s,s'} C_p[s][s'] b_s b_s' on the matrix cores
      // the three matrices are independent: one stream each (k_cgemm is
      // bound by load latency, so three in flight overlap well)
      // (a C-matrix slice per stream; one slice when the work stays on one
      // stream)
      e.fork(3);
      const size_t ns = (size_t)e.forked + 1;
      DBuf<L> cm(&pool, ns * nb * 1024);
      for (int m = 0; m < 3; m++) {
        e.on(m);
        L* cmm = cm.p + (size_t)(m % ns) * nb * 1024;
        e.launch("k_cmat", (double)nb * 1024 * sizeof(L), 0, (k_cmat<T>), dim3(grid_for(nb, 8, 4096)), dim3(256),
                 (const uint32_t*)mr[m].ptr.p, (const uint32_t*)mr[m].idx.p, (const L*)mr[m].val.p, w, cmm,
                 (uint64_t)nb);
        e.cgemm(cmm, sums.p + m * nb * E, nb);
      }
      e.join();
    } else {
      // the same factorization with 64-bit limbs: C_p, then the products
      // b_s b_s' on the matrix cores with exact digit classes (k_cgemm64)
      DBuf<L> cm(&pool, nb * 1024);
      for (int m = 0; m < 3; m++) {
        e.launch("k_cmat", (double)nb * 1024 * sizeof(L), 0, (k_cmat<T>), dim3(grid_for(nb, 8, 4096)), dim3(256),
                 (const uint32_t*)mr[m].ptr.p, (const uint32_t*)mr[m].idx.p, (const L*)mr[m].val.p, w, cm.p,
                 (uint64_t)nb);
        e.cgemm64(cm.p, sums.p + m * nb * E, (uint64_t)nb);
      }
    }
    // 4. c_hat = pack_r2(psi(c_sum)) (general.rs:410-414); a_hat, b_hat = phi(slots)
    e.psi(sums.p + 2 * nb * E, abc_slots.p + 2 * N, nb);
    // hats holds a_hat, b_hat, c_hat and then e_mult: the mult check folds
    // the four tables at one stride
    DBuf<L> hats(&pool, 4 * nb * E);
    e.pack(abc_slots.p, hats.p, 3 * nb, 1);
    // 5. e_A, e_B, e_C = sums - hats (stacked error), e_mult = a_hat b_hat - c_hat
    DBuf<L> estack(&pool, 3 * nb * E);
    L* emult = hats.p + 3 * nb * E;
    e.sub(sums.p, hats.p, estack.p, 3 * nb * E);
    e.template mul_sub<16>(hats.p, hats.p + nb * E, hats.p + 2 * nb * E, emult, nb);
    sums.release();
    stg.mark("reduce_matvec");
    // 6. commits (general.rs:434-451)
    // the four commitments are independent: one stream each
    e.fork(4);
    e.on(1);
    CommitDev<T> ces = e.commit(pe, estack.p, 3 * pe.P.k1, estack.p);
    e.on(2);
    CommitDev<T> cabc = e.commit(pabc, abc_slots.p, 3 * pabc.P.k1, hats.p);
    e.on(3);
    CommitDev<T> cem = e.commit(pe, emult, pe.P.k1, emult);
    e.on(0);
    CommitDev<T> cw = e.commit(pw, w, pw.P.k1, rho.p);
    e.join();
    e.resolve_roots({&cw, &cabc, &cem, &ces});
    stg.mark("commit");
    Transcript tr("general");
    tr.root(cw.root);
    tr.root(cabc.root);
    tr.root(cem.root);
    tr.root(ces.root);
    StdRng rs = tr.sampler("mult_check_target");
    std::vector<G> rt = e.sample(rs, log_b);
    // 7. mult check, sub form, on (a_hat, b_hat, c_hat, e_mult)
    tr.label("mult_check_sumcheck");
    const L* mt[4] = {hats.p, hats.p + nb * E, hats.p + 2 * nb * E, emult};
    auto mp = e.mult_check(mt, 4, log_b, rt, tr);
    tr.elems("mult/finals", mp.finals.data(), 4);
    G gam = e.challenge(tr, "linear_check_gamma");
    G gam2 = gam * gam;
    stg.mark("mult_sumcheck");
    // proof layout (general.rs:120-150): roots, mult, lin, open_emult,
    // open_abc, open_estack, open_rho -- sections sized up front
    size_t s_mult, s_lin, s_em, s_abc, s_es, s_rho;
    const size_t total = sections(s_mult, s_lin, s_em, s_abc, s_es, s_rho);
    uint8_t* buf = cx.proofbuf.get(total);
    uint8_t* o_mult = buf + 128;
    uint8_t* o_lin = o_mult + s_mult;
    uint8_t* o_em = o_lin + s_lin;
    uint8_t* o_abc = o_em + s_em;
    uint8_t* o_es = o_abc + s_abc;
    uint8_t* o_rho = o_es + s_es;
    // 8. openings at s1 (general.rs:475-478)
    G wts[3] = {G::one(), gam, gam2};
    Sink k_es(o_es, s_es), k_abc(o_abc, s_abc), k_em(o_em, s_em);
    // estack's value enters the linear claim: it is formed first on the
    // prover stream (the rest of its opening in the background).  The other
    // two values enter only the final self-check: those openings run on the
    // background stream, beside the layer rows and the linear check, and are
    // enqueued after the layer rows (the host's launches are the serial
    // resource here, so the critical path goes first).
    DBuf<L> es_y = e.open_dev(pe, ces, mp.chal.data(), wts, 3, k_es, true);
    stg.mark("open_s1");
    // 9. layer rows (general.rs:233-271) as sum over cells of
    //    {eq_s1, gamma eq_s1, gamma^2 eq_s1}[p] * D[cell]
    DBuf<L> eqs(&pool, 3 * nb * E);
    DBuf<L> rows(&pool, 32 * nb * E);
    {
      DBuf<L> eq1 = e.eq_table(mp.chal.data(), log_b);
      DBuf<L> dg = e.put(&gam, 1), dg2 = e.put(&gam2, 1);
      layer_rows_from(e, eq1.p, dg.p, dg2.p, eqs.p, rows.p);
    }
    eqs.release();
    stg.mark("layer_rows");
    e.bg_begin();
    DBuf<L> abc_y = e.open_dev(pabc, cabc, mp.chal.data(), wts, 3, k_abc);
    DBuf<L> em_y = e.open_dev(pe, cem, mp.chal.data(), nullptr, 1, k_em);
    e.bg_end();
    const G estack_val = e.get(es_y.p);
    // 10. linear check over the 32 (rho_l, row_l) pairs; rho is low-degree
    G claim_lin = mp.finals[0] + gam * mp.finals[1] + gam2 * mp.finals[2] + estack_val;
    tr.label("linear_check_sumcheck");
    auto lp = e.template linear<16>(rho.p, nb, rows.p, nb, 32, log_b, claim_lin, tr);
    rows.release();
    stg.mark("linear_sumcheck");
    // 11. rotated openings of w at s2 (pcs.rs:973-1040)
    Sink k_rho(o_rho, s_rho);
    DBuf<L> ys_d = e.rotated_open_dev(pw, cw, rho.p, nb, lp.chal.data(), 32, k_rho);
    stg.mark("open_rho");
    // the opening values of the self-check, one synchronization
    e.bg_join();
    std::vector<G> late = e.pick_many({{abc_y.p, 1, 1}, {em_y.p, 1, 1}, {ys_d.p, 1, 32}});
    const G abc_val = late[0], em_val = late[1];
    const G* ys = late.data() + 2;
    Sink hd(buf, 128 + s_mult + s_lin);
    for (auto* r : {cw.root, cabc.root, cem.root, ces.root}) hd.bytes(r, 32);
    hd.u64(mp.chal.size());
    for (auto& g : mp.rounds) hd.gr(g);
    for (int i = 0; i < 4; i++) hd.gr(mp.finals[i]);
    hd.gr(claim_lin);
    hd.u64(lp.chal.size());
    for (auto& g : lp.rounds) hd.gr(g);
    for (int l = 0; l < 32; l++) hd.gr(lp.finals[32 + l]);  // final_row
    for (int l = 0; l < 32; l++) hd.gr(lp.finals[l]);       // final_rho
    if (!hd.full() || !k_es.full() || !k_abc.full() || !k_em.full() || !k_rho.full())
      throw Error(NENYA_E_DEVICE, "proof size mismatch");
    bool ok = em_val == mp.finals[3] &&
              abc_val == mp.finals[0] + gam * mp.finals[1] + gam2 * mp.finals[2];
    for (int l = 0; l < 32; l++) ok = ok && ys[l] == lp.finals[l];
    if (!ok && !skip_selfcheck()) throw Error(NENYA_E_DEVICE, "internal opening self-check failed");
    e.finish_downloads();
    cx.last_timing = stg.collect();
    if (prof.on) prof.collect();
    note_footprint(pool.peak - pool_base);
    return {buf, total};
  }

  ProofView prove_host(Ctx<T>& cx, const L* w, size_t n) {
    if (n != N) throw Error(NENYA_E_SHAPE, "witness length must be 32 * 2^log_blocks");
    DBuf<L> dw(&cx.pool, N);
    cx.eng->h2d(dw.p, w, N * sizeof(L));
    return prove_device(cx, dw.p);
  }
};

}  // namespace nenya
