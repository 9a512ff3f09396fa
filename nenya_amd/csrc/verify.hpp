// Verifiers (verify_simd / verify_mimd / verify_general, snark-ring/src/
// simd.rs:380-522, mimd.rs:403-551, general.rs:545-698) over serialized
// proofs.  The transcript replay, the sumcheck checks and the PCS checks
// (m-bar evaluation, query derivation, RAA re-encoding, column combination,
// Merkle paths, psi-kernel test) run on the host; the linear-size row
// recompute (the M_gamma row at s1 dotted with eq(s2)) runs on the GPU with the
// params' own matrices and kernels.
// Verdicts: 0 Ok, then VerifyError in error.rs:5-16 order (1 MultSumcheck,
// 2 LinSumcheck, 3 LinClaimMismatch, 4 LinARowMismatch, 5 PcsW, 6 PcsE,
// 7 KernelCheck), and 8 for bytes that do not parse as a proof of these
// params (the reference verifies proof structs, so it has no such case).
#pragma once
#include <thread>

#include "snark.hpp"

namespace nenya {

enum { V_OK = 0, V_MULT = 1, V_LIN = 2, V_LINCLAIM = 3, V_AROW = 4, V_PCSW = 5, V_PCSE = 6, V_KERNEL = 7, V_PARSE = 8 };

// ---------------------------------------------------------------------------
// Host arithmetic the verifier needs beyond Gr<T>
// ---------------------------------------------------------------------------
template <class T> struct HostRing {
  using L = typename T::L;
  using C = typename T::C;
  using G = Gr<T>;
  // Teichmuller lift as on the device: sigma^-(k-1) (the host matrix) then
  // k-1 squarings.
  static G lift(const G& x) {
    const Consts<T>& K = Consts<T>::get();
    G y = G::zero();
    for (int j = 0; j < E; j++) {
      const C xj = (C)x.v[j];
      if (!xj) continue;
      for (int o = 0; o < E; o++) y.v[o] = (L)(y.v[o] + (L)((C)K.frob_t[(size_t)j * E + o] * xj));
    }
    for (int i = 0; i < T::K - 1; i++) y = y * y;
    return y;
  }
  // R2Sampler::sample (sampling.rs:57-68)
  static std::vector<G> sample(StdRng& rng, size_t cnt) {
    std::vector<G> out(cnt);
    for (size_t i = 0; i < cnt; i++)
      for (int l = 0; l < E; l++) out[i].v[l] = (L)rng.limb(T::K);
    for (auto& g : out) g = lift(g);
    return out;
  }
  static G challenge(Transcript& tr, const char* label) {
    StdRng r = tr.sampler(label);
    return sample(r, 1)[0];
  }
  // build_eq_table (eq.rs:35-58): r[0] is the most significant index bit
  static std::vector<G> eq_table(const G* r, size_t n) {
    std::vector<G> t(size_t(1) << n);
    t[0] = G::one();
    size_t len = 1;
    for (size_t l = 0; l < n; l++) {
      for (size_t k = len; k-- > 0;) {
        const G hi = t[k] * r[l];
        t[2 * k] = t[k] - hi;
        t[2 * k + 1] = hi;
      }
      len *= 2;
    }
    return t;
  }
  static G eq_one(const G& t, const G& r) { return t * r + (G::one() - t) * (G::one() - r); }
  // x * s for an R0 scalar s
  static G scale(const G& x, L s) {
    G o;
    for (int i = 0; i < E; i++) o.v[i] = (L)((C)x.v[i] * (C)s);
    return o;
  }
  // psi_r2 (embedding.rs:247-261) flattened: 16 evaluations at the T1
  // points, each split into its values at alpha = 0 and alpha = 1
  static bool in_kernel(const G& x) {
    const Consts<T>& K = Consts<T>::get();
    for (int pt = 0; pt < 16; pt++) {
      C raw[7] = {0, 0, 0, 0, 0, 0, 0};
      for (int k = 0; k < 32; k++) {
        C a[4], b[4], p[7];
        for (int c = 0; c < 4; c++) {
          a[c] = (C)x.v[4 * k + c];
          b[c] = (C)K.psi_pw[((size_t)pt * 32 + k) * 4 + c];
        }
        g4_raw(a, b, p);
        for (int c = 0; c < 7; c++) raw[c] += p[c];
      }
      C g[4];
      g4_reduce(raw, g);
      if ((L)g[0] != 0 || (L)(g[0] + g[1] + g[2] + g[3]) != 0) return false;
    }
    return true;
  }
};

// ---------------------------------------------------------------------------
// Proof reader
// ---------------------------------------------------------------------------
template <class T> struct Reader {
  using L = typename T::L;
  using G = Gr<T>;
  const uint8_t* p;
  size_t n, o = 0;
  bool bad = false;
  Reader(const uint8_t* b, size_t len) : p(b), n(len) {}
  const uint8_t* take(size_t k) {
    if (bad || n - o < k) {
      bad = true;
      return nullptr;
    }
    const uint8_t* r = p + o;
    o += k;
    return r;
  }
  uint64_t u64() {
    const uint8_t* b = take(8);
    uint64_t x = 0;
    if (b)
      for (int i = 0; i < 8; i++) x |= (uint64_t)b[i] << (8 * i);
    return x;
  }
  uint32_t u32() {
    const uint8_t* b = take(4);
    uint32_t x = 0;
    if (b)
      for (int i = 0; i < 4; i++) x |= (uint32_t)b[i] << (8 * i);
    return x;
  }
  G gr() {
    G g = G::zero();
    const uint8_t* b = take(sizeof g.v);
    if (b) std::memcpy(g.v, b, sizeof g.v);
    return g;
  }
  // a count of items of at least `unit` bytes that can still fit
  bool sane(uint64_t cnt, size_t unit) {
    if (bad || cnt > (n - o) / (unit ? unit : 1) + 1) bad = true;
    return !bad;
  }
};

// A serialized sumcheck proof: mult (mult_check.rs:127-172, 398-424), SIMD /
// MIMD linear (linear_check_simd.rs:87-131, linear_check_mimd.rs:39-83) or
// general linear (linear_check_general.rs:142-190).
template <class T> struct ScProof {
  Gr<T> claimed = Gr<T>::zero();
  std::vector<Gr<T>> rounds, fin, fin2;
  size_t n = 0;
  void read(Reader<T>& r, bool with_claim, size_t nfin, size_t nfin2 = 0) {
    if (with_claim) claimed = r.gr();
    const uint64_t cnt = r.u64();
    if (!r.sane(cnt, 3 * sizeof(Gr<T>))) return;
    n = cnt;
    rounds.resize(3 * n);
    for (auto& g : rounds) g = r.gr();
    fin.resize(nfin);
    for (auto& g : fin) g = r.gr();
    fin2.resize(nfin2);
    for (auto& g : fin2) g = r.gr();
  }
};

// OpeningProof / R2OpeningProof (pcs.rs:357-406, r2_pcs.rs:109-160) and the
// rotated batch (pcs.rs:426-471)
template <class T> struct Opening {
  using L = typename T::L;
  size_t nmbar = 0, k2 = 0, t = 0, col_len = 0, plen = 0;
  int col_elem = 32;
  std::vector<Gr<T>> mbar;
  std::vector<uint32_t> idx;
  std::vector<L> cols;
  std::vector<uint8_t> paths;  // t * plen * 32
  void read(Reader<T>& r, bool batch, int celem) {
    col_elem = celem;
    if (batch) {
      const uint64_t nl = r.u64();
      if (!r.sane(nl, 8)) return;
      nmbar = nl;
      for (size_t l = 0; l < nl && !r.bad; l++) {
        const uint64_t kk = r.u64();
        if (!r.sane(kk, sizeof(Gr<T>))) return;
        if (l == 0) {
          k2 = kk;
          mbar.resize(nl * kk);
        } else if (kk != k2) {
          r.bad = true;
          return;
        }
        for (size_t j = 0; j < kk; j++) mbar[l * kk + j] = r.gr();
      }
    } else {
      nmbar = 1;
      const uint64_t kk = r.u64();
      if (!r.sane(kk, sizeof(Gr<T>))) return;
      k2 = kk;
      mbar.resize(kk);
      for (auto& g : mbar) g = r.gr();
    }
    const uint64_t tt = r.u64();
    if (!r.sane(tt, 4)) return;
    t = tt;
    idx.resize(t);
    for (auto& x : idx) x = r.u32();
    if (r.u64() != t) {
      r.bad = true;
      return;
    }
    for (size_t q = 0; q < t && !r.bad; q++) {
      const uint64_t cl = r.u64();
      if (!r.sane(cl, (size_t)celem * sizeof(L))) return;
      if (q == 0) {
        col_len = cl;
        cols.resize(t * cl * celem);
      } else if (cl != col_len) {
        r.bad = true;
        return;
      }
      const uint8_t* b = r.take(cl * celem * sizeof(L));
      if (b) std::memcpy(&cols[q * cl * celem], b, cl * celem * sizeof(L));
    }
    if (r.u64() != t) {
      r.bad = true;
      return;
    }
    for (size_t q = 0; q < t && !r.bad; q++) {
      const uint64_t pl = r.u64();
      if (!r.sane(pl, 32)) return;
      if (q == 0) {
        plen = pl;
        paths.resize(t * pl * 32);
      } else if (pl != plen) {
        r.bad = true;
        return;
      }
      const uint8_t* b = r.take(pl * 32);
      if (b) std::memcpy(&paths[q * pl * 32], b, pl * 32);
    }
  }
};

// ---------------------------------------------------------------------------
// PCS checks
// ---------------------------------------------------------------------------
template <class T> struct PcsCheck {
  using L = typename T::L;
  using C = typename T::C;
  using G = Gr<T>;
  using H = HostRing<T>;

  static void words_of(const uint8_t* b, uint32_t w[8]) {
    for (int i = 0; i < 8; i++) w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) |
                                        ((uint32_t)b[4 * i + 2] << 16) | ((uint32_t)b[4 * i + 3] << 24);
  }
  // MerkleTree::verify_path (merkle.rs:115-127)
  static bool merkle_ok(const uint8_t leaf[32], size_t idx, const uint8_t* path, size_t plen, const uint8_t root[32]) {
    uint32_t cur[8], sib[8], o[8];
    words_of(leaf, cur);
    for (size_t i = 0; i < plen; i++) {
      words_of(path + 32 * i, sib);
      if ((idx & 1) == 0) b3::hash64(cur, sib, o);
      else b3::hash64(sib, cur, o);
      std::memcpy(cur, o, sizeof o);
      idx >>= 1;
    }
    uint32_t r[8];
    words_of(root, r);
    return std::memcmp(cur, r, sizeof r) == 0;
  }
  // hash_packed_column / hash_r2_column: BLAKE3 of the column's limbs
  static void leaf_of(const Opening<T>& o, size_t q, uint8_t out[32]) {
    b3::Hasher h;
    h.update(&o.cols[q * o.col_len * o.col_elem], o.col_len * o.col_elem * sizeof(L));
    h.finalize(0, out, 32);
  }
  // RaaParams::encode (raa.rs:67-98) of k2 ring elements
  static std::vector<G> encode(const PcsParams& P, const G* msg) {
    const size_t n = P.n;
    std::vector<G> t(n), out(n);
    for (size_t i = 0; i < n; i++) t[i] = H::scale(msg[P.perm1[i] / P.q], (L)P.d1[i]);
    for (size_t i = 1; i < n; i++) t[i] = t[i] + t[i - 1];
    for (size_t i = 0; i < n; i++) out[i] = H::scale(t[P.perm2[i]], (L)P.d2[i]);
    for (size_t i = 1; i < n; i++) out[i] = out[i] + out[i - 1];
    return out;
  }
  static std::vector<uint32_t> queries(const PcsParams& P, const uint8_t root[32], const G* mbar, size_t cnt) {
    return derive_queries(P.packed ? "pcs-ring/fs/queries" : "snark-ring/r2-pcs/fs/queries", root, mbar,
                          cnt * sizeof(G), P.n, P.ell);
  }
  // sum_i e0s[i] * column entry i: packed entries are 32 slots lifted by
  // pack_r2 (rotated by `rot` slots), R2 entries are ring elements.  For
  // packed columns: V[j] = sum_i col[i][j] e0s[i], then
  // sum_s b_s V[(s + rot) % 32] with b_{2i+h} = l_beta_i u_h (u_0 = 1 - alpha,
  // u_1 = alpha), i.e. 16 products by l_beta_i per rotation.
  static void combine(const Opening<T>& o, size_t q, const G* e0s, size_t rows, const int* rots, size_t nrot,
                      G* acc) {
    if (o.col_elem == E) {
      G a = G::zero();
      for (size_t i = 0; i < rows; i++) {
        G c;
        std::memcpy(c.v, &o.cols[(q * rows + i) * E], sizeof c.v);
        a = a + e0s[i] * c;
      }
      acc[0] = a;
      return;
    }
    const Consts<T>& K = Consts<T>::get();
    G V[32];
    for (int j = 0; j < 32; j++) V[j] = G::zero();
    for (size_t i = 0; i < rows; i++)
      for (int j = 0; j < 32; j++) {
        const L s = o.cols[(q * rows + i) * 32 + j];
        if (s) V[j] = V[j] + H::scale(e0s[i], s);
      }
    for (size_t r = 0; r < nrot; r++) {
      G a = G::zero();
      for (int i = 0; i < 16; i++) {
        const G& v0 = V[(2 * i + rots[r]) % 32];
        const G& v1 = V[(2 * i + 1 + rots[r]) % 32];
        // u_0 v0 + u_1 v1 per R1 coefficient: (1 - alpha) v0 + alpha v1
        G u;
        for (int d = 0; d < 32; d++) {
          C x[4], y[4];
          for (int c = 0; c < 4; c++) {
            x[c] = (C)v0.v[4 * d + c];
            y[c] = (C)v1.v[4 * d + c];
          }
          // (1 - alpha) x = (x0 + x3, x1 - x0 + x3, x2 - x1, x3 - x2); alpha y = (-y3, y0 - y3, y1, y2)
          u.v[4 * d + 0] = (L)(x[0] + x[3] - y[3]);
          u.v[4 * d + 1] = (L)(x[1] - x[0] + x[3] + y[0] - y[3]);
          u.v[4 * d + 2] = (L)(x[2] - x[1] + y[1]);
          u.v[4 * d + 3] = (L)(x[3] - x[2] + y[2]);
        }
        a = a + K.lbeta[i] * u;
      }
      acc[r] = a;
    }
  }
  template <class F> static void par_for(size_t n, F f) {
    const size_t nt = std::max<size_t>(1, std::min<size_t>({(size_t)std::thread::hardware_concurrency(), 16, n}));
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    for (size_t w = 0; w < nt; w++)
      th.emplace_back([&] {
        for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
      });
    for (auto& x : th) x.join();
  }
  // Pcs::verify / verify_stacked (pcs.rs:594-649, 782-837) and R2Pcs::verify /
  // verify_stacked (r2_pcs.rs:340-393, 562-618): weights w (s of them) or none
  static bool weighted(const PcsParams& P, const uint8_t root[32], const G* z, const G* w, size_t s, const G& val,
                       const Opening<T>& o) {
    const size_t rows = s * P.k1;
    if (o.nmbar != 1 || o.k2 != P.k2 || o.t != P.ell || o.col_len != rows) return false;
    const size_t lk1 = (size_t)__builtin_ctzll(P.k1), lk2 = (size_t)__builtin_ctzll(P.k2);
    std::vector<G> e0 = H::eq_table(z, lk1), e1 = H::eq_table(z + lk1, lk2);
    std::vector<G> e0s(rows);
    for (size_t t = 0; t < s; t++)
      for (size_t i = 0; i < P.k1; i++) e0s[t * P.k1 + i] = w ? w[t] * e0[i] : e0[i];
    G y = G::zero();
    for (size_t j = 0; j < P.k2; j++) y = y + o.mbar[j] * e1[j];
    if (!(y == val)) return false;
    if (queries(P, root, o.mbar.data(), P.k2) != o.idx) return false;
    std::vector<G> ubar = encode(P, o.mbar.data());
    std::atomic<bool> ok{true};
    const int rot0 = 0;
    par_for(P.ell, [&](size_t q) {
      if (!ok) return;
      const size_t j = o.idx[q];
      G acc;
      combine(o, q, e0s.data(), rows, &rot0, 1, &acc);
      uint8_t leaf[32];
      leaf_of(o, q, leaf);
      if (!(acc == ubar[j]) || !merkle_ok(leaf, j, &o.paths[q * o.plen * 32], o.plen, root)) ok = false;
    });
    return ok;
  }
  // Pcs::rotated_verify_batch (pcs.rs:1044-1115): one query set over nl m-bars
  static bool rotated(const PcsParams& P, const uint8_t root[32], const G* z, const G* ys, size_t nl,
                      const Opening<T>& o) {
    if (o.nmbar != nl || o.k2 != P.k2 || o.t != P.ell || o.col_len != P.k1) return false;
    const size_t lk1 = (size_t)__builtin_ctzll(P.k1), lk2 = (size_t)__builtin_ctzll(P.k2);
    std::vector<G> e0 = H::eq_table(z, lk1), e1 = H::eq_table(z + lk1, lk2);
    for (size_t l = 0; l < nl; l++) {
      G y = G::zero();
      for (size_t j = 0; j < P.k2; j++) y = y + o.mbar[l * P.k2 + j] * e1[j];
      if (!(y == ys[l])) return false;
    }
    if (queries(P, root, o.mbar.data(), nl * P.k2) != o.idx) return false;
    std::vector<std::vector<G>> ubar(nl);
    par_for(nl, [&](size_t l) { ubar[l] = encode(P, &o.mbar[l * P.k2]); });
    std::vector<int> rots(nl);
    for (size_t l = 0; l < nl; l++) rots[l] = (int)l;
    std::atomic<bool> ok{true};
    par_for(P.ell, [&](size_t q) {
      if (!ok) return;
      const size_t j = o.idx[q];
      uint8_t leaf[32];
      leaf_of(o, q, leaf);
      if (!merkle_ok(leaf, j, &o.paths[q * o.plen * 32], o.plen, root)) {
        ok = false;
        return;
      }
      std::vector<G> acc(nl);
      combine(o, q, e0.data(), P.k1, rots.data(), nl, acc.data());
      for (size_t l = 0; l < nl; l++)
        if (!(acc[l] == ubar[l][j])) ok = false;
    });
    return ok;
  }
  // columns_in_kernel (r2_pcs.rs:182-188)
  static bool in_kernel(const Opening<T>& o) {
    std::atomic<bool> ok{true};
    par_for(o.t, [&](size_t q) {
      for (size_t i = 0; i < o.col_len && ok; i++) {
        G c;
        std::memcpy(c.v, &o.cols[(q * o.col_len + i) * E], sizeof c.v);
        if (!H::in_kernel(c)) ok = false;
      }
    });
    return ok;
  }
  static bool idx_in_range(const PcsParams& P, const Opening<T>& o) {
    for (uint32_t j : o.idx)
      if (j >= P.n) return false;
    return true;
  }
};

// ---------------------------------------------------------------------------
// Sumcheck checks (Fiat-Shamir replay)
// ---------------------------------------------------------------------------
template <class T> struct ScCheck {
  using G = Gr<T>;
  using H = HostRing<T>;
  static G next(const G q[3], const G& r) { return q[0] + r * (q[1] + r * q[2]); }
  // verify_mult_check (mult_check.rs:327-377) and its sub form (517-549):
  // returns false on a failed round or final check; chal receives the point
  static bool mult(const ScProof<T>& pf, const std::vector<G>& rt, bool sub, const G& start, Transcript& tr,
                   std::vector<G>& chal) {
    const size_t nv = rt.size();
    if (pf.n != nv) return false;
    G c = start;
    chal.resize(nv);
    for (size_t r = 0; r < nv; r++) {
      const G* q = &pf.rounds[3 * r];
      if (!(q[0] + rt[r] * (q[1] + q[2]) == c)) return false;
      tr.elems("sc/round", q, 3);
      chal[r] = H::challenge(tr, "sc/challenge");
      c = H::eq_one(rt[r], chal[r]) * next(q, chal[r]);
    }
    G ef = G::one();
    for (size_t i = 0; i < nv; i++) ef = ef * H::eq_one(rt[i], chal[i]);
    G prod = pf.fin[0] * pf.fin[1];
    if (sub) prod = prod - pf.fin[2] - pf.fin[3];
    return c == ef * prod;
  }
  // linear checks: s(0) + s(1) = 2 a0 + a1 + a2 per round, then the final
  // product (linear_check_simd.rs verify, linear_check_mimd.rs:229-259,
  // linear_check_general.rs verify)
  static bool linear(const ScProof<T>& pf, Transcript& tr, std::vector<G>& chal, bool general) {
    if (pf.n == 0) return false;
    G c = pf.claimed;
    chal.resize(pf.n);
    for (size_t r = 0; r < pf.n; r++) {
      const G* q = &pf.rounds[3 * r];
      if (!(q[0] + q[0] + q[1] + q[2] == c)) return false;
      tr.elems("sc/round", q, 3);
      chal[r] = H::challenge(tr, "sc/challenge");
      c = next(q, chal[r]);
    }
    G f = G::zero();
    if (general) {
      for (int l = 0; l < 32; l++) f = f + pf.fin[l] * pf.fin2[l];
    } else {
      f = pf.fin[0] * pf.fin[1];
    }
    return c == f;
  }
};

// ---------------------------------------------------------------------------
// SNARK verifiers
// ---------------------------------------------------------------------------
template <class T> struct Verifier {
  using L = typename T::L;
  using G = Gr<T>;
  using H = HostRing<T>;
  using P = PcsCheck<T>;
  using S = ScCheck<T>;

  static std::vector<G> target(Transcript& tr, size_t n) {
    StdRng r = tr.sampler("mult_check_target");
    return H::sample(r, n);
  }
  static G inner(const std::vector<G>& a, const std::vector<G>& b) {
    G s = G::zero();
    for (size_t i = 0; i < a.size() && i < b.size(); i++) s = s + a[i] * b[i];
    return s;
  }

  // verify_simd (simd.rs:380-522)
  static int simd(SimdParams<T>& sp, const HostCoo& A, const HostCoo& B, const HostCoo& Cm, const uint8_t* proof,
                  size_t len) {
    const size_t n = size_t(1) << sp.log_n, m = size_t(1) << sp.log_m;
    check_coo(A, n, m);
    check_coo(B, n, m);
    check_coo(Cm, n, m);
    Reader<T> rd(proof, len);
    uint8_t root_w[32] = {0}, root_e[32] = {0};
    if (const uint8_t* b = rd.take(32)) std::memcpy(root_w, b, 32);
    if (const uint8_t* b = rd.take(32)) std::memcpy(root_e, b, 32);
    ScProof<T> mp, lp;
    Opening<T> ow, oe;
    mp.read(rd, true, 2);
    lp.read(rd, true, 2);
    ow.read(rd, false, 32);
    oe.read(rd, false, E);
    if (rd.bad || rd.o != rd.n) return V_PARSE;
    const PcsParams& PW = sp.pw.P;
    const PcsParams& PQ = sp.pq.P;
    Transcript tr("simd");
    tr.root(root_w);
    tr.root(root_e);
    std::vector<G> rt = target(tr, sp.log_n);
    const size_t lk1 = (size_t)__builtin_ctzll(PQ.k1);
    if (oe.k2 != PQ.k2) return V_PCSE;
    G e_at_r = inner(oe.mbar, H::eq_table(rt.data() + lk1, sp.log_n - lk1));
    if (!P::idx_in_range(PQ, oe) || !P::weighted(PQ, root_e, rt.data(), nullptr, 1, e_at_r, oe)) return V_PCSE;
    if (!P::in_kernel(oe)) return V_KERNEL;
    const G claim_mult = mp.claimed, cw_at_r = claim_mult - e_at_r;
    tr.elems("claim_mult", &claim_mult, 1);
    tr.label("mult_check_sumcheck");
    std::vector<G> s1, s2;
    if (!S::mult(mp, rt, false, claim_mult, tr, s1)) return V_MULT;
    tr.elems("mult/finals", mp.fin.data(), 2);
    const G gam = H::challenge(tr, "linear_check_gamma"), gam2 = gam * gam;
    const G claim_lin = mp.fin[0] + gam * mp.fin[1] + gam2 * cw_at_r;
    if (!(claim_lin == lp.claimed)) return V_LINCLAIM;
    tr.elems("claim_lin", &claim_lin, 1);
    tr.label("linear_check_sumcheck");
    if (!S::linear(lp, tr, s2, false) || lp.n != (size_t)sp.log_m) return V_LIN;
    // M_gamma row at s2 on the device (simd.rs:539-580)
    G at_s2;
    {
      Ctx<T>& cx = sp.cx0;
      Engine<T>& e = *cx.eng;
      e.prof = nullptr;
      Pool& pool = cx.pool;
      Csr<T> ac, bc, cc;
      {
        DevSetup<T> ds(e);
        DevCoo<T> coo(e, {&A, &B, &Cm});
        ds.csr(ac, coo.c.p, coo.r.p, coo.v.p, A.nnz, m);
        ds.csr(bc, coo.c.p + coo.off[1], coo.r.p + coo.off[1], coo.v.p + coo.off[1], B.nnz, m);
        ds.csr(cc, coo.c.p + coo.off[2], coo.r.p + coo.off[2], coo.v.p + coo.off[2], Cm.nnz, m);
      }
      DBuf<L> eq_s1 = e.eq_table(s1.data(), sp.log_n), eq_r = e.eq_table(rt.data(), sp.log_n);
      DBuf<L> eq_s2 = e.eq_table(s2.data(), sp.log_m);
      DBuf<L> geq(&pool, n * E), g2eq(&pool, n * E), mrow(&pool, m * E);
      DBuf<L> dg = e.put(&gam, 1), dg2 = e.put(&gam2, 1);
      e.scale(dg.p, eq_s1.p, geq.p, n);
      e.scale(dg2.p, eq_r.p, g2eq.p, n);
      spmv_gr1(e, ac, (const L*)eq_s1.p, mrow.p, (uint64_t)m, 0);
      spmv_gr1(e, bc, (const L*)geq.p, mrow.p, (uint64_t)m, 1);
      spmv_gr1(e, cc, (const L*)g2eq.p, mrow.p, (uint64_t)m, 1);
      at_s2 = e.template dot<32>(eq_s2.p, mrow.p, m);
    }
    if (!(at_s2 == lp.fin[0])) return V_AROW;
    tr.elems("lin/finals", lp.fin.data(), 2);
    if (!P::idx_in_range(PW, ow) || !P::weighted(PW, root_w, s2.data(), nullptr, 1, lp.fin[1], ow)) return V_PCSW;
    return V_OK;
  }

  // verify_mimd (mimd.rs:403-551)
  static int mimd(MimdParams<T>& mp_, const uint8_t* proof, size_t len) {
    const size_t n = mp_.n, m = mp_.m;
    Reader<T> rd(proof, len);
    uint8_t roots[4][32] = {{0}};
    for (int i = 0; i < 4; i++)
      if (const uint8_t* b = rd.take(32)) std::memcpy(roots[i], b, 32);
    ScProof<T> mp, lp;
    Opening<T> ow, oem, oabc, oes;
    mp.read(rd, false, 4);
    lp.read(rd, true, 2);
    ow.read(rd, false, 32);
    oem.read(rd, false, E);
    oabc.read(rd, false, 32);
    oes.read(rd, false, E);
    if (rd.bad || rd.o != rd.n) return V_PARSE;
    const PcsParams &PW = mp_.pw.P, &PA = mp_.pabc.P, &PE = mp_.pe.P;
    Transcript tr("mimd");
    for (auto& r : roots) tr.root(r);
    std::vector<G> rt = target(tr, mp_.log_n);
    tr.label("mult_check_sumcheck");
    std::vector<G> s1, s2;
    if (!S::mult(mp, rt, true, G::zero(), tr, s1)) return V_MULT;
    tr.elems("mult/finals", mp.fin.data(), 4);
    const G gam = H::challenge(tr, "linear_check_gamma"), gam2 = gam * gam;
    tr.label("linear_check_sumcheck");
    if (!S::linear(lp, tr, s2, false) || lp.n != (size_t)mp_.log_m) return V_LIN;
    const size_t lke = (size_t)__builtin_ctzll(PE.k1), lka = (size_t)__builtin_ctzll(PA.k1);
    if (oes.k2 != PE.k2 || oem.k2 != PE.k2 || oabc.k2 != PA.k2) return V_LINCLAIM;
    const std::vector<G> e1e = H::eq_table(s1.data() + lke, mp_.log_n - lke);
    const std::vector<G> e1a = H::eq_table(s1.data() + lka, mp_.log_n - lka);
    const G estack_val = inner(oes.mbar, e1e), em_s1 = inner(oem.mbar, e1e), abc_val = inner(oabc.mbar, e1a);
    const G abc_exp = mp.fin[0] + gam * mp.fin[1] + gam2 * mp.fin[2];
    if (!(abc_val == abc_exp) || !(em_s1 == mp.fin[3])) return V_LINCLAIM;
    if (!(lp.claimed == abc_exp + estack_val)) return V_LINCLAIM;
    // M_gamma row at s2 on the device (mimd.rs:489-503)
    G at_s2;
    {
      Ctx<T>& cx = mp_.cx0;
      Engine<T>& e = *cx.eng;
      e.prof = nullptr;
      Pool& pool = cx.pool;
      DBuf<L> eqs(&pool, 3 * n * E), mrow(&pool, m * E);
      {
        DBuf<L> eq1 = e.eq_table(s1.data(), mp_.log_n);
        e.d2d(eqs.p, eq1.p, n * E * sizeof(L));
        DBuf<L> dg = e.put(&gam, 1), dg2 = e.put(&gam2, 1);
        e.scale(dg.p, eq1.p, eqs.p + n * E, n);
        e.scale(dg2.p, eq1.p, eqs.p + 2 * n * E, n);
      }
      mp_.m_row(e, eqs.p, mrow.p);
      DBuf<L> eq_s2 = e.eq_table(s2.data(), mp_.log_m);
      at_s2 = e.template dot<32>(eq_s2.p, mrow.p, m);
    }
    if (!(at_s2 == lp.fin[0])) return V_AROW;
    const G wts[3] = {G::one(), gam, gam2};
    if (!P::idx_in_range(PW, ow) || !P::weighted(PW, roots[0], s2.data(), nullptr, 1, lp.fin[1], ow)) return V_PCSW;
    if (!P::idx_in_range(PE, oem) || !P::weighted(PE, roots[2], s1.data(), nullptr, 1, em_s1, oem)) return V_PCSE;
    if (!P::idx_in_range(PA, oabc) || !P::weighted(PA, roots[1], s1.data(), wts, 3, abc_val, oabc)) return V_PCSW;
    if (!P::idx_in_range(PE, oes) || !P::weighted(PE, roots[3], s1.data(), wts, 3, estack_val, oes)) return V_PCSE;
    if (!P::in_kernel(oem) || !P::in_kernel(oes)) return V_KERNEL;
    return V_OK;
  }

  // verify_general (general.rs:545-698)
  static int general(GeneralParams<T>& gp, const uint8_t* proof, size_t len) {
    const size_t nb = gp.nb;
    Reader<T> rd(proof, len);
    uint8_t roots[4][32] = {{0}};
    for (int i = 0; i < 4; i++)
      if (const uint8_t* b = rd.take(32)) std::memcpy(roots[i], b, 32);
    ScProof<T> mp, lp;
    Opening<T> oem, oabc, oes, orho;
    mp.read(rd, false, 4);
    lp.read(rd, true, 32, 32);
    oem.read(rd, false, E);
    oabc.read(rd, false, 32);
    oes.read(rd, false, E);
    orho.read(rd, true, 32);
    if (rd.bad || rd.o != rd.n) return V_PARSE;
    const PcsParams &PW = gp.pw.P, &PA = gp.pabc.P, &PE = gp.pe.P;
    Transcript tr("general");
    for (auto& r : roots) tr.root(r);
    std::vector<G> rt = target(tr, gp.log_b);
    tr.label("mult_check_sumcheck");
    std::vector<G> s1, s2;
    if (!S::mult(mp, rt, true, G::zero(), tr, s1)) return V_MULT;
    tr.elems("mult/finals", mp.fin.data(), 4);
    const G gam = H::challenge(tr, "linear_check_gamma"), gam2 = gam * gam;
    tr.label("linear_check_sumcheck");
    if (!S::linear(lp, tr, s2, true) || lp.n != (size_t)gp.log_b) return V_LIN;
    const size_t lke = (size_t)__builtin_ctzll(PE.k1), lka = (size_t)__builtin_ctzll(PA.k1);
    if (oes.k2 != PE.k2 || oem.k2 != PE.k2 || oabc.k2 != PA.k2) return V_LINCLAIM;
    const std::vector<G> e1e = H::eq_table(s1.data() + lke, gp.log_b - lke);
    const std::vector<G> e1a = H::eq_table(s1.data() + lka, gp.log_b - lka);
    const G estack_val = inner(oes.mbar, e1e), em_s1 = inner(oem.mbar, e1e), abc_val = inner(oabc.mbar, e1a);
    const G abc_exp = mp.fin[0] + gam * mp.fin[1] + gam2 * mp.fin[2];
    if (!(abc_val == abc_exp) || !(em_s1 == mp.fin[3])) return V_LINCLAIM;
    if (!(lp.claimed == abc_exp + estack_val)) return V_LINCLAIM;
    // the 32 layer rows at s1 (general.rs:233-271) dotted with eq(s2), on the device
    std::vector<G> at_s2;
    {
      Ctx<T>& cx = gp.cx0;
      Engine<T>& e = *cx.eng;
      e.prof = nullptr;
      Pool& pool = cx.pool;
      DBuf<L> eqs(&pool, 3 * nb * E), rows(&pool, 32 * nb * E);
      {
        DBuf<L> eq1 = e.eq_table(s1.data(), gp.log_b);
        DBuf<L> dg = e.put(&gam, 1), dg2 = e.put(&gam2, 1);
        gp.layer_rows_from(e, eq1.p, dg.p, dg2.p, eqs.p, rows.p);
      }
      DBuf<L> eq_s2 = e.eq_table(s2.data(), gp.log_b);
      at_s2 = e.template dots<32>(rows.p, nb, 32, eq_s2.p, nb);
    }
    for (int l = 0; l < 32; l++)
      if (!(at_s2[l] == lp.fin[l])) return V_AROW;
    const G wts[3] = {G::one(), gam, gam2};
    if (!P::idx_in_range(PE, oem) || !P::weighted(PE, roots[2], s1.data(), nullptr, 1, em_s1, oem)) return V_PCSE;
    if (!P::idx_in_range(PA, oabc) || !P::weighted(PA, roots[1], s1.data(), wts, 3, abc_val, oabc)) return V_PCSW;
    if (!P::idx_in_range(PE, oes) || !P::weighted(PE, roots[3], s1.data(), wts, 3, estack_val, oes)) return V_PCSE;
    if (orho.nmbar != 32) return V_PCSW;
    if (!P::idx_in_range(PW, orho) || !P::rotated(PW, roots[0], s2.data(), lp.fin2.data(), 32, orho)) return V_PCSW;
    if (!P::in_kernel(oem) || !P::in_kernel(oes)) return V_KERNEL;
    return V_OK;
  }
};

}  // namespace nenya
