// Host-side building blocks: device memory pool, per-width constants,
// PCS parameters, Fiat-Shamir transcript.
#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>

#include "blake3.hpp"
#include "common.hpp"
#include "rng.hpp"

namespace nenya {

// ---------------------------------------------------------------------------
// Caching device allocator: blocks are returned to per-size-class free lists
// and reused by later proofs (no hipMalloc in steady state).
// ---------------------------------------------------------------------------
// Debug switch: NENYA_POISON=1 fills every block the pool hands out with
// 0xA5 bytes, so a kernel that reads memory nobody wrote shows up as a
// wrong proof instead of depending on what the block held before.
inline bool pool_poison() {
  static const bool on = [] {
    const char* e = std::getenv("NENYA_POISON");
    return e && e[0] == '1';
  }();
  return on;
}

// Debug switch: NENYA_GUARD=1 surrounds every pool block with 64 KiB guard
// zones of 0x5A bytes, checked when the block is returned (an out-of-bounds
// write is reported with the block size and offset; an out-of-bounds read
// sees 0x5A instead of whatever neighbours the block).
inline bool pool_guard() {
  static const bool on = [] {
    const char* e = std::getenv("NENYA_GUARD");
    return e && e[0] == '1';
  }();
  return on;
}
constexpr size_t GUARD = 1 << 16;

// Launch shapes picked by the autotuner (tune.hpp), by limb width (index
// log2(k) - 3) and the log2 of the launch's work items; 0 keeps the built-in
// heuristic.  Written by nenya_autotune, read by every later launch.
struct Tuning {
  static constexpr int NL = 40;
  // k_fold_mx: 1 split mode (a block per 32-element tile), 2 bulk mode;
  // fold_cap: bulk-mode workgroups (by log2 of the tile count)
  std::atomic<int> fold_mode[4][NL], fold_cap[4][NL];
  // k_gram_lin workgroups (by log2 of the 32-element chunk count)
  std::atomic<int> gram_blocks[4][NL];
  // k_fold (64-bit limbs, vector ALUs): workgroup cap (by log2 of the elements)
  std::atomic<int> vfold_cap[NL];
  Tuning() {
    for (int k = 0; k < 4; k++)
      for (int i = 0; i < NL; i++) fold_mode[k][i] = fold_cap[k][i] = gram_blocks[k][i] = 0;
    for (int i = 0; i < NL; i++) vfold_cap[i] = 0;
  }
  static Tuning& get() {
    static Tuning t;
    return t;
  }
  static int kidx(int k) { return k == 8 ? 0 : k == 16 ? 1 : k == 32 ? 2 : 3; }
  static int lg(uint64_t n) { return n <= 1 ? 0 : std::min(NL - 1, 64 - __builtin_clzll(n - 1)); }
};

// Number of live batched-proof servers in the process.
inline std::atomic<int>& live_servers() {
  static std::atomic<int> n{0};
  return n;
}

// Turns on the device.  A single-proof call pipelines its sumcheck rounds
// (kernels wait for challenges the calling thread posts) only while no other
// call of the library runs: another thread's runtime calls could wait for
// the device while holding locks the posting thread needs.  Such a call
// holds device_turns() exclusively; every other call (proofs that do not
// pipeline, server jobs, verifiers) holds it shared, so they overlap each
// other but never a pipelining proof.
inline std::shared_mutex& device_turns() {
  static std::shared_mutex m;
  return m;
}
inline bool& solo_call() {
  static thread_local bool v = false;
  return v;
}
struct CallTurn {
  bool excl = false;
  // want_solo: the call would like to pipeline its rounds; it does when it
  // gets the device to itself right away (and no server exists)
  explicit CallTurn(bool want_solo) {
    if (want_solo && device_turns().try_lock()) {
      if (live_servers().load() == 0) {
        excl = true;
        solo_call() = true;
        return;
      }
      device_turns().unlock();
    }
    device_turns().lock_shared();
  }
  CallTurn(const CallTurn&) = delete;
  CallTurn& operator=(const CallTurn&) = delete;
  ~CallTurn() {
    if (excl) {
      solo_call() = false;
      device_turns().unlock();
    } else {
      device_turns().unlock_shared();
    }
  }
};

// Allocation size classes: powers of two up to 64 KiB, then four classes per
// octave (a block is at most 25 % larger than the request), so blocks freed
// by one proof serve later proofs of other shapes and widths.
inline size_t size_class(size_t bytes) {
  if (bytes <= 256) return 256;
  if (bytes <= (size_t(1) << 16)) {
    size_t c = 512;
    while (c < bytes) c <<= 1;
    return c;
  }
  const int e = 63 - __builtin_clzll(bytes);  // 2^e <= bytes < 2^(e+1)
  const size_t q = size_t(1) << (e - 2);
  return (bytes + q - 1) / q * q;
}

// A fixed device range handed out first-fit, for calls that must not
// allocate device memory: blocks are 256-byte aligned offsets into
// [base, base + size), free space is a sorted list of at most MAXB ranges in
// this object's own storage (no heap use), and neighbouring free ranges
// merge when a block comes back.
struct Arena {
  static constexpr int MAXB = 2048;
  struct Range {
    size_t off, len;
  };
  uint8_t* base = nullptr;
  size_t size = 0, used = 0, peak = 0;
  int nfree = 0;
  Range fr[MAXB];
  void reset(void* b, size_t s) {
    base = (uint8_t*)b;
    size = s;
    used = peak = 0;
    nfree = 1;
    fr[0] = {0, s};
  }
  static size_t round(size_t bytes) { return (std::max<size_t>(bytes, 1) + 255) & ~size_t(255); }
  bool owns(const void* p) const { return base && p >= base && p < base + size; }
  void* get(size_t bytes) {
    bytes = round(bytes);
    for (int i = 0; i < nfree; i++) {
      if (fr[i].len < bytes) continue;
      const size_t off = fr[i].off;
      fr[i].off += bytes;
      fr[i].len -= bytes;
      if (fr[i].len == 0) {
        for (int j = i; j + 1 < nfree; j++) fr[j] = fr[j + 1];
        nfree--;
      }
      used += bytes;
      peak = std::max(peak, used);
      return base + off;
    }
    throw Error(NENYA_E_OOM, "no-alloc scratch exhausted: " + std::to_string(bytes) + " bytes requested with " +
                                 std::to_string(size - used) + " of " + std::to_string(size) + " free");
  }
  void put(const void* p, size_t bytes) {
    bytes = round(bytes);
    const size_t off = (size_t)((const uint8_t*)p - base);
    int i = 0;
    while (i < nfree && fr[i].off < off) i++;
    const bool left = i > 0 && fr[i - 1].off + fr[i - 1].len == off;
    const bool right = i < nfree && off + bytes == fr[i].off;
    if (left && right) {
      fr[i - 1].len += bytes + fr[i].len;
      for (int j = i; j + 1 < nfree; j++) fr[j] = fr[j + 1];
      nfree--;
    } else if (left) {
      fr[i - 1].len += bytes;
    } else if (right) {
      fr[i].off = off;
      fr[i].len += bytes;
    } else {
      if (nfree == MAXB) throw Error(NENYA_E_DEVICE, "no-alloc scratch free list full");
      for (int j = nfree; j > i; j--) fr[j] = fr[j - 1];
      fr[i] = {off, bytes};
      nfree++;
    }
    used -= bytes;
  }
};

struct Pool {
  std::multimap<size_t, void*> free_;
  std::vector<std::pair<void*, size_t>> all_;
  size_t live = 0, peak = 0;
  std::mutex mu;
  // While deferring (work on several streams), returned blocks wait here
  // until the streams have joined: one stream may still read a block
  // another has let go of.
  int defer = 0;
  std::vector<std::pair<void*, size_t>> deferred;
  void defer_frees(bool on) {
    std::lock_guard<std::mutex> g(mu);
    if (on) {
      defer++;
      return;
    }
    if (defer == 0 || --defer > 0) return;
    for (auto& d : deferred) {
      free_.insert({d.second, d.first});
      live -= d.second;
    }
    deferred.clear();
  }
  // While set, every block comes from this arena (the no-alloc entry point);
  // blocks of the arena go back to it, others to the free lists.
  Arena* arena = nullptr;
  // Test switch of the server (NENYA_INJECT_OOM): when > 0, counts down per
  // request and the request that reaches 0 fails as if the device were full.
  long fail_in = 0;
  void* get(size_t bytes) {
    if (fail_in > 0 && --fail_in == 0) throw Error(NENYA_E_OOM, "injected device allocation failure");
    if (arena) {
      std::lock_guard<std::mutex> g(mu);
      return arena->get(bytes);
    }
    bytes = size_class(bytes);
    std::lock_guard<std::mutex> g(mu);
    auto it = free_.find(bytes);
    if (it != free_.end()) {
      void* p = it->second;
      free_.erase(it);
      live += bytes;
      peak = std::max(peak, live);
      if (pool_poison()) poison(p, bytes);
      if (pool_guard()) fill_guards(p, bytes);
      return p;
    }
    void* p = nullptr;
    if (pool_guard()) {
      void* base = nullptr;
      if (hipMalloc(&base, bytes + 2 * GUARD) != hipSuccess) throw Error(NENYA_E_OOM, "guarded allocation failed");
      p = (uint8_t*)base + GUARD;
      all_.push_back({base, bytes});
      live += bytes;
      peak = std::max(peak, live);
      fill_guards(p, bytes);
      return p;
    }
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      // release cached blocks and retry once
      for (auto& kv : free_) {
        (void)hipFree(kv.second);
        for (auto& a : all_)
          if (a.first == kv.second) a.first = nullptr;
      }
      free_.clear();
      (void)hipGetLastError();
      e = hipMalloc(&p, bytes);
      if (e != hipSuccess) throw Error(NENYA_E_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    all_.push_back({p, bytes});
    live += bytes;
    peak = std::max(peak, live);
    if (pool_poison()) poison(p, bytes);
    return p;
  }
  static void fill_guards(void* p, size_t bytes) {
    if (hipDeviceSynchronize() != hipSuccess || hipMemset((uint8_t*)p - GUARD, 0x5A, GUARD) != hipSuccess ||
        hipMemset((uint8_t*)p + bytes, 0x5A, GUARD) != hipSuccess || hipMemset(p, 0xA5, bytes) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
      throw Error(NENYA_E_DEVICE, "guard fill failed");
  }
  static void check_guards(void* p, size_t bytes) {
    std::vector<uint8_t> g(GUARD);
    if (hipDeviceSynchronize() != hipSuccess) throw Error(NENYA_E_DEVICE, "guard check sync failed");
    for (int side = 0; side < 2; side++) {
      const uint8_t* at = side ? (const uint8_t*)p + bytes : (const uint8_t*)p - GUARD;
      if (hipMemcpy(g.data(), at, GUARD, hipMemcpyDeviceToHost) != hipSuccess) throw Error(NENYA_E_DEVICE, "guard read failed");
      for (size_t i = 0; i < GUARD; i++)
        if (g[i] != 0x5A) {
          std::fprintf(stderr, "NENYA_GUARD: %s overrun of a %zu-byte block at guard offset %zu (byte 0x%02x)\n",
                       side ? "upper" : "lower", bytes, i, g[i]);
          std::fflush(stderr);
          break;
        }
    }
  }
  static void poison(void* p, size_t bytes) {
    // device-wide: orders after every stream's pending work on the block
    if (hipDeviceSynchronize() != hipSuccess || hipMemset(p, 0xA5, bytes) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
      throw Error(NENYA_E_DEVICE, "poison fill failed");
  }
  // restart peak tracking; returns the bytes live now
  size_t mark() {
    std::lock_guard<std::mutex> g(mu);
    peak = live;
    return live;
  }
  void put(void* p, size_t bytes) {
    if (!p) return;
    if (arena && arena->owns(p)) {
      std::lock_guard<std::mutex> g(mu);
      arena->put(p, bytes);
      return;
    }
    bytes = size_class(bytes);
    if (pool_guard()) check_guards(p, bytes);
    std::lock_guard<std::mutex> g(mu);
    if (defer) {
      deferred.push_back({p, bytes});
      return;
    }
    free_.insert({bytes, p});
    live -= bytes;
  }
  ~Pool() {
    for (auto& a : all_)
      if (a.first) (void)hipFree(a.first);
  }
};

template <class X> struct DBuf {
  Pool* pool = nullptr;
  X* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(Pool* pl, size_t cnt) : pool(pl), n(cnt) { p = (X*)pool->get(cnt * sizeof(X)); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : pool(o.pool), p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    release();
    pool = o.pool; p = o.p; n = o.n;
    o.p = nullptr; o.n = 0;
    return *this;
  }
  void release() {
    if (p) pool->put(p, n * sizeof(X));
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
  X* get() const { return p; }
};

// ---------------------------------------------------------------------------
// Per-width packing constants (reduction/src/constants.rs:198-302):
//   T1 (teichmuller.rs:65-76), l_beta_i(t) = 1 - t^15 (i = 0) or
//   15^-1 prod_{j != i} (t - t_j) evaluated at beta, and the 32 images
//   basis[2i] = l_beta_i (1 - alpha), basis[2i+1] = l_beta_i alpha
//   (= pack_r2(e_n), pcs.rs:49-55).  All coefficients lie in R1 and the
//   polynomials have degree <= 15 in beta, so evaluation at beta needs no
//   reduction mod f2: slot c of l_beta_i is the t^c coefficient.
// ---------------------------------------------------------------------------
template <class T> struct G4 {
  using C = typename T::C;
  C v[4];
  static G4 mul(const G4& a, const G4& b) {
    C p[7], o[4];
    g4_raw(a.v, b.v, p);
    g4_reduce(p, o);
    G4 r;
    for (int i = 0; i < 4; i++) r.v[i] = (typename T::L)o[i];
    return r;
  }
};

// Balanced base-256 digits of v mod 2^(8 nd): byte p is digit p as int8.
inline uint32_t bal_digits_host(uint32_t v, int nd) {
  if (nd == 1) return v & 0xffu;
  if (nd == 2) return ((v + 0x8080u) ^ 0x8080u) & 0xffffu;
  return (v + 0x80808080u) ^ 0x80808080u;
}

// int8 MFMA A-fragments (v_mfma_i32_32x32x32_i8) of a 128 x 32*kblocks matrix
// over Z/2^(8 nd) given by at(row, k): fragment (p, kb, rb) = digit plane p,
// k-block kb, row-block rb; lane (row%32) + 32h, byte b carries
// k = 32 kb + 8(b/4) + 4h + b%4.  Layout [p][kb][rb][lane][16 bytes].
template <class F> std::vector<uint32_t> mx_frags(int nd, int kblocks, F at) {
  std::vector<uint32_t> fr((size_t)nd * kblocks * 4 * 64 * 4, 0);
  for (int kb = 0; kb < kblocks; kb++)
    for (int rb = 0; rb < 4; rb++)
      for (int lane = 0; lane < 64; lane++)
        for (int wd = 0; wd < 4; wd++) {
          const int h = lane >> 5, row = 32 * rb + (lane & 31);
          uint32_t z[4];
          for (int c = 0; c < 4; c++) z[c] = bal_digits_host(at(row, 32 * kb + 8 * wd + 4 * h + c), nd);
          for (int p = 0; p < nd; p++) {
            uint32_t word = 0;
            for (int c = 0; c < 4; c++) word |= ((z[c] >> (8 * p)) & 0xffu) << (8 * c);
            fr[((((size_t)p * kblocks + kb) * 4 + rb) * 64 + lane) * 4 + wd] = word;
          }
        }
  return fr;
}

// Matrix of x -> r x over Z/2^k: M[row][k] = (r e_k)[row], e_{4i+c} = X^c Y^i,
// built from Y^i r (Y^32 = -Y^3 - Y - alpha) and X (X^4 = -X - 1).
template <class T> std::vector<typename T::C> mulmat_host(const Gr<T>& r) {
  using C = typename T::C;
  std::vector<C> M((size_t)E * E);
  C v[E], nv[E];
  for (int l = 0; l < E; l++) v[l] = r.v[l];
  for (int i = 0; i < 32; i++) {
    for (int t = 0; t < 32; t++) {
      C a[4] = {v[4 * t], v[4 * t + 1], v[4 * t + 2], v[4 * t + 3]};
      for (int c = 0; c < 4; c++) {
        for (int cp = 0; cp < 4; cp++) M[(size_t)(4 * t + cp) * E + 4 * i + c] = a[cp];
        C n0 = 0 - a[3], n1 = a[0] - a[3], n2 = a[1], n3 = a[2];
        a[0] = n0; a[1] = n1; a[2] = n2; a[3] = n3;
      }
    }
    // v <- Y v
    const C h0 = v[124], h1 = v[125], h2 = v[126], h3 = v[127];
    for (int l = E - 1; l >= 4; l--) nv[l] = v[l - 4];
    nv[0] = h3; nv[1] = h3 - h0; nv[2] = 0 - h1; nv[3] = 0 - h2;
    for (int c = 0; c < 4; c++) { nv[4 + c] -= v[124 + c]; nv[12 + c] -= v[124 + c]; }
    for (int l = 0; l < E; l++) v[l] = nv[l];
  }
  return M;
}

template <class T> struct Consts {
  using L = typename T::L;
  using C = typename T::C;
  G4<T> t1[16];
  Gr<T> lbeta[16];
  Gr<T> basis[32];
  std::vector<L> psi_pw;  // [16][32][4]: t_i^k
  std::vector<L> frob_t;  // [j][o]: matrix of sigma^-(k-1), transposed
  std::vector<uint32_t> b2frag;  // MFMA fragments of [b_s b_s'] (128 x 1024), k <= 32
  std::vector<uint32_t> lbfrag;  // MFMA fragments of x -> l_beta_i x for i < 16, k <= 32
  std::vector<uint32_t> packfrag;  // MFMA fragments of pack_r2 o rot_L (128 x 32) for L < 32, k <= 32
  std::vector<L> b2plain;          // [b_s b_s'] as 1024 elements [32 s + s'][limb], k = 64
  L inv15;

  Consts() {
    const int K = T::K;
    // inverse of 15 by 6 Newton doublings (limb.rs:91-101)
    L x = 1;
    for (int i = 0; i < 6; i++) x = (L)(x * (L)(2 - (L)(15 * x)));
    inv15 = x;
    // T1: iterate x -> x^16 K times from the natural lifts of alpha^(i-1)
    G4<T> alpha{{0, 1, 0, 0}}, pw{{1, 0, 0, 0}};
    t1[0] = G4<T>{{0, 0, 0, 0}};
    for (int i = 1; i < 16; i++) {
      G4<T> y = pw;
      for (int it = 0; it < K; it++)
        for (int s = 0; s < 4; s++) y = G4<T>::mul(y, y);
      t1[i] = y;
      pw = G4<T>::mul(pw, alpha);
    }
    // Lagrange basis polynomials in t with R1 coefficients
    for (int i = 0; i < 16; i++) {
      std::vector<G4<T>> poly(17, G4<T>{{0, 0, 0, 0}});
      if (i == 0) {
        poly[0].v[0] = 1;
        poly[15].v[0] = (L)(0 - 1);
      } else {
        poly[0].v[0] = 1;
        int deg = 0;
        for (int j = 0; j < 16; j++) {
          if (j == i) continue;
          // poly *= (t - t_j)
          std::vector<G4<T>> np(17, G4<T>{{0, 0, 0, 0}});
          for (int d = 0; d <= deg; d++) {
            for (int c = 0; c < 4; c++) np[d + 1].v[c] = (L)(np[d + 1].v[c] + poly[d].v[c]);
            G4<T> p = G4<T>::mul(poly[d], t1[j]);
            for (int c = 0; c < 4; c++) np[d].v[c] = (L)(np[d].v[c] - p.v[c]);
          }
          poly = np;
          deg++;
        }
        for (auto& g : poly)
          for (int c = 0; c < 4; c++) g.v[c] = (L)(g.v[c] * inv15);
      }
      Gr<T> e = Gr<T>::zero();
      for (int d = 0; d < 16; d++)
        for (int c = 0; c < 4; c++) e.v[4 * d + c] = (L)poly[d].v[c];
      lbeta[i] = e;
    }
    G4<T> la0{{1, (C)(L)(0 - 1), 0, 0}}, la1{{0, 1, 0, 0}};
    for (int i = 0; i < 16; i++)
      for (int h = 0; h < 2; h++) {
        Gr<T> b = Gr<T>::zero();
        for (int d = 0; d < 32; d++) {
          G4<T> s{{lbeta[i].v[4 * d], lbeta[i].v[4 * d + 1], lbeta[i].v[4 * d + 2], lbeta[i].v[4 * d + 3]}};
          G4<T> p = G4<T>::mul(s, h ? la1 : la0);
          for (int c = 0; c < 4; c++) b.v[4 * d + c] = (L)p.v[c];
        }
        basis[2 * i + h] = b;
      }
    psi_pw.resize(16 * 32 * 4);
    for (int i = 0; i < 16; i++) {
      G4<T> p{{1, 0, 0, 0}};
      for (int k = 0; k < 32; k++) {
        for (int c = 0; c < 4; c++) psi_pw[(i * 32 + k) * 4 + c] = (L)p.v[c];
        p = G4<T>::mul(p, t1[i]);
      }
    }
    frobenius_setup();
    if (K == 64) {
      // the products b_s b_s' for the vector-ALU C-matrix matvec
      b2plain.assign((size_t)1024 * E, 0);
      for (int a = 0; a < 32; a++)
        for (int b = a; b < 32; b++) {
          const Gr<T> p = basis[a] * basis[b];
          for (int l = 0; l < E; l++) b2plain[(size_t)(32 * a + b) * E + l] = b2plain[(size_t)(32 * b + a) * E + l] = p.v[l];
        }
    }
    if (K <= 32) {
      // products b_s b_s' of the packing basis, column k = 32 s + s'
      std::vector<Gr<T>> b2(1024);
      for (int a = 0; a < 32; a++)
        for (int b = a; b < 32; b++) b2[32 * a + b] = b2[32 * b + a] = basis[a] * basis[b];
      b2frag = mx_frags(K / 8, 32, [&](int row, int k) { return (uint32_t)b2[k].v[row]; });
      // pack_r2(rot_L(x)) = sum_s x[(s + L) % 32] b_s = P_L x, P_L[row][k] = b_{(k - L) % 32}[row]
      for (int L = 0; L < 32; L++) {
        std::vector<uint32_t> f = mx_frags(K / 8, 1, [&](int row, int k) { return (uint32_t)basis[(k - L) & 31].v[row]; });
        packfrag.insert(packfrag.end(), f.begin(), f.end());
      }
      {  // self-check of mulmat_host against the product on a fixed element
        Gr<T> x;
        uint64_t sd = 0xD1B54A32D192ED03ull;
        for (int l = 0; l < E; l++) {
          sd = sd * 6364136223846793005ull + 1442695040888963407ull;
          x.v[l] = (L)(sd >> 23);
        }
        std::vector<C> M = mulmat_host<T>(lbeta[3]);
        Gr<T> want = lbeta[3] * x, got;
        for (int o = 0; o < E; o++) {
          C acc = 0;
          for (int k2 = 0; k2 < E; k2++) acc += M[(size_t)o * E + k2] * (C)x.v[k2];
          got.v[o] = (L)acc;
        }
        if (!(got == want)) throw Error(NENYA_E_DEVICE, "multiplication matrix failed its self-check");
      }
      for (int i = 0; i < 16; i++) {
        std::vector<C> M = mulmat_host<T>(lbeta[i]);
        std::vector<uint32_t> f = mx_frags(K / 8, 4, [&](int row, int k) { return (uint32_t)M[(size_t)row * E + k]; });
        lbfrag.insert(lbfrag.end(), f.begin(), f.end());
      }
    }
  }
  static const Consts& get() {
    static Consts c;
    return c;
  }

  // Matrix of sigma^-n, n = k-1, where sigma is the Frobenius automorphism
  // (sigma(t) = t^2 on Teichmuller elements, sigma^128 = id).  The elements
  // b_j = Teich(X)^c Teich(Y)^i (limb j = 4i+c) are Teichmuller and congruent
  // to the standard basis mod 2, so B = [b_j] is invertible and
  // sigma^-n = P B^-1 with P = [sigma^-n b_j] = [b_j^(2^(128-n))].
  void frobenius_setup() {
    const int n = T::K - 1;
    using G = Gr<T>;
    auto lift128 = [](G g) {
      for (int s = 0; s < 128; s++) g = g * g;
      return g;
    };
    G tx = G::zero(), ty = G::zero();
    tx.v[1] = 1;  // alpha
    ty.v[4] = 1;  // beta
    tx = lift128(tx);
    ty = lift128(ty);
    G px = tx, py = ty;
    for (int s = 0; s < 128 - n; s++) { px = px * px; py = py * py; }
    std::vector<C> Bm(E * E), Pm(E * E);  // [o][j]
    G yi = G::one(), qi = G::one();
    for (int i = 0; i < 32; i++) {
      G b = yi, p = qi;
      for (int c = 0; c < 4; c++) {
        for (int o = 0; o < E; o++) {
          Bm[o * E + 4 * i + c] = b.v[o];
          Pm[o * E + 4 * i + c] = p.v[o];
        }
        b = b * tx;
        p = p * px;
      }
      yi = yi * ty;
      qi = qi * py;
    }
    auto matmul = [](const std::vector<C>& A, const std::vector<C>& Bq) {
      std::vector<C> R(E * E, 0);
      for (int r = 0; r < E; r++)
        for (int m = 0; m < E; m++) {
          C a = A[r * E + m];
          if (!a) continue;
          for (int c = 0; c < E; c++) R[r * E + c] += a * Bq[m * E + c];
        }
      return R;
    };
    // B^-1 by Newton's iteration X <- X (2I - B X) from X = I (B = I mod 2);
    // each step doubles the 2-adic precision.
    std::vector<C> X(E * E, 0);
    for (int i = 0; i < E; i++) X[i * E + i] = 1;
    for (int it = 0; (1 << it) < 2 * T::K; it++) {
      std::vector<C> BX = matmul(Bm, X);
      for (int i = 0; i < E * E; i++) BX[i] = (C)(0 - BX[i]);
      for (int i = 0; i < E; i++) BX[i * E + i] += 2;
      X = matmul(X, BX);
    }
    std::vector<C> M = matmul(Pm, X);
    frob_t.assign(E * E, 0);
    for (int o = 0; o < E; o++)
      for (int j = 0; j < E; j++) frob_t[j * E + o] = (L)M[o * E + j];
    // self-check against the plain definition on a fixed element
    G x;
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (int l = 0; l < E; l++) {
      s = s * 6364136223846793005ull + 1442695040888963407ull;
      x.v[l] = (L)(s >> 17);
    }
    G z = G::zero();
    for (int o = 0; o < E; o++) {
      C acc = 0;
      for (int j = 0; j < E; j++) acc += (C)frob_t[j * E + o] * (C)x.v[j];
      z.v[o] = (L)acc;
    }
    for (int q = 0; q < n; q++) z = z * z;
    if (!(z == lift128(x))) throw Error(NENYA_E_DEVICE, "Frobenius lift matrix failed its self-check");
  }
};

// ---------------------------------------------------------------------------
// PCS parameters (pcs.rs:45-293; r2_pcs.rs:41-62)
// ---------------------------------------------------------------------------
inline size_t recommended_query_size(double delta, size_t lambda) {
  double lam = (double)lambda, bt = 1.0 - delta / 3.0, bo = 1.0 - 2.0 * delta / 3.0;
  size_t t = (size_t)std::ceil(lam / (-std::log2(bt)));
  double target = std::exp2(-lam);
  for (;;) {
    double pt = 1.0, po = 1.0;  // f64::powi: repeated products
    for (size_t i = 0; i < t; i++) { pt *= bt; po *= bo; }
    if (pt + 2.0 * po > target) t++; else break;
  }
  return t;
}
inline size_t proof_size_bytes_for(size_t k1, size_t k2, size_t n, size_t t, size_t mb, size_t cb) {
  size_t np = 1;
  while (np < n) np <<= 1;
  size_t pl = (size_t)__builtin_ctzll(np);
  return 32 + k2 * mb + t * (8 + k1 * cb) + t * (8 + pl * 32) + t * 4;
}
inline bool best_shape(size_t nv, size_t q, size_t t, size_t mb, size_t cb, bool floor_t, size_t& k1, size_t& k2) {
  bool found = false;
  size_t best = 0;
  for (size_t l = 1; l < nv; l++) {
    size_t a = size_t(1) << l, b = size_t(1) << (nv - l);
    if (floor_t && q * b < t) continue;
    size_t s = proof_size_bytes_for(a, b, q * b, t, mb, cb);
    if (!found || s < best) { best = s; k1 = a; k2 = b; found = true; }
  }
  return found;
}

struct PcsParams {
  size_t k1 = 0, k2 = 0, q = 4, n = 0, ell = 0;
  bool packed = true;
  std::vector<uint32_t> perm1, perm2;
  std::vector<uint64_t> d1, d2;

  // RaaParams::setup (raa.rs:44-64): two Fisher-Yates permutations then 2n odd units
  void raa_setup(uint64_t seed, int K) {
    n = k2 * q;
    StdRng r = StdRng::seed_from_u64(seed);
    for (auto* p : {&perm1, &perm2}) {
      p->resize(n);
      for (size_t i = 0; i < n; i++) (*p)[i] = (uint32_t)i;
      for (size_t i = n - 1; i >= 1; i--) {
        size_t j = (size_t)r.gen_range_incl(0, i);
        std::swap((*p)[i], (*p)[j]);
      }
    }
    uint64_t mask = K == 64 ? ~0ull : ((1ull << K) - 1);
    d1.resize(n);
    d2.resize(n);
    for (size_t i = 0; i < n; i++) d1[i] = (r.next_u64() | 1) & mask;
    for (size_t i = 0; i < n; i++) d2[i] = (r.next_u64() | 1) & mask;
  }
  static PcsParams make(size_t k1, size_t k2, size_t q, size_t ell, uint64_t seed, bool packed, int K) {
    PcsParams p;
    p.k1 = k1; p.k2 = k2; p.q = q; p.ell = ell; p.packed = packed;
    p.raa_setup(seed, K);
    return p;
  }
  // Pcs::setup / setup_batched / setup_stacked and R2Pcs::setup
  static PcsParams setup(size_t nv, uint64_t seed, int K) {
    size_t lb = K / 8, t = recommended_query_size(0.19, 100), k1 = 0, k2 = 0;
    best_shape(nv, 4, t, 128 * lb, 32 * lb, false, k1, k2);
    return make(k1, k2, 4, t, seed, true, K);
  }
  static PcsParams setup_batched(size_t nv, uint64_t seed, size_t nopen, int K) {
    size_t lb = K / 8, t = recommended_query_size(0.19, 100), k1 = 0, k2 = 0;
    if (!best_shape(nv, 4, t, nopen * 128 * lb, 32 * lb, true, k1, k2)) best_shape(nv, 4, t, 128 * lb, 32 * lb, false, k1, k2);
    return make(k1, k2, 4, t, seed, true, K);
  }
  static PcsParams setup_stacked(size_t nv, uint64_t seed, size_t stack, int K) {
    size_t lb = K / 8, t = recommended_query_size(0.19, 100), k1 = 0, k2 = 0;
    if (!best_shape(nv, 4, t, 128 * lb, stack * 32 * lb, true, k1, k2))
      best_shape(nv, 4, t, 128 * lb, stack * 32 * lb, false, k1, k2);
    return make(k1, k2, 4, t, seed, true, K);
  }
  static PcsParams setup_r2(size_t nv, uint64_t seed, int K) {
    size_t lb = K / 8, t = recommended_query_size(0.19, 100), k1 = 0, k2 = 0;
    best_shape(nv, 4, t, 128 * lb, 128 * lb, false, k1, k2);
    return make(k1, k2, 4, t, seed, false, K);
  }
};

// ---------------------------------------------------------------------------
// Fiat-Shamir transcript (snark-ring/src/transcript.rs:23-105)
// ---------------------------------------------------------------------------
struct Transcript {
  b3::Hasher h;
  explicit Transcript(const char* label) {
    h.update("snark-ring/v0", 13);
    u32(std::strlen(label));
    h.update(label, std::strlen(label));
  }
  void u32(uint32_t x) {
    uint8_t b[4] = {(uint8_t)x, (uint8_t)(x >> 8), (uint8_t)(x >> 16), (uint8_t)(x >> 24)};
    h.update(b, 4);
  }
  void label(const char* l) {
    h.update("L", 1);
    u32(std::strlen(l));
    h.update(l, std::strlen(l));
  }
  void root(const uint8_t r[32]) {
    h.update("R", 1);
    h.update(r, 32);
  }
  template <class T> void elems(const char* l, const Gr<T>* xs, size_t n) {
    label(l);
    u32((uint32_t)n);
    for (size_t i = 0; i < n; i++) {
      h.update("E", 1);
      h.update(xs[i].v, sizeof xs[i].v);
    }
  }
  // challenge_sampler: squeeze 8 bytes of XOF from a clone, absorb "S"||bytes
  StdRng sampler(const char* l) {
    label(l);
    uint8_t b[8];
    h.finalize(0, b, 8);
    uint64_t seed = 0;
    for (int i = 0; i < 8; i++) seed |= (uint64_t)b[i] << (8 * i);
    h.update("S", 1);
    h.update(b, 8);
    return StdRng::seed_from_u64(seed);
  }
};

// XOF query indices (pcs.rs:1130-1156; r2_pcs.rs:220-246): BLAKE3(domain || root
// || m-bar limbs), 4-byte LE words masked to next_pow2(n)-1, kept if < n.
// `root_node` are the inputs of the message's root compression (cv, block,
// blen, flags), computed on the device; the XOF runs that node with counters.
inline std::vector<uint32_t> queries_from_root(const uint32_t root_node[26], size_t n, size_t t) {
  size_t np = 1;
  while (np < n) np <<= 1;
  uint32_t mask = (uint32_t)(np - 1);
  std::vector<uint32_t> out;
  out.reserve(t);
  for (uint64_t blk = 0; out.size() < t; blk++) {
    uint32_t o[16];
    b3::compress(root_node, root_node + 8, blk, root_node[24], root_node[25] | b3::ROOT, o);
    for (int i = 0; i < 16 && out.size() < t; i++) {
      uint32_t v = o[i] & mask;
      if (v < (uint32_t)n) out.push_back(v);
    }
  }
  return out;
}

inline std::vector<uint32_t> derive_queries(const char* domain, const uint8_t root[32], const void* mbar_bytes,
                                            size_t nbytes, size_t n, size_t t) {
  b3::Hasher h;
  h.update(domain, std::strlen(domain));
  h.update(root, 32);
  h.update(mbar_bytes, nbytes);
  size_t np = 1;
  while (np < n) np <<= 1;
  uint32_t mask = (uint32_t)(np - 1);
  std::vector<uint32_t> out;
  out.reserve(t);
  uint64_t pos = 0;
  while (out.size() < t) {
    uint8_t buf[256];
    h.finalize(pos, buf, sizeof buf);
    pos += sizeof buf;
    for (size_t o = 0; o + 4 <= sizeof buf && out.size() < t; o += 4) {
      uint32_t v = ((uint32_t)buf[o] | ((uint32_t)buf[o + 1] << 8) | ((uint32_t)buf[o + 2] << 16) |
                    ((uint32_t)buf[o + 3] << 24)) & mask;
      if (v < (uint32_t)n) out.push_back(v);
    }
  }
  return out;
}

}  // namespace nenya
