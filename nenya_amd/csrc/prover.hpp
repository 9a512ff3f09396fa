// Prover orchestration: device-resident tables, host-side Fiat-Shamir,
// proof serialization in the reference's byte order.
//
// Reference entry points mirrored: prove_simd (snark-ring/src/simd.rs:212-373)
// and prove_general (snark-ring/src/general.rs:367-537).  Every value is exact
// arithmetic mod 2^k, so the algebraically equal device schedules below emit
// the reference's bytes (see the identities cited at each step).
#pragma once
#include <chrono>
#include <exception>
#include <initializer_list>
#include <map>
#include <thread>

#include "fsdev.hpp"
#include "host.hpp"
#include "kernels.hpp"
#include "mfma.hpp"

namespace nenya {

inline unsigned grid_for(uint64_t items, uint64_t per_block, unsigned cap = 16384) {
  uint64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

// Stage timer (GeneralProveTiming / SimdProveTiming analogue); synchronizes the
// stream at stage boundaries only when enabled.
// Stage times of a proof from HIP events on the prover stream: mark() records
// where the stream has got to, without a synchronization, so the stages keep
// the proof's overlap (work on other streams shows only where the prover
// stream waits for it, and host delays where the stream runs dry);
// collect() turns the marks into milliseconds once the proof is done.
struct Stages {
  bool on = false;
  hipStream_t st = nullptr;
  std::vector<hipEvent_t> ev;
  std::vector<std::string> names;
  void rec(const char* name) {
    hipEvent_t e = nullptr;
    HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(e, st));
    ev.push_back(e);
    names.push_back(name);
  }
  void start() {
    if (on) rec("");
  }
  void mark(const char* name) {
    if (on) rec(name);
  }
  std::vector<std::pair<std::string, double>> collect() {
    std::vector<std::pair<std::string, double>> t;
    if (!on || ev.empty()) return t;
    HIP_CHECK(hipEventSynchronize(ev.back()));
    for (size_t i = 1; i < ev.size(); i++) {
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, ev[i - 1], ev[i]));
      t.push_back({names[i], (double)ms});
    }
    return t;
  }
  ~Stages() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

// Per-kernel HIP-event timing with the algorithmic bytes and limb-MACs of
// each launch (for roofline figures); off unless enabled on the params.
struct KProf {
  bool on = false;
  std::string only;  // when non-empty, only launches of this kernel are timed
  struct Rec { std::string name; hipEvent_t a, b; double bytes, macs; };
  struct Stat { double ms = 0, bytes = 0, macs = 0; long count = 0; };
  std::vector<Rec> recs;
  std::map<std::string, Stat> stats;
  void collect() {
    for (auto& r : recs) {
      (void)hipEventSynchronize(r.b);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, r.a, r.b);
      Stat& s = stats[r.name];
      s.ms += ms;
      s.bytes += r.bytes;
      s.macs += r.macs;
      s.count++;
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    recs.clear();
  }
};

// Set on a thread while it runs a no-alloc call: a pinned buffer that would
// have to grow fails the call instead of allocating.
inline bool& pinned_no_grow() {
  static thread_local bool v = false;
  return v;
}

// Pinned host buffer the proof is assembled in (grow-only, owned by params).
struct Pinned {
  uint8_t* p = nullptr;
  size_t cap = 0;
  unsigned flags = hipHostMallocDefault;  // hipHostMallocCoherent for kernel-shared areas
  Pinned() = default;
  Pinned(const Pinned&) = delete;
  Pinned& operator=(const Pinned&) = delete;
  Pinned(Pinned&& o) noexcept : p(o.p), cap(o.cap), flags(o.flags) {
    o.p = nullptr;
    o.cap = 0;
  }
  Pinned& operator=(Pinned&& o) noexcept {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
    std::swap(flags, o.flags);
    return *this;
  }
  uint8_t* get(size_t n) {
    if (n > cap) {
      if (pinned_no_grow())
        throw Error(NENYA_E_OOM, "a pinned staging buffer of " + std::to_string(cap) + " bytes would have to grow to " +
                                     std::to_string(n) + " in a no-alloc call");
      if (p) (void)hipHostFree(p);
      p = nullptr;
      cap = 0;
      const size_t c = size_class(n);  // headroom for a slightly larger next proof
      HIP_CHECK(hipHostMalloc((void**)&p, c, flags));
      cap = c;
    }
    return p;
  }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

// memcpy of a large host block on up to 8 threads (one core's copy rate is
// well under the host link's).
inline void par_memcpy(uint8_t* dst, const uint8_t* src, size_t n) {
  const size_t PART = size_t(1) << 20;
  const int nt = (int)std::min<size_t>(8, n / PART);
  if (nt < 2) {
    std::memcpy(dst, src, n);
    return;
  }
  const size_t per = (n / nt + 63) & ~size_t(63);
  std::vector<std::thread> th;
  size_t done = std::min(n, per);  // parts [0, done) copied here, the rest by threads
  try {
    for (int i = 1; i < nt; i++) {
      const size_t a = std::min(n, i * per), e = std::min(n, a + per);
      if (e > a) th.emplace_back([=] { std::memcpy(dst + a, src + a, e - a); });
    }
  } catch (...) {  // a thread that could not start: copy its part and the rest here
    const size_t from = std::min(n, (th.size() + 1) * per);
    std::memcpy(dst + from, src + from, n - from);
  }
  std::memcpy(dst, src, done);
  for (auto& t : th) t.join();
}

// Workgroups of the copy kernel that moves proof sections to pinned host
// memory; NENYA_DL_BLOCKS overrides the default.
inline unsigned dl_blocks() {
  static const unsigned v = [] {
    const char* e = std::getenv("NENYA_DL_BLOCKS");
    return e ? (unsigned)std::max(1, std::atoi(e)) : 32u;
  }();
  return v;
}

// Write cursor over a pre-sized section of the proof buffer.
struct Sink {
  uint8_t* p;
  uint8_t* end;
  Sink(uint8_t* b, size_t n) : p(b), end(b + n) {}
  uint8_t* take(size_t n) {
    if ((size_t)(end - p) < n) throw Error(NENYA_E_DEVICE, "proof section overflow");
    uint8_t* r = p;
    p += n;
    return r;
  }
  void u64(uint64_t x) {
    uint8_t* d = take(8);
    for (int i = 0; i < 8; i++) d[i] = (uint8_t)(x >> (8 * i));
  }
  void u32(uint32_t x) {
    uint8_t* d = take(4);
    for (int i = 0; i < 4; i++) d[i] = (uint8_t)(x >> (8 * i));
  }
  void bytes(const void* s, size_t n) { std::memcpy(take(n), s, n); }
  template <class G> void gr(const G& g) { bytes(g.v, sizeof g.v); }
  bool full() const { return p == end; }
};

// Serialized size of an opening (OpeningProof / R2OpeningProof, pcs.rs:357-406,
// r2_pcs.rs:109-160) or a rotated batch of nl m-bars (pcs.rs:426-471).
inline size_t opening_bytes(const PcsParams& P, size_t rows, size_t limb_bytes, size_t nl) {
  size_t W = P.packed ? 32 : 128, eb = 128 * limb_bytes, plen = (size_t)__builtin_ctzll(P.n);
  size_t mb = nl == 0 ? 8 + P.k2 * eb : 8 + nl * (8 + P.k2 * eb);
  return mb + 8 + 4 * P.ell + 8 + P.ell * (8 + rows * W * limb_bytes) + 8 + P.ell * (8 + plen * 32);
}

template <class T> struct PcsDev {
  using L = typename T::L;
  PcsParams P;
  DBuf<uint32_t> perm1, perm2;
  DBuf<L> d1, d2;
  template <class Eng> void upload(Eng& e) {
    perm1 = DBuf<uint32_t>(e.pool, P.n);
    perm2 = DBuf<uint32_t>(e.pool, P.n);
    d1 = DBuf<L>(e.pool, P.n);
    d2 = DBuf<L>(e.pool, P.n);
    std::vector<L> a(P.n), b(P.n);
    for (size_t i = 0; i < P.n; i++) { a[i] = (L)P.d1[i]; b[i] = (L)P.d2[i]; }
    e.h2d(perm1.p, P.perm1.data(), P.n * 4);
    e.h2d(perm2.p, P.perm2.data(), P.n * 4);
    e.h2d(d1.p, a.data(), P.n * sizeof(L));
    e.h2d(d2.p, b.data(), P.n * sizeof(L));
    e.sync();
  }
};

template <class T> struct CommitDev {
  using L = typename T::L;
  uint8_t root[32];
  const uint8_t* root_stage = nullptr;  // pinned copy in flight (Engine::resolve_roots fills root)
  size_t rows = 0;
  DBuf<L> enc;          // rows x n x W
  DBuf<uint32_t> nodes;  // (2n-1) x 8 words, level order
  const L* msg = nullptr;  // rows x k2 lifted message (borrowed)
};

struct OpenOut {
  std::vector<uint8_t> bytes;
};

// The round hub of a server (hub.hpp): served contexts hand a sumcheck's
// small tail rounds to it.
struct Hub;
template <class T> struct HubTask;
template <class T> void hub_run(Hub* h, HubTask<T>& t, hipStream_t st);
constexpr uint64_t HUB_FOLD_MAX = 4096;  // rounds whose folds touch fewer elements go to the hub
// Rounds whose folds touch fewer elements (latency-bound) run as one tail
// kernel (k_tail_mult / k_tail_lin: fold, marginalization, the next round's
// sums and its post) after the challenge's lift.
constexpr uint64_t TAIL_FOLD_MAX = 4096;
// Block caps of the tail kernels (one block per pair, or per (table, pair)):
// past these, the matrix-core folds and Gram sums of the ordinary rounds are
// faster (measured at 2^20 constraints: a 512-block linear tail 41 us
// against 22 us for k_fold_mx + k_gram_lin + k_sum_post).
constexpr uint64_t TAIL_MULT_BLOCKS = 256, TAIL_LIN_BLOCKS = 64;

template <class T> struct Engine {
  using L = typename T::L;
  using C = typename T::C;
  using G = Gr<T>;
  Pool* pool;
  hipStream_t st;
  const Consts<T>& K;
  DBuf<L> d_basis, d_pw;
  DBuf<L> d_one;
  DBuf<L> d_frob;  // sigma^-(k-1), transposed [j][o]
  DBuf<uint32_t> d_b2frag;  // MFMA fragments of the packing-basis products
  DBuf<uint32_t> d_lbfrag;  // MFMA fragments of x -> l_beta_i x
  DBuf<uint32_t> d_packfrag;  // MFMA fragments of pack_r2 o rot_L
  DBuf<L> d_b2;               // basis products b_s b_s' as limbs (k = 64)
  KProf* prof = nullptr;

  // Host-device mailbox for the sumcheck rounds (coherent pinned memory):
  // word 0 aborts waiting kernels, word 1 is set by a wait that ran out of
  // time (the host then fails the proof), then a flag word per slot and direction,
  // then per slot the round polynomial (3 elements, device -> host) and the
  // natural challenge sample (1 element, host -> device).  Flags take
  // strictly increasing sequence values, so slots need no reset.
  static constexpr int MB_SLOTS = 64;
  Pinned mbox;
  uint8_t* mb_h = nullptr;  // host view
  uint8_t* mb_d = nullptr;  // device view
  uint32_t mseq = 0;
  uint64_t clock_khz = 0;  // device wall clock
  // Device-side limit of a wait for a posted challenge: 10 s, or
  // NENYA_WAIT_LIMIT_US (a debug switch that lets a test force the time-out
  // path from the host).
  uint64_t spin_ticks() const {
    uint64_t us = 10000000;
    if (const char* e = std::getenv("NENYA_WAIT_LIMIT_US")) us = std::strtoull(e, nullptr, 10);
    return us * clock_khz / 1000;
  }
  uint64_t wait_ticks = 0;  // spin_ticks() for the running sumcheck
  // Pipelined rounds (a kernel waits for the host's challenge) only where
  // one thread drives the device: allocation and free calls of other threads
  // can wait for the device while holding runtime locks this thread's
  // launches need.  So only a call that holds the device to itself
  // (CallTurn, solo_call()) pipelines; server contexts, calls that overlap
  // others and the pool's synchronizing debug modes post each challenge
  // before the lift that reads it is enqueued, so no kernel waits on a host
  // thread.
  bool pipelined = true;
  // Served jobs: a cancel request this engine checks between phases
  // (checkpoint()); the job's proof then fails with NENYA_E_CANCELED.
  const std::atomic<bool>* cancel = nullptr;
  void checkpoint() const {
    if (cancel && cancel->load(std::memory_order_relaxed)) throw Error(NENYA_E_CANCELED, "job canceled");
  }
  // Contexts that do not pipeline may draw their round challenges on the
  // device (fsdev.hpp): no host round trip per round, at the price of the
  // transcript's BLAKE3 work in the lift kernel (servers created with
  // NENYA_DEVICE_FS=1).
  bool devfs = false;
  // Served contexts of a server with a round hub: the sumchecks' rounds use
  // the device transcript, and once a round's folds touch fewer than
  // HUB_FOLD_MAX elements the remaining rounds run on the hub, fused with
  // other proofs' rounds (hev: the hand-over's ready / done events).
  Hub* hub = nullptr;
  hipEvent_t hev[2] = {};
  void hub_events() {
    if (!hev[0]) {
      HIP_CHECK(hipEventCreateWithFlags(&hev[0], hipEventDisableTiming));
      HIP_CHECK(hipEventCreateWithFlags(&hev[1], hipEventDisableTiming));
    }
  }
  // No-alloc calls: one stream only (no forks, downloads in order on the
  // prover stream), so no stream or event is created during the call.
  bool serial = false;
  bool pipeline_rounds() const { return pipelined && solo_call() && !pool_poison() && !pool_guard(); }
  DBuf<uint32_t> d_count;  // last-block counter of k_sum_post
  DBuf<L> d_sums;          // reduced round sums (3 elements)

  // Proof sections leave the device on a second stream (default priority,
  // created on first use) while the prover's later stages run; their device
  // buffers are held until finish_downloads(), which every proof calls
  // before it returns (it also waits for the prover stream, whose kernels and
  // copies write parts of the proof too) and again before it starts (after a
  // failed proof).
  hipStream_t cst = nullptr;
  hipEvent_t cev = nullptr;
  std::vector<std::pair<void*, size_t>> held;

  // Launch on the engine's stream, bracketed by events when profiling.
  template <class Kern, class... Args>
  void launch(const char* name, double bytes, double macs, Kern kern, dim3 grid, dim3 block, Args... args) {
    hipEvent_t a = nullptr, b = nullptr;
    const bool timed = prof && prof->on && (prof->only.empty() || prof->only == name);
    if (timed) {
      HIP_CHECK(hipEventCreate(&a));
      HIP_CHECK(hipEventCreate(&b));
      HIP_CHECK(hipEventRecord(a, st));
    }
    hipLaunchKernelGGL(kern, grid, block, 0, st, args...);
    HIP_CHECK(hipGetLastError());
    if (timed) {
      HIP_CHECK(hipEventRecord(b, st));
      prof->recs.push_back({name, a, b, bytes, macs});
    }
  }
  static constexpr double EB = E * sizeof(typename T::L);  // bytes per element
  static constexpr double PM = 512.0;                      // limb-MACs per beta term of a group product

  Engine(Pool* p, hipStream_t s) : pool(p), st(s), K(Consts<T>::get()) {
    d_basis = DBuf<L>(pool, 32 * E);
    d_pw = DBuf<L>(pool, K.psi_pw.size());
    d_one = DBuf<L>(pool, E);
    d_frob = DBuf<L>(pool, E * E);
    std::vector<L> b(32 * E);
    for (int i = 0; i < 32; i++) std::memcpy(&b[i * E], K.basis[i].v, E * sizeof(L));
    G one = G::one();
    upbuf.flags = resbuf.flags = mbox.flags = hipHostMallocCoherent;
    mb_h = mbox.get(mb_bytes());
    std::memset(mb_h, 0, mb_bytes());
    HIP_CHECK(hipHostGetDevicePointer((void**)&mb_d, mb_h, 0));
    {
      int dev = 0, khz = 0;
      HIP_CHECK(hipGetDevice(&dev));
      HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
      clock_khz = khz > 0 ? (uint64_t)khz : 100000;
    }
    d_count = DBuf<uint32_t>(pool, 1);
    d_sums = DBuf<L>(pool, 3 * E);
    HIP_CHECK(hipMemsetAsync(d_count.p, 0, 4, st));
    h2d(d_basis.p, b.data(), b.size() * sizeof(L));
    h2d(d_pw.p, K.psi_pw.data(), K.psi_pw.size() * sizeof(L));
    h2d(d_one.p, one.v, E * sizeof(L));
    h2d(d_frob.p, K.frob_t.data(), E * E * sizeof(L));
    if (!K.b2plain.empty()) {
      d_b2 = DBuf<L>(pool, K.b2plain.size());
      h2d(d_b2.p, K.b2plain.data(), K.b2plain.size() * sizeof(L));
    }
    if (!K.b2frag.empty()) {
      d_b2frag = DBuf<uint32_t>(pool, K.b2frag.size());
      h2d(d_b2frag.p, K.b2frag.data(), K.b2frag.size() * 4);
      d_packfrag = DBuf<uint32_t>(pool, K.packfrag.size());
      h2d(d_packfrag.p, K.packfrag.data(), K.packfrag.size() * 4);
      d_lbfrag = DBuf<uint32_t>(pool, K.lbfrag.size());
      h2d(d_lbfrag.p, K.lbfrag.data(), K.lbfrag.size() * 4);
    }
    if constexpr (kMx64<T>) {
      // x -> l_beta_i x fragments (8 digit planes each), made on the device
      DBuf<L> lb(pool, 16 * E);
      for (int i = 0; i < 16; i++) h2d(lb.p + (size_t)i * E, K.lbeta[i].v, E * sizeof(L));
      d_lbfrag = DBuf<uint32_t>(pool, (size_t)16 * 8 * 4096);
      hipLaunchKernelGGL(k_mulmat64, dim3(16, 4), dim3(256), 0, st, (const uint64_t*)lb.p, d_lbfrag.p);
      HIP_CHECK(hipGetLastError());
      // the basis products' fragments (8 digit planes, 1 MiB) for k_cgemm64
      d_b2frag = DBuf<uint32_t>(pool, (size_t)8 * 32 * 4 * 64 * 4);
      hipLaunchKernelGGL(k_b2frag64, dim3(32 * 4 * 64 * 4 / 256), dim3(256), 0, st, (const uint64_t*)d_b2.p,
                         d_b2frag.p);
      HIP_CHECK(hipGetLastError());
      // pack_r2 o rot_L fragments (8 digit planes, 32 rotations, 1 MiB) for k_pack_mx64
      d_packfrag = DBuf<uint32_t>(pool, (size_t)32 * 8 * 4 * 64 * 4);
      hipLaunchKernelGGL(k_packfrag64, dim3(32 * 4 * 64 * 4 / 256), dim3(256), 0, st, (const uint64_t*)d_basis.p,
                         d_packfrag.p);
      HIP_CHECK(hipGetLastError());
      sync();
    }
    sync();
  }

  void check() { HIP_CHECK(hipGetLastError()); }
  void sync() {
    // the sumcheck rounds wait here for a few hundred bytes about 30 times
    // per proof; other contexts keep the device busy meanwhile
    HIP_CHECK(hipStreamSynchronize(st));
  }

  // Host-device transfers never touch pageable memory (pageable copies stage
  // through buffers shared by the whole process): small uploads go through a
  // coherent pinned ring, larger ones and all downloads through a pinned
  // bounce buffer.  A ring slot is recycled only after a synchronization
  // that follows its use (the ring rotates; it is not reset per round).
  static constexpr size_t UPCAP = 1 << 16;
  Pinned upbuf, xfer, xfer2;
  hipEvent_t xev[2] = {};
  size_t up_off = 0;
  uint8_t* ring_slot(size_t bytes) {
    if (up_off + bytes > UPCAP) {
      // every stream that may still copy from the ring
      sync();
      if (forked) {
        HIP_CHECK(hipStreamSynchronize(st_main));
        for (int i = 0; i < forked; i++) HIP_CHECK(hipStreamSynchronize(sst[i]));
      }
      if (bg_live) {
        HIP_CHECK(hipStreamSynchronize(st_fg));
        HIP_CHECK(hipStreamSynchronize(sbg));
      }
      up_off = 0;
    }
    uint8_t* slot = upbuf.get(UPCAP) + up_off;
    up_off += (bytes + 255) & ~size_t(255);
    return slot;
  }
  // Copies src into the pinned ring and returns its device-side address, for
  // kernels that read it directly.
  const void* stage_host(const void* src, size_t bytes) {
    if (bytes > UPCAP) throw Error(NENYA_E_DEVICE, "staging request too large");
    uint8_t* slot = ring_slot(bytes);
    std::memcpy(slot, src, bytes);
    uint8_t* base = upbuf.p;
    if (base != up_hbase) {
      HIP_CHECK(hipHostGetDevicePointer((void**)&up_dbase, base, 0));
      up_hbase = base;
    }
    return up_dbase + (slot - base);
  }
  uint8_t *up_hbase = nullptr, *up_dbase = nullptr;  // the ring's host and device addresses
  // Commitment roots in flight to the host (see commit / resolve_roots).
  static constexpr int ROOT_SLOTS = 16;
  Pinned rootbuf;
  int root_used = 0;
  // Pinned area (up to 64 elements) that small result kernels write directly;
  // the host reads it after sync().
  Pinned resbuf;
  L* res_host() { return (L*)resbuf.get(64 * E * sizeof(L)); }
  L* res_dev() {
    void* dptr = nullptr;
    HIP_CHECK(hipHostGetDevicePointer(&dptr, res_host(), 0));
    return (L*)dptr;
  }
  void put_async(void* dst, const void* src, size_t bytes) { h2d(dst, src, bytes); }
  // Sizes every pinned staging area a device-witness proof uses (small
  // uploads, downloads of up to 1 MiB, commitment roots, results, picks), so
  // a no-alloc call finds them allocated.
  void reserve_staging() {
    upbuf.get(UPCAP);
    xfer.get(size_t(1) << 20);
    rootbuf.get(ROOT_SLOTS * 32);
    res_host();
    pickbuf.get(128 * E * sizeof(L));
  }

  // Upload (stream-ordered; the host buffer may be reused on return).
  void h2d(void* d, const void* h, size_t bytes) {
    if (!bytes) return;
    if (bytes <= UPCAP / 4) {
      // the copy kernel reads the ring slot through its device address
      const void* src = stage_host(h, bytes);
      launch("k_copy_small", 2.0 * bytes, 0, k_copy_small, dim3((unsigned)std::min<size_t>((bytes + 4095) / 4096, 4)),
             dim3(256), (const uint8_t*)src, (uint8_t*)d, (uint64_t)bytes);
      return;
    }
    // bounce in chunks through two pinned buffers, so the host fills one
    // while the other's copy runs (a buffer is refilled once its last copy
    // has ended)
    const size_t CH = size_t(8) << 20;
    uint8_t* bb[2] = {xfer.get(std::min(bytes, CH)), bytes > CH ? xfer2.get(CH) : nullptr};
    if (bytes > CH && !xev[0]) {
      HIP_CHECK(hipEventCreateWithFlags(&xev[0], hipEventDisableTiming));
      HIP_CHECK(hipEventCreateWithFlags(&xev[1], hipEventDisableTiming));
    }
    int k = 0;
    for (size_t off = 0; off < bytes; off += CH, k ^= 1) {
      const size_t c = std::min(CH, bytes - off);
      if (off >= 2 * CH) HIP_CHECK(hipEventSynchronize(xev[k]));
      par_memcpy(bb[k], (const uint8_t*)h + off, c);
      HIP_CHECK(hipMemcpyAsync((uint8_t*)d + off, bb[k], c, hipMemcpyHostToDevice, st));
      if (bytes > CH) HIP_CHECK(hipEventRecord(xev[k], st));
    }
    sync();
  }
  // Several uploads with one synchronization: small ones share the pinned
  // bounce buffer (a synchronization only when it fills), larger ones take
  // the chunked path.
  struct Up {
    void* d;
    const void* h;
    size_t bytes;
  };
  void h2d_many(const std::vector<Up>& ups) {
    const size_t CH = size_t(8) << 20;
    uint8_t* b = xfer.get(CH);
    size_t off = 0;
    for (const Up& u : ups) {
      if (!u.bytes) continue;
      const size_t sz = (u.bytes + 255) & ~size_t(255);
      if (sz > CH / 2) {
        if (off) sync();
        off = 0;
        h2d(u.d, u.h, u.bytes);  // synchronizes at its end
        continue;
      }
      if (off + sz > CH) {
        sync();
        off = 0;
      }
      par_memcpy(b + off, (const uint8_t*)u.h, u.bytes);
      HIP_CHECK(hipMemcpyAsync(u.d, b + off, u.bytes, hipMemcpyHostToDevice, st));
      off += sz;
    }
    sync();
  }
  void d2h(void* h, const void* d, size_t bytes) {
    if (!bytes) return;
    uint8_t* b = xfer.get(bytes);
    HIP_CHECK(hipMemcpyAsync(b, d, bytes, hipMemcpyDeviceToHost, st));
    sync();
    std::memcpy(h, b, bytes);
  }
  void d2d(void* d, const void* s, size_t bytes) { HIP_CHECK(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, st)); }
  G get(const L* d) {
    G g;
    d2h(g.v, d, sizeof g.v);
    return g;
  }
  DBuf<L> put(const G* xs, size_t n) {
    DBuf<L> b(pool, n * E);
    h2d(b.p, xs, n * sizeof(G));
    return b;
  }

  // ---------------- element kernels ----------------
  void pack(const L* slots, L* out, uint64_t n, int nrot, int rot0 = 0) {
    if constexpr (kMx<T>) {
      const uint64_t ntiles = (n + 31) / 32;
      launch("k_pack_mx", n * (32.0 * sizeof(L) + nrot * EB), n * nrot * 4096.0, (k_pack_mx<T>),
             dim3((unsigned)std::min<uint64_t>(ntiles, 8192)), dim3(256), (const uint32_t*)d_packfrag.p, slots, n, nrot,
             rot0, out);
    } else {
      const uint64_t ntiles = (n + 31) / 32;
      launch("k_pack_mx64", n * (32.0 * sizeof(L) + nrot * EB), n * nrot * 4096.0, k_pack_mx64,
             dim3((unsigned)std::min<uint64_t>(ntiles, 8192)), dim3(256), (const uint32_t*)d_packfrag.p,
             (const uint64_t*)slots, n, nrot, rot0, (uint64_t*)out);
    }
  }
  void psi(const L* x, L* slots, uint64_t n) {
    launch("k_psi", n * (EB + 32.0 * sizeof(L)), n * 16 * 32 * 16.0, (k_psi<T>), dim3(grid_for(n, GPB)), dim3(GB), x,
           d_pw.p, slots, n);
  }
  template <int NA> void mul_sub(const L* A, const L* B, const L* S, L* out, uint64_t n) {
    launch("k_mul_sub", n * 4 * EB, n * NA * PM, (k_mul_sub<T, NA>), dim3(grid_for(n, GPB)), dim3(GB), A, B, S, out, n);
  }
  // out[i] = s X[i]: with k <= 32 and many elements, x -> s x as
  // matrix-core fragments (k_mulmat) applied to every X[i] (one fixed
  // matrix, k_mulfix_mx); otherwise group products on the vector ALUs.
  void scale(const L* s, const L* X, L* out, uint64_t n) {
    if constexpr (kMx<T>) {
      if (n >= 1024) {
        DBuf<uint32_t> fr(pool, (size_t)kND<T> * 4096);
        launch("k_mulmat", EB + kND<T> * 16384.0, 0, (k_mulmat<T>), dim3(1), dim3(256), s, fr.p);
        const uint64_t ntiles = (n + 31) / 32;
        launch("k_scale_mx", n * 2 * EB, (double)n * E * E, (k_mulfix_mx<T>), dim3(grid_for(ntiles, 4, 1024), 1),
               dim3(256), (const uint32_t*)fr.p, X, n, out, 0);
        return;
      }
    }
    if constexpr (kMx64<T>) {
      if (n >= 512) {
        DBuf<uint32_t> fr(pool, (size_t)8 * 4096);
        mulmat64(s, fr.p);
        launch("k_scale_mx64", n * 2 * EB, (double)n * E * E, k_mulfix_mx64, dim3(grid_for((n + 31) / 32, 4, 256), 1),
               dim3(256), (const uint32_t*)fr.p, (const uint64_t*)X, (uint64_t)n, (uint64_t*)out, 0);
        return;
      }
    }
    launch("k_scale", n * 2 * EB, n * 32 * PM, (k_scale<T>), dim3(grid_for(n, GPB)), dim3(GB), s, X, out, n);
  }
  // G[e][i] = l_beta_i x[e] (i < 16) for nx elements: the products of the
  // packing basis b_{2i+h} = l_beta_i u_h that the layer rows and M_gamma
  // rows gather (a fixed-matrix product per i on the matrix cores)
  void mulfix_lb(const L* x, uint64_t nx, L* G) {
    const uint64_t ntiles = (nx + 31) / 32;
    if constexpr (kMx64<T>)
      launch("k_mulfix_mx64", nx * 17.0 * EB, nx * 16.0 * E * E, k_mulfix_mx64, dim3(grid_for(ntiles, 4, 16), 16),
             dim3(256), (const uint32_t*)d_lbfrag.p, (const uint64_t*)x, nx, (uint64_t*)G, 0);
    else
      launch("k_mulfix_mx", nx * 17.0 * EB, nx * 16.0 * E * E, (k_mulfix_mx<T>), dim3(grid_for(ntiles, 4, 32), 16),
             dim3(256), (const uint32_t*)d_lbfrag.p, x, nx, G, 0);
  }
  // out[p] = sum_{s,s'} C_p[s][s'] b_s b_s' for nb block rows (k <= 32):
  // the basis-product GEMM split four ways over its 1024-long sum (a slice
  // of the fragments in each block's LDS), then the slices added
  void cgemm(const L* cm, L* out, uint64_t nb) {
    const uint64_t ntiles = (nb + 31) / 32;
    DBuf<L> part(pool, 4 * nb * E);
    launch("k_cgemm", (double)nb * (1024 * sizeof(L) + 4 * EB), (double)nb * 1024 * E, (k_cgemm_sk<T>),
           dim3(grid_for(ntiles, 4, kND<T> <= 2 ? 256 : 64), 4), dim3(256), (const uint32_t*)d_b2frag.p, cm, part.p, nb);
    launch("k_sum_split", 5.0 * nb * EB, 0, (k_sum_split<T>), dim3(grid_for(nb * E, 256, 4096)), dim3(256),
           (const L*)part.p, (uint64_t)4, nb, out);
  }
  // The same GEMM with 64-bit limbs (exact digit classes, no split).
  void cgemm64(const L* cm, L* out, uint64_t nb) {
    launch("k_cgemm64", (double)nb * (1024 * sizeof(L) + EB), (double)nb * 1024 * E, k_cgemm64,
           dim3(grid_for((nb + 31) / 32, 1, 1024)), dim3(256), (const uint32_t*)d_b2frag.p, (const uint64_t*)cm,
           (uint64_t*)out, nb);
  }
  void sub(const L* a, const L* b, L* o, uint64_t nl) {
    launch("k_sub", nl * 3.0 * sizeof(L), 0, (k_sub<T>), dim3(grid_for(nl, 256, 1u << 20)), dim3(256), a, b, o, nl);
  }
  void fold(const L* r, const L* src, L* dst, uint64_t m, uint64_t ntab, uint64_t sstride, uint64_t dstride) {
    const int cap = Tuning::get().vfold_cap[Tuning::lg(m * ntab)].load(std::memory_order_relaxed);
    launch("k_fold", m * ntab * 3 * EB, m * ntab * 32 * PM, (k_fold<T>), dim3(grid_for(m * ntab, GPB, cap ? cap : 16384)),
           dim3(GB), r, src, dst, m, ntab, sstride, dstride);
  }
  // x -> r*x as matrix-core fragments (k <= 32); folds by the same
  // challenge share one.
  struct MulMat {
    DBuf<uint32_t> frag;         // owned fragments, or
    const uint32_t* fp = nullptr;  // the fragments the fold reads
    const L* r = nullptr;
  };
  MulMat mulmat(const L* r) {
    MulMat mm;
    mm.r = r;
    if constexpr (kMx<T>) {
      mm.frag = DBuf<uint32_t>(pool, (size_t)kND<T> * 4096);
      mm.fp = mm.frag.p;
      launch("k_mulmat", EB + kND<T> * 16384.0, 0, (k_mulmat<T>), dim3(1), dim3(256), r, mm.frag.p);
    } else if constexpr (kMx64<T>) {
      mm.frag = DBuf<uint32_t>(pool, (size_t)8 * 4096);
      mm.fp = mm.frag.p;
      mulmat64(r, mm.frag.p);
    }
    return mm;
  }
  // The fragment slot of sumcheck round r for a fold of fold_elems elements
  // (64-bit limbs: none below MX64_MIN, where the fold stays on the vector
  // ALUs and the fragments would not be read).
  static uint32_t* round_frag(uint32_t* frags, size_t r, uint64_t fold_elems) {
    if constexpr (kMx64<T>) return fold_elems >= MX64_MIN ? frags + r * 8 * 4096 : nullptr;
    else return frags + (kMx<T> ? r * kND<T> * 4096 : 0);
  }
  // 64-bit limbs: folds of at least MX64_MIN elements use the matrix cores
  // (below that the fragments, 128 KB, cost more than they save)
  static constexpr uint64_t MX64_MIN = 4096;
  void mulmat64(const L* r, uint32_t* frag) {
    launch("k_mulmat64", EB + 8 * 16384.0, 0, k_mulmat64, dim3(1, 4), dim3(256), (const uint64_t*)r, frag);
  }
  // A scalar update that needs nothing from the fold it accompanies: on the
  // matrix-core path it runs in an extra block of the fold launch, otherwise
  // as a kernel of its own after the folds.
  struct Side {
    int kind = 0;  // 0 none, 1 so = eq_one(sa, sb), 2 so = lin_claim(q = sa, r = sb)
    const L* sa = nullptr;
    const L* sb = nullptr;
    L* so = nullptr;
  };
  void side_kernel(const Side& sd) {
    if (sd.kind == 1)
      launch("k_eq_one", 3 * EB, 32 * PM, (k_eq_one<T>), dim3(1), dim3(64), sd.sa, sd.sb, sd.so);
    else if (sd.kind == 2)
      launch("k_lin_claim", 5 * EB, 64 * PM, (k_lin_claim<T>), dim3(1), dim3(64), sd.sa, sd.sb, sd.so);
  }
  void fold(const MulMat& mm, const L* src, L* dst, uint64_t m, uint64_t ntab, uint64_t sstride, uint64_t dstride,
            const Side& sd = Side{}) {
    fold2(mm, m, src, dst, ntab, sstride, dstride, nullptr, nullptr, 0, 0, 0, sd);
  }
  // Two groups of tables folded by the same challenge in one launch.
  void fold2(const MulMat& mm, uint64_t m, const L* s0, L* d0, uint64_t n0, uint64_t ss0, uint64_t ds0, const L* s1,
             L* d1, uint64_t n1, uint64_t ss1, uint64_t ds1, const Side& sd = Side{}) {
    if constexpr (kMx<T>) {
      FoldJob<T> J{{s0, s1 ? s1 : s0}, {d0, d1 ? d1 : d0}, {n0, s1 ? n1 : 0}, {ss0, ss1}, {ds0, ds1}, m,
                   sd.kind, sd.sa, sd.sb, sd.so};
      const uint64_t ntiles = (m * (J.ntab[0] + J.ntab[1]) + 31) / 32;
      // few tiles: 4 waves per tile, fragments from L2 (k_fold_mx); many:
      // persistent blocks with loads in flight (k_fold_mx_bulk); the
      // autotuner's picks, when it ran, replace the built-in choice
      const Tuning& tu = Tuning::get();
      const int lt = Tuning::lg(ntiles), kx = Tuning::kidx(T::K);
      const int tmode = tu.fold_mode[kx][lt].load(std::memory_order_relaxed);
      const int tcap = tu.fold_cap[kx][lt].load(std::memory_order_relaxed);
      const int split = tmode ? tmode == 1 : ntiles <= 256;
      const double fb = m * (J.ntab[0] + J.ntab[1]) * 3 * EB, fm = m * (J.ntab[0] + J.ntab[1]) * 32 * PM;
      if (split)
        launch("k_fold_mx", fb, fm, (k_fold_mx<T>), dim3((unsigned)ntiles + (sd.kind ? 1 : 0)), dim3(256), mm.fp, J);
      else
        launch("k_fold_mx", fb, fm, (k_fold_mx_bulk<T>),
               dim3((unsigned)std::min<uint64_t>(ntiles, tcap ? (unsigned)tcap : 512u) + (sd.kind ? 1 : 0)), dim3(256),
               mm.fp, J);
    } else {
      if constexpr (kMx64<T>) {
        const uint64_t tot = m * (n0 + (s1 ? n1 : 0));
        if (mm.fp && tot >= MX64_MIN) {
          FoldJob<T> J{{s0, s1 ? s1 : s0}, {d0, d1 ? d1 : d0}, {n0, s1 ? n1 : 0}, {ss0, ss1}, {ds0, ds1}, m,
                       sd.kind, sd.sa, sd.sb, sd.so};
          // one block per CU at most (each keeps the 128 KB of fragments in LDS)
          const unsigned nb = grid_for((tot + 31) / 32, 4, 256);
          launch("k_fold_mx64", tot * 3 * EB, tot * 32 * PM, (k_fold_mx64<T>), dim3(nb + (sd.kind ? 1 : 0)), dim3(256),
                 mm.fp, J);
          return;
        }
      }
      fold(mm.r, s0, d0, m, n0, ss0, ds0);
      if (s1) fold(mm.r, s1, d1, m, n1, ss1, ds1);
      side_kernel(sd);
    }
  }
  // Linear-check round sums {sum f_lo g_lo, sum df dg} as two Gram matrices
  // on the matrix cores; x rows past `xrows` limbs are zero.
  // Partials [block][2] of the two sums into `part` (room for 2 * 256
  // elements); returns the block count.
  static constexpr unsigned GRAM_BLOCKS = 256;      // built-in workgroup count
  static constexpr unsigned GRAM_BLOCKS_MAX = 1024;  // room for the autotuner's picks
  unsigned gram_lin(const L* f, uint64_t fs, const L* g, uint64_t gs, uint64_t ntab, uint64_t m, int xrows, L* part) {
    const uint64_t nch = (m * ntab + 31) / 32;
    const int tb = Tuning::get().gram_blocks[Tuning::kidx(T::K)][Tuning::lg(nch)].load(std::memory_order_relaxed);
    const unsigned nwg = (unsigned)std::min<uint64_t>(nch, tb ? (unsigned)tb : GRAM_BLOCKS);
    launch("k_gram_lin", m * ntab * 4.0 * EB, m * ntab * 2.0 * E * E, (k_gram_lin2<T>), dim3(nwg), dim3(512), f, fs, g,
           gs, ntab, m, m, xrows, part, 2, 0);
    return nwg;
  }
  // The same with 64-bit limbs (k_gram_lin64: one block per CU, its LDS
  // holds a chunk's digit planes).
  unsigned gram_lin64(const L* f, uint64_t fs, const L* g, uint64_t gs, uint64_t ntab, uint64_t m, int xrows, L* part) {
    const uint64_t nch = (m * ntab + 31) / 32;
    const unsigned nwg = (unsigned)std::min<uint64_t>(nch, GRAM_BLOCKS);
    launch("k_gram_lin64", m * ntab * 4.0 * EB, m * ntab * 2.0 * E * E, k_gram_lin64, dim3(nwg), dim3(512),
           (const uint64_t*)f, fs, (const uint64_t*)g, gs, ntab, m, m, xrows, (uint64_t*)part, 2, 0);
    return nwg;
  }
  void marg(const L* s, L* eg, uint64_t m) {
    launch("k_marg", m * 3 * EB, m * 32 * PM, (k_marg<T>), dim3(grid_for(m, GPB)), dim3(GB), s, eg, m);
  }
  DBuf<L> reduce(const DBuf<L>& partials, unsigned nblocks, int ns) {
    DBuf<L> out(pool, ns * E);
    launch("k_sum_partials", (double)nblocks * ns * EB, 0, (k_sum_partials<T>), dim3(ns * 4), dim3(1024), partials.p,
           (uint64_t)nblocks, ns, out.p);
    return out;
  }
  // ntab inner products <A_y, B> (A_y = A + y*astride elements), left on
  // the device (ntab contiguous elements).
  template <int NA> DBuf<L> dots_dev(const L* A, uint64_t astride, int ntab, const L* B, uint64_t n) {
    if (kFrag<T> && ntab == 1 && n >= DOT_MX_MIN) {
      // one long inner product: the Gram-matrix round kernel with no hi
      // half (only its first sum is formed), contracted once
      DBuf<L> part(pool, (size_t)GRAM_BLOCKS * E);
      const int xrows = NA == 16 ? 64 : 128;
      unsigned nwg;
      if constexpr (kMx64<T>) {
        nwg = (unsigned)std::min<uint64_t>((n + 31) / 32, GRAM_BLOCKS);
        launch("k_dot_gram64", n * 2.0 * EB, n * 2.0 * E * E, k_gram_lin64, dim3(nwg), dim3(512), (const uint64_t*)A,
               (uint64_t)0, (const uint64_t*)B, (uint64_t)0, (uint64_t)1, n, (uint64_t)0, xrows, (uint64_t*)part.p, 1,
               0);
      } else {
        nwg = (unsigned)std::min<uint64_t>((n + 31) / 32, GRAM_BLOCKS);
        launch("k_dot_gram", n * 2.0 * EB, n * 2.0 * E * E, (k_gram_lin2<T>), dim3(nwg), dim3(512), A, (uint64_t)0, B,
               (uint64_t)0, (uint64_t)1, n, (uint64_t)0, xrows, part.p, 1, 0);
      }
      return reduce(part, nwg, 1);
    }
    unsigned nb = grid_for(n, GPB, ntab > 1 ? 256 : 2048);
    DBuf<L> part(pool, (size_t)nb * ntab * E);
    launch("k_dot", n * (ntab + 1.0) * EB, n * ntab * NA * PM, (k_dot<T, NA>), dim3(nb, ntab), dim3(GB), A, astride, B, n,
           part.p);
    return reduce(part, nb, ntab);
  }
  // the same, one D2H
  template <int NA> std::vector<G> dots(const L* A, uint64_t astride, int ntab, const L* B, uint64_t n) {
    DBuf<L> r = dots_dev<NA>(A, astride, ntab, B, n);
    std::vector<G> out(ntab);
    d2h(out.data(), r.p, ntab * sizeof(G));
    return out;
  }
  template <int NA> G dot(const L* A, const L* B, uint64_t n) { return dots<NA>(A, 0, 1, B, n)[0]; }
  // Several picks (element 0 of cnt tables spaced `stride` elements apart;
  // stride 1 reads cnt consecutive elements) with one synchronization: the
  // pick kernels write straight into a pinned staging area.
  struct PickSeg {
    const L* src;
    uint64_t stride, cnt;
  };
  Pinned pickbuf;
  std::vector<G> pick_many(std::initializer_list<PickSeg> segs) {
    uint64_t tot = 0;
    for (const PickSeg& g : segs) tot += g.cnt;
    std::vector<G> out(tot);
    if (!tot) return out;
    L* h = (L*)pickbuf.get(tot * E * sizeof(L));
    L* d = nullptr;
    HIP_CHECK(hipHostGetDevicePointer((void**)&d, h, 0));
    uint64_t off = 0;
    for (const PickSeg& g : segs) {
      if (g.cnt)
        launch("k_pick", 0, 0, (k_pick<T>), dim3(grid_for(g.cnt * E, 256)), dim3(256), g.src, g.stride, g.cnt,
               d + off * E);
      off += g.cnt;
    }
    sync();
    std::memcpy(out.data(), h, tot * sizeof(G));
    return out;
  }
  // element 0 of ntab tables spaced `stride` elements apart, one D2H
  std::vector<G> pick(const L* src, uint64_t stride, int ntab) {
    DBuf<L> o(pool, (size_t)ntab * E);
    launch("k_pick", 0, 0, (k_pick<T>), dim3(grid_for((uint64_t)ntab * E, 256)), dim3(256), src, stride, (uint64_t)ntab, o.p);
    std::vector<G> out(ntab);
    d2h(out.data(), o.p, ntab * sizeof(G));
    return out;
  }

  // Gruen rounds over at least this many element pairs form their sums on
  // the matrix cores
  static constexpr uint64_t MULT_MX_MIN = 8192;
  // inner products of at least this many terms use the matrix cores
  static constexpr uint64_t DOT_MX_MIN = 8192;
  // outer-product steps of at least this many entries use the matrix cores
  static constexpr size_t EQ_MX_MIN = 8192;
  // build_eq_table (eq.rs:35-58) for host points r[0..n): 2^n entries
  DBuf<L> eq_table(const G* r, size_t n) {
    if (n == 0) {
      DBuf<L> a(pool, E);
      d2d(a.p, d_one.p, E * sizeof(L));
      return a;
    }
    DBuf<L> rr = put(r, n);
    // chunk tables in one launch, then outer products from the last chunk leftwards
    const int k = (int)((n + EQ_HEAD - 1) / EQ_HEAD);
    std::vector<size_t> sz(k), off(k);
    size_t tot = 0;
    for (int j = 0; j < k; j++) {
      sz[j] = n / k + ((size_t)j < n % k ? 1 : 0);
      off[j] = tot;
      tot += size_t(1) << sz[j];
    }
    DBuf<L> heads(pool, tot * E);
    launch("k_eq_chunks", 0, 0, (k_eq_chunks<T>), dim3(k), dim3(1024), (const L*)rr.p, (int)n, k, heads.p);
    if (k == 1) return heads;
    DBuf<L> cur;
    const L* tail = heads.p + off[k - 1] * E;
    size_t len = size_t(1) << sz[k - 1];
    for (int j = k - 2; j >= 0; j--) {
      const size_t na = size_t(1) << sz[j];
      DBuf<L> nx(pool, na * len * E);
      const L* hj = heads.p + off[j] * E;
      if (kFrag<T> && na * len >= EQ_MX_MIN) {
        // each head entry is a fixed multiplier of the whole tail: its
        // fragments, then one fixed-matrix product per head entry
        const uint64_t ntiles = (len + 31) / 32;
        const unsigned per = (unsigned)std::max<size_t>(1, 512 / na);  // ~512 blocks in all
        DBuf<uint32_t> fr(pool, na * kND<T> * 4096);
        if constexpr (kMx64<T>) {
          launch("k_mulmat64", na * (EB + 8 * 16384.0), 0, k_mulmat64, dim3((unsigned)na, 4), dim3(256),
                 (const uint64_t*)hj, fr.p);
          launch("k_eq_outer_mx64", na * len * EB, (double)na * len * E * E, k_mulfix_mx64,
                 dim3(grid_for(ntiles, 4, per), (unsigned)na), dim3(256), (const uint32_t*)fr.p, (const uint64_t*)tail,
                 (uint64_t)len, (uint64_t*)nx.p, 1);
        } else if constexpr (kMx<T>) {
          launch("k_mulmat", na * (EB + kND<T> * 16384.0), 0, (k_mulmat<T>), dim3((unsigned)na), dim3(256), hj, fr.p);
          launch("k_eq_outer_mx", na * len * EB, (double)na * len * E * E, (k_mulfix_mx<T>),
                 dim3(grid_for(ntiles, 4, per), (unsigned)na), dim3(256), (const uint32_t*)fr.p, tail, (uint64_t)len,
                 nx.p, 1);
        }
        cur = std::move(nx);
        tail = cur.p;
        len *= na;
        continue;
      }
      launch("k_eq_outer", na * len * EB, na * len * 32 * PM, (k_eq_outer<T>), dim3(grid_for(na * len, GPB)), dim3(GB),
             hj, (uint64_t)na, tail, (uint64_t)len, nx.p);
      cur = std::move(nx);
      tail = cur.p;
      len *= na;
    }
    return cur;
  }


  // ---------------- proof downloads ----------------
  // Enqueue the copy of a finished device buffer into the pinned proof
  // buffer at host address `dst` (after the prover stream's work so far);
  // the buffer is held until finish_downloads().
  //
  // The copy kernel runs dl_blocks() workgroups: that already fills the host
  // link, and a copy with a workgroup per CU queues so many posted writes
  // that every other queue's packet and kernel-argument reads from host
  // memory wait behind them (a 64 MiB copy with 256 workgroups held a
  // one-block kernel on another stream for 1.1 ms; with 16, for 70 us, at
  // the same copy time).

  // Sections of at least dl_dma_min() bytes (256 KiB) go through the DMA engines
  // (hipMemcpyAsync into the pinned buffer): a copy kernel's writes to host
  // memory slow the HBM kernels it overlaps (a 2 GB device copy beside a
  // 64 MiB copy-kernel download took 1.3-3.3x its time, beside the DMA copy
  // 1.0x), while small sections keep the copy kernel (lower per-call cost).
  // NENYA_DL_DMA_MIN overrides the threshold (bytes).
  static size_t dl_dma_min() {
    static const size_t v = [] {
      const char* e = std::getenv("NENYA_DL_DMA_MIN");
      return e ? (size_t)std::strtoull(e, nullptr, 10) : (size_t)(256 << 10);
    }();
    return v;
  }
  template <class X> void download(uint8_t* dst, DBuf<X>& src, size_t bytes) {
    if (!bytes) return;
    uint8_t* ddst = nullptr;
    HIP_CHECK(hipHostGetDevicePointer((void**)&ddst, dst, 0));
    const dim3 grid((unsigned)std::min<uint64_t>((bytes + 4095) / 4096, dl_blocks()));
    auto copy = [&](hipStream_t s, bool dma) {
      if (dma) {
        HIP_CHECK(hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, s));
      } else {
        hipLaunchKernelGGL(k_copy_out, grid, dim3(256), 0, s, (const uint8_t*)src.p, ddst, (uint64_t)bytes);
        HIP_CHECK(hipGetLastError());
      }
    };
    if (!pipelined || serial) {
      // contexts of a multi-worker server and no-alloc calls: in order on the
      // prover stream (the workers' streams already overlap each other); the
      // no-alloc calls keep the copy kernel for every section
      copy(st, !serial && bytes >= dl_dma_min());
      return;
    }
    if (!cst) {
      HIP_CHECK(hipStreamCreate(&cst));
      HIP_CHECK(hipEventCreateWithFlags(&cev, hipEventDisableTiming));
    }
    HIP_CHECK(hipEventRecord(cev, st));
    HIP_CHECK(hipStreamWaitEvent(cst, cev, 0));
    copy(cst, bytes >= dl_dma_min());
    held.push_back({src.p, src.n * sizeof(X)});
    src.p = nullptr;
    src.n = 0;
  }
  // ---------------- concurrent phases ----------------
  // Independent work of a proof (its commitments) runs on up to four
  // streams: fork() makes the side streams wait for the prover stream's work
  // so far, on(i) sends launches to stream i (0: the prover stream), join()
  // makes the prover stream wait for the side streams.  Blocks freed in
  // between go back to the pool at join().  Server contexts stay on one
  // stream (their workers already overlap), and so do the debug pools.
  static constexpr int NSIDE = 3;
  hipStream_t sst[NSIDE] = {};
  hipEvent_t fev = nullptr, jev[NSIDE] = {};
  hipStream_t st_main = nullptr;
  int forked = 0;
  void fork(int ns) {
    st_main = st;
    forked = (pipelined && !serial && !pool_poison() && !pool_guard()) ? std::min(ns - 1, NSIDE) : 0;
    if (forked <= 0) {
      forked = 0;
      return;
    }
    if (!fev) {
      HIP_CHECK(hipEventCreateWithFlags(&fev, hipEventDisableTiming));
      for (int i = 0; i < NSIDE; i++) {
        HIP_CHECK(hipEventCreateWithFlags(&jev[i], hipEventDisableTiming));
        HIP_CHECK(hipStreamCreate(&sst[i]));
      }
    }
    HIP_CHECK(hipEventRecord(fev, st));
    for (int i = 0; i < forked; i++) HIP_CHECK(hipStreamWaitEvent(sst[i], fev, 0));
    pool->defer_frees(true);
  }
  void on(int i) {
    if (forked) st = i == 0 ? st_main : sst[(i - 1) % forked];
  }
  // A background stream for work whose results are needed only at the end
  // of a proof (opening tails, openings whose values enter only the final
  // self-check): bg_begin() makes it wait for the prover stream's work so
  // far and sends launches to it until the matching bg_end(); bg_join()
  // makes the prover stream wait for it.  Regions may nest and may reopen
  // after bg_end().  Frees are deferred from the first bg_begin() to
  // bg_join().  Contexts that stay on one stream run the same work in place.
  hipStream_t sbg = nullptr, st_fg = nullptr;
  hipEvent_t bgev = nullptr, bgjev = nullptr;
  bool bg_live = false;
  int bg_nest = 0;
  void bg_begin() {
    if (forked || !pipelined || serial || pool_poison() || pool_guard()) return;
    if (bg_live && st == sbg) {
      bg_nest++;
      return;
    }
    if (!sbg) {
      HIP_CHECK(hipStreamCreate(&sbg));
      HIP_CHECK(hipEventCreateWithFlags(&bgev, hipEventDisableTiming));
      HIP_CHECK(hipEventCreateWithFlags(&bgjev, hipEventDisableTiming));
    }
    HIP_CHECK(hipEventRecord(bgev, st));
    HIP_CHECK(hipStreamWaitEvent(sbg, bgev, 0));
    if (!bg_live) {
      st_fg = st;
      bg_live = true;
      pool->defer_frees(true);
    }
    st = sbg;
  }
  void bg_end() {
    if (!bg_live || st != sbg) return;
    if (bg_nest > 0) {
      bg_nest--;
      return;
    }
    st = st_fg;
  }
  void bg_join() {
    if (!bg_live) return;
    HIP_CHECK(hipEventRecord(bgjev, sbg));
    HIP_CHECK(hipStreamWaitEvent(st_fg, bgjev, 0));
    st = st_fg;
    bg_live = false;
    bg_nest = 0;
    pool->defer_frees(false);
  }
  void join() {
    if (!forked) return;
    for (int i = 0; i < forked; i++) {
      HIP_CHECK(hipEventRecord(jev[i], sst[i]));
      HIP_CHECK(hipStreamWaitEvent(st_main, jev[i], 0));
    }
    st = st_main;
    forked = 0;
    pool->defer_frees(false);
  }
  void finish_downloads() {
    // a proof that failed between fork() and join() left the streams forked
    if (forked) {
      st = st_main;
      for (int i = 0; i < forked; i++) (void)hipStreamSynchronize(sst[i]);
      forked = 0;
      pool->defer_frees(false);
    }
    // ... or its background work running
    if (bg_live) {
      st = st_fg;
      (void)hipStreamSynchronize(sbg);
      bg_live = false;
      bg_nest = 0;
      pool->defer_frees(false);
    }
    // the prover stream's own writes into the proof buffer (server contexts
    // copy their sections on it; k_queries writes indices directly) and any
    // commitment roots a failed proof left unresolved
    sync();
    root_used = 0;
    if (!cst) return;
    HIP_CHECK(hipStreamSynchronize(cst));
    for (auto& h : held) pool->put(h.first, h.second);
    held.clear();
  }
  ~Engine() {
    for (auto& e : hev)
      if (e) (void)hipEventDestroy(e);
    if (sbg) {
      (void)hipStreamSynchronize(sbg);
      (void)hipStreamDestroy(sbg);
      (void)hipEventDestroy(bgev);
      (void)hipEventDestroy(bgjev);
    }
    for (int i = 0; i < NSIDE; i++)
      if (sst[i]) {
        (void)hipStreamSynchronize(sst[i]);
        (void)hipStreamDestroy(sst[i]);
      }
    if (fev) (void)hipEventDestroy(fev);
    for (auto& x : xev)
      if (x) (void)hipEventDestroy(x);
    for (int i = 0; i < NSIDE; i++)
      if (jev[i]) (void)hipEventDestroy(jev[i]);
    if (cst) {
      (void)hipStreamSynchronize(cst);
      for (auto& h : held) pool->put(h.first, h.second);
      (void)hipStreamDestroy(cst);
    }
    if (cev) (void)hipEventDestroy(cev);
  }

  // ---------------- mailbox ----------------
  static constexpr size_t mb_bytes() { return 1024 + (size_t)MB_SLOTS * 4 * E * sizeof(L); }
  static constexpr size_t mb_down_flag(int s) { return 64 + 4 * (size_t)s; }
  static constexpr size_t mb_up_flag(int s) { return 64 + 4 * ((size_t)MB_SLOTS + s); }
  static constexpr size_t mb_down(int s) { return 1024 + (size_t)s * 3 * E * sizeof(L); }
  static constexpr size_t mb_up(int s) { return 1024 + ((size_t)MB_SLOTS * 3 + s) * E * sizeof(L); }
  template <class X> X* mbh(size_t off) { return reinterpret_cast<X*>(mb_h + off); }
  template <class X> X* mbd(size_t off) { return reinterpret_cast<X*>(mb_d + off); }

  // Enqueue: reduce a round's partial sums and post q to slot s (flag value
  // returned).  mode 0 mult check, 1 linear check (claim c, q also to qd).
  uint32_t post_round(const L* part, uint64_t nparts, int ns, int mode, const L* c, L* qd, int s) {
    const uint32_t v = ++mseq;
    // few partials: one block, no cross-block handoff
    const unsigned grid = nparts <= 64 ? 1 : ns * 4;
    launch("k_sum_post", (double)nparts * ns * EB, 0, (k_sum_post<T>), dim3(grid), dim3(1024), part, nparts, ns, mode, c,
           d_sums.p, qd, mbd<L>(mb_down(s)), d_count.p, mbd<uint32_t>(mb_down_flag(s)), v);
    return v;
  }
  // Enqueue: reduce a round's partial sums into q at qh (device), for the
  // device transcript (no host posting).
  void post_round_dev(const L* part, uint64_t nparts, int ns, int mode, const L* c, L* qd, L* qh) {
    const unsigned grid = nparts <= 64 ? 1 : ns * 4;
    launch("k_sum_post", (double)nparts * ns * EB, 0, (k_sum_post<T>), dim3(grid), dim3(1024), part, nparts, ns, mode, c,
           d_sums.p, qd, qh, d_count.p, (uint32_t*)nullptr, 0u);
  }
  // Enqueue: absorb q (device) into the device transcript st, draw the
  // round challenge there and lift it into d (with k <= 32 its fragments go
  // to frag and mm points at them).
  void lift_device_fs(b3::State* st, const L* q, L* d, uint32_t* frag, MulMat* mm) {
    const FsArg<T> f{st, q};
    const double lm = E * E + (T::K - 1) * 16.5 * PM;
    mm->r = d;
    if constexpr (kMx<T>) {
      mm->fp = frag;
      if (frag)
        launch("k_lift_mulmat_fs", 5 * EB, lm, (k_lift_mulmat_fs<T>), dim3(1), dim3(256), f, d, d_frob.p, T::K - 1, frag);
      else  // a tail round's fold needs no fragments
        launch("k_lift_fs", 5 * EB, lm, (k_lift_fs<T>), dim3(1), dim3(256), f, d, d_frob.p, T::K - 1);
    } else {
      launch("k_lift_fs", 5 * EB, lm, (k_lift_fs<T>), dim3(1), dim3(256), f, d, d_frob.p, T::K - 1);
      if constexpr (kMx64<T>) {
        if (frag) {
          mm->fp = frag;
          mulmat64(d, frag);
        }
      }
    }
  }
  // The device transcript of a sumcheck: the host transcript's state
  // uploaded before the rounds, fetched back with the sumcheck's tail.
  Pinned trbuf;
  void fetch_transcript_async(const b3::State* d) {
    uint8_t* h = trbuf.get(sizeof(b3::State));
    HIP_CHECK(hipMemcpyAsync(h, d, sizeof(b3::State), hipMemcpyDeviceToHost, st));
  }
  void adopt_transcript(Transcript& tr) { std::memcpy(static_cast<b3::State*>(&tr.h), trbuf.p, sizeof(b3::State)); }

  // Host: wait for slot s's round polynomial (bounded; device errors raise).
  void wait_round(int s, uint32_t v, G q[3]) {
    const uint32_t* f = mbh<uint32_t>(mb_down_flag(s));
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t it = 1; __atomic_load_n(f, __ATOMIC_ACQUIRE) != v; it++) {
      __builtin_ia32_pause();
      if ((it & 4095) != 0) continue;
      check_waits();
      const hipError_t e = hipStreamQuery(st);
      if (e != hipSuccess && e != hipErrorNotReady) throw Error(NENYA_E_DEVICE, hipGetErrorString(e));
      if (e == hipSuccess && __atomic_load_n(f, __ATOMIC_ACQUIRE) != v)
        throw Error(NENYA_E_DEVICE, "sumcheck round values were not posted");
      if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60))
        throw Error(NENYA_E_DEVICE, "device did not post sumcheck round values");
      std::this_thread::yield();
    }
    std::memcpy(q, mbh<L>(mb_down(s)), 3 * sizeof(G));
    check_waits();
  }
  // A device wait that ran out of time went on with a stale sample: the
  // proof fails (the caller's guard releases the other waits).  Also called
  // once the round loop's work has finished.
  void check_waits() {
    if (__atomic_load_n(mbh<uint32_t>(4), __ATOMIC_ACQUIRE) != 0u)
      throw Error(NENYA_E_DEVICE, "a device wait for a posted challenge ran out of time");
  }
  // Enqueue: lift the challenge the host will post to slot s (flag value v)
  // into d; with k <= 32 its multiplication fragments go to `frag`
  // (kND * 4096 words, allocated by the caller) and mm points at them.
  void lift_posted(int s, uint32_t v, L* d, uint32_t* frag, MulMat* mm) {
    const MbWait w{mbd<uint32_t>(mb_up_flag(s)), v, mbd<uint32_t>(0), wait_ticks};
    const L* src = mbd<L>(mb_up(s));
    const double lm = E * E + (T::K - 1) * 16.5 * PM;
    mm->r = d;
    if constexpr (kMx<T>) {
      mm->fp = frag;
      if (frag)
        launch("k_lift_mulmat", 2 * EB, lm, (k_lift_mulmat<T>), dim3(LIFT_MM_BLOCKS), dim3(256), src, d, d_frob.p,
               T::K - 1, frag, w);
      else  // a tail round's fold needs no fragments
        launch("k_lift", 2 * EB, lm, (k_lift_from<T>), dim3(1), dim3(256), src, d, d_frob.p, T::K - 1, w);
    } else if constexpr (kMx64<T>) {
      if (frag) {
        mm->fp = frag;
        launch("k_lift_mulmat64", 2 * EB, lm, k_lift_mulmat64, dim3(LIFT_MM_BLOCKS), dim3(256), (const uint64_t*)src,
               (uint64_t*)d, (const uint64_t*)d_frob.p, T::K - 1, frag, w);
      } else {
        launch("k_lift", 2 * EB, lm, (k_lift_from<T>), dim3(1), dim3(256), src, d, d_frob.p, T::K - 1, w);
      }
    } else {
      launch("k_lift", 2 * EB, lm, (k_lift_from<T>), dim3(1), dim3(256), src, d, d_frob.p, T::K - 1, w);
    }
  }
  // Host: sample the challenge (the transcript's sampler) and post it.
  void post_challenge(Transcript& tr, const char* label, int s, uint32_t v) {
    StdRng rng = tr.sampler(label);
    L* up = mbh<L>(mb_up(s));
    for (int l = 0; l < E; l++) up[l] = (L)rng.limb(T::K);
    __atomic_store_n(mbh<uint32_t>(mb_up_flag(s)), v, __ATOMIC_RELEASE);
  }
  // Kernels left waiting when the host leaves a round loop by an exception
  // are released through the abort word before the stack unwinds further.
  struct MbGuard {
    Engine& e;
    int n0 = std::uncaught_exceptions();
    explicit MbGuard(Engine& en) : e(en) {}
    ~MbGuard() {
      if (std::uncaught_exceptions() <= n0) return;
      __atomic_store_n(e.mbh<uint32_t>(0), 1u, __ATOMIC_RELEASE);
      (void)hipStreamSynchronize(e.st);
      __atomic_store_n(e.mbh<uint32_t>(0), 0u, __ATOMIC_RELEASE);
      __atomic_store_n(e.mbh<uint32_t>(4), 0u, __ATOMIC_RELEASE);
    }
  };

  // ---------------- Fiat-Shamir challenges ----------------
  // R2Sampler::sample (sampling.rs:57-68): 128 L::random draws then the
  // Teichmuller lift; the lift runs on the device.
  std::vector<G> sample(StdRng& rng, size_t cnt, DBuf<L>* dev_out = nullptr, MulMat* mm = nullptr) {
    std::vector<G> nat(cnt);
    for (size_t i = 0; i < cnt; i++)
      for (int l = 0; l < E; l++) nat[i].v[l] = (L)rng.limb(T::K);
    DBuf<L> d = put(nat.data(), cnt);
    const double lm = E * E + (T::K - 1) * 16.5 * PM;
    bool fused = false;
    if constexpr (kMx<T>) {
      if (mm && cnt == 1) {
        mm->frag = DBuf<uint32_t>(pool, (size_t)kND<T> * 4096);
        mm->fp = mm->frag.p;
        // in place (src = d): one block (see k_lift_mulmat)
        launch("k_lift_mulmat", 2 * EB, lm, (k_lift_mulmat<T>), dim3(1), dim3(256), (const L*)d.p, d.p, d_frob.p,
               T::K - 1, mm->frag.p, MbWait{});
        fused = true;
      }
    }
    if (!fused)
      launch("k_lift", cnt * 2 * EB, cnt * lm, (k_lift<T>), dim3(cnt), dim3(256), d.p, d_frob.p, T::K - 1);
    d2h(nat.data(), d.p, cnt * sizeof(G));
    if (mm) {
      mm->r = d.p;
      if (!fused) *mm = mulmat(d.p);
    }
    if (dev_out) *dev_out = std::move(d);
    return nat;
  }
  // Samples one challenge and lifts it in device memory at d without waiting
  // for it; mm (if given) receives its multiplication fragments.
  void challenge_into(Transcript& tr, const char* label, L* d, MulMat* mm) {
    StdRng rng = tr.sampler(label);
    G nat;
    for (int l = 0; l < E; l++) nat.v[l] = (L)rng.limb(T::K);
    // the lift kernel reads the natural sample from pinned memory (no copy)
    const L* src = (const L*)stage_host(nat.v, sizeof nat.v);
    const double lm = E * E + (T::K - 1) * 16.5 * PM;
    bool fused = false;
    if constexpr (kMx<T>) {
      if (mm) {
        mm->frag = DBuf<uint32_t>(pool, (size_t)kND<T> * 4096);
        mm->fp = mm->frag.p;
        launch("k_lift_mulmat", 2 * EB, lm, (k_lift_mulmat<T>), dim3(LIFT_MM_BLOCKS), dim3(256), src, d, d_frob.p,
               T::K - 1, mm->frag.p, MbWait{});
        fused = true;
      }
    } else if constexpr (kMx64<T>) {
      if (mm) {
        mm->frag = DBuf<uint32_t>(pool, (size_t)8 * 4096);
        mm->fp = mm->frag.p;
        launch("k_lift_mulmat64", 2 * EB, lm, k_lift_mulmat64, dim3(LIFT_MM_BLOCKS), dim3(256), (const uint64_t*)src,
               (uint64_t*)d, (const uint64_t*)d_frob.p, T::K - 1, mm->frag.p, MbWait{});
        fused = true;
      }
    }
    if (!fused)
      launch("k_lift", 2 * EB, lm, (k_lift_from<T>), dim3(1), dim3(256), src, d, d_frob.p, T::K - 1, MbWait{});
    if (mm) {
      mm->r = d;
      if (!fused) *mm = mulmat(d);
    }
  }
  G challenge(Transcript& tr, const char* label, DBuf<L>* dev = nullptr, MulMat* mm = nullptr) {
    StdRng rng = tr.sampler(label);
    return sample(rng, 1, dev, mm)[0];
  }

  // ---------------- PCS ----------------
  // Commit (pcs.rs:489-543, 670-730; r2_pcs.rs:259-294, 452-494): encode the
  // `rows` message rows (k2 entries of W limbs, contiguous), hash the n
  // columns, build the Merkle tree.
  CommitDev<T> commit(const PcsDev<T>& pc, const L* rowsbuf, size_t rows, const L* lifted) {
    checkpoint();
    const PcsParams& P = pc.P;
    const int W = P.packed ? 32 : E;
    CommitDev<T> c;
    c.rows = rows;
    c.msg = lifted;
    DBuf<L> tmp(pool, rows * P.n * W);
    c.enc = DBuf<L>(pool, rows * P.n * W);
    if (P.packed) raa<32>(pc, rowsbuf, rows, tmp.p, c.enc.p);
    else raa<E>(pc, rowsbuf, rows, tmp.p, c.enc.p);
    size_t n = P.n;  // a power of two
    c.nodes = DBuf<uint32_t>(pool, (2 * n - 1) * 8);
    // leaves: thread per (column, 1 KiB chunk), then the per-column merge
    const int esl = __builtin_ctzll((uint64_t)W * sizeof(L)), ln = __builtin_ctzll((uint64_t)n);
    const uint64_t colb = (uint64_t)rows * W * sizeof(L), nch = (colb + 1023) / 1024;
    DBuf<uint32_t> cvs(pool, nch > 1 ? n * nch * 8 : 8);
    launch("k_col_chunks", (double)rows * n * W * sizeof(L), 0, k_col_chunks, dim3(grid_for(n * nch, 64, 1u << 16)),
           dim3(64), (const uint8_t*)c.enc.p, (uint64_t)rows, ln, esl, nch, cvs.p, c.nodes.p);
    if (nch > 1)
      launch("k_col_merge", (double)n * nch * 32, 0, k_col_merge, dim3(grid_for(n, 64)), dim3(64), cvs.p, (uint64_t)n,
             nch, c.nodes.p);
    size_t start = 0, len = n;
    while (len > (size_t)MERKLE_TOP) {
      launch("k_merkle_level", len * 32.0 + len * 16.0, 0, k_merkle_level, dim3(grid_for(len / 2, 128)), dim3(128),
             c.nodes.p, (uint64_t)start, (uint64_t)len);
      start += len;
      len /= 2;
    }
    if (len > 1)
      launch("k_merkle_top", len * 64.0, 0, k_merkle_top, dim3(1), dim3(1024), c.nodes.p, (uint64_t)start, (uint64_t)len);
    // the root lands in a pinned slot without a synchronization; the
    // caller's resolve_roots() waits once for a group of commitments
    if (root_used >= ROOT_SLOTS) throw Error(NENYA_E_DEVICE, "too many unresolved commitment roots");
    uint8_t* slot = rootbuf.get(ROOT_SLOTS * 32) + 32 * root_used++;
    uint8_t* dslot = nullptr;
    HIP_CHECK(hipHostGetDevicePointer((void**)&dslot, slot, 0));
    launch("k_copy_small", 64, 0, k_copy_small, dim3(1), dim3(256), (const uint8_t*)(c.nodes.p + (2 * n - 2) * 8), dslot,
           (uint64_t)32);
    c.root_stage = slot;
    return c;
  }
  // One synchronization for the roots of the given commitments.
  void resolve_roots(std::initializer_list<CommitDev<T>*> cs) {
    sync();
    for (CommitDev<T>* c : cs) {
      if (!c->root_stage) continue;
      std::memcpy(c->root, c->root_stage, 32);
      c->root_stage = nullptr;
    }
    root_used = 0;
  }

  // RAA encode of `rows` rows (raa.rs:106-153 / 67-98), W limbs per entry.
  template <int W> void raa(const PcsDev<T>& pc, const L* msg, size_t rows, L* tmp, L* out) {
    const PcsParams& P = pc.P;
    if (P.q & (P.q - 1)) throw Error(NENYA_E_SHAPE, "RAA repetition must be a power of two");
    const int qs = __builtin_ctzll((uint64_t)P.q);
    uint64_t nseg = (P.n + Raa<T, W>::SEG - 1) / Raa<T, W>::SEG;
    DBuf<C> seg(pool, rows * nseg * W);
    dim3 grid((unsigned)nseg, (unsigned)rows);
    double eb = W * sizeof(L), rowb = (double)rows * P.n * eb;
    for (int pass = 0; pass < 2; pass++) {
      const L* src = pass == 0 ? msg : tmp;
      L* dst = pass == 0 ? tmp : out;
      launch("k_raa_gather", rowb * 2.0, 0, (k_raa_gather<T, W>), grid, dim3(RAA_TB), src,
             (uint64_t)(pass == 0 ? P.k2 : P.n), (const uint32_t*)(pass == 0 ? pc.perm1.p : pc.perm2.p),
             (const L*)(pass == 0 ? pc.d1.p : pc.d2.p), pass == 0 ? qs : 0, (uint64_t)P.n, dst, seg.p, nseg);
      launch("k_raa_segscan", 0, 0, (k_raa_segscan<T, W>), dim3((unsigned)rows), dim3(1024), seg.p, nseg);
      launch("k_raa_scan", rowb * 2.0, 0, (k_raa_scan<T, W>), grid, dim3(RAA_TB), dst, (uint64_t)P.n, (const C*)seg.p, nseg);
    }
  }

  // BLAKE3 root node of prefix || device bytes, hashed on the device
  // (chunk chaining values in parallel, then the left-complete merge levels).
  // With didx (device, t entries) and pidx (host address of the proof's
  // index section, t words) given and a power-of-two codeword length n, the
  // indices are expanded on the device (k_queries) and node is not filled:
  // no host round trip.
  void hash_root_node(const uint8_t* prefix, size_t plen, const void* dbytes, size_t dlen, uint32_t node[26],
                      uint32_t* didx = nullptr, uint8_t* pidx = nullptr, uint64_t t = 0, uint64_t n = 0) {
    auto expand = [&](const uint32_t* src, int parent) {
      uint32_t* dp = nullptr;
      HIP_CHECK(hipHostGetDevicePointer((void**)&dp, pidx, 0));
      launch("k_queries", 0, 0, k_queries, dim3((unsigned)((t + 255) / 256)), dim3(256), src, parent, t,
             (uint32_t)(n - 1), didx, dp);
    };
    size_t len = plen + dlen;
    DBuf<uint8_t> msg(pool, len + 64);
    h2d(msg.p, prefix, plen);
    d2d(msg.p + plen, dbytes, dlen);
    uint64_t nch = len == 0 ? 1 : (len + 1023) / 1024;
    DBuf<uint32_t> cv(pool, nch * 8), cv2(pool, (nch / 2 + 1) * 8), rn(pool, 32);
    launch("k_b3_chunks", (double)len, 0, k_b3_chunks, dim3(grid_for(nch, 64, 1u << 20)), dim3(64), (const uint8_t*)msg.p,
           (uint64_t)len, nch, cv.p, rn.p);
    if (nch == 1) {
      if (didx) expand(rn.p, 0);
      else d2h(node, rn.p, 26 * 4);
      return;
    }
    uint32_t* a = cv.p;
    uint32_t* b = cv2.p;
    uint64_t cnt = nch;
    while (cnt > (uint64_t)MERKLE_TOP) {
      launch("k_b3_merge", cnt * 32.0, 0, k_b3_merge, dim3(grid_for(cnt / 2 + 1, 64, 1u << 20)), dim3(64),
             (const uint32_t*)a, cnt, b);
      std::swap(a, b);
      cnt = (cnt + 1) / 2;
    }
    if (cnt > 2) {
      launch("k_b3_merge_top", cnt * 32.0, 0, k_b3_merge_top, dim3(1), dim3(1024), (const uint32_t*)a, cnt, b);
      std::swap(a, b);
    }
    if (didx) {
      expand(a, 1);
      return;
    }
    uint32_t kids[16];
    d2h(kids, a, 64);
    b3::iv(node);
    for (int i = 0; i < 16; i++) node[8 + i] = kids[i];
    node[24] = 64;
    node[25] = b3::PARENT;
  }

  // Shared tail of every opening: query indices (derived from BLAKE3 of
  // domain || root || m-bar), columns, Merkle paths; D2H straight into the
  // pinned proof buffer.
  void open_tail(const PcsDev<T>& pc, const CommitDev<T>& c, const L* d_mbar, size_t nmbar, Sink& out) {
    checkpoint();
    const PcsParams& P = pc.P;
    const int W = P.packed ? 32 : E;
    const char* dom = P.packed ? "pcs-ring/fs/queries" : "snark-ring/r2-pcs/fs/queries";
    std::vector<uint8_t> prefix(dom, dom + std::strlen(dom));
    prefix.insert(prefix.end(), c.root, c.root + 32);
    size_t nq = P.ell;
    DBuf<uint32_t> didx(pool, nq);
    out.u64(nq);
    if ((P.n & (P.n - 1)) == 0) {
      // indices expanded on the device, straight into the proof
      hash_root_node(prefix.data(), prefix.size(), d_mbar, nmbar * E * sizeof(L), nullptr, didx.p, out.take(nq * 4),
                     nq, P.n);
    } else {
      uint32_t node[26];
      hash_root_node(prefix.data(), prefix.size(), d_mbar, nmbar * E * sizeof(L), node);
      std::vector<uint32_t> idx = queries_from_root(node, P.n, P.ell);
      for (uint32_t j : idx) out.u32(j);
      h2d(didx.p, idx.data(), idx.size() * 4);
    }
    size_t per = c.rows * W;
    size_t colb = nq * (8 + per * sizeof(L));
    DBuf<uint8_t> cols(pool, colb);
    launch("k_gather_cols", nq * per * 2.0 * sizeof(L), 0, (k_gather_cols<T>), dim3(grid_for(nq * per, 256, 1u << 20)),
           dim3(256), (const L*)c.enc.p, (const uint32_t*)didx.p, (uint64_t)nq, (uint64_t)c.rows, (uint64_t)P.n, W,
           cols.p);
    size_t n = P.n;
    int plen = __builtin_ctzll(n);
    size_t pathb = nq * (8 + (size_t)plen * 32);
    DBuf<uint8_t> paths(pool, std::max<size_t>(pathb, 8));
    if (plen > 0)
      launch("k_merkle_paths", 0, 0, k_merkle_paths, dim3(grid_for(nq * plen, 64, 1u << 20)), dim3(64),
             (const uint32_t*)c.nodes.p, (uint64_t)n, (const uint32_t*)didx.p, (uint64_t)nq, plen, paths.p);
    out.u64(nq);
    download(out.take(colb), cols, colb);
    out.u64(nq);
    if (plen > 0) {
      download(out.take(pathb), paths, pathb);
    } else {
      for (size_t q = 0; q < nq; q++) out.u64(0);
    }
  }

  // Pcs::open / open_stacked (pcs.rs:547-591, 734-779) and R2Pcs::open /
  // open_stacked (r2_pcs.rs:297-338, 512-559).  weights: s elements or null (s=1).
  // out[l][j] = sum_{i < rows} e[i] X[l * lstride + i * k2 + j] on the matrix
  // cores: fragments of every e[i], then one pass over X.
  void mbar_mx(const L* e, uint64_t rows, const L* X, uint64_t k2, uint64_t lstride, uint64_t nl, L* out) {
    DBuf<uint32_t> fr(pool, (size_t)rows * kND<T> * 4096);
    launch("k_mulmat", rows * (EB + kND<T> * 16384.0), 0, (k_mulmat<T>), dim3((unsigned)rows), dim3(256), e, fr.p);
    const uint64_t tiles = nl * ((k2 + 31) / 32), quads = (tiles + 3) / 4;
    // split the row range eight ways (one share per XCD), and further while
    // there are fewer than ~512 blocks
    uint64_t nsplit = std::min<uint64_t>(rows, 8);
    while (nsplit * 2 <= rows && quads * nsplit < 512) nsplit *= 2;
    const uint64_t nqb = std::min<uint64_t>(quads, 65535 / nsplit);
    DBuf<L> part(pool, nsplit > 1 ? nsplit * nl * k2 * E : 1);
    launch("k_mbar_mx", (double)nl * rows * k2 * EB + nsplit * nl * k2 * EB, (double)nl * rows * k2 * E * E,
           (k_mbar_mx<T>), dim3((unsigned)(nqb * nsplit)), dim3(256), (const uint32_t*)fr.p, X, rows, k2, lstride, nl,
           nsplit, nsplit > 1 ? part.p : out);
    if (nsplit > 1)
      launch("k_sum_split", (nsplit + 1.0) * nl * k2 * EB, 0, (k_sum_split<T>), dim3(grid_for(nl * k2 * E, 256, 4096)),
             dim3(256), (const L*)part.p, nsplit, nl * k2, out);
  }

  // mbar_mx with 64-bit limbs (k_mbar_mx64): the rows are split eight ways
  // (one share of the fragments per XCD's L2) when their fragments exceed
  // 2 MB, and further while there are fewer than 1024 blocks.
  void mbar_mx64(const L* e, uint64_t rows, const L* X, uint64_t k2, uint64_t lstride, uint64_t nl, L* out) {
    DBuf<uint32_t> fr(pool, (size_t)rows * 8 * 4096);
    launch("k_mulmat64", rows * (EB + 8 * 16384.0), 0, k_mulmat64, dim3((unsigned)rows, 4), dim3(256),
           (const uint64_t*)e, fr.p);
    const uint64_t tiles = nl * ((k2 + 31) / 32);
    uint64_t nsplit = rows * 131072 > (2u << 20) ? std::min<uint64_t>(rows, 8) : 1;
    while (nsplit * 2 <= rows && tiles * nsplit < 1024) nsplit *= 2;
    const uint64_t ntb = std::min<uint64_t>(tiles, 65535 / nsplit);
    DBuf<L> part(pool, nsplit > 1 ? nsplit * nl * k2 * E : 1);
    launch("k_mbar_mx64", (double)nl * rows * k2 * EB + nsplit * nl * k2 * EB, (double)nl * rows * k2 * E * E,
           k_mbar_mx64, dim3((unsigned)(ntb * nsplit)), dim3(256), (const uint32_t*)fr.p, (const uint64_t*)X, rows, k2,
           lstride, nl, nsplit, (uint64_t*)(nsplit > 1 ? part.p : out));
    if (nsplit > 1)
      launch("k_sum_split", (nsplit + 1.0) * nl * k2 * EB, 0, (k_sum_split<T>), dim3(grid_for(nl * k2 * E, 256, 4096)),
             dim3(256), (const L*)part.p, nsplit, nl * k2, out);
  }

  G open(const PcsDev<T>& pc, const CommitDev<T>& c, const G* z, const G* weights, size_t s, Sink& out,
         bool bg_tail = false) {
    DBuf<L> y = open_dev(pc, c, z, weights, s, out, bg_tail);
    G v;
    d2h(&v, y.p, sizeof(G));
    return v;
  }
  // The opening with its value y = <m-bar, e1> left on the device (no
  // synchronization; callers that need y only later fetch it then).
  // bg_tail: the m-bar's hashing, the query columns, the paths and the
  // downloads go to the background stream (y does not depend on them).
  DBuf<L> open_dev(const PcsDev<T>& pc, const CommitDev<T>& c, const G* z, const G* weights, size_t s, Sink& out,
                   bool bg_tail = false) {
    const PcsParams& P = pc.P;
    size_t lk1 = (size_t)__builtin_ctzll(P.k1), lk2 = (size_t)__builtin_ctzll(P.k2);
    DBuf<L> e0 = eq_table(z, lk1), e1 = eq_table(z + lk1, lk2);
    DBuf<L> e0s(pool, s * P.k1 * E);
    for (size_t t = 0; t < s; t++) {
      if (!weights) d2d(e0s.p + t * P.k1 * E, e0.p, P.k1 * E * sizeof(L));
      else {
        DBuf<L> w = put(&weights[t], 1);
        scale(w.p, e0.p, e0s.p + t * P.k1 * E, P.k1);
      }
    }
    DBuf<L> mbar(pool, P.k2 * E);
    uint64_t rows = s * P.k1;
    double mbb = (rows * P.k2 + P.k2) * EB;
    if constexpr (kMx<T>) {
      mbar_mx(e0s.p, rows, c.msg, P.k2, 0, 1, mbar.p);
    } else if constexpr (kMx64<T>) {
      mbar_mx64(e0s.p, rows, c.msg, P.k2, 0, 1, mbar.p);
    } else if (P.packed)
      launch("k_mbar", mbb, rows * P.k2 * 16 * PM, (k_mbar<T, 16>), dim3(grid_for(P.k2, GPB)), dim3(GB), (const L*)e0s.p,
             c.msg, rows, (uint64_t)P.k2, mbar.p, (uint64_t)0);
    else
      launch("k_mbar", mbb, rows * P.k2 * 32 * PM, (k_mbar<T, 32>), dim3(grid_for(P.k2, GPB)), dim3(GB), (const L*)e0s.p,
             c.msg, rows, (uint64_t)P.k2, mbar.p, (uint64_t)0);
    DBuf<L> y = dots_dev<32>(mbar.p, 0, 1, e1.p, P.k2);
    out.u64(P.k2);
    uint8_t* mdst = out.take(P.k2 * E * sizeof(L));
    if (bg_tail) bg_begin();
    open_tail(pc, c, mbar.p, P.k2, out);
    download(mdst, mbar, P.k2 * E * sizeof(L));
    if (bg_tail) bg_end();
    return y;
  }

  // Pcs::rotated_open_batch (pcs.rs:973-1040) over the unfolded rho layers
  // (rho_l[q] = pack_r2(rot_l(slots_q)), the lifted messages of the rotations).
  std::vector<G> rotated_open(const PcsDev<T>& pc, const CommitDev<T>& c, const L* rho, uint64_t nb, const G* z, int nl,
                              Sink& out) {
    DBuf<L> y = rotated_open_dev(pc, c, rho, nb, z, nl, out);
    std::vector<G> ys(nl);
    d2h(ys.data(), y.p, (size_t)nl * sizeof(G));
    return ys;
  }
  // The rotated opening with its nl values left on the device.
  DBuf<L> rotated_open_dev(const PcsDev<T>& pc, const CommitDev<T>& c, const L* rho, uint64_t nb, const G* z, int nl,
                           Sink& out) {
    const PcsParams& P = pc.P;
    size_t lk1 = (size_t)__builtin_ctzll(P.k1), lk2 = (size_t)__builtin_ctzll(P.k2);
    DBuf<L> e0 = eq_table(z, lk1), e1 = eq_table(z + lk1, lk2);
    DBuf<L> mbar(pool, (size_t)nl * P.k2 * E);
    if constexpr (kMx<T>)
      mbar_mx(e0.p, P.k1, rho, P.k2, nb, nl, mbar.p);
    else if constexpr (kMx64<T>)
      mbar_mx64(e0.p, P.k1, rho, P.k2, nb, nl, mbar.p);
    else
      launch("k_rot_mbar", (double)nl * P.k1 * P.k2 * EB, (double)nl * P.k1 * P.k2 * 16 * PM, (k_rot_mbar<T>),
             dim3(grid_for((uint64_t)nl * P.k2, GPB)), dim3(GB), (const L*)e0.p, rho, nb, (uint64_t)P.k1, (uint64_t)P.k2,
             nl, mbar.p);
    DBuf<L> ys = dots_dev<32>(mbar.p, P.k2, nl, e1.p, P.k2);
    out.u64(nl);
    // the nl length-prefixed m-bars, laid out on the device, one download
    const uint64_t per = P.k2 * E * sizeof(L) / 8, recb = (size_t)nl * (per + 1) * 8;
    DBuf<uint64_t> rec(pool, (size_t)nl * (per + 1));
    launch("k_len_prefixed", 2.0 * recb, 0, k_len_prefixed, dim3(grid_for((uint64_t)nl * (per + 1), 256, 4096)),
           dim3(256), (const uint64_t*)mbar.p, (uint64_t)nl, per, (uint64_t)P.k2, rec.p);
    download(out.take(recb), rec, recb);
    open_tail(pc, c, mbar.p, (size_t)nl * P.k2, out);
    return ys;
  }

  // ---------------- sumchecks ----------------
  static G eq_one(const G& t, const G& r) { return t * r + (G::one() - t) * (G::one() - r); }

  // Gruen mult check (mult_check.rs:180-241) and its sub form (426-514).
  // tabs: a, b (, c, e) device tables of 2^nv entries; round 0 folds
  // out-of-place into `work` so the caller's tables survive.
  struct SumcheckOut {
    std::vector<G> rounds;  // 3 per round
    std::vector<G> chal;
    std::vector<G> finals;
  };
  // Rounds go through the mailbox (slot r carries round r).  Pipelined:
  // before the host waits for round r's polynomial, the stream already holds
  // round r's lift (waiting for the challenge the host posts), its folds and
  // round r+1's sums, so no stream synchronization or launch sits between a
  // round's values and the next round's work.  Otherwise the host posts the
  // challenge first and then enqueues the round's work.
  SumcheckOut mult_check(const L* const* tabs, int ntab, size_t nv, const std::vector<G>& rt, Transcript& tr) {
    SumcheckOut o;
    if (nv > (size_t)MB_SLOTS) throw Error(NENYA_E_ARG, "too many sumcheck rounds");
    size_t N = size_t(1) << nv;
    G one = G::one();
    DBuf<L> eg = nv > 1 ? eq_table(rt.data() + 1, nv - 1) : put(&one, 1);
    DBuf<L> work(pool, (size_t)ntab * (N / 2 > 0 ? N / 2 : 1) * E);
    // challenges and the point rt stay on the device
    DBuf<L> drt = put(rt.data(), nv), dch(pool, nv * E), dsc(pool, E);
    size_t len = N;
    const L* cur[4];
    for (int t = 0; t < ntab; t++) cur[t] = tabs[t];
    // everything the pipelined rounds use is allocated before the first
    // waiting kernel is enqueued (no allocator call can then wait on them)
    auto blocks = [](size_t m) -> unsigned {
      // small rounds are latency-bound: spread each element's three sums over three blocks
      return m <= 4096 ? 3 * grid_for(m, GPB, 682) : grid_for(m, GPB, 2048);
    };
    unsigned nbk_max = 1;
    for (size_t m = N / 2; m >= 1; m /= 2) nbk_max = std::max(nbk_max, blocks(m));
    DBuf<L> part(pool, (size_t)std::max<unsigned>(nbk_max, GRAM_BLOCKS) * 3 * E);
    // the rounds' element products (matrix-core form of the sums)
    DBuf<L> prods(pool, kFrag<T> && N / 2 >= MULT_MX_MIN ? 3 * (N / 2) * E : 1);
    DBuf<uint32_t> frags(pool, kFrag<T> ? nv * kND<T> * 4096 : 1);
    MbGuard guard(*this);
    const bool pipe = pipeline_rounds();
    // contexts that do not pipeline: Fiat-Shamir of the rounds on the device
    const bool use_hub = hub && !pipe;
    const bool dfs = !pipe && (devfs || use_hub);
    DBuf<L> drounds(pool, dfs ? 3 * nv * E : 1);
    DBuf<b3::State> dtr(pool, 1);
    DBuf<C> hsums(pool, use_hub ? (size_t)SUM_COPIES * 3 * E : 1);
    size_t hub_from = nv;  // first round run on the hub
    // tail rounds: their sums accumulate in tsums, tcount finds the last block
    DBuf<C> tsums(pool, (size_t)SUM_COPIES * 3 * E);
    DBuf<uint32_t> tcount(pool, 1);
    HIP_CHECK(hipMemsetAsync(tsums.p, 0, (size_t)SUM_COPIES * 3 * E * sizeof(C), st));
    HIP_CHECK(hipMemsetAsync(tcount.p, 0, sizeof(uint32_t), st));
    if (dfs) h2d(dtr.p, static_cast<const b3::State*>(&tr.h), sizeof(b3::State));
    wait_ticks = spin_ticks();
    // enqueue the sums of round r over tables of 2m entries; returns the flag value
    auto round_sums = [&](size_t r, size_t m) -> uint32_t {
      const bool split = m <= 4096;
      const unsigned nbk = blocks(m);
      const L* c3 = ntab > 2 ? cur[2] : (const L*)nullptr;
      const L* e3 = ntab > 2 ? cur[3] : (const L*)nullptr;
      if (kFrag<T> && m >= MULT_MX_MIN) {
        // the element products P_s on the vector ALUs, then the three sums
        // sum_i eg_i P_s[i] as Gram contractions on the matrix cores
        launch("k_mult_prod", m * (2.0 * ntab + 3) * EB, m * 3 * 32 * PM, (k_mult_prod<T>),
               dim3(grid_for(3 * m, GPB, 4096)), dim3(GB), cur[0], cur[1], c3, e3, (uint64_t)m, prods.p);
        const unsigned nwg = (unsigned)std::min<uint64_t>((m + 31) / 32, GRAM_BLOCKS);
        for (int s = 0; s < 3; s++) {
          if constexpr (kMx64<T>)
            launch("k_mult_gram64", m * 2.0 * EB, m * 1.0 * E * E, k_gram_lin64, dim3(nwg), dim3(512),
                   (const uint64_t*)eg.p, (uint64_t)0, (const uint64_t*)(prods.p + (size_t)s * m * E), (uint64_t)0,
                   (uint64_t)1, (uint64_t)m, (uint64_t)0, 128, (uint64_t*)part.p, 3, s);
          else
            launch("k_mult_gram", m * 2.0 * EB, m * 1.0 * E * E, (k_gram_lin2<T>), dim3(nwg), dim3(512), (const L*)eg.p,
                   (uint64_t)0, (const L*)(prods.p + (size_t)s * m * E), (uint64_t)0, (uint64_t)1, (uint64_t)m,
                   (uint64_t)0, 128, part.p, 3, s);
        }
        if (dfs) {
          post_round_dev(part.p, nwg, 3, 0, nullptr, nullptr, drounds.p + r * 3 * E);
          return 0u;
        }
        return post_round(part.p, nwg, 3, 0, nullptr, nullptr, (int)r);
      }
      if (split)
        launch("k_mult_round", m * (2.0 * ntab + 1) * EB, m * 6 * 32 * PM, (k_mult_round<T, true>), dim3(nbk),
               dim3(GB), (const L*)eg.p, cur[0], cur[1], c3, e3, (uint64_t)m, part.p);
      else
        launch("k_mult_round", m * (2.0 * ntab + 1) * EB, m * 6 * 32 * PM, (k_mult_round<T, false>), dim3(nbk),
               dim3(GB), (const L*)eg.p, cur[0], cur[1], c3, e3, (uint64_t)m, part.p);
      if (dfs) {
        post_round_dev(part.p, nbk, 3, 0, nullptr, nullptr, drounds.p + r * 3 * E);
        return 0u;
      }
      return post_round(part.p, nbk, 3, 0, nullptr, nullptr, (int)r);
    };
    uint32_t vdown = round_sums(0, len / 2);
    for (size_t r = 0; r < nv; r++) {
      size_t m = len / 2;
      L* rc = dch.p + r * E;
      MulMat mr;
      const uint32_t vup = ++mseq;
      // host side of round r: its polynomial, the transcript, the challenge
      auto host_round = [&] {
        checkpoint();
        G q[3];
        wait_round((int)r, vdown, q);
        tr.elems("sc/round", q, 3);
        for (int i = 0; i < 3; i++) o.rounds.push_back(q[i]);
        post_challenge(tr, "sc/challenge", (int)r, vup);
      };
      // a small round: its fold, the marginalization and round r+1's sums
      // and post in one tail kernel (the fold needs no fragments)
      const bool tail = !use_hub && (uint64_t)m * ntab < TAIL_FOLD_MAX && m / 2 <= TAIL_MULT_BLOCKS && r + 1 < nv;
      uint32_t* rfrag = tail ? nullptr : round_frag(frags.p, r, (uint64_t)m * ntab);
      if (dfs) {
        checkpoint();
        lift_device_fs(dtr.p, drounds.p + r * 3 * E, rc, rfrag, &mr);
      } else {
        if (!pipe) host_round();
        lift_posted((int)r, vup, rc, rfrag, &mr);
      }
      if (tail) {
        TailM<T> A{};
        for (int t = 0; t < ntab; t++) {
          A.src[t] = cur[t];
          A.dst[t] = work.p + (size_t)t * (N / 2) * E;
        }
        A.m = m;
        A.rc = rc;
        A.rt = drt.p + r * E;
        A.eg = eg.p;
        A.sums = tsums.p;
        A.count = tcount.p;
        uint32_t vnext = 0;
        if (dfs) {
          A.qh = drounds.p + (r + 1) * 3 * E;
        } else {
          vnext = ++mseq;
          A.qh = mbd<L>(mb_down((int)r + 1));
          A.flag = mbd<uint32_t>(mb_down_flag((int)r + 1));
          A.val = vnext;
        }
        const double tb = (m * ntab * 3.0 + m * 1.5) * EB;
        // one block per pair of round r+1
        if (ntab == 4)
          launch("k_tail_mult", tb, m * (2.0 * ntab + 4) * 32 * PM, (k_tail_mult<T, 4>), dim3((unsigned)(m / 2)),
                 dim3(32 * 9), A);
        else
          launch("k_tail_mult", tb, m * (2.0 * ntab + 4) * 32 * PM, (k_tail_mult<T, 2>), dim3((unsigned)(m / 2)),
                 dim3(32 * 5), A);
        for (int t = 0; t < ntab; t++) cur[t] = work.p + (size_t)t * (N / 2) * E;
        len = m;
        if (pipe) host_round();
        vdown = vnext;
        continue;
      }
      // the next round's eq factor eq_one(rt[r], r_r) rides along with the fold
      Side sd;
      if (r + 1 < nv) sd = Side{1, drt.p + r * E, rc, dsc.p};
      // tables at a common stride fold in one launch
      bool uniform = true;
      uniform = ntab > 1 && cur[1] > cur[0] && (cur[1] - cur[0]) % E == 0;
      for (int t = 2; t < ntab; t++) uniform = uniform && cur[t] == cur[0] + (cur[1] - cur[0]) * t;
      if (uniform && ntab > 1) {
        fold(mr, cur[0], work.p, m, ntab, (uint64_t)(cur[1] - cur[0]) / E, N / 2, sd);
      } else {
        for (int t = 0; t < ntab; t++) fold(mr, cur[t], work.p + (size_t)t * (N / 2) * E, m, 1, 0, 0, t ? Side{} : sd);
      }
      for (int t = 0; t < ntab; t++) cur[t] = work.p + (size_t)t * (N / 2) * E;
      len = m;
      uint32_t vnext = 0;
      if (r + 1 < nv) {
        marg(dsc.p, eg.p, m / 2);
        if (use_hub && (m / 2) * ntab < HUB_FOLD_MAX) {
          hub_from = r + 1;
          break;
        }
        vnext = round_sums(r + 1, m / 2);
      }
      if (pipe) host_round();
      vdown = vnext;
    }
    if (hub_from < nv) {
      checkpoint();
      HubTask<T> t;
      t.type = 0;
      t.r = (int)hub_from;
      t.nv = (int)nv;
      t.m = len / 2;
      t.ntab = (uint64_t)ntab;
      t.work = work.p;
      t.stride = N / 2;
      t.eg = eg.p;
      t.drt = drt.p;
      t.dsc = dsc.p;
      t.dch = dch.p;
      t.drounds = drounds.p;
      t.dtr = dtr.p;
      t.sums = hsums.p;
      t.frob = d_frob.p;
      t.cancel = cancel;
      hub_events();
      t.ready = hev[0];
      t.done = hev[1];
      HIP_CHECK(hipMemsetAsync(hsums.p, 0, (size_t)SUM_COPIES * 3 * E * sizeof(C), st));
      HIP_CHECK(hipEventRecord(t.ready, st));
      hub_run<T>(hub, t, st);
    }
    // challenges and finals (and, with the device transcript, the round
    // polynomials and the transcript state) with one synchronization
    if (dfs) fetch_transcript_async(dtr.p);
    std::vector<G> tail =
        pick_many({{dch.p, 1, nv}, {work.p, N / 2, (uint64_t)ntab}, {drounds.p, 1, dfs ? 3 * (uint64_t)nv : 0}});
    check_waits();
    o.chal.assign(tail.begin(), tail.begin() + nv);
    o.finals.assign(tail.begin() + nv, tail.begin() + nv + ntab);
    if (dfs) {
      o.rounds.assign(tail.begin() + nv + ntab, tail.end());
      adopt_transcript(tr);
    }
    return o;
  }

  // {0,1,inf} linear sumcheck over `ntab` pairs (f_t, g_t)
  // (linear_check_simd.rs:196-286 for ntab = 1, linear_check_general.rs:214-350
  // for ntab = 32).  Round 0 reads the caller's tables and folds into fresh
  // buffers, so both inputs survive.  f0 may be low-degree (NAF0 = 16).
  template <int NAF0>
  SumcheckOut linear(const L* f0, uint64_t fstride0, const L* g0, uint64_t gstride0, int ntab, size_t nv, G claim,
                     Transcript& tr) {
    SumcheckOut o;
    if (nv > (size_t)MB_SLOTS) throw Error(NENYA_E_ARG, "too many sumcheck rounds");
    size_t N = size_t(1) << nv;
    size_t H = N / 2 > 0 ? N / 2 : 1;
    DBuf<L> fw(pool, (size_t)ntab * H * E), gw(pool, (size_t)ntab * H * E);
    const L* f = f0;
    const L* g = g0;
    uint64_t fs = fstride0, gs = gstride0;
    // the running claim and the challenges stay on the device (pipelined
    // through the mailbox as in mult_check)
    DBuf<L> dc = put(&claim, 1), dq(pool, 3 * E), dch(pool, nv * E);
    size_t len = N;
    // allocated before the first waiting kernel (see mult_check)
    const size_t nparts = kFrag<T> ? GRAM_BLOCKS_MAX : grid_for((uint64_t)(N / 2 > 0 ? N / 2 : 1) * ntab, GPB, 2048);
    DBuf<L> part(pool, nparts * 2 * E);
    DBuf<uint32_t> frags(pool, kFrag<T> ? nv * kND<T> * 4096 : 1);
    MbGuard guard(*this);
    const bool pipe = pipeline_rounds();
    const bool use_hub = hub && !pipe;  // as in mult_check
    const bool dfs = !pipe && (devfs || use_hub);
    DBuf<L> drounds(pool, dfs ? 3 * nv * E : 1);
    DBuf<b3::State> dtr(pool, 1);
    DBuf<C> hsums(pool, use_hub ? (size_t)SUM_COPIES * 3 * E : 1);
    size_t hub_from = nv;
    DBuf<C> tsums(pool, (size_t)SUM_COPIES * 3 * E);  // tail rounds (as in mult_check)
    DBuf<uint32_t> tcount(pool, 1);
    HIP_CHECK(hipMemsetAsync(tsums.p, 0, (size_t)SUM_COPIES * 3 * E * sizeof(C), st));
    HIP_CHECK(hipMemsetAsync(tcount.p, 0, sizeof(uint32_t), st));
    if (dfs) h2d(dtr.p, static_cast<const b3::State*>(&tr.h), sizeof(b3::State));
    wait_ticks = spin_ticks();
    auto post = [&](size_t r, unsigned nparts) -> uint32_t {
      if (dfs) {
        post_round_dev(part.p, nparts, 2, 1, dc.p, dq.p, drounds.p + r * 3 * E);
        return 0u;
      }
      return post_round(part.p, nparts, 2, 1, dc.p, dq.p, (int)r);
    };
    auto round_sums = [&](size_t r, size_t m) -> uint32_t {
      if constexpr (kMx<T>) {
        const unsigned nwg = gram_lin(f, fs, g, gs, (uint64_t)ntab, (uint64_t)m, (r == 0 && NAF0 == 16) ? 64 : 128, part.p);
        return post(r, nwg);
      } else if constexpr (kMx64<T>) {
        const unsigned nwg = gram_lin64(f, fs, g, gs, (uint64_t)ntab, (uint64_t)m, (r == 0 && NAF0 == 16) ? 64 : 128,
                                        part.p);
        return post(r, nwg);
      } else {
        const unsigned nbk = grid_for(m * ntab, GPB, 2048);
        double lb = m * ntab * 4.0 * EB;
        if (r == 0 && NAF0 == 16)
          launch("k_lin_round", lb, m * ntab * 2 * 16 * PM, (k_lin_round<T, 16>), dim3(nbk), dim3(GB), f, fs, g, gs,
                 (uint64_t)ntab, (uint64_t)m, part.p);
        else
          launch("k_lin_round", lb, m * ntab * 2 * 32 * PM, (k_lin_round<T, 32>), dim3(nbk), dim3(GB), f, fs, g, gs,
                 (uint64_t)ntab, (uint64_t)m, part.p);
        return post(r, nbk);
      }
    };
    uint32_t vdown = round_sums(0, len / 2);
    for (size_t r = 0; r < nv; r++) {
      size_t m = len / 2;
      L* rc = dch.p + r * E;
      MulMat mr;
      const uint32_t vup = ++mseq;
      auto host_round = [&] {
        checkpoint();
        G q[3];
        wait_round((int)r, vdown, q);
        tr.elems("sc/round", q, 3);
        for (int i = 0; i < 3; i++) o.rounds.push_back(q[i]);
        post_challenge(tr, "sc/challenge", (int)r, vup);
      };
      const bool tail =
          !use_hub && 2 * m * (uint64_t)ntab < TAIL_FOLD_MAX && (uint64_t)ntab * (m / 2) <= TAIL_LIN_BLOCKS && r + 1 < nv;
      uint32_t* rfrag = tail ? nullptr : round_frag(frags.p, r, (uint64_t)m * 2 * ntab);
      if (dfs) {
        checkpoint();
        lift_device_fs(dtr.p, drounds.p + r * 3 * E, rc, rfrag, &mr);
      } else {
        if (!pipe) host_round();
        lift_posted((int)r, vup, rc, rfrag, &mr);
      }
      if (tail) {
        TailL<T> A{};
        A.f = f;
        A.gt = g;
        A.fs = fs;
        A.gs = gs;
        A.fw = fw.p;
        A.gw = gw.p;
        A.H = H;
        A.ntab = (uint64_t)ntab;
        A.m = m;
        A.rc = rc;
        A.dq = dq.p;
        A.dc = dc.p;
        A.sums = tsums.p;
        A.count = tcount.p;
        uint32_t vnext = 0;
        if (dfs) {
          A.qh = drounds.p + (r + 1) * 3 * E;
        } else {
          vnext = ++mseq;
          A.qh = mbd<L>(mb_down((int)r + 1));
          A.flag = mbd<uint32_t>(mb_down_flag((int)r + 1));
          A.val = vnext;
        }
        // one block per (table, pair) of round r+1
        launch("k_tail_lin", m * ntab * 6.0 * EB, m * ntab * 5.0 * 32 * PM, (k_tail_lin<T>),
               dim3((unsigned)((uint64_t)ntab * (m / 2))), dim3(128), A);
        f = fw.p;
        g = gw.p;
        fs = gs = H;
        len = m;
        if (pipe) host_round();
        vdown = vnext;
        continue;
      }
      // the next claim q_r(r_r) rides along with the folds
      fold2(mr, m, f, fw.p, ntab, fs, H, g, gw.p, ntab, gs, H, Side{2, dq.p, rc, dc.p});
      f = fw.p;
      g = gw.p;
      fs = gs = H;
      len = m;
      if (r + 1 < nv && use_hub && 2 * (m / 2) * (uint64_t)ntab < HUB_FOLD_MAX) {
        hub_from = r + 1;
        break;
      }
      const uint32_t vnext = r + 1 < nv ? round_sums(r + 1, m / 2) : 0;
      if (pipe) host_round();
      vdown = vnext;
    }
    if (hub_from < nv) {
      checkpoint();
      HubTask<T> t;
      t.type = 1;
      t.r = (int)hub_from;
      t.nv = (int)nv;
      t.m = len / 2;
      t.ntab = (uint64_t)ntab;
      t.fw = fw.p;
      t.gw = gw.p;
      t.H = H;
      t.dc = dc.p;
      t.dq = dq.p;
      t.dch = dch.p;
      t.drounds = drounds.p;
      t.dtr = dtr.p;
      t.sums = hsums.p;
      t.frob = d_frob.p;
      t.cancel = cancel;
      hub_events();
      t.ready = hev[0];
      t.done = hev[1];
      HIP_CHECK(hipMemsetAsync(hsums.p, 0, (size_t)SUM_COPIES * 3 * E * sizeof(C), st));
      HIP_CHECK(hipEventRecord(t.ready, st));
      hub_run<T>(hub, t, st);
    }
    if (dfs) fetch_transcript_async(dtr.p);
    std::vector<G> tail = pick_many({{dch.p, 1, nv}, {f, fs, (uint64_t)ntab}, {g, gs, (uint64_t)ntab},
                                     {drounds.p, 1, dfs ? 3 * (uint64_t)nv : 0}});
    check_waits();
    o.chal.assign(tail.begin(), tail.begin() + nv);
    o.finals.assign(tail.begin() + nv, tail.begin() + nv + 2 * (size_t)ntab);
    if (dfs) {
      o.rounds.assign(tail.begin() + nv + 2 * (size_t)ntab, tail.end());
      adopt_transcript(tr);
    }
    return o;
  }
};

}  // namespace nenya

#include "hub.hpp"  // the round hub's kernels and thread (uses the declarations above)
