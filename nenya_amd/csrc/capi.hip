// extern "C" boundary of the MI355X Nenya prover (include/nenya.h).
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <unordered_map>

#include "../../include/nenya.h"
#include "container.hpp"
#include "server.hpp"
#include "tune.hpp"
#include "verify.hpp"

using namespace nenya;

struct nenya_params {
  ParamsBase* impl = nullptr;
  std::mutex mu;
};

static thread_local std::string g_err;

template <class F> static int guard(F&& f) {
  try {
    g_err.clear();
    return f();
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return NENYA_E_OOM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return NENYA_E_DEVICE;
  }
}

static void need_gpu() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    throw Error(NENYA_E_NOGPU, "no HIP device visible");
  }
}
static void need_width(int k) {
  if (k != 8 && k != 16 && k != 32 && k != 64) throw Error(NENYA_E_ARG, "k must be 8, 16, 32 or 64");
}
// Owned proofs: the proof is assembled straight in a pinned buffer that is
// then handed to the caller (no copy); nenya_free takes it back into a small
// pool of spare buffers, so a caller that frees each proof before the next
// one allocates nothing in steady state.
struct Handout {
  std::mutex mu;
  std::unordered_map<const void*, Pinned> out;  // buffers the callers hold
  std::vector<Pinned> spare;                    // returned ones, largest last
  static constexpr size_t SPARE_MAX = 4;
  static Handout& get() {
    static Handout* h = new Handout;  // never destroyed: callers may free at exit
    return *h;
  }
  Pinned take() {
    std::lock_guard<std::mutex> g(mu);
    if (spare.empty()) return Pinned();
    Pinned b = std::move(spare.back());
    spare.pop_back();
    return b;
  }
  uint8_t* give(Pinned&& b) {
    std::lock_guard<std::mutex> g(mu);
    uint8_t* p = b.p;
    out.emplace(p, std::move(b));
    return p;
  }
  // true when p was one of ours
  bool back(const void* p) {
    std::lock_guard<std::mutex> g(mu);
    auto it = out.find(p);
    if (it == out.end()) return false;
    spare.push_back(std::move(it->second));
    out.erase(it);
    std::sort(spare.begin(), spare.end(), [](const Pinned& a, const Pinned& b) { return a.cap < b.cap; });
    if (spare.size() > SPARE_MAX) spare.erase(spare.begin());  // drops (frees) the smallest
    return true;
  }
};

// Runs `prove` (which returns a view into the context's proof buffer) with
// a handout buffer swapped in as that buffer, then gives the buffer to the
// caller as an owned proof.
template <class F> static void prove_owned(CtxBase& cx, F&& prove, uint8_t** proof, size_t* len) {
  Pinned b = Handout::get().take();
  std::swap(cx.proofbuf, b);
  struct Restore {
    CtxBase& cx;
    Pinned& b;
    ~Restore() { std::swap(cx.proofbuf, b); }
  };
  ProofView v;
  {
    Restore r{cx, b};
    v = prove();
  }
  if (v.first != b.p) throw Error(NENYA_E_DEVICE, "proof not in its buffer");
  *len = v.second;
  *proof = Handout::get().give(std::move(b));
}

template <class T>
static ProofView simd_prove_t(ParamsBase* b, const HostCoo* m3, const void* wit) {
  auto* sp = static_cast<SimdParams<T>*>(b);
  return sp->prove(sp->cx0, m3[0], m3[1], m3[2], (const typename T::L*)wit);
}

template <class T> static ProofView mimd_t(ParamsBase* b, const void* wit) {
  auto* mp = static_cast<MimdParams<T>*>(b);
  return mp->prove(mp->cx0, (const typename T::L*)wit);
}

template <class T> static ProofView general_t(ParamsBase* b, const void* w, size_t n, bool dev) {
  auto* gp = static_cast<GeneralParams<T>*>(b);
  if (dev) {
    if (n != gp->N) throw Error(NENYA_E_SHAPE, "witness length must be 32 * 2^log_blocks");
    return gp->prove_device(gp->cx0, (const typename T::L*)w);
  }
  return gp->prove_host(gp->cx0, (const typename T::L*)w, n);
}

static int general_common(const nenya_params* p, const void* w, size_t n, uint8_t** proof, size_t* len, bool dev,
                          const uint8_t** view = nullptr) {
  return guard([&] {
    if (!p || !p->impl || p->impl->kind != 1 || !w || !len || (!proof && !view)) throw Error(NENYA_E_ARG, "bad argument");
    auto* np = const_cast<nenya_params*>(p);
    std::lock_guard<std::mutex> g(np->mu);
    CallTurn turn(true);
    auto run = [&]() -> ProofView {
      switch (p->impl->k) {
        case 8: return general_t<W<8>>(p->impl, w, n, dev);
        case 16: return general_t<W<16>>(p->impl, w, n, dev);
        case 32: return general_t<W<32>>(p->impl, w, n, dev);
        default: return general_t<W<64>>(p->impl, w, n, dev);
      }
    };
    if (view) {
      ProofView v = run();
      *view = v.first;
      *len = v.second;
    } else {
      prove_owned(p->impl->ctx0(), run, proof, len);
    }
    return NENYA_OK;
  });
}

// ---------------- no-alloc General proofs ----------------
// Device scratch of nenya_prove_general_device_noalloc: every device buffer
// of such a proof is carved out of this array (Arena, first fit).  Each GPU
// the library runs on holds its own copy of it until the process ends.  In
// host code `scratch` only names the symbol: its device address on the
// current GPU comes from hipGetSymbolAddress (NoAllocDev::ptr), and only
// that address is ever handed to kernels or runtime calls.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(256) << 20;
constexpr int NOALLOC_MAX_LOG_BLOCKS = 10;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

struct NoAllocDev {
  std::mutex mu;         // calls on one GPU share its scratch: one at a time
  void* ptr = nullptr;   // device address of `scratch` on this GPU
  Arena arena;
};
// The state of GPU `dev`; the first call on a GPU (which is then the current
// device) looks up the scratch address there, once.
static NoAllocDev* noalloc_dev(int dev) {
  static std::mutex mu;
  static NoAllocDev* devs[64] = {};  // kept until the process ends, like the arrays
  if (dev < 0 || dev >= 64) throw Error(NENYA_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> g(mu);
  if (!devs[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      throw Error(NENYA_E_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    auto* d = new NoAllocDev;
    d->ptr = ptr;
    devs[dev] = d;
  }
  return devs[dev];
}

// prove_device with every device buffer taken from the GPU's scratch, one
// stream, and the pinned areas the params reserved at setup.
template <class T> static ProofView general_noalloc_t(ParamsBase* b, NoAllocDev* d, const void* w) {
  auto* gp = static_cast<GeneralParams<T>*>(b);
  Ctx<T>& cx = gp->cx0;
  Engine<T>& e = *cx.eng;
  e.finish_downloads();  // an earlier proof's copies release their (non-scratch) blocks
  d->arena.reset(d->ptr, NOALLOC_SCRATCH_BYTES);
  struct Scope {
    Ctx<T>& cx;
    bool prof, timing;
    int n0 = std::uncaught_exceptions();
    Scope(Ctx<T>& c, Arena* a) : cx(c), prof(c.prof.on), timing(c.timing) {
      cx.pool.arena = a;
      cx.eng->serial = true;
      pinned_no_grow() = true;
      cx.prof.on = false;
      cx.timing = false;
    }
    ~Scope() {
      // a failed proof may leave kernels on the stream that use scratch blocks
      if (std::uncaught_exceptions() > n0) (void)hipStreamSynchronize(cx.st);
      cx.pool.arena = nullptr;
      cx.eng->serial = false;
      pinned_no_grow() = false;
      cx.prof.on = prof;
      cx.timing = timing;
    }
  } scope(cx, &d->arena);
  ProofView v = gp->prove_device(cx, (const typename T::L*)w);
  if (d->arena.used != 0) throw Error(NENYA_E_DEVICE, "scratch blocks outlived a no-alloc proof");
  return v;
}

template <class F> static int verify_common(const nenya_params* p, int kind, int* verdict, F&& f) {
  return guard([&] {
    if (!p || !p->impl || p->impl->kind != kind || !verdict) throw Error(NENYA_E_ARG, "bad argument");
    auto* np = const_cast<nenya_params*>(p);
    std::lock_guard<std::mutex> g(np->mu);
    CallTurn turn(false);
    *verdict = f(p->impl);
    return NENYA_OK;
  });
}

extern "C" {

int nenya_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

const char* nenya_last_error(void) { return g_err.c_str(); }

int nenya_device_copy(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    if ((!dst || !src) && bytes) throw Error(NENYA_E_ARG, "bad argument");
    need_gpu();
    if (!bytes) return NENYA_OK;
    CallTurn turn(false);
    const unsigned grid = (unsigned)std::min<size_t>(2048, (bytes + 4095) / 4096);
    hipLaunchKernelGGL(k_copy_out, dim3(grid), dim3(256), 0, nullptr, (const uint8_t*)src, (uint8_t*)dst,
                       (uint64_t)bytes);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    return NENYA_OK;
  });
}

void nenya_free(void* p) {
  if (!p) return;
  CallTurn turn(false);  // may free a pinned buffer
  if (!Handout::get().back(p)) std::free(p);
}

void nenya_params_free(nenya_params* p) {
  if (!p) return;
  CallTurn turn(false);
  delete p->impl;
  delete p;
}

int nenya_params_simd_secure(int k, int log_n, int log_m, uint64_t seed_w, uint64_t seed_q, nenya_params** out) {
  return guard([&] {
    if (!out) throw Error(NENYA_E_ARG, "null out");
    need_width(k);
    if (log_n < 2 || log_m < 2 || log_n > 30 || log_m > 30) throw Error(NENYA_E_SHAPE, "log_n/log_m out of range");
    need_gpu();
    CallTurn turn(false);
    auto* p = new nenya_params;
    try {
      switch (k) {
        case 8: p->impl = new SimdParams<W<8>>(log_n, log_m, seed_w, seed_q); break;
        case 16: p->impl = new SimdParams<W<16>>(log_n, log_m, seed_w, seed_q); break;
        case 32: p->impl = new SimdParams<W<32>>(log_n, log_m, seed_w, seed_q); break;
        default: p->impl = new SimdParams<W<64>>(log_n, log_m, seed_w, seed_q); break;
      }
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
    return NENYA_OK;
  });
}

static int simd_common(const nenya_params* p, const uint32_t* ar, const uint32_t* ac, const void* av, size_t an,
                       const uint32_t* br, const uint32_t* bc, const void* bv, size_t bn, const uint32_t* cr,
                       const uint32_t* cc, const void* cv, size_t cn, const void* witnesses, uint8_t** proof,
                       size_t* proof_len, const uint8_t** view) {
  return guard([&] {
    if (!p || !p->impl || p->impl->kind != 0 || !witnesses || (!proof && !view) || !proof_len)
      throw Error(NENYA_E_ARG, "bad argument");
    auto* np = const_cast<nenya_params*>(p);
    std::lock_guard<std::mutex> g(np->mu);
    CallTurn turn(true);
    ParamsBase* b = p->impl;
    int ln = 0, lm = 0;
    switch (b->k) {
      case 8: ln = static_cast<SimdParams<W<8>>*>(b)->log_n; lm = static_cast<SimdParams<W<8>>*>(b)->log_m; break;
      case 16: ln = static_cast<SimdParams<W<16>>*>(b)->log_n; lm = static_cast<SimdParams<W<16>>*>(b)->log_m; break;
      case 32: ln = static_cast<SimdParams<W<32>>*>(b)->log_n; lm = static_cast<SimdParams<W<32>>*>(b)->log_m; break;
      default: ln = static_cast<SimdParams<W<64>>*>(b)->log_n; lm = static_cast<SimdParams<W<64>>*>(b)->log_m; break;
    }
    size_t n = size_t(1) << ln, m = size_t(1) << lm;
    if ((an && (!ar || !ac || !av)) || (bn && (!br || !bc || !bv)) || (cn && (!cr || !cc || !cv)))
      throw Error(NENYA_E_ARG, "null matrix array");
    HostCoo m3[3] = {{n, m, an, ar, ac, av}, {n, m, bn, br, bc, bv}, {n, m, cn, cr, cc, cv}};
    auto run = [&]() -> ProofView {
      switch (b->k) {
        case 8: return simd_prove_t<W<8>>(b, m3, witnesses);
        case 16: return simd_prove_t<W<16>>(b, m3, witnesses);
        case 32: return simd_prove_t<W<32>>(b, m3, witnesses);
        default: return simd_prove_t<W<64>>(b, m3, witnesses);
      }
    };
    if (view) {
      ProofView v = run();
      *view = v.first;
      *proof_len = v.second;
    } else {
      prove_owned(b->ctx0(), run, proof, proof_len);
    }
    return NENYA_OK;
  });
}

int nenya_prove_simd(const nenya_params* p, const uint32_t* ar, const uint32_t* ac, const void* av, size_t an,
                     const uint32_t* br, const uint32_t* bc, const void* bv, size_t bn, const uint32_t* cr,
                     const uint32_t* cc, const void* cv, size_t cn, const void* witnesses, uint8_t** proof,
                     size_t* proof_len) {
  return simd_common(p, ar, ac, av, an, br, bc, bv, bn, cr, cc, cv, cn, witnesses, proof, proof_len, nullptr);
}

int nenya_prove_simd_view(const nenya_params* p, const uint32_t* ar, const uint32_t* ac, const void* av, size_t an,
                          const uint32_t* br, const uint32_t* bc, const void* bv, size_t bn, const uint32_t* cr,
                          const uint32_t* cc, const void* cv, size_t cn, const void* witnesses, const uint8_t** proof,
                          size_t* proof_len) {
  return simd_common(p, ar, ac, av, an, br, bc, bv, bn, cr, cc, cv, cn, witnesses, nullptr, proof_len, proof);
}

int nenya_params_mimd_secure(int k, int log_n, int log_m, const size_t* a_off, const uint32_t* ar, const uint32_t* ac,
                             const void* av, const size_t* b_off, const uint32_t* br, const uint32_t* bc, const void* bv,
                             const size_t* c_off, const uint32_t* cr, const uint32_t* cc, const void* cv,
                             uint64_t seed_w, uint64_t seed_e, nenya_params** out) {
  return guard([&] {
    if (!out || !a_off || !b_off || !c_off) throw Error(NENYA_E_ARG, "null argument");
    need_width(k);
    if (log_n < 1 || log_n > 26 || log_m < 1 || log_m > 26) throw Error(NENYA_E_SHAPE, "log_n / log_m out of range");
    const size_t* offs[3] = {a_off, b_off, c_off};
    const uint32_t* rs[3] = {ar, br, cr};
    const uint32_t* cs[3] = {ac, bc, cc};
    const void* vs[3] = {av, bv, cv};
    MimdSet sets[3];
    for (int w = 0; w < 3; w++) {
      if (offs[w][0] != 0) throw Error(NENYA_E_ARG, "instance offsets must start at 0");
      if (offs[w][32] && (!rs[w] || !cs[w] || !vs[w])) throw Error(NENYA_E_ARG, "null matrix array");
      sets[w] = {offs[w], rs[w], cs[w], vs[w]};
    }
    need_gpu();
    CallTurn turn(false);
    auto* p = new nenya_params;
    try {
      switch (k) {
        case 8: p->impl = new MimdParams<W<8>>(log_n, log_m, sets, seed_w, seed_e); break;
        case 16: p->impl = new MimdParams<W<16>>(log_n, log_m, sets, seed_w, seed_e); break;
        case 32: p->impl = new MimdParams<W<32>>(log_n, log_m, sets, seed_w, seed_e); break;
        default: p->impl = new MimdParams<W<64>>(log_n, log_m, sets, seed_w, seed_e); break;
      }
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
    return NENYA_OK;
  });
}

static int mimd_common(const nenya_params* p, const void* witnesses, uint8_t** proof, size_t* len,
                       const uint8_t** view) {
  return guard([&] {
    if (!p || !p->impl || p->impl->kind != 2 || !witnesses || (!proof && !view) || !len)
      throw Error(NENYA_E_ARG, "bad argument");
    auto* np = const_cast<nenya_params*>(p);
    std::lock_guard<std::mutex> g(np->mu);
    CallTurn turn(true);
    auto run = [&]() -> ProofView {
      switch (p->impl->k) {
        case 8: return mimd_t<W<8>>(p->impl, witnesses);
        case 16: return mimd_t<W<16>>(p->impl, witnesses);
        case 32: return mimd_t<W<32>>(p->impl, witnesses);
        default: return mimd_t<W<64>>(p->impl, witnesses);
      }
    };
    if (view) {
      ProofView v = run();
      *view = v.first;
      *len = v.second;
    } else {
      prove_owned(p->impl->ctx0(), run, proof, len);
    }
    return NENYA_OK;
  });
}

int nenya_prove_mimd(const nenya_params* p, const void* witnesses, uint8_t** proof, size_t* len) {
  return mimd_common(p, witnesses, proof, len, nullptr);
}

int nenya_prove_mimd_view(const nenya_params* p, const void* witnesses, const uint8_t** proof, size_t* len) {
  return mimd_common(p, witnesses, nullptr, len, proof);
}

int nenya_params_general_secure(int k, int log_blocks, const uint32_t* ar, const uint32_t* ac, const void* av, size_t an,
                                const uint32_t* br, const uint32_t* bc, const void* bv, size_t bn, const uint32_t* cr,
                                const uint32_t* cc, const void* cv, size_t cn, uint64_t seed_w, uint64_t seed_e,
                                nenya_params** out) {
  return guard([&] {
    if (!out) throw Error(NENYA_E_ARG, "null out");
    need_width(k);
    if (log_blocks < 2 || log_blocks > 26) throw Error(NENYA_E_SHAPE, "log_blocks out of range");
    if ((an && (!ar || !ac || !av)) || (bn && (!br || !bc || !bv)) || (cn && (!cr || !cc || !cv)))
      throw Error(NENYA_E_ARG, "null matrix array");
    need_gpu();
    CallTurn turn(false);
    size_t n = size_t(32) << log_blocks;
    HostCoo m3[3] = {{n, n, an, ar, ac, av}, {n, n, bn, br, bc, bv}, {n, n, cn, cr, cc, cv}};
    auto* p = new nenya_params;
    try {
      switch (k) {
        case 8: p->impl = new GeneralParams<W<8>>(log_blocks, m3, seed_w, seed_e); break;
        case 16: p->impl = new GeneralParams<W<16>>(log_blocks, m3, seed_w, seed_e); break;
        case 32: p->impl = new GeneralParams<W<32>>(log_blocks, m3, seed_w, seed_e); break;
        default: p->impl = new GeneralParams<W<64>>(log_blocks, m3, seed_w, seed_e); break;
      }
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
    return NENYA_OK;
  });
}

int nenya_prove_general(const nenya_params* p, const void* w, size_t n, uint8_t** proof, size_t* len) {
  return general_common(p, w, n, proof, len, false);
}
int nenya_prove_general_device(const nenya_params* p, const void* d_w, size_t n, uint8_t** proof, size_t* len) {
  return general_common(p, d_w, n, proof, len, true);
}
int nenya_prove_general_device_view(const nenya_params* p, const void* d_w, size_t n, const uint8_t** proof,
                                    size_t* len) {
  return general_common(p, d_w, n, nullptr, len, true, proof);
}

int nenya_prove_general_view(const nenya_params* p, const void* w, size_t n, const uint8_t** proof, size_t* len) {
  return general_common(p, w, n, nullptr, len, false, proof);
}

int nenya_prove_general_device_noalloc(const nenya_params* p, const void* d_w, size_t n, uint8_t* proof,
                                       size_t proof_cap, size_t* proof_len) {
  return guard([&] {
    if (!p || !p->impl || p->impl->kind != 1 || !d_w || !proof_len) throw Error(NENYA_E_ARG, "bad argument");
    ParamsBase* b = p->impl;
    size_t need = 0, N = 0;
    int log_b = 0;
    auto shape = [&](auto* gp) {
      need = gp->proof_size();
      N = gp->N;
      log_b = gp->log_b;
    };
    switch (b->k) {
      case 8: shape(static_cast<GeneralParams<W<8>>*>(b)); break;
      case 16: shape(static_cast<GeneralParams<W<16>>*>(b)); break;
      case 32: shape(static_cast<GeneralParams<W<32>>*>(b)); break;
      default: shape(static_cast<GeneralParams<W<64>>*>(b)); break;
    }
    if (log_b > NOALLOC_MAX_LOG_BLOCKS)
      throw Error(NENYA_E_SHAPE, "log_blocks " + std::to_string(log_b) + " exceeds the no-alloc limit of " +
                                     std::to_string(NOALLOC_MAX_LOG_BLOCKS) + " (256 MiB device scratch)");
    if (n != N) throw Error(NENYA_E_SHAPE, "witness length must be 32 * 2^log_blocks");
    *proof_len = need;
    if (!proof || proof_cap < need)
      throw Error(NENYA_E_ARG, "proof buffer of " + std::to_string(proof_cap) + " bytes is smaller than the proof (" +
                                   std::to_string(need) + " bytes, written to *proof_len)");
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    NoAllocDev* d = noalloc_dev(dev);
    std::lock_guard<std::mutex> g1(d->mu);
    auto* np = const_cast<nenya_params*>(p);
    std::lock_guard<std::mutex> g2(np->mu);
    CallTurn turn(true);
    ProofView v;
    switch (b->k) {
      case 8: v = general_noalloc_t<W<8>>(b, d, d_w); break;
      case 16: v = general_noalloc_t<W<16>>(b, d, d_w); break;
      case 32: v = general_noalloc_t<W<32>>(b, d, d_w); break;
      default: v = general_noalloc_t<W<64>>(b, d, d_w); break;
    }
    if (v.second != need) throw Error(NENYA_E_DEVICE, "proof size mismatch");
    std::memcpy(proof, v.first, v.second);
    *proof_len = v.second;
    return NENYA_OK;
  });
}

int nenya_verify_simd(const nenya_params* p, const uint32_t* ar, const uint32_t* ac, const void* av, size_t an,
                      const uint32_t* br, const uint32_t* bc, const void* bv, size_t bn, const uint32_t* cr,
                      const uint32_t* cc, const void* cv, size_t cn, const uint8_t* proof, size_t len, int* verdict) {
  return verify_common(p, 0, verdict, [&](ParamsBase* b) -> int {
    if (!proof && len) throw Error(NENYA_E_ARG, "null proof");
    if (!container::unwrap_for(*b, proof, len)) return 8;
    if ((an && (!ar || !ac || !av)) || (bn && (!br || !bc || !bv)) || (cn && (!cr || !cc || !cv)))
      throw Error(NENYA_E_ARG, "null matrix array");
    auto run = [&](auto* sp) {
      size_t n = size_t(1) << sp->log_n, m = size_t(1) << sp->log_m;
      HostCoo A{n, m, an, ar, ac, av}, B{n, m, bn, br, bc, bv}, Cm{n, m, cn, cr, cc, cv};
      using T = typename std::remove_pointer_t<decltype(sp)>::Width;
      return Verifier<T>::simd(*sp, A, B, Cm, proof, len);
    };
    switch (b->k) {
      case 8: return run(static_cast<SimdParams<W<8>>*>(b));
      case 16: return run(static_cast<SimdParams<W<16>>*>(b));
      case 32: return run(static_cast<SimdParams<W<32>>*>(b));
      default: return run(static_cast<SimdParams<W<64>>*>(b));
    }
  });
}

int nenya_verify_mimd(const nenya_params* p, const uint8_t* proof, size_t len, int* verdict) {
  return verify_common(p, 2, verdict, [&](ParamsBase* b) -> int {
    if (!proof && len) throw Error(NENYA_E_ARG, "null proof");
    if (!container::unwrap_for(*b, proof, len)) return 8;
    switch (b->k) {
      case 8: return Verifier<W<8>>::mimd(*static_cast<MimdParams<W<8>>*>(b), proof, len);
      case 16: return Verifier<W<16>>::mimd(*static_cast<MimdParams<W<16>>*>(b), proof, len);
      case 32: return Verifier<W<32>>::mimd(*static_cast<MimdParams<W<32>>*>(b), proof, len);
      default: return Verifier<W<64>>::mimd(*static_cast<MimdParams<W<64>>*>(b), proof, len);
    }
  });
}

int nenya_verify_general(const nenya_params* p, const uint8_t* proof, size_t len, int* verdict) {
  return verify_common(p, 1, verdict, [&](ParamsBase* b) -> int {
    if (!proof && len) throw Error(NENYA_E_ARG, "null proof");
    if (!container::unwrap_for(*b, proof, len)) return 8;
    switch (b->k) {
      case 8: return Verifier<W<8>>::general(*static_cast<GeneralParams<W<8>>*>(b), proof, len);
      case 16: return Verifier<W<16>>::general(*static_cast<GeneralParams<W<16>>*>(b), proof, len);
      case 32: return Verifier<W<32>>::general(*static_cast<GeneralParams<W<32>>*>(b), proof, len);
      default: return Verifier<W<64>>::general(*static_cast<GeneralParams<W<64>>*>(b), proof, len);
    }
  });
}

int nenya_set_timing(nenya_params* p, int enabled) {
  if (!p || !p->impl) return NENYA_E_ARG;
  p->impl->ctx0().timing = enabled != 0;
  return NENYA_OK;
}
int nenya_set_profile(nenya_params* p, int enabled) {
  if (!p || !p->impl) return NENYA_E_ARG;
  p->impl->ctx0().prof.on = enabled != 0;
  p->impl->ctx0().prof.stats.clear();
  return NENYA_OK;
}
int nenya_set_profile_only(nenya_params* p, const char* kernel) {
  if (!p || !p->impl) return NENYA_E_ARG;
  p->impl->ctx0().prof.only = kernel ? kernel : "";
  return NENYA_OK;
}
int nenya_kernel_stats(const nenya_params* p, int cap, char* names, size_t names_cap, double* ms, double* count,
                       double* bytes, double* macs) {
  if (!p || !p->impl) return 0;
  std::string all;
  int i = 0;
  for (const auto& kv : p->impl->ctx0().prof.stats) {
    if (i >= cap) break;
    if (ms) ms[i] = kv.second.ms;
    if (count) count[i] = (double)kv.second.count;
    if (bytes) bytes[i] = kv.second.bytes;
    if (macs) macs[i] = kv.second.macs;
    all += kv.first;
    all += '\n';
    i++;
  }
  if (names && names_cap) {
    size_t c = std::min(all.size(), names_cap - 1);
    std::memcpy(names, all.data(), c);
    names[c] = 0;
  }
  return i;
}
static int export_timing(const std::vector<std::pair<std::string, double>>& t, double* ms, int cap, char* names,
                         size_t names_cap) {
  int n = (int)std::min<size_t>(t.size(), cap > 0 ? (size_t)cap : 0);
  std::string all;
  for (int i = 0; i < n; i++) {
    if (ms) ms[i] = t[i].second;
    all += t[i].first;
    all += '\n';
  }
  if (names && names_cap) {
    size_t c = std::min(all.size(), names_cap - 1);
    std::memcpy(names, all.data(), c);
    names[c] = 0;
  }
  return n;
}
int nenya_last_timing(const nenya_params* p, double* ms, int cap, char* names, size_t names_cap) {
  if (!p || !p->impl) return 0;
  return export_timing(p->impl->ctx0().last_timing, ms, cap, names, names_cap);
}
int nenya_setup_timing(const nenya_params* p, double* ms, int cap, char* names, size_t names_cap) {
  if (!p || !p->impl) return 0;
  return export_timing(p->impl->setup_timing, ms, cap, names, names_cap);
}

// ---------------- batched / served proofs ----------------
struct nenya_server {
  Server* impl = nullptr;
};

int nenya_server_create(int workers, size_t mem_budget, nenya_server** out) {
  return guard([&] {
    if (!out || workers < 1 || workers > 64) throw Error(NENYA_E_ARG, "workers must be 1..64");
    need_gpu();
    CallTurn turn(false);
    auto* s = new nenya_server;
    try {
      s->impl = new Server(workers, mem_budget);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
    return NENYA_OK;
  });
}

int nenya_submit_general(nenya_server* s, const nenya_params* p, const void* w, size_t n, uint64_t* job) {
  return guard([&] {
    if (!s || !s->impl || !p || !p->impl || p->impl->kind != 1 || !w || !job) throw Error(NENYA_E_ARG, "bad argument");
    std::unique_ptr<Job> j(new Job);
    j->params = p->impl;
    j->kind = 1;
    j->w = w;
    j->n = n;
    *job = s->impl->submit(std::move(j));
    return NENYA_OK;
  });
}

int nenya_submit_simd(nenya_server* s, const nenya_params* p, const uint32_t* ar, const uint32_t* ac, const void* av,
                      size_t an, const uint32_t* br, const uint32_t* bc, const void* bv, size_t bn, const uint32_t* cr,
                      const uint32_t* cc, const void* cv, size_t cn, const void* witnesses, uint64_t* job) {
  return guard([&] {
    if (!s || !s->impl || !p || !p->impl || p->impl->kind != 0 || !witnesses || !job)
      throw Error(NENYA_E_ARG, "bad argument");
    if ((an && (!ar || !ac || !av)) || (bn && (!br || !bc || !bv)) || (cn && (!cr || !cc || !cv)))
      throw Error(NENYA_E_ARG, "null matrix array");
    ParamsBase* b = p->impl;
    int ln = 0, lm = 0;
    switch (b->k) {
      case 8: ln = static_cast<SimdParams<W<8>>*>(b)->log_n; lm = static_cast<SimdParams<W<8>>*>(b)->log_m; break;
      case 16: ln = static_cast<SimdParams<W<16>>*>(b)->log_n; lm = static_cast<SimdParams<W<16>>*>(b)->log_m; break;
      case 32: ln = static_cast<SimdParams<W<32>>*>(b)->log_n; lm = static_cast<SimdParams<W<32>>*>(b)->log_m; break;
      default: ln = static_cast<SimdParams<W<64>>*>(b)->log_n; lm = static_cast<SimdParams<W<64>>*>(b)->log_m; break;
    }
    size_t n = size_t(1) << ln, m = size_t(1) << lm;
    std::unique_ptr<Job> j(new Job);
    j->params = b;
    j->kind = 0;
    j->w = witnesses;
    j->m3[0] = {n, m, an, ar, ac, av};
    j->m3[1] = {n, m, bn, br, bc, bv};
    j->m3[2] = {n, m, cn, cr, cc, cv};
    *job = s->impl->submit(std::move(j));
    return NENYA_OK;
  });
}

int nenya_poll(nenya_server* s, uint64_t job) {
  int done = 0;
  int rc = guard([&] {
    if (!s || !s->impl) throw Error(NENYA_E_ARG, "bad argument");
    done = s->impl->poll(job) ? 1 : 0;
    return NENYA_OK;
  });
  return rc == NENYA_OK ? done : -rc;
}

int nenya_submit_mimd(nenya_server* s, const nenya_params* p, const void* witnesses, uint64_t* job) {
  return guard([&] {
    if (!s || !s->impl || !p || !p->impl || p->impl->kind != 2 || !witnesses || !job)
      throw Error(NENYA_E_ARG, "bad argument");
    std::unique_ptr<Job> j(new Job);
    j->params = p->impl;
    j->kind = 2;
    j->w = witnesses;
    *job = s->impl->submit(std::move(j));
    return NENYA_OK;
  });
}

int nenya_wait(nenya_server* s, uint64_t job, const uint8_t** proof, size_t* proof_len, double* queued_ms,
               double* run_ms) {
  return guard([&] {
    if (!s || !s->impl || !proof || !proof_len) throw Error(NENYA_E_ARG, "bad argument");
    Job* j = s->impl->wait(job);
    if (queued_ms) *queued_ms = j->t_start - j->t_submit;
    if (run_ms) *run_ms = j->t_end - j->t_start;
    if (j->rc != NENYA_OK) throw Error(j->rc, j->err);
    *proof = j->proof.p;
    *proof_len = j->len;
    return NENYA_OK;
  });
}

int nenya_job_release(nenya_server* s, uint64_t job) {
  return guard([&] {
    if (!s || !s->impl) throw Error(NENYA_E_ARG, "bad argument");
    s->impl->release(job);
    return NENYA_OK;
  });
}

int nenya_proof_container(const nenya_params* p, const uint8_t* proof, size_t len, uint8_t** out, size_t* out_len) {
  return guard([&] {
    if (!p || !p->impl || (!proof && len) || !out || !out_len) throw Error(NENYA_E_ARG, "bad argument");
    if (container::is_container(proof, len)) throw Error(NENYA_E_ARG, "already a container");
    uint8_t* b = (uint8_t*)std::malloc(container::HEADER + len);
    if (!b) throw std::bad_alloc();
    container::wrap(*p->impl, proof, len, b);
    *out = b;
    *out_len = container::HEADER + len;
    return NENYA_OK;
  });
}

int nenya_proof_payload(const uint8_t* c, size_t len, const uint8_t** payload, size_t* payload_len, int* model,
                        int* k) {
  return guard([&] {
    if (!c || !payload || !payload_len) throw Error(NENYA_E_ARG, "bad argument");
    container::Info in;
    if (!container::parse(c, len, in, payload)) throw Error(NENYA_E_ARG, "not an intact proof container");
    *payload_len = (size_t)in.len;
    if (model) *model = in.model;
    if (k) *k = in.k;
    return NENYA_OK;
  });
}

int nenya_autotune(int k_mask, int log_min, int log_max, char* report, size_t report_cap) {
  return guard([&] {
    if (log_min < 8 || log_max > 22 || log_min > log_max || (k_mask & ~15) || !k_mask)
      throw Error(NENYA_E_ARG, "autotune: k_mask in 1..15, 8 <= log_min <= log_max <= 22");
    need_gpu();
    CallTurn turn(false);
    const auto t0 = std::chrono::steady_clock::now();
    std::ostringstream js;
    js << "{";
    bool first = true;
    for (int b = 0; b < 4; b++) {
      if (!(k_mask & (1 << b))) continue;
      if (!first) js << ",";
      first = false;
      switch (b) {
        case 0: tune_width<W<8>>(js, log_min, log_max); break;
        case 1: tune_width<W<16>>(js, log_min, log_max); break;
        case 2: tune_width<W<32>>(js, log_min, log_max); break;
        default: tune_width<W<64>>(js, log_min, log_max); break;
      }
    }
    js << ",\"seconds\":" << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << "}";
    const std::string out = js.str();
    if (report && report_cap) {
      const size_t c = std::min(out.size(), report_cap - 1);
      std::memcpy(report, out.data(), c);
      report[c] = 0;
    }
    return NENYA_OK;
  });
}

int nenya_cancel(nenya_server* s, uint64_t job) {
  return guard([&] {
    if (!s || !s->impl) throw Error(NENYA_E_ARG, "bad argument");
    s->impl->cancel(job);
    return NENYA_OK;
  });
}

void nenya_server_free(nenya_server* s) {
  if (!s) return;
  CallTurn turn(false);
  delete s->impl;
  delete s;
}

}  // extern "C"
