// Round hub: the latency-bound tail rounds of many proofs' sumchecks fused
// into shared launches.  The reference proves one instance per call and runs
// a sumcheck's rounds one after another (mult_check.rs:180-241,
// linear_check_general.rs:214-350); a server runs several proofs at once,
// and once a round's tables are small, each of its kernels (sums, the
// challenge's Fiat-Shamir and lift, the fold, the eq table's
// marginalization) does little work and mostly waits for dispatch.  Workers
// therefore hand such sumchecks to the hub, whose thread advances every
// sumcheck it holds by one round per step: one launch per kernel type and
// limb width covers all of them, each task owning a contiguous range of the
// launch's workgroups (its descriptor travels in the kernel arguments).
// Tasks join and leave between steps.  The transcript is the device one
// (fsdev.hpp), so no step waits on the host, and the values are the ones the
// worker's own round kernels would produce: round sums are taken mod 2^32 or
// 2^64 with integer atomics (2^k divides both), and the fold is the group
// product form of k_fold.
#pragma once
#include <condition_variable>
#include <mutex>
#include <thread>

#include "prover.hpp"

namespace nenya {

constexpr int HUB_MAXT = 16;  // tasks per fused launch (descriptors in the kernel arguments)

template <class D> struct HubMulti {
  D d[HUB_MAXT];
  uint32_t start[HUB_MAXT + 1];  // task t owns workgroups [start[t], start[t+1])
  int n;
};

// The task that owns workgroup b (uniform over the block).
template <class D> __device__ __forceinline__ int hub_task(const HubMulti<D>& M, unsigned b) {
  int t = 0;
  while (t + 1 < M.n && b >= M.start[t + 1]) t++;
  return __builtin_amdgcn_readfirstlane(t);
}

// Gruen round sums (as k_mult_round<T, true>) into d.sums[3].
template <class T> struct HubSumsM {
  const LT<T>* eg;
  const LT<T>* a;
  const LT<T>* b;
  const LT<T>* c;
  const LT<T>* e;
  uint64_t m;
  CT<T>* sums;
};
template <class T> __global__ void __launch_bounds__(GB) k_hub_mult_sums(HubMulti<HubSumsM<T>> M) {
  const int t = hub_task(M, blockIdx.x);
  const HubSumsM<T>& d = M.d[t];
  const unsigned lb = blockIdx.x - M.start[t];
  mult_round_body<T, true>(d.eg, d.a, d.b, d.c, d.e, d.m, nullptr, d.sums + (size_t)(lb % SUM_COPIES) * 3 * E, lb,
                           M.start[t + 1] - M.start[t]);
}

// Linear-check round sums (as k_gram_lin / k_gram_lin64) into d.sums[2].
template <class T> struct HubSumsL {
  const LT<T>* f;
  const LT<T>* g;
  uint64_t fs, gs, ntab, m;
  int xrows;
  CT<T>* sums;
};
template <class T> __global__ void __launch_bounds__(1024) k_hub_lin_sums(HubMulti<HubSumsL<T>> M) {
  const int t = hub_task(M, blockIdx.x);
  const HubSumsL<T>& d = M.d[t];
  const unsigned lb = blockIdx.x - M.start[t];
  gram_lin_body<T>(d.f, d.fs, d.g, d.gs, d.ntab, d.m, d.m, d.xrows, nullptr, 2, 0,
                   d.sums + (size_t)(lb % SUM_COPIES) * 3 * E, lb, M.start[t + 1] - M.start[t]);
}
__global__ void __launch_bounds__(512) k_hub_lin_sums64(HubMulti<HubSumsL<W<64>>> M) {
  const int t = hub_task(M, blockIdx.x);
  const HubSumsL<W<64>>& d = M.d[t];
  const unsigned lb = blockIdx.x - M.start[t];
  gram_lin64_body(d.f, d.fs, d.g, d.gs, d.ntab, d.m, d.m, d.xrows, nullptr, 2, 0,
                  d.sums + (size_t)(lb % SUM_COPIES) * 3 * E, lb, M.start[t + 1] - M.start[t]);
}

// One block per task: the round polynomial from the sums (k_sum_post's
// forms; mode 0 mult check, 1 linear check with claim c), written to qh
// (and qd); the sums cleared for the next round; then the transcript
// absorbs it, draws the challenge and lifts it into x (k_lift_fs).
template <class T> struct HubPost {
  CT<T>* sums;
  const LT<T>* c;
  LT<T>* qd;
  LT<T>* qh;
  b3::State* st;
  LT<T>* x;
  const LT<T>* frob;
  int mode;
};
template <class T> __global__ void __launch_bounds__(256) k_hub_post_lift(HubMulti<HubPost<T>> M) {
  using C = CT<T>;
  using L = LT<T>;
  __shared__ __attribute__((aligned(16))) C sx[E];
  __shared__ __attribute__((aligned(16))) L nat[E];
  __shared__ __attribute__((aligned(16))) L q[3 * E];
  __shared__ fsd::RoundLds<T> fl;
  __shared__ C ssum[3][E];
  const int t = hub_task(M, blockIdx.x);
  const HubPost<T>& d = M.d[t];
  gather_sums<T>(d.sums, 3, ssum);
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * E; i += blockDim.x) {
    const int k = i / E, li = i % E;
    C v;
    if (d.mode == 0) v = k == 1 ? (C)(ssum[1][li] - ssum[0][li] - ssum[2][li]) : ssum[k][li];
    else v = k == 0 ? ssum[0][li] : k == 1 ? (C)((C)d.c[li] - ssum[0][li] - ssum[0][li] - ssum[1][li]) : ssum[1][li];
    q[i] = (L)v;
    d.qh[i] = (L)v;
    if (d.qd) d.qd[i] = (L)v;
  }
  __syncthreads();
  if (threadIdx.x < 64) fsd::round_challenge<T>(d.st, q, fl, nat);
  __syncthreads();
  teich_lift_block<T>(nat, d.x, d.frob, T::K - 1, sx);
}

// Folds by the round challenge r on the vector ALUs (k_fold's group
// product) over up to two groups of tables (FoldJob), in place or not; the
// job's side update runs on one extra block of the task's range.
template <class T> struct HubFold {
  const LT<T>* r;
  FoldJob<T> J;
};
template <class T> __global__ void __launch_bounds__(GB) k_hub_fold(HubMulti<HubFold<T>> M) {
  __shared__ __attribute__((aligned(16))) LT<T> sr[E];
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][E];
  const int tk = hub_task(M, blockIdx.x);
  const HubFold<T>& d = M.d[tk];
  const FoldJob<T>& J = d.J;
  const unsigned lb = blockIdx.x - M.start[tk], nb = M.start[tk + 1] - M.start[tk];
  const unsigned nfold = nb - (J.side ? 1u : 0u);
  if (J.side && lb == nfold) {
    if (J.side == 1) eq_one_block<T>(J.sa, J.sb, J.so);
    else lin_claim_block<T>(J.sa, J.sb, J.so);
    return;
  }
  const int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  if (threadIdx.x < 32) stage<T>(sr, d.r, t);
  const uint64_t total = J.m * (J.ntab[0] + J.ntab[1]);
  for (uint64_t base = (uint64_t)lb * GPB; base < total; base += (uint64_t)nfold * GPB) {
    const uint64_t w = base + g;
    const LT<T>* lo = nullptr;
    LT<T>* out = nullptr;
    const bool ok = w < total && fold_locate<T>(J, w, lo, out);
    CT<T> a[4] = {0, 0, 0, 0};
    if (ok) {
      CT<T> h[4];
      ld4<T>(lo + 4 * t, a);
      ld4<T>(lo + J.m * E + 4 * t, h);
      CT<T> dd[4] = {h[0] - a[0], h[1] - a[1], h[2] - a[2], h[3] - a[3]};
      st4<T>(sh[g] + 4 * t, dd);
    }
    __syncthreads();
    Acc<T> acc;
    acc.clear();
    if (ok) acc_mul<T>(acc, sr, sh[g], t);
    CT<T> o[4];
    acc_final<T>(acc, t, o);
    if (ok) {
      for (int c = 0; c < 4; c++) o[c] += a[c];
      st4<T>(out + 4 * t, o);
    }
    __syncthreads();
  }
}

// eg[i] = scale (eg[i] + eg[i+m]) (k_marg).
template <class T> struct HubMarg {
  const LT<T>* scale;
  LT<T>* eg;
  uint64_t m;
};
template <class T> __global__ void __launch_bounds__(GB) k_hub_marg(HubMulti<HubMarg<T>> M) {
  const int t = hub_task(M, blockIdx.x);
  const HubMarg<T>& d = M.d[t];
  marg_body<T>(d.scale, d.eg, d.m, blockIdx.x - M.start[t], M.start[t + 1] - M.start[t]);
}

// A sumcheck's remaining rounds, handed over by its worker with every value
// on the device.  Mult check (type 0): tables work + t*stride (t < ntab),
// eq table eg, point drt, eq factor dsc.  Linear check (type 1): tables fw,
// gw (ntab each, stride H), claim dc, round polynomial dq.  Both: challenges
// dch, round polynomials drounds, device transcript dtr, sums (SUM_COPIES x
// 3 elements of the compute type, zero on entry).
template <class T> struct HubTask {
  using L = typename T::L;
  using C = typename T::C;
  int type = 0;
  int r = 0, nv = 0;  // next round, rounds in all
  uint64_t m = 0;     // half the current table length
  uint64_t ntab = 0;
  L* work = nullptr;
  uint64_t stride = 0;
  L* eg = nullptr;
  const L* drt = nullptr;
  L* dsc = nullptr;
  L* fw = nullptr;
  L* gw = nullptr;
  uint64_t H = 0;
  L* dc = nullptr;
  L* dq = nullptr;
  L* dch = nullptr;
  L* drounds = nullptr;
  b3::State* dtr = nullptr;
  C* sums = nullptr;
  const L* frob = nullptr;
  hipEvent_t ready = nullptr;  // recorded by the worker after the tables are in place
  hipEvent_t done = nullptr;   // recorded by the hub after the task's last launch
  const std::atomic<bool>* cancel = nullptr;
  bool finished = false;
  int rc = NENYA_OK;
  std::string err;
};

// One width's lane of the hub: its own thread and stream, so the widths
// advance independently (a u64 lift does not hold up the u8 tasks).
template <class T> struct HubLane {
  using L = typename T::L;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  int device = 0;
  hipStream_t st = nullptr;
  std::thread th;
  std::vector<HubTask<T>*> pending, active;
  hipEvent_t sev[3] = {};  // the last steps' ends (at most two steps in flight)
  uint64_t nstep = 0;
  std::atomic<uint64_t>* launches;
  std::atomic<uint64_t>* task_launches;

  HubLane(std::atomic<uint64_t>* nl, std::atomic<uint64_t>* ntl) : launches(nl), task_launches(ntl) {
    HIP_CHECK(hipGetDevice(&device));
    HIP_CHECK(hipStreamCreate(&st));
    for (auto& e : sev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    th = std::thread([this] { loop(); });
  }
  ~HubLane() {
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
    }
    cv_work.notify_all();
    th.join();
    (void)hipStreamSynchronize(st);
    for (auto& e : sev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(st);
  }

  // Worker side: hand the task over and wait until its last launch is
  // enqueued (the caller's stream then waits on task.done).
  void run(HubTask<T>& t) {
    {
      std::lock_guard<std::mutex> g(mu);
      if (stop) throw Error(NENYA_E_DEVICE, "round hub stopped");
      pending.push_back(&t);
    }
    cv_work.notify_all();
    std::unique_lock<std::mutex> g(mu);
    cv_done.wait(g, [&] { return t.finished; });
    if (t.rc != NENYA_OK) throw Error(t.rc, t.err);
  }

  void finish(HubTask<T>* t, int rc, const std::string& err) {
    (void)hipEventRecord(t->done, st);
    std::lock_guard<std::mutex> g(mu);
    t->rc = rc;
    t->err = err;
    t->finished = true;
  }

  template <class D, class K> void launch_multi(K kern, std::vector<std::pair<D, uint32_t>>& items, unsigned threads) {
    for (size_t i0 = 0; i0 < items.size(); i0 += HUB_MAXT) {
      HubMulti<D> M{};
      const size_t n = std::min<size_t>(HUB_MAXT, items.size() - i0);
      uint32_t acc = 0;
      for (size_t i = 0; i < n; i++) {
        M.d[i] = items[i0 + i].first;
        M.start[i] = acc;
        acc += items[i0 + i].second;
      }
      M.start[n] = acc;
      M.n = (int)n;
      hipLaunchKernelGGL(kern, dim3(acc), dim3(threads), 0, st, M);
      HIP_CHECK(hipGetLastError());
      launches->fetch_add(1, std::memory_order_relaxed);
      task_launches->fetch_add(n, std::memory_order_relaxed);
    }
  }

  // One round of every active task.
  void step() {
    // canceled tasks leave first; the rest stay listed in `active` until
    // the step's launches are enqueued (a failure then fails them all)
    std::vector<HubTask<T>*> act;
    for (HubTask<T>* t : active) {
      if (t->cancel && t->cancel->load(std::memory_order_relaxed)) finish(t, NENYA_E_CANCELED, "job canceled");
      else act.push_back(t);
    }
    active = act;
    if (act.empty()) return;
    std::vector<std::pair<HubSumsM<T>, uint32_t>> ms;
    std::vector<std::pair<HubSumsL<T>, uint32_t>> ls;
    std::vector<std::pair<HubPost<T>, uint32_t>> ps;
    std::vector<std::pair<HubFold<T>, uint32_t>> fs;
    std::vector<std::pair<HubMarg<T>, uint32_t>> gs;
    for (HubTask<T>* t : act) {
      const uint64_t m = t->m;
      L* rc = t->dch + (size_t)t->r * E;
      const bool more = t->r + 1 < t->nv;
      if (t->type == 0) {
        L* w = t->work;
        const uint64_t s = t->stride * E;
        ms.push_back({HubSumsM<T>{t->eg, w, w + s, t->ntab > 2 ? w + 2 * s : nullptr, t->ntab > 2 ? w + 3 * s : nullptr,
                                  m, t->sums},
                      3 * grid_for(m, GPB, 682)});
        ps.push_back({HubPost<T>{t->sums, nullptr, nullptr, t->drounds + (size_t)t->r * 3 * E, t->dtr, rc, t->frob, 0},
                      1});
        FoldJob<T> J{{w, w}, {w, w}, {t->ntab, 0}, {t->stride, t->stride}, {t->stride, t->stride}, m,
                     more ? 1 : 0, t->drt + (size_t)t->r * E, rc, t->dsc};
        fs.push_back({HubFold<T>{rc, J}, grid_for(m * t->ntab, GPB) + (more ? 1u : 0u)});
        if (more) gs.push_back({HubMarg<T>{t->dsc, t->eg, m / 2}, grid_for(m / 2, GPB)});
      } else {
        const uint64_t nch = (m * t->ntab + 31) / 32;
        ls.push_back({HubSumsL<T>{t->fw, t->gw, t->H, t->H, t->ntab, m, 128, t->sums},
                      (uint32_t)std::min<uint64_t>(nch, 256)});
        ps.push_back(
            {HubPost<T>{t->sums, t->dc, t->dq, t->drounds + (size_t)t->r * 3 * E, t->dtr, rc, t->frob, 1}, 1});
        FoldJob<T> J{{t->fw, t->gw}, {t->fw, t->gw}, {t->ntab, t->ntab}, {t->H, t->H}, {t->H, t->H}, m, 2, t->dq, rc,
                     t->dc};
        fs.push_back({HubFold<T>{rc, J}, grid_for(2 * m * t->ntab, GPB) + 1u});
      }
    }
    if (!ms.empty()) launch_multi<HubSumsM<T>>(k_hub_mult_sums<T>, ms, GB);
    if (!ls.empty()) {
      if constexpr (T::K == 64) launch_multi<HubSumsL<T>>(k_hub_lin_sums64, ls, 512);
      else launch_multi<HubSumsL<T>>(k_hub_lin_sums<T>, ls, 1024);
    }
    launch_multi<HubPost<T>>(k_hub_post_lift<T>, ps, 256);
    launch_multi<HubFold<T>>(k_hub_fold<T>, fs, GB);
    if (!gs.empty()) launch_multi<HubMarg<T>>(k_hub_marg<T>, gs, GB);
    active.clear();
    for (HubTask<T>* t : act) {
      t->r++;
      t->m /= 2;
      if (t->r == t->nv) finish(t, NENYA_OK, "");
      else active.push_back(t);
    }
  }

  void loop() {
    (void)hipSetDevice(device);
    for (;;) {
      {
        std::unique_lock<std::mutex> g(mu);
        cv_work.wait(g, [&] { return stop || !pending.empty() || !active.empty(); });
        if (pending.empty() && active.empty()) return;  // stopping
        for (HubTask<T>* t : pending) {
          // a wait that cannot be enqueued fails the step below
          (void)hipStreamWaitEvent(st, t->ready, 0);
          active.push_back(t);
        }
        pending.clear();
      }
      bool ok = true;
      std::string err;
      try {
        step();
        // at most two steps in flight, so tasks that arrive join soon
        if (nstep >= 2) HIP_CHECK(hipEventSynchronize(sev[(nstep - 2) % 3]));
        HIP_CHECK(hipEventRecord(sev[nstep % 3], st));
        nstep++;
      } catch (const std::exception& e) {
        ok = false;
        err = e.what();
      }
      if (!ok) {
        for (HubTask<T>* t : active) finish(t, NENYA_E_DEVICE, err);
        active.clear();
      }
      cv_done.notify_all();
    }
  }
};

struct Hub {
  std::mutex mu;  // lane creation
  std::unique_ptr<HubLane<W<8>>> l8;
  std::unique_ptr<HubLane<W<16>>> l16;
  std::unique_ptr<HubLane<W<32>>> l32;
  std::unique_ptr<HubLane<W<64>>> l64;
  // fused launches and the tasks they covered (the mean fusion width)
  std::atomic<uint64_t> launches{0}, task_launches{0};

  template <class T> std::unique_ptr<HubLane<T>>& slot() {
    if constexpr (T::K == 8) return l8;
    else if constexpr (T::K == 16) return l16;
    else if constexpr (T::K == 32) return l32;
    else return l64;
  }
  // the width's lane, started on first use
  template <class T> HubLane<T>& lane() {
    std::lock_guard<std::mutex> g(mu);
    auto& p = slot<T>();
    if (!p) p.reset(new HubLane<T>(&launches, &task_launches));
    return *p;
  }
  template <class T> void run(HubTask<T>& t) { lane<T>().run(t); }
};

// Runs the task's rounds on the hub; the worker's stream st then waits for
// them (also when the task failed, so its buffers outlive the hub's reads).
template <class T> void hub_run(Hub* h, HubTask<T>& t, hipStream_t st) {
  try {
    h->run(t);
  } catch (...) {
    (void)hipStreamWaitEvent(st, t.done, 0);
    throw;
  }
  HIP_CHECK(hipStreamWaitEvent(st, t.done, 0));
}

}  // namespace nenya
