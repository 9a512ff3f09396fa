// Batched and served proofs on one GPU.  The reference proves one instance
// per call (snark-ring/src/general.rs:367-372, simd.rs:212-217,
// mimd.rs:235-240); here
// independent proofs are queued and drained by worker threads, each with its
// own execution contexts (HIP stream, stream-ordered pool, pinned proof
// buffer) per limb width, so the latency-bound phases of one proof (sumcheck
// round trips, challenge lifts) overlap the throughput phases of others.
// Admission: the queue is FIFO and its head starts only when its device
// footprint fits the remaining budget (or nothing else is running).
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "snark.hpp"

namespace nenya {

inline double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
  uint64_t id = 0;
  ParamsBase* params = nullptr;
  int kind = 1;              // 0 simd, 1 general, 2 mimd
  const void* w = nullptr;   // general: witness (host); simd, mimd: 32 x m witnesses
  size_t n = 0;
  HostCoo m3[3];             // simd matrices (borrowed until the job ends)
  Pinned proof;
  size_t len = 0;
  size_t fp = 0;
  int rc = NENYA_OK;
  std::string err;
  double t_submit = 0, t_start = 0, t_end = 0;
  bool done = false, started = false;
  std::atomic<bool> cancel{false};
};

// One stream and one device pool per worker, shared by its per-width
// contexts: the runtime maps streams onto a few hardware queues
// (GPU_MAX_HW_QUEUES, 4 by default), and two streams on one queue run in
// order, so a stream per width would let two workers' proofs serialize on a
// shared queue; one pool lets blocks freed by a proof of one width serve the
// next proof of another.  A worker runs one proof at a time, so its contexts
// never overlap each other.
struct Worker {
  Pool pool;  // destroyed after the contexts
  hipStream_t st = nullptr;
  std::unique_ptr<CtxBase> ctx[4];  // per width: 8, 16, 32, 64
  bool devfs = false;               // round challenges drawn on the device
  Hub* hub = nullptr;               // the server's round hub, if any
  Worker(bool dfs, Hub* h) : devfs(dfs), hub(h) { HIP_CHECK(hipStreamCreate(&st)); }
  ~Worker() {
    for (auto& c : ctx) c.reset();
    if (st) (void)hipStreamDestroy(st);
  }
  template <class T> Ctx<T>& get() {
    const int i = T::K == 8 ? 0 : T::K == 16 ? 1 : T::K == 32 ? 2 : 3;
    if (!ctx[i]) {
      ctx[i].reset(new Ctx<T>(st, pool));
      // several workers drive the device: no kernel waits on a host thread
      static_cast<Ctx<T>&>(*ctx[i]).eng->pipelined = false;
      static_cast<Ctx<T>&>(*ctx[i]).eng->devfs = devfs;
      static_cast<Ctx<T>&>(*ctx[i]).eng->hub = hub;
      // NENYA_SERVER_TIMING=1: stage times of every served proof to stderr
      const char* tm = std::getenv("NENYA_SERVER_TIMING");
      static_cast<Ctx<T>&>(*ctx[i]).timing = tm && tm[0] == '1';
    }
    return static_cast<Ctx<T>&>(*ctx[i]);
  }
};

struct Server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  std::vector<Pinned> spare;  // proof buffers returned by released jobs
  // declared before the workers: destroyed after their threads have ended
  std::unique_ptr<Hub> hub;
  std::vector<std::unique_ptr<Worker>> workers;
  std::vector<std::thread> threads;
  size_t budget = 0, in_use = 0;
  int running = 0;
  uint64_t next_id = 1;
  bool stop = false;
  int device = 0;
  // test switch NENYA_INJECT_OOM=<job>:<n>: the n-th device allocation of
  // job <job> fails
  uint64_t inject_job = 0;
  long inject_at = 0;

  Server(int nworkers, size_t budget_bytes) {
    HIP_CHECK(hipGetDevice(&device));
    if (budget_bytes == 0) {
      size_t fr = 0, tot = 0;
      HIP_CHECK(hipMemGetInfo(&fr, &tot));
      budget_bytes = fr / 10 * 8;
    }
    budget = budget_bytes;
    if (const char* e = std::getenv("NENYA_INJECT_OOM")) {
      unsigned long long jid = 0;
      long at = 0;
      if (std::sscanf(e, "%llu:%ld", &jid, &at) == 2 && jid > 0 && at > 0) {
        inject_job = jid;
        inject_at = at;
      }
    }
    const char* dfs = std::getenv("NENYA_DEVICE_FS");
    const char* hb = std::getenv("NENYA_HUB");
    if (hb && hb[0] == '1') hub.reset(new Hub());
    for (int i = 0; i < nworkers; i++) workers.emplace_back(new Worker(dfs && dfs[0] == '1', hub.get()));
    live_servers()++;  // before any job can be submitted
    for (int i = 0; i < nworkers; i++) threads.emplace_back([this, i] { loop(i); });
  }
  ~Server() {
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
    }
    cv_work.notify_all();
    for (auto& t : threads) t.join();
    workers.clear();
    live_servers()--;
  }

  uint64_t submit(std::unique_ptr<Job> j) {
    std::lock_guard<std::mutex> g(mu);
    j->fp = j->params->footprint;
    if (j->fp > budget)
      throw Error(NENYA_E_OOM, "job footprint of " + std::to_string(j->fp) + " device bytes exceeds the server's budget of " +
                                   std::to_string(budget));
    j->id = next_id++;
    j->t_submit = now_ms();
    Job* raw = j.get();
    jobs[raw->id] = std::move(j);
    queue.push_back(raw);
    cv_work.notify_all();
    return raw->id;
  }
  Job* find(uint64_t id) {
    auto it = jobs.find(id);
    if (it == jobs.end()) throw Error(NENYA_E_ARG, "unknown job");
    return it->second.get();
  }
  Job* wait(uint64_t id) {
    std::unique_lock<std::mutex> g(mu);
    Job* j = find(id);
    cv_done.wait(g, [&] { return j->done; });
    return j;
  }
  bool poll(uint64_t id) {
    std::lock_guard<std::mutex> g(mu);
    return find(id)->done;
  }
  // A queued job leaves the queue done (canceled); a running one is flagged
  // and stops at its next checkpoint.
  void cancel(uint64_t id) {
    {
      std::lock_guard<std::mutex> g(mu);
      Job* j = find(id);
      if (j->done) throw Error(NENYA_E_ARG, "job already done");
      j->cancel.store(true);
      if (j->started) return;
      for (auto it = queue.begin(); it != queue.end(); ++it)
        if (*it == j) {
          queue.erase(it);
          break;
        }
      j->rc = NENYA_E_CANCELED;
      j->err = "job canceled before it started";
      j->t_start = j->t_end = now_ms();
      j->done = true;
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
  void release(uint64_t id) {
    std::lock_guard<std::mutex> g(mu);
    Job* j = find(id);
    if (!j->done) throw Error(NENYA_E_ARG, "job still running");
    if (j->proof.p) {
      spare.push_back(std::move(j->proof));
      // keep a few more buffers than there are workers: drop the smallest
      if (spare.size() > 2 * workers.size() + 2) {
        size_t lo = 0;
        for (size_t i = 1; i < spare.size(); i++)
          if (spare[i].cap < spare[lo].cap) lo = i;
        spare.erase(spare.begin() + lo);
      }
    }
    jobs.erase(id);
  }

  template <class T> void run_t(Worker& wk, Job& j) {
    Ctx<T>& cx = wk.get<T>();
    CallTurn turn(false);
    struct Cancel {
      Engine<T>& e;
      Cancel(Engine<T>& en, const std::atomic<bool>* c) : e(en) { e.cancel = c; }
      ~Cancel() { e.cancel = nullptr; }
    } can(*cx.eng, &j.cancel);
    ProofView v;
    if (j.kind == 1) {
      auto* gp = static_cast<GeneralParams<T>*>(j.params);
      v = gp->prove_host(cx, (const typename T::L*)j.w, j.n);
    } else if (j.kind == 2) {
      auto* mp = static_cast<MimdParams<T>*>(j.params);
      v = mp->prove(cx, (const typename T::L*)j.w);
    } else {
      auto* sp = static_cast<SimdParams<T>*>(j.params);
      v = sp->prove(cx, j.m3[0], j.m3[1], j.m3[2], (const typename T::L*)j.w);
    }
    if (cx.timing) {
      std::string line = "job " + std::to_string(j.id) + " kind " + std::to_string(j.kind) + " k " +
                         std::to_string(T::K) + " start " + std::to_string(now_ms() - j.t_start) + ":";
      for (auto& st : cx.last_timing) line += " " + st.first + " " + std::to_string(st.second);
      std::fprintf(stderr, "%s\n", line.c_str());
    }
    j.len = v.second;
    size_t cur = j.params->proof_len.load();
    while (v.second > cur && !j.params->proof_len.compare_exchange_weak(cur, v.second)) {
    }
    // the proof was assembled in cx.proofbuf: hand that buffer to the job
    std::swap(j.proof, cx.proofbuf);
  }

  void loop(int wi) {
    (void)hipSetDevice(device);
    Worker& wk = *workers[wi];
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> g(mu);
        cv_work.wait(g, [&] {
          return stop || (!queue.empty() && (running == 0 || in_use + queue.front()->fp <= budget));
        });
        if (stop) return;
        j = queue.front();
        queue.pop_front();
        in_use += j->fp;
        running++;
        j->started = true;
        j->t_start = now_ms();
        // a recycled pinned buffer becomes the context's proof buffer: the
        // smallest that holds the params' last proof, or (before their first
        // proof) the largest; none if none is large enough, so a too-small
        // buffer is not freed only to be allocated again
        if (!spare.empty()) {
          const size_t want = j->params->proof_len.load();
          size_t pick = spare.size();
          for (size_t i = 0; i < spare.size(); i++) {
            if (want ? spare[i].cap >= want && (pick == spare.size() || spare[i].cap < spare[pick].cap)
                     : pick == spare.size() || spare[i].cap > spare[pick].cap)
              pick = i;
          }
          if (pick < spare.size()) {
            j->proof = std::move(spare[pick]);
            spare.erase(spare.begin() + pick);
          }
        }
      }
      int rc = NENYA_OK;
      std::string err;
      wk.pool.fail_in = j->id == inject_job ? inject_at : 0;
      try {
        switch (j->params->k) {
          case 8: swap_in<W<8>>(wk, *j); run_t<W<8>>(wk, *j); break;
          case 16: swap_in<W<16>>(wk, *j); run_t<W<16>>(wk, *j); break;
          case 32: swap_in<W<32>>(wk, *j); run_t<W<32>>(wk, *j); break;
          default: swap_in<W<64>>(wk, *j); run_t<W<64>>(wk, *j); break;
        }
      } catch (const Error& e) {
        rc = e.code;
        err = e.what();
      } catch (const std::exception& e) {
        rc = NENYA_E_DEVICE;
        err = e.what();
      }
      wk.pool.fail_in = 0;
      if (rc == NENYA_OK && j->cancel.load()) {
        rc = NENYA_E_CANCELED;
        err = "job canceled while it ran";
      }
      {
        std::lock_guard<std::mutex> g(mu);
        in_use -= j->fp;
        running--;
        j->rc = rc;
        j->err = err;
        j->t_end = now_ms();
        j->done = true;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
  // the job's (recycled) buffer becomes the context's proof buffer
  template <class T> void swap_in(Worker& wk, Job& j) { std::swap(j.proof, wk.get<T>().proofbuf); }
};

}  // namespace nenya
