// Autotuner for the launch shapes of the sumcheck's throughput kernels.
//
// For each limb width and each power-of-two work size it times the candidate
// shapes on synthetic tables (the kernels' cost does not depend on the
// values) and records the fastest in Tuning, which every later launch reads:
//   * folds (k <= 32): split mode (k_fold_mx: a workgroup per 32-element
//     tile, the fragments read from L2) or bulk mode (k_fold_mx_bulk:
//     persistent workgroups, a wave per row-block, loads two tiles ahead)
//     with 256 / 512 / 1024 workgroups;
//   * k_gram_lin2 (k <= 32): 128 / 256 / 512 / 1024 workgroups;
//   * k_fold (k = 64, vector ALUs): a workgroup cap of 2048 .. 16384.
// The report lists, per width and size, the pick with its time and the
// built-in choice's time.
#pragma once
#include <sstream>

#include "snark.hpp"

namespace nenya {

// ms of one launch: the least of three means over `reps` launches each,
// after one warm-up (the least mean is the least disturbed by other work).
template <class F> double time_launch(hipStream_t st, F&& go, int reps = 4) {
  hipEvent_t a, b;
  HIP_CHECK(hipEventCreate(&a));
  HIP_CHECK(hipEventCreate(&b));
  go();
  double best = 1e30;
  for (int trial = 0; trial < 3; trial++) {
    HIP_CHECK(hipEventRecord(a, st));
    for (int i = 0; i < reps; i++) go();
    HIP_CHECK(hipEventRecord(b, st));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, (double)ms / reps);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

template <class T> void tune_width(std::ostringstream& js, int lmin, int lmax) {
  using L = typename T::L;
  Ctx<T> cx;
  Engine<T>& e = *cx.eng;
  const size_t NE = size_t(1) << lmax;
  DBuf<L> src(&cx.pool, 2 * NE * E), dst(&cx.pool, NE * E);
  HIP_CHECK(hipMemsetAsync(src.p, 0x5b, 2 * NE * E * sizeof(L), cx.st));
  Tuning& tu = Tuning::get();
  const int kx = Tuning::kidx(T::K);
  js << "\"k" << T::K << "\":{";
  if constexpr (kMx<T>) {
    DBuf<uint32_t> fr(&cx.pool, (size_t)kND<T> * 4096);
    hipLaunchKernelGGL(k_mulmat<T>, dim3(1), dim3(256), 0, cx.st, (const L*)e.d_one.p, fr.p);
    HIP_CHECK(hipGetLastError());
    DBuf<L> part(&cx.pool, (size_t)Engine<T>::GRAM_BLOCKS_MAX * 2 * E);
    js << "\"fold\":{";
    for (int lg = lmin; lg <= lmax; lg++) {
      const uint64_t m = uint64_t(1) << lg, ntiles = m / 32;
      FoldJob<T> J{{src.p, src.p}, {dst.p, dst.p}, {1, 0}, {0, 0}, {0, 0}, m, 0, nullptr, nullptr, nullptr};
      auto run = [&](int split, unsigned cap) {
        return time_launch(cx.st, [&] {
          if (split)
            hipLaunchKernelGGL(k_fold_mx<T>, dim3((unsigned)ntiles), dim3(256), 0, cx.st, (const uint32_t*)fr.p, J);
          else
            hipLaunchKernelGGL(k_fold_mx_bulk<T>, dim3((unsigned)std::min<uint64_t>(ntiles, cap)), dim3(256), 0,
                               cx.st, (const uint32_t*)fr.p, J);
        });
      };
      // the built-in choice (split up to 256 tiles, else 512 workgroups) is
      // one of the candidates; its own time is reported beside the pick
      const bool def_split = ntiles <= 256;
      double def = 0;
      int bmode = 0;
      unsigned bcap = 0;
      double best = 1e30;
      if (ntiles <= 4096) {
        const double t = run(1, 0);
        if (def_split) def = t;
        if (t < best) { best = t; bmode = 1; bcap = 0; }
      }
      for (unsigned cap : {256u, 512u, 1024u}) {
        const double t = run(0, cap);
        if (!def_split && cap == 512u) def = t;
        if (t < best) { best = t; bmode = 2; bcap = cap; }
      }
      const int lt = Tuning::lg(ntiles);
      tu.fold_mode[kx][lt] = bmode;
      tu.fold_cap[kx][lt] = (int)bcap;
      js << (lg > lmin ? "," : "") << "\"2^" << lg << "\":{\"pick\":\"" << (bmode == 1 ? "split" : "bulk")
         << "\",\"workgroups\":" << (bmode == 1 ? ntiles : std::min<uint64_t>(ntiles, bcap)) << ",\"us\":" << best * 1e3
         << ",\"builtin_us\":" << def * 1e3 << "}";
    }
    js << "},\"gram\":{";
    for (int lg = lmin; lg <= lmax; lg++) {
      const uint64_t m = uint64_t(1) << (lg - 1), nch = m / 32;  // 2^lg elements read per table pair half
      auto run = [&](unsigned nwg) {
        return time_launch(cx.st, [&] {
          hipLaunchKernelGGL(k_gram_lin2<T>, dim3(nwg), dim3(512), 0, cx.st, (const L*)src.p, (uint64_t)0,
                             (const L*)src.p, (uint64_t)0, (uint64_t)1, m, m, 128, part.p, 2, 0);
        });
      };
      const unsigned defn = (unsigned)std::min<uint64_t>(nch, Engine<T>::GRAM_BLOCKS);
      double def = 0;
      unsigned bn = 0;
      double best = 1e30;
      for (unsigned nwg : {128u, 256u, 512u, 1024u}) {
        if (nwg > nch && nwg != 128u) continue;
        const unsigned n = (unsigned)std::min<uint64_t>(nch, nwg);
        const double t = run(n);
        if (n == defn) def = t;
        if (t < best) { best = t; bn = n; }
      }
      tu.gram_blocks[kx][Tuning::lg(nch)] = (int)bn;
      js << (lg > lmin ? "," : "") << "\"2^" << lg << "\":{\"workgroups\":" << bn << ",\"us\":" << best * 1e3
         << ",\"builtin_us\":" << def * 1e3 << "}";
    }
    js << "}";
  } else {
    DBuf<L> r(&cx.pool, E);
    HIP_CHECK(hipMemcpyAsync(r.p, e.d_one.p, E * sizeof(L), hipMemcpyDeviceToDevice, cx.st));
    js << "\"fold\":{";
    for (int lg = lmin; lg <= lmax; lg++) {
      const uint64_t m = uint64_t(1) << lg;
      auto run = [&](unsigned cap) {
        return time_launch(cx.st, [&] {
          hipLaunchKernelGGL(k_fold<T>, dim3(grid_for(m, GPB, cap)), dim3(GB), 0, cx.st, (const L*)r.p,
                             (const L*)src.p, dst.p, m, (uint64_t)1, (uint64_t)0, (uint64_t)0);
        });
      };
      double def = 0;
      unsigned bc = 0;
      double best = 1e30;
      for (unsigned cap : {2048u, 4096u, 8192u, 16384u}) {
        const double t = run(cap);
        if (cap == 16384u) def = t;
        if (t < best) { best = t; bc = cap; }
      }
      tu.vfold_cap[Tuning::lg(m)] = (int)bc;
      js << (lg > lmin ? "," : "") << "\"2^" << lg << "\":{\"workgroups\":" << grid_for(m, GPB, bc) << ",\"us\":"
         << best * 1e3 << ",\"builtin_us\":" << def * 1e3 << "}";
    }
    js << "}";
  }
  HIP_CHECK(hipStreamSynchronize(cx.st));
  js << "}";
}

}  // namespace nenya
