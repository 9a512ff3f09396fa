// Matrix packing on the device (setup, not prover time): the groupings of
// pack_general_matrices / pack_mimd_matrices (reduce.rs:249-273, 372-438,
// 618-651) and the per-proof CSRs of the SIMD matvecs, built from COO
// entries uploaded once.
//
// Every grouping is a stable radix sort of (key, entry index) pairs
// (rocPRIM's LSD radix sort is stable), so entries keep their input order
// within a group exactly as the host's counting sorts and stable sorts left
// them; group pointers are lower bounds of the group ids in the sorted keys;
// merges of equal keys (the MIMD union-support cells) are runs of the sorted
// keys, numbered by a scan of run heads.
#pragma once
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "prover.hpp"

// Included by snark.hpp after HostCoo and Csr.
namespace nenya {

__global__ void __launch_bounds__(256) k_iota(uint32_t* out, uint64_t n, uint32_t base) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = base + (uint32_t)i;
}

// ptr[g] = the first position of sorted `keys` whose key is >= g, g <= ngroups
template <class K, class P>
__global__ void __launch_bounds__(256) k_group_ptr(const K* keys, uint64_t n, uint64_t ngroups, P base, P* ptr) {
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g <= ngroups;
       g += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if ((uint64_t)keys[mid] < g) lo = mid + 1;
      else hi = mid;
    }
    ptr[g] = base + (P)lo;
  }
}

// CSR arrays from the sorted permutation: idx = minor[perm], val = v[perm]
template <class L>
__global__ void __launch_bounds__(256) k_csr_gather(const uint32_t* perm, const uint32_t* minor, const L* v, uint64_t n,
                                                    uint32_t* idx, L* val) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t e = perm[i];
    idx[i] = minor[e];
    val[i] = v[e];
  }
}

// Majors with at least `heavy` entries, in ascending order: a flag scan
// (pos, exclusive) tells each its slot; maxcnt gets the largest count.
__global__ void __launch_bounds__(256) k_heavy_flags(const uint32_t* ptr, uint64_t nmaj, uint32_t heavy, uint32_t* flag) {
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < nmaj; o += (uint64_t)gridDim.x * blockDim.x)
    flag[o] = ptr[o + 1] - ptr[o] >= heavy ? 1u : 0u;
}
__global__ void __launch_bounds__(256) k_heavy_place(const uint32_t* ptr, const uint32_t* flag, const uint32_t* pos,
                                                     uint64_t nmaj, uint32_t* list, uint32_t* maxcnt) {
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < nmaj; o += (uint64_t)gridDim.x * blockDim.x) {
    if (!flag[o]) continue;
    list[pos[o]] = (uint32_t)o;
    atomicMax(maxcnt, ptr[o + 1] - ptr[o]);
  }
}

// Layer-row entry keys of General (matrix-core widths): output
// (l = (c - r) mod 32, q = c / 32) of entry (r, c), as l * nb + q.
__global__ void __launch_bounds__(256) k_layer_keys(const uint32_t* r, const uint32_t* c, uint64_t n, uint64_t nb,
                                                    uint32_t* key) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t rr = r[i], cc = c[i];
    key[i] = (uint32_t)((uint64_t)((cc - rr) & 31u) * nb + cc / 32);
  }
}
// ... and their codes p << 7 | w << 5 | s with values, from the sorted
// permutation (entry e of matrix w when off[w] <= e < off[w + 1])
template <class L>
__global__ void __launch_bounds__(256) k_layer_place(const uint32_t* perm, const uint32_t* r, const L* v, uint64_t n,
                                                     uint64_t off1, uint64_t off2, uint32_t* code, L* val) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t e = perm[i];
    const uint32_t w = e >= off2 ? 2u : e >= off1 ? 1u : 0u;
    const uint32_t rr = r[e];
    code[i] = ((rr / 32) << 7) | (w << 5) | (rr % 32);
    val[i] = v[e];
  }
}

// MIMD cell keys r * m + c
__global__ void __launch_bounds__(256) k_rc_keys(const uint32_t* r, const uint32_t* c, uint64_t n, uint64_t m,
                                                 uint64_t* key) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    key[i] = (uint64_t)r[i] * m + c[i];
}

// Run heads of sorted keys (1 where a key differs from its predecessor).
template <class K>
__global__ void __launch_bounds__(256) k_run_heads(const K* key, uint64_t n, uint32_t* head) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    head[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

// Cells from runs: run j (heads scanned inclusively: run id = id[i] - 1)
// owns positions [start_j, start_{j+1}); the thread at a run head sums the
// run's values into the cell's 32 slots (slot of entry e: slot[e], its
// instance) and records the cell's key.  slots must start zeroed.
template <class L>
__global__ void __launch_bounds__(256) k_cell_sum(const uint64_t* skey, const uint32_t* perm, const uint32_t* id,
                                                  uint64_t n, const uint8_t* slot, const L* v, uint64_t* ckey, L* slots) {
  using C = typename std::conditional<sizeof(L) == 8, uint64_t, uint32_t>::type;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (i != 0 && skey[i] == skey[i - 1]) continue;
    const uint64_t cell = id[i] - 1;
    ckey[cell] = skey[i];
    L* out = slots + cell * 32;
    for (uint64_t j = i; j < n && skey[j] == skey[i]; j++) {
      const uint32_t e = perm[j];
      L& x = out[slot[e]];
      x = (L)((C)x + (C)v[e]);
    }
  }
}

// MIMD cells: row and column of each cell from its key r * m + c
__global__ void __launch_bounds__(256) k_cell_rc(const uint64_t* ckey, uint64_t ncell, uint64_t m, uint32_t* row,
                                                 uint32_t* col) {
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ncell; g += (uint64_t)gridDim.x * blockDim.x) {
    row[g] = (uint32_t)(ckey[g] / m);
    col[g] = (uint32_t)(ckey[g] % m);
  }
}
// Nonzero instance values of the cells, in (cell, slot) order: flag, then
// (after an exclusive scan) their column keys and (cell, slot) positions
template <class L>
__global__ void __launch_bounds__(256) k_nz_flags(const L* cval, uint64_t n, uint32_t* flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    flag[i] = cval[i] ? 1u : 0u;
}
__global__ void __launch_bounds__(256) k_nz_place(const uint32_t* flag, const uint32_t* pos, const uint32_t* col,
                                                  uint64_t n, uint32_t* key, uint32_t* at) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (!flag[i]) continue;
    key[pos[i]] = col[i / 32];
    at[pos[i]] = (uint32_t)i;
  }
}
// MIMD M_gamma-row entries (matrix cores): code row << 7 | w << 5 | slot
template <class L>
__global__ void __launch_bounds__(256) k_nz_codes(const uint32_t* perm, const uint32_t* at, const uint32_t* row,
                                                  const L* cval, uint64_t n, uint64_t c1, uint64_t c2, uint32_t* code,
                                                  L* val) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t x = at[perm[i]], g = x / 32, sl = x % 32;
    const uint32_t w = g >= c2 ? 2u : g >= c1 ? 1u : 0u;
    code[i] = (row[g] << 7) | (w << 5) | sl;
    val[i] = cval[x];
  }
}
// MIMD entries' instance numbers from the per-instance offsets
__global__ void __launch_bounds__(256) k_inst_of(const uint64_t* off, uint64_t n, uint8_t* inst) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    int lo = 0;
#pragma unroll
    for (int s = 16; s > 0; s >>= 1)
      if (off[lo + s] <= i) lo += s;
    inst[i] = (uint8_t)lo;
  }
}

inline int bits_for(uint64_t maxval) {  // bits that hold every key <= maxval
  int b = 1;
  while (b < 64 && (maxval >> b) != 0) b++;
  return b;
}

// Host-side driver of the device groupings, on an engine's stream and pool.
template <class T> struct DevSetup {
  using L = typename T::L;
  Engine<T>& e;
  explicit DevSetup(Engine<T>& en) : e(en) {}
  static unsigned grid(uint64_t n) { return grid_for(std::max<uint64_t>(n, 1), 256, 1u << 16); }

  // stable sort of (key, value) pairs (keys below 2^end_bit)
  template <class K>
  void sort_pairs(const K* kin, K* kout, const uint32_t* vin, uint32_t* vout, uint64_t n, int end_bit) {
    if (n == 0) return;
    size_t tb = 0;
    HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, kin, kout, vin, vout, (size_t)n, 0, end_bit, e.st));
    DBuf<uint8_t> tmp(e.pool, std::max<size_t>(tb, 1));
    HIP_CHECK(rocprim::radix_sort_pairs((void*)tmp.p, tb, kin, kout, vin, vout, (size_t)n, 0, end_bit, e.st));
  }
  // out[i] = sum_{j <= i} in[j] (inclusive) or sum_{j < i} (exclusive)
  void scan(const uint32_t* in, uint32_t* out, uint64_t n, bool inclusive) {
    if (n == 0) return;
    size_t tb = 0;
    if (inclusive) {
      HIP_CHECK(rocprim::inclusive_scan(nullptr, tb, in, out, (size_t)n, rocprim::plus<uint32_t>(), e.st));
      DBuf<uint8_t> tmp(e.pool, std::max<size_t>(tb, 1));
      HIP_CHECK(rocprim::inclusive_scan((void*)tmp.p, tb, in, out, (size_t)n, rocprim::plus<uint32_t>(), e.st));
    } else {
      HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), e.st));
      DBuf<uint8_t> tmp(e.pool, std::max<size_t>(tb, 1));
      HIP_CHECK(rocprim::exclusive_scan((void*)tmp.p, tb, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), e.st));
    }
  }
  uint32_t last_of(const uint32_t* d, uint64_t n) {
    uint32_t x = 0;
    if (n) e.d2h(&x, d + n - 1, 4);
    return x;
  }
  DBuf<uint32_t> iota(uint64_t n, uint32_t base = 0) {
    DBuf<uint32_t> v(e.pool, std::max<uint64_t>(n, 1));
    e.launch("k_iota", 4.0 * n, 0, k_iota, dim3(grid(n)), dim3(256), v.p, n, base);
    return v;
  }

  // The CSR of n entries grouped by major (maj < nmaj), minor index and value
  // per entry, with the heavy majors listed as Csr::prepare lists them.
  void csr(Csr<T>& m, const uint32_t* maj, const uint32_t* minor, const L* v, uint64_t n, uint64_t nmaj) {
    m.nnz = n;
    m.ptr = DBuf<uint32_t>(e.pool, nmaj + 1);
    m.idx = DBuf<uint32_t>(e.pool, std::max<uint64_t>(n, 1));
    m.val = DBuf<L>(e.pool, std::max<uint64_t>(n, 1));
    DBuf<uint32_t> skey(e.pool, std::max<uint64_t>(n, 1)), perm(e.pool, std::max<uint64_t>(n, 1));
    {
      DBuf<uint32_t> id = iota(n);
      sort_pairs<uint32_t>(maj, skey.p, id.p, perm.p, n, bits_for(nmaj));
    }
    e.launch("k_group_ptr", 4.0 * nmaj, 0, (k_group_ptr<uint32_t, uint32_t>), dim3(grid(nmaj + 1)), dim3(256),
             (const uint32_t*)skey.p, n, nmaj, 0u, m.ptr.p);
    e.launch("k_csr_gather", 12.0 * n, 0, (k_csr_gather<L>), dim3(grid(n)), dim3(256), (const uint32_t*)perm.p, minor, v,
             n, m.idx.p, m.val.p);
    // heavy majors (Csr::HEAVY or more entries), ascending
    DBuf<uint32_t> flag(e.pool, std::max<uint64_t>(nmaj, 1)), pos(e.pool, std::max<uint64_t>(nmaj, 1));
    e.launch("k_heavy_flags", 4.0 * nmaj, 0, k_heavy_flags, dim3(grid(nmaj)), dim3(256), (const uint32_t*)m.ptr.p, nmaj,
             Csr<T>::HEAVY, flag.p);
    scan(flag.p, pos.p, nmaj, false);
    const uint32_t nh = last_of(pos.p, nmaj) + last_of(flag.p, nmaj);
    m.nheavy = nh;
    m.maxheavy = 0;
    if (nh) {
      m.heavy = DBuf<uint32_t>(e.pool, nh);
      DBuf<uint32_t> mx(e.pool, 1);
      HIP_CHECK(hipMemsetAsync(mx.p, 0, 4, e.st));
      e.launch("k_heavy_place", 8.0 * nmaj, 0, k_heavy_place, dim3(grid(nmaj)), dim3(256), (const uint32_t*)m.ptr.p,
               (const uint32_t*)flag.p, (const uint32_t*)pos.p, nmaj, m.heavy.p, mx.p);
      e.d2h(&m.maxheavy, mx.p, 4);
    }
  }
};

// COO matrices uploaded once (concatenated r, c, v with per-matrix offsets).
template <class T> struct DevCoo {
  using L = typename T::L;
  DBuf<uint32_t> r, c;
  DBuf<L> v;
  uint64_t off[4] = {0, 0, 0, 0};
  DevCoo(Engine<T>& e, std::initializer_list<const HostCoo*> ms) {
    uint64_t tot = 0;
    int i = 0;
    for (const HostCoo* m : ms) {
      off[i++] = tot;
      tot += m->nnz;
    }
    for (; i < 4; i++) off[i] = tot;
    r = DBuf<uint32_t>(e.pool, std::max<uint64_t>(tot, 1));
    c = DBuf<uint32_t>(e.pool, std::max<uint64_t>(tot, 1));
    v = DBuf<L>(e.pool, std::max<uint64_t>(tot, 1));
    std::vector<typename Engine<T>::Up> ups;
    i = 0;
    for (const HostCoo* m : ms) {
      ups.push_back({r.p + off[i], m->r, m->nnz * 4});
      ups.push_back({c.p + off[i], m->c, m->nnz * 4});
      ups.push_back({v.p + off[i], m->v, m->nnz * sizeof(L)});
      i++;
    }
    e.h2d_many(ups);
  }
};

}  // namespace nenya
