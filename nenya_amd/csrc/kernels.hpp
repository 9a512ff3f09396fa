// gfx950 kernels for the Nenya prover, templated on the limb width.
//
// Core primitive: a 32-lane group computes GR(2^k,128) products.  Lane t owns
// R1 coefficient t (4 limbs) of the result.  For a = sum a_i Y^i and
// b = sum b_j Y^j it accumulates the raw alpha-convolutions of
//   circ_t = sum_i a_i b_{(t-i) mod 32}      (all 32 terms, uniform per lane)
//   hi_t   = sum_{i>t} a_i b_{t+32-i}        (the wrap-around part = Y^(32+t))
// Sums of products (round polynomials, matvecs, m-bar) accumulate these raw
// values and reduce once: alpha-reduce (X^4 = -X - 1), lo = circ - hi, then
// Y^32 = -Y^3 - Y - alpha with the cascade of positions 61, 62 into 32, 33.
// The result equals the reference's schoolbook + top-down reduction
// (galois-ring/src/gr128.rs:215-296) because both compute the unique
// representative of a*b mod f2.
//
// Operands are staged in LDS by the group; every kernel uses 64-thread
// blocks (one wave, two groups) so __syncthreads is a single-wave barrier.
#pragma once
#include "blake3.hpp"
#include "common.hpp"

namespace nenya {

constexpr int GB = 64;  // threads per block for group kernels
constexpr int GPB = GB / 32;

template <class T> using LT = typename T::L;
template <class T> using CT = typename T::C;

template <class T> __device__ __forceinline__ void ld4(const LT<T>* p, CT<T> o[4]) {
  using L = LT<T>;
  if constexpr (sizeof(L) == 4) {
    uint4 v = *reinterpret_cast<const uint4*>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  } else if constexpr (sizeof(L) == 8) {
    const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
    ulonglong2 a = q[0], b = q[1];
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
  } else if constexpr (sizeof(L) == 2) {
    uint2 v = *reinterpret_cast<const uint2*>(p);
    o[0] = v.x & 0xffffu; o[1] = v.x >> 16; o[2] = v.y & 0xffffu; o[3] = v.y >> 16;
  } else {
    uint32_t v = *reinterpret_cast<const uint32_t*>(p);
    o[0] = v & 0xff; o[1] = (v >> 8) & 0xff; o[2] = (v >> 16) & 0xff; o[3] = v >> 24;
  }
}
template <class T> __device__ __forceinline__ void st4(LT<T>* p, const CT<T> v[4]) {
  using L = LT<T>;
  if constexpr (sizeof(L) == 4) {
    *reinterpret_cast<uint4*>(p) = make_uint4(v[0], v[1], v[2], v[3]);
  } else if constexpr (sizeof(L) == 8) {
    ulonglong2* q = reinterpret_cast<ulonglong2*>(p);
    q[0] = make_ulonglong2(v[0], v[1]);
    q[1] = make_ulonglong2(v[2], v[3]);
  } else if constexpr (sizeof(L) == 2) {
    *reinterpret_cast<uint2*>(p) = make_uint2((v[0] & 0xffffu) | (v[1] << 16), (v[2] & 0xffffu) | (v[3] << 16));
  } else {
    *reinterpret_cast<uint32_t*>(p) = (v[0] & 0xff) | ((v[1] & 0xff) << 8) | ((v[2] & 0xff) << 16) | (v[3] << 24);
  }
}

template <class C> __device__ __forceinline__ void d_raw(const C a[4], const C b[4], C p[7]) {
  p[0] = a[0] * b[0];
  p[1] = a[0] * b[1] + a[1] * b[0];
  p[2] = a[0] * b[2] + a[1] * b[1] + a[2] * b[0];
  p[3] = a[0] * b[3] + a[1] * b[2] + a[2] * b[1] + a[3] * b[0];
  p[4] = a[1] * b[3] + a[2] * b[2] + a[3] * b[1];
  p[5] = a[2] * b[3] + a[3] * b[2];
  p[6] = a[3] * b[3];
}
template <class C> __device__ __forceinline__ void d_red(const C p[7], C o[4]) {
  o[0] = p[0] - p[4];
  o[1] = p[1] - p[4] - p[5];
  o[2] = p[2] - p[5] - p[6];
  o[3] = p[3] - p[6];
}

template <class T> struct Acc {
  CT<T> c[7];
  CT<T> h[7];
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int i = 0; i < 7; i++) c[i] = h[i] = 0;
  }
};

// acc += A * B for the group; A has beta-degree < NA (16 for packed images).
template <class T, int NA = 32>
__device__ __forceinline__ void acc_mul(Acc<T>& acc, const LT<T>* A, const LT<T>* B, int t) {
  using C = CT<T>;
#pragma unroll 4
  for (int i = 0; i < NA; i++) {
    C a[4], b[4], p[7];
    ld4<T>(A + 4 * i, a);
    ld4<T>(B + 4 * ((t - i) & 31), b);
    d_raw(a, b, p);
    const C m = (i > t) ? (C)~(C)0 : (C)0;
#pragma unroll
    for (int c = 0; c < 7; c++) {
      acc.c[c] += p[c];
      acc.h[c] += p[c] & m;
    }
  }
}

// acc += the pair terms i0 <= i < i1 of A * B (lane t of a group).
template <class T>
__device__ __forceinline__ void acc_mul_range(Acc<T>& acc, const LT<T>* A, const LT<T>* B, int t, int i0, int i1) {
  using C = CT<T>;
#pragma unroll 4
  for (int i = i0; i < i1; i++) {
    C a[4], b[4], p[7];
    ld4<T>(A + 4 * i, a);
    ld4<T>(B + 4 * ((t - i) & 31), b);
    d_raw(a, b, p);
    const C m = (i > t) ? (C)~(C)0 : (C)0;
#pragma unroll
    for (int c = 0; c < 7; c++) {
      acc.c[c] += p[c];
      acc.h[c] += p[c] & m;
    }
  }
}

// Reduce the group's accumulated raw sums; lane t gets coefficient t.
template <class T> __device__ __forceinline__ void acc_final(const Acc<T>& acc, int t, CT<T> out[4]) {
  using C = CT<T>;
  C cl[4], H[4];
  d_red(acc.c, cl);
  d_red(acc.h, H);
  C h[4], hm1[4], hm3[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    C x = __shfl(H[c], (t < 2) ? t + 29 : t, 32);
    h[c] = (t < 2) ? H[c] - x : H[c];
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    C a = __shfl(h[c], (t >= 1) ? t - 1 : 0, 32);
    C b = __shfl(h[c], (t >= 3) ? t - 3 : 0, 32);
    hm1[c] = (t >= 1) ? a : (C)0;
    hm3[c] = (t >= 3) ? b : (C)0;
  }
  C lo[4];
#pragma unroll
  for (int c = 0; c < 4; c++) lo[c] = cl[c] - H[c] - hm1[c] - hm3[c];
  out[0] = lo[0] + h[3];
  out[1] = lo[1] - (h[0] - h[3]);
  out[2] = lo[2] - h[1];
  out[3] = lo[3] - h[2];
}

// Stage one element (128 limbs) into group LDS: lane t copies coefficient t.
template <class T> __device__ __forceinline__ void stage(LT<T>* dst, const LT<T>* src, int t) {
  CT<T> v[4];
  ld4<T>(src + 4 * t, v);
  st4<T>(dst + 4 * t, v);
}

// ============================================================================
// Element-wise group kernels
// ============================================================================

// out[i] = A[i]*B[i] - (S ? S[i] : 0)   (S may alias out)
template <class T, int NA>
__global__ void __launch_bounds__(GB) k_mul_sub(const LT<T>* A, const LT<T>* B, const LT<T>* S, LT<T>* out, uint64_t n) {
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][2][E];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < n; base += (uint64_t)gridDim.x * GPB) {
    uint64_t i = base + g;
    bool ok = i < n;
    if (ok) { stage<T>(sh[g][0], A + i * E, t); stage<T>(sh[g][1], B + i * E, t); }
    __syncthreads();
    Acc<T> acc;
    acc.clear();
    if (ok) acc_mul<T, NA>(acc, sh[g][0], sh[g][1], t);
    CT<T> o[4];
    acc_final<T>(acc, t, o);
    if (ok) {
      if (S) {
        CT<T> s[4];
        ld4<T>(S + i * E + 4 * t, s);
        for (int c = 0; c < 4; c++) o[c] -= s[c];
      }
      st4<T>(out + i * E + 4 * t, o);
    }
    __syncthreads();
  }
}

// out[i] = s * X[i]  (s one shared element)
template <class T>
__global__ void __launch_bounds__(GB) k_scale(const LT<T>* s, const LT<T>* X, LT<T>* out, uint64_t n) {
  __shared__ __attribute__((aligned(16))) LT<T> ss[E];
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][E];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  if (threadIdx.x < 32) stage<T>(ss, s, t);
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < n; base += (uint64_t)gridDim.x * GPB) {
    uint64_t i = base + g;
    bool ok = i < n;
    if (ok) stage<T>(sh[g], X + i * E, t);
    __syncthreads();
    Acc<T> acc;
    acc.clear();
    if (ok) acc_mul<T>(acc, ss, sh[g], t);
    CT<T> o[4];
    acc_final<T>(acc, t, o);
    if (ok) st4<T>(out + i * E + 4 * t, o);
    __syncthreads();
  }
}

// Mle::fold_in_place (mle.rs:42-54): dst[i] = lo[i] + r*(hi[i] - lo[i]),
// lo = src[i], hi = src[i+m].  dst may equal src (each i reads then writes i).
template <class T>
__global__ void __launch_bounds__(GB) k_fold(const LT<T>* r, const LT<T>* src, LT<T>* dst, uint64_t m, uint64_t ntab,
                                             uint64_t src_stride, uint64_t dst_stride) {
  __shared__ __attribute__((aligned(16))) LT<T> sr[E];
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][E];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  if (threadIdx.x < 32) stage<T>(sr, r, t);
  uint64_t total = m * ntab;
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < total; base += (uint64_t)gridDim.x * GPB) {
    uint64_t w = base + g;
    bool ok = w < total;
    uint64_t tab = ok ? w / m : 0, i = ok ? w % m : 0;
    const LT<T>* s = src + tab * src_stride * E;
    CT<T> lo[4], hi[4];
    if (ok) {
      ld4<T>(s + i * E + 4 * t, lo);
      ld4<T>(s + (i + m) * E + 4 * t, hi);
      CT<T> d[4] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], hi[3] - lo[3]};
      st4<T>(sh[g] + 4 * t, d);
    }
    __syncthreads();
    Acc<T> acc;
    acc.clear();
    if (ok) acc_mul<T>(acc, sr, sh[g], t);
    CT<T> o[4];
    acc_final<T>(acc, t, o);
    if (ok) {
      for (int c = 0; c < 4; c++) o[c] += lo[c];
      st4<T>(dst + tab * dst_stride * E + i * E + 4 * t, o);
    }
    __syncthreads();
  }
}

// marginalize_and_scale (mult_check.rs:302-313): eg[i] = scale*(eg[i]+eg[i+m])
// (block blk of nblk; the kernel runs it as block blockIdx.x of gridDim.x)
template <class T>
__device__ __forceinline__ void marg_body(const LT<T>* scale, LT<T>* eg, uint64_t m, unsigned blk, unsigned nblk) {
  __shared__ __attribute__((aligned(16))) LT<T> ss[E];
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][E];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  if (threadIdx.x < 32) stage<T>(ss, scale, t);
  for (uint64_t base = (uint64_t)blk * GPB; base < m; base += (uint64_t)nblk * GPB) {
    uint64_t i = base + g;
    bool ok = i < m;
    if (ok) {
      CT<T> a[4], b[4];
      ld4<T>(eg + i * E + 4 * t, a);
      ld4<T>(eg + (i + m) * E + 4 * t, b);
      for (int c = 0; c < 4; c++) a[c] += b[c];
      st4<T>(sh[g] + 4 * t, a);
    }
    __syncthreads();
    Acc<T> acc;
    acc.clear();
    if (ok) acc_mul<T>(acc, ss, sh[g], t);
    CT<T> o[4];
    acc_final<T>(acc, t, o);
    if (ok) st4<T>(eg + i * E + 4 * t, o);
    __syncthreads();
  }
}
template <class T>
__global__ void __launch_bounds__(GB) k_marg(const LT<T>* scale, LT<T>* eg, uint64_t m) {
  marg_body<T>(scale, eg, m, blockIdx.x, gridDim.x);
}

// build_eq_table (eq.rs:35-58) by chunks: the table over variables
// r_0..r_{n-1} (r_0 the most significant index bit) is the outer product of
// the tables over consecutive chunks of at most EQ_HEAD variables.
constexpr int EQ_HEAD = 6;

// The chunk tables, one block each: chunk j of k (n/k variables, one more
// for the first n % k chunks) is written to out + off_j * E, off_j the
// entries of the chunks before it.  Within a chunk, level l (as one doubling
// level of build_eq_table: next[2p+1] = prev[p] r_l, next[2p] = prev[p] -
// next[2p+1]) has 2^l products; a product takes S = min(8, 32 / 2^l)
// 32-lane groups, group q of them summing the pair terms
// [32q/S, 32(q+1)/S), so the early levels run at a fraction of one product's
// latency.  The entries stay in LDS between levels (each level reads all of
// them before it writes any); only the last level goes to memory.
template <class T>
__global__ void __launch_bounds__(1024) k_eq_chunks(const LT<T>* r, int n, int k, LT<T>* out) {
  using C = CT<T>;
  __shared__ __attribute__((aligned(16))) LT<T> sr[EQ_HEAD][E];
  __shared__ __attribute__((aligned(16))) LT<T> sh[32][E];
  __shared__ C red[32][14][32];
  const int g = threadIdx.x >> 5, t = threadIdx.x & 31, j = blockIdx.x;
  const int q0 = n / k, rem = n % k;
  const int nl = q0 + (j < rem ? 1 : 0), start = j * q0 + (j < rem ? j : rem);
  uint64_t off = 0;
  for (int i = 0; i < j; i++) off += (uint64_t)1 << (q0 + (i < rem ? 1 : 0));
  out += off * E;
  if (g < nl) stage<T>(sr[g], r + (uint64_t)(start + g) * E, t);
  if (g == 0) {
    const C one[4] = {(C)(t == 0 ? 1 : 0), 0, 0, 0};
    st4<T>(sh[0] + 4 * t, one);
  }
  for (int l = 0; l < nl; l++) {
    const int P = 1 << l, S = P >= 32 ? 1 : (32 / P < 8 ? 32 / P : 8), span = 32 / S;
    const int p = g / S, q = g % S;
    const bool active = p < P;
    __syncthreads();  // the previous level's entries (or the start values) are in place
    C pv[4];
    if (active && q == 0) ld4<T>(sh[p] + 4 * t, pv);
    Acc<T> acc;
    acc.clear();
    if (active) acc_mul_range<T>(acc, sr[l], sh[p], t, q * span, (q + 1) * span);
    if (active && q > 0)
#pragma unroll
      for (int c = 0; c < 7; c++) {
        red[g][c][t] = acc.c[c];
        red[g][7 + c][t] = acc.h[c];
      }
    __syncthreads();  // every read of this level's entries is done
    if (active && q == 0) {
      for (int kk = 1; kk < S; kk++)
#pragma unroll
        for (int c = 0; c < 7; c++) {
          acc.c[c] += red[g + kk][c][t];
          acc.h[c] += red[g + kk][7 + c][t];
        }
      C o[4];
      acc_final<T>(acc, t, o);
      C lo[4] = {pv[0] - o[0], pv[1] - o[1], pv[2] - o[2], pv[3] - o[3]};
      if (l == nl - 1) {
        st4<T>(out + (2 * (uint64_t)p + 1) * E + 4 * t, o);
        st4<T>(out + (2 * (uint64_t)p) * E + 4 * t, lo);
      } else {
        st4<T>(sh[2 * p + 1] + 4 * t, o);
        st4<T>(sh[2 * p] + 4 * t, lo);
      }
    }
  }
}

// out[a * nb + b] = A[a] * B[b]: the table over (A's variables, B's
// variables) from the two tables.
template <class T>
__global__ void __launch_bounds__(GB) k_eq_outer(const LT<T>* A, uint64_t na, const LT<T>* B, uint64_t nb, LT<T>* out) {
  __shared__ __attribute__((aligned(16))) LT<T> sa[GPB][E], sb[GPB][E];
  const int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  const uint64_t total = na * nb;
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < total; base += (uint64_t)gridDim.x * GPB) {
    const uint64_t x = base + g;
    const bool ok = x < total;
    if (ok) {
      stage<T>(sa[g], A + (x / nb) * E, t);
      stage<T>(sb[g], B + (x % nb) * E, t);
    }
    __syncthreads();
    Acc<T> acc;
    acc.clear();
    if (ok) acc_mul<T>(acc, sa[g], sb[g], t);
    CT<T> o[4];
    acc_final<T>(acc, t, o);
    if (ok) st4<T>(out + x * E + 4 * t, o);
    __syncthreads();
  }
}

// ============================================================================
// Sumcheck round kernels: per-block raw partial sums -> reduced partials.
// ============================================================================

// Write the block's reduced sum of nsum accumulators.  Groups' raw sums are
// added through LDS (linear), then group 0 reduces.
// Sums that many blocks add atomically are spread over SUM_COPIES copies
// (block b adds into copy b % SUM_COPIES): same-address atomics serialize at
// the L2, so one copy would make every block wait for all the others.
constexpr int SUM_COPIES = 16;

// Integer atomic add (wrap-around) for 32- and 64-bit words.
__device__ __forceinline__ void atom_add(uint32_t* p, uint32_t v) { atomicAdd(p, v); }
__device__ __forceinline__ void atom_add(uint64_t* p, uint64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

// The block's NS sums: written as partials of block blk, or (sums != null)
// added with integer atomics into sums[s] (exact: the sums are taken mod
// 2^32 or 2^64, and 2^k divides both); only_s >= 0 emits that sum alone
// (the others are zero).
// One sum of a block, as sum s of NSO: partials of block blk (the other
// NSO - 1 sums written as zero), or added atomically into sums[s].
template <class T>
__device__ __forceinline__ void emit_one(const Acc<T>& acc, int s, int nso, LT<T>* partials, int t, int g,
                                         unsigned blk, CT<T>* sums) {
  __shared__ CT<T> red[GPB][14][32];
  for (int c = 0; c < 7; c++) {
    red[g][c][t] = acc.c[c];
    red[g][7 + c][t] = acc.h[c];
  }
  __syncthreads();
  if (g == 0) {
    Acc<T> a = acc;
    for (int gg = 1; gg < GPB; gg++)
      for (int c = 0; c < 7; c++) { a.c[c] += red[gg][c][t]; a.h[c] += red[gg][7 + c][t]; }
    CT<T> o[4];
    acc_final<T>(a, t, o);
    if (sums) {
      for (int c = 0; c < 4; c++) atom_add(sums + s * E + 4 * t + c, o[c]);
    } else {
      const CT<T> z[4] = {0, 0, 0, 0};
      for (int so = 0; so < nso; so++) st4<T>(partials + ((uint64_t)blk * nso + so) * E + 4 * t, so == s ? o : z);
    }
  }
}

template <class T, int NS>
__device__ __forceinline__ void block_emit(Acc<T> (&acc)[NS], LT<T>* partials, int t, int g, unsigned blk,
                                           CT<T>* sums = nullptr, int only_s = -1) {
  __shared__ CT<T> red[GPB][NS][14][32];
  for (int s = 0; s < NS; s++)
    for (int c = 0; c < 7; c++) {
      red[g][s][c][t] = acc[s].c[c];
      red[g][s][7 + c][t] = acc[s].h[c];
    }
  __syncthreads();
  if (g == 0) {
    for (int s = 0; s < NS; s++) {
      if (only_s >= 0 && s != only_s) continue;
      Acc<T> a = acc[s];
      for (int gg = 1; gg < GPB; gg++)
        for (int c = 0; c < 7; c++) { a.c[c] += red[gg][s][c][t]; a.h[c] += red[gg][s][7 + c][t]; }
      CT<T> o[4];
      acc_final<T>(a, t, o);
      if (sums) {
        for (int c = 0; c < 4; c++) atom_add(sums + s * E + 4 * t + c, o[c]);
      } else {
        st4<T>(partials + ((uint64_t)blk * NS + s) * E + 4 * t, o);
      }
    }
  }
}

// Gruen mult-check round (mult_check.rs:250-296, sub form 426-461):
//   s0 += eg*(a_lo b_lo - c_lo - e_lo), s1 += eg*(a_hi b_hi - c_hi - e_hi),
//   sinf += eg*(da db).  c, e may be null (plain form).
// SPLIT (small m, latency-bound): block b computes only sum s = b % 3 over
// element groups b / 3, so an element's three product chains run on three
// blocks at once; the other two partial sums of the block are zero.
// (block blk of nblk, a multiple of 3 with SPLIT; sums != null: the block's
// sums are added atomically into sums[3] instead of written as partials)
template <class T, bool SPLIT>
__device__ __forceinline__ void mult_round_body(const LT<T>* eg, const LT<T>* a, const LT<T>* b, const LT<T>* cc,
                                                const LT<T>* ee, uint64_t m, LT<T>* partials, CT<T>* sums,
                                                unsigned blk, unsigned nblk) {
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][3][E];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  const int sblk = SPLIT ? (int)(blk % 3) : 0;
  const uint64_t bid = SPLIT ? blk / 3 : blk, nbid = SPLIT ? nblk / 3 : nblk;
  // SPLIT: one accumulator, for sum sblk (a runtime index into three would
  // put them in scratch memory)
  constexpr int NACC = SPLIT ? 1 : 3;
  Acc<T> acc[NACC];
  for (int s = 0; s < NACC; s++) acc[s].clear();
  for (uint64_t base = bid * GPB; base < m; base += nbid * GPB) {
    uint64_t i = base + g;
    bool ok = i < m;
    CT<T> alo[4] = {0, 0, 0, 0}, ahi[4] = {0, 0, 0, 0}, blo[4] = {0, 0, 0, 0}, bhi[4] = {0, 0, 0, 0};
    if (ok) {
      if (!SPLIT || sblk != 1) {
        ld4<T>(a + i * E + 4 * t, alo);
        ld4<T>(b + i * E + 4 * t, blo);
      }
      if (!SPLIT || sblk != 0) {
        ld4<T>(a + (i + m) * E + 4 * t, ahi);
        ld4<T>(b + (i + m) * E + 4 * t, bhi);
      }
      stage<T>(sh[g][2], eg + i * E, t);
    }
    // three products of the varying tables, each then weighted by eg
    for (int s = 0; s < 3; s++) {
      if (SPLIT && s != sblk) continue;  // uniform over the block
      CT<T> x[4], y[4];
      for (int c = 0; c < 4; c++) {
        x[c] = s == 0 ? alo[c] : s == 1 ? ahi[c] : ahi[c] - alo[c];
        y[c] = s == 0 ? blo[c] : s == 1 ? bhi[c] : bhi[c] - blo[c];
      }
      st4<T>(sh[g][0] + 4 * t, x);
      st4<T>(sh[g][1] + 4 * t, y);
      __syncthreads();
      Acc<T> pa;
      pa.clear();
      if (ok) acc_mul<T>(pa, sh[g][0], sh[g][1], t);
      CT<T> p[4];
      acc_final<T>(pa, t, p);
      if (ok && s < 2 && cc) {
        CT<T> c4[4], e4[4];
        uint64_t off = (s == 0 ? i : i + m) * E + 4 * t;
        ld4<T>(cc + off, c4);
        ld4<T>(ee + off, e4);
        for (int c = 0; c < 4; c++) p[c] -= c4[c] + e4[c];
      }
      __syncthreads();
      st4<T>(sh[g][0] + 4 * t, p);
      __syncthreads();
      if (ok) acc_mul<T>(acc[SPLIT ? 0 : s], sh[g][2], sh[g][0], t);
      __syncthreads();
    }
  }
  if constexpr (SPLIT) emit_one<T>(acc[0], sblk, 3, partials, t, g, blk, sums);
  else block_emit<T, 3>(acc, partials, t, g, blk, sums);
}
template <class T, bool SPLIT>
__global__ void __launch_bounds__(GB) k_mult_round(const LT<T>* eg, const LT<T>* a, const LT<T>* b, const LT<T>* cc,
                                                   const LT<T>* ee, uint64_t m, LT<T>* partials) {
  mult_round_body<T, SPLIT>(eg, a, b, cc, ee, m, partials, nullptr, blockIdx.x, gridDim.x);
}

// The element products of a Gruen round, for the matrix-core form of its
// sums (sum_i eg_i P_s[i] as Gram contractions):
//   P[0][i] = a_lo b_lo - c_lo - e_lo, P[1][i] = a_hi b_hi - c_hi - e_hi,
//   P[2][i] = (a_hi - a_lo)(b_hi - b_lo)    (c, e may be null).
template <class T>
__global__ void __launch_bounds__(GB) k_mult_prod(const LT<T>* a, const LT<T>* b, const LT<T>* cc, const LT<T>* ee,
                                                  uint64_t m, LT<T>* P) {
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][2][E];
  const int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < 3 * m; base += (uint64_t)gridDim.x * GPB) {
    const uint64_t x = base + g;
    const bool ok = x < 3 * m;
    const int s = ok ? (int)(x / m) : 0;
    const uint64_t i = ok ? x % m : 0;
    if (ok) {
      CT<T> lo[4], hi[4], xv[4], yv[4];
      if (s != 1) {
        ld4<T>(a + i * E + 4 * t, lo);
        ld4<T>(b + i * E + 4 * t, yv);
      }
      if (s != 0) {
        ld4<T>(a + (i + m) * E + 4 * t, hi);
        ld4<T>(b + (i + m) * E + 4 * t, xv);
      }
      CT<T> f[4], h[4];
      for (int c = 0; c < 4; c++) {
        f[c] = s == 0 ? lo[c] : s == 1 ? hi[c] : hi[c] - lo[c];
        h[c] = s == 0 ? yv[c] : s == 1 ? xv[c] : xv[c] - yv[c];
      }
      st4<T>(sh[g][0] + 4 * t, f);
      st4<T>(sh[g][1] + 4 * t, h);
    }
    __syncthreads();
    Acc<T> pa;
    pa.clear();
    if (ok) acc_mul<T>(pa, sh[g][0], sh[g][1], t);
    CT<T> p[4];
    acc_final<T>(pa, t, p);
    if (ok) {
      if (s < 2 && cc) {
        CT<T> c4[4], e4[4];
        const uint64_t off = (s == 0 ? i : i + m) * E + 4 * t;
        ld4<T>(cc + off, c4);
        ld4<T>(ee + off, e4);
        for (int c = 0; c < 4; c++) p[c] -= c4[c] + e4[c];
      }
      st4<T>(P + ((uint64_t)s * m + i) * E + 4 * t, p);
    }
    __syncthreads();
  }
}

// Linear-check round over `ntab` table pairs (linear_check_simd.rs:264-286,
// linear_check_general.rs:289-334): s0 += f_lo g_lo, sinf += df dg.
// f, g: ntab tables of stride `stride` elements, current length 2m.
template <class T, int NAF>
__global__ void __launch_bounds__(GB) k_lin_round(const LT<T>* f, uint64_t fstride, const LT<T>* gt, uint64_t gstride,
                                                  uint64_t ntab, uint64_t m, LT<T>* partials) {
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][2][E];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  Acc<T> acc[2];
  acc[0].clear();
  acc[1].clear();
  uint64_t total = m * ntab;
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < total; base += (uint64_t)gridDim.x * GPB) {
    uint64_t w = base + g;
    bool ok = w < total;
    uint64_t tab = ok ? w / m : 0, i = ok ? w % m : 0;
    const LT<T>* F = f + tab * fstride * E;
    const LT<T>* G = gt + tab * gstride * E;
    CT<T> flo[4] = {0, 0, 0, 0}, fhi[4] = {0, 0, 0, 0}, glo[4] = {0, 0, 0, 0}, ghi[4] = {0, 0, 0, 0};
    if (ok) {
      ld4<T>(F + i * E + 4 * t, flo);
      ld4<T>(F + (i + m) * E + 4 * t, fhi);
      ld4<T>(G + i * E + 4 * t, glo);
      ld4<T>(G + (i + m) * E + 4 * t, ghi);
    }
    st4<T>(sh[g][0] + 4 * t, flo);
    st4<T>(sh[g][1] + 4 * t, glo);
    __syncthreads();
    if (ok) acc_mul<T, NAF>(acc[0], sh[g][0], sh[g][1], t);
    __syncthreads();
    for (int c = 0; c < 4; c++) { fhi[c] -= flo[c]; ghi[c] -= glo[c]; }
    st4<T>(sh[g][0] + 4 * t, fhi);
    st4<T>(sh[g][1] + 4 * t, ghi);
    __syncthreads();
    if (ok) acc_mul<T, NAF>(acc[1], sh[g][0], sh[g][1], t);
    __syncthreads();
  }
  block_emit<T, 2>(acc, partials, t, g, blockIdx.x);
}

// partial sums of sum_i A_y[i]*B[i] for tables y = blockIdx.y (A_y = A + y*astride
// elements; A low-degree if NA = 16); partials laid out [block][y].
template <class T, int NA>
__global__ void __launch_bounds__(GB) k_dot(const LT<T>* A, uint64_t astride, const LT<T>* B, uint64_t n, LT<T>* partials) {
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][2][E];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  const LT<T>* Ay = A + (uint64_t)blockIdx.y * astride * E;
  Acc<T> acc[1];
  acc[0].clear();
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < n; base += (uint64_t)gridDim.x * GPB) {
    uint64_t i = base + g;
    bool ok = i < n;
    if (ok) {
      stage<T>(sh[g][0], Ay + i * E, t);
      stage<T>(sh[g][1], B + i * E, t);
    }
    __syncthreads();
    if (ok) acc_mul<T, NA>(acc[0], sh[g][0], sh[g][1], t);
    __syncthreads();
  }
  __shared__ CT<T> red[GPB][14][32];
  for (int c = 0; c < 7; c++) { red[g][c][t] = acc[0].c[c]; red[g][7 + c][t] = acc[0].h[c]; }
  __syncthreads();
  if (g == 0) {
    Acc<T> a = acc[0];
    for (int gg = 1; gg < GPB; gg++)
      for (int c = 0; c < 7; c++) { a.c[c] += red[gg][c][t]; a.h[c] += red[gg][7 + c][t]; }
    CT<T> o[4];
    acc_final<T>(a, t, o);
    st4<T>(partials + ((uint64_t)blockIdx.x * gridDim.y + blockIdx.y) * E + 4 * t, o);
  }
}

// out[y] = src[y * stride]  (element copies, thread per limb)
template <class T>
__global__ void k_pick(const LT<T>* src, uint64_t stride, uint64_t cnt, LT<T>* out) {
  uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= cnt * E) return;
  uint64_t y = id / E, l = id % E;
  out[id] = src[y * stride * E + l];
}

// out[s] = sum_b partials[b][s]: block (s, 32-limb slice) of 1024 threads,
// 32 slices of the partial list per limb, then an LDS tree.
template <class T>
__global__ void __launch_bounds__(1024) k_sum_partials(const LT<T>* partials, uint64_t nblocks, int ns, LT<T>* out) {
  __shared__ CT<T> red[32][33];
  int s = blockIdx.x / 4, l0 = (blockIdx.x % 4) * 32;
  int l = threadIdx.x & 31, sl = threadIdx.x >> 5;
  // eight independent loads in flight per thread
  CT<T> a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t b = sl;
  for (; b + 7 * 32 < nblocks; b += 8 * 32)
#pragma unroll
    for (int u = 0; u < 8; u++) a[u] += partials[((b + 32 * u) * ns + s) * E + l0 + l];
  for (; b < nblocks; b += 32) a[0] += partials[(b * ns + s) * E + l0 + l];
  CT<T> acc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  red[sl][l] = acc;
  __syncthreads();
  if (sl == 0) {
    for (int k = 1; k < 32; k++) acc += red[k][l];
    out[s * E + l0 + l] = (LT<T>)acc;
  }
}

// ============================================================================
// Host-device mailbox (coherent pinned memory, no caching on the device)
// ============================================================================

// A kernel that needs a value the host has not posted yet waits on a flag
// word.  Every wait ends: on the posted value, on the abort word (abort[0]),
// or after `ticks` of the constant-rate wall clock, which it reports in
// abort[1] so the host fails the proof.
struct MbWait {
  const uint32_t* flag;  // null: the value is already in place
  uint32_t val;
  uint32_t* abort;
  uint64_t ticks;
};

__device__ __forceinline__ void mb_wait(const MbWait& w) {
  if (threadIdx.x == 0) {
    const uint64_t t0 = (uint64_t)wall_clock64();
    while (__hip_atomic_load(w.flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) != w.val) {
      if (__hip_atomic_load(w.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) break;
      if ((uint64_t)wall_clock64() - t0 > w.ticks) {
        __hip_atomic_store(w.abort + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

// After mb_wait: the `words` posted 32-bit words into LDS through
// system-scope loads.
__device__ __forceinline__ void mb_fetch(const uint32_t* src, int words, uint32_t* dst) {
  for (int i = threadIdx.x; i < words; i += blockDim.x)
    dst[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __syncthreads();
}

// Sumcheck round values posted to the host.  The partial sums are reduced as
// in k_sum_partials: one block per (sum, 32 limbs) piece into `sums`
// (device), the last block to finish taking over (cross-block handoff
// through `count`, which the last block resets), or with a single block
// (gridDim.x == 1) all pieces in turn, with no handoff.  The finishing
// block forms the round polynomial q, writes it to the mailbox slot `qh`
// (and to `qd` on the device when given), then raises `flag` to `val`.
//   mode 0 (mult check, ns = 3):   q = (s0, s1 - s0 - s2, s2)
//                                   (mult_check.rs:250-296)
//   mode 1 (linear check, ns = 2): q = (s0, c - 2 s0 - s1, s1)
//                                   (linear_check_general.rs:289-334)
// The round polynomial q from the round's sums ssum (LDS) by the whole
// block (modes as below), written to qh (and qd); then the flag rises
// (none: the round's Fiat-Shamir runs on the device).
template <class T>
__device__ __forceinline__ void post_q(const CT<T> (*ssum)[E], int mode, const LT<T>* c, LT<T>* qd, LT<T>* qh,
                                       uint32_t* flag, uint32_t val) {
  using C = CT<T>;
  for (int i = threadIdx.x; i < 3 * E; i += blockDim.x) {
    const int k = i / E, li = i % E;
    C v;
    if (mode == 0) v = k == 1 ? (C)(ssum[1][li] - ssum[0][li] - ssum[2][li]) : ssum[k][li];
    else v = k == 0 ? ssum[0][li] : k == 1 ? (C)((C)c[li] - ssum[0][li] - ssum[0][li] - ssum[1][li]) : ssum[1][li];
    if (qd) qd[i] = (LT<T>)v;
    qh[i] = (LT<T>)v;
  }
  // each writer's stores have completed before one system-scope release
  // orders them ahead of the flag
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0 && flag) {
    __threadfence_system();
    __hip_atomic_store(flag, val, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

template <class T>
__global__ void __launch_bounds__(1024) k_sum_post(const LT<T>* partials, uint64_t nblocks, int ns, int mode,
                                                  const LT<T>* c, LT<T>* sums, LT<T>* qd, LT<T>* qh, uint32_t* count,
                                                  uint32_t* flag, uint32_t val) {
  using C = CT<T>;
  __shared__ C red[32][33];
  __shared__ C ssum[3][E];
  __shared__ bool last;
  const bool single = gridDim.x == 1;
  const int l = threadIdx.x & 31, sl = threadIdx.x >> 5;
  if (single) {
    // every (sum, limb) column at once: S slices of the blocks per column
    const int cols = ns * E, S = 1024 / cols, o = threadIdx.x % cols, sv = threadIdx.x / cols;
    C a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (sv < S) {
      const int s = o / E, li = o % E;
      uint64_t bb = sv;
      for (; bb + 7 * (uint64_t)S < nblocks; bb += 8 * (uint64_t)S)
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] += partials[((bb + (uint64_t)S * u) * ns + s) * E + li];
      for (; bb < nblocks; bb += S) a[0] += partials[(bb * ns + s) * E + li];
    }
    C acc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    C* part = &red[0][0];  // 33 * 32 >= 1024 slots
    part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < cols) {
      for (int k = 1; k < S; k++) acc += part[threadIdx.x + k * cols];
      ssum[threadIdx.x / E][threadIdx.x % E] = acc;
    }
    __syncthreads();
  }
  for (int piece = (int)blockIdx.x; !single && piece < (int)blockIdx.x + 1; piece++) {
    const int s = piece / 4, l0 = (piece % 4) * 32;
    C a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t b = sl;
    for (; b + 7 * 32 < nblocks; b += 8 * 32)
#pragma unroll
      for (int u = 0; u < 8; u++) a[u] += partials[((b + 32 * u) * ns + s) * E + l0 + l];
    for (; b < nblocks; b += 32) a[0] += partials[(b * ns + s) * E + l0 + l];
    C acc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    red[sl][l] = acc;
    __syncthreads();
    if (sl == 0) {
      for (int k = 1; k < 32; k++) acc += red[k][l];
      sums[s * E + l0 + l] = (LT<T>)acc;
      __threadfence();  // the writing wave publishes its piece
    }
    __syncthreads();
  }
  if (!single) {
    if (threadIdx.x == 0) last = atomicAdd(count, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    if (threadIdx.x == 0) __threadfence();  // acquire the other blocks' pieces (same CU: one wave suffices)
    __syncthreads();
    for (int i = threadIdx.x; i < ns * E; i += blockDim.x) ssum[i / E][i % E] = (C)sums[i];
    __syncthreads();
  }
  if (!single && threadIdx.x == 0) *count = 0;
  post_q<T>(ssum, mode, c, qd, qh, flag, val);
}

// ============================================================================
// Small sumcheck rounds in one kernel: the fold of round r, the eq table's
// marginalization and round r+1's sums (then posted by the last block), so a
// latency-bound round costs this kernel and the challenge's lift.  A block
// takes pairs i of round r+1 (half mh = m/2) in turn; its 32-lane groups
// each run one group product of a pair at a time, so a pair costs a few
// product latencies, not the sum of its products: the folds of positions
// i and i+mh of every table (from positions i, i+m, i+mh, i+mh+m; each
// position is read and written by this block alone, so in place is safe)
// run side by side, then the products of the round's sums.  Block sums go
// to `sums` with integer atomics (mod 2^32 or 2^64); the block that
// finishes last (count) forms q_{r+1}, clears sums and count, and posts.
// ============================================================================

// out = c (hi - lo) + lo for one group (lane t: coefficients 4t..4t+3); sc
// the challenge, shd the group's scratch element (LDS); the caller syncs
// the block between calls that share shd.
template <class T>
__device__ __forceinline__ void fold_one(const LT<T>* sc, LT<T>* shd, const LT<T>* src, uint64_t lo_at, uint64_t hi_at,
                                         int t, CT<T> out[4]) {
  using C = CT<T>;
  C lo[4], hi[4];
  ld4<T>(src + lo_at * E + 4 * t, lo);
  ld4<T>(src + hi_at * E + 4 * t, hi);
  C d[4] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], hi[3] - lo[3]};
  st4<T>(shd + 4 * t, d);
  __builtin_amdgcn_wave_barrier();  // the group's 32 lanes sit in one wave
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  Acc<T> pa;
  pa.clear();
  acc_mul<T>(pa, sc, shd, t);
  acc_final<T>(pa, t, out);
  for (int c = 0; c < 4; c++) out[c] += lo[c];
}

// ns sums from their SUM_COPIES atomic copies (each 3 elements apart) into
// ssum (LDS), the copies cleared for the next launch; the whole block calls
// it, and syncs before ssum is read.
template <class T> __device__ __forceinline__ void gather_sums(CT<T>* sums, int ns, CT<T> (*ssum)[E]) {
  using C = CT<T>;
  for (int k = threadIdx.x; k < ns * E; k += blockDim.x) {
    // all copies' loads in flight at once (one L2 latency), then the sum
    C v[SUM_COPIES];
#pragma unroll
    for (int cp = 0; cp < SUM_COPIES; cp++)
      v[cp] = __hip_atomic_load(sums + (size_t)cp * 3 * E + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    C acc = 0;
#pragma unroll
    for (int cp = 0; cp < SUM_COPIES; cp++) acc += v[cp];
    ssum[k / E][k % E] = acc;
  }
  __syncthreads();  // every copy read before any is cleared
  for (int k = threadIdx.x; k < SUM_COPIES * 3 * E; k += blockDim.x) sums[k] = 0;
}

// Last block of a tail launch: true for the block that sees every other
// block's atomic sums in.
__device__ __forceinline__ bool tail_last(uint32_t* count, bool* last) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) *last = atomicAdd(count, 1u) == gridDim.x - 1;
  __syncthreads();
  return *last;
}

template <class T> struct TailM {
  const LT<T>* src[4];  // the round's tables (2m entries each)
  LT<T>* dst[4];        // the folded tables (m entries; dst may be src)
  uint64_t m;
  const LT<T>* rc;  // the round's challenge
  const LT<T>* rt;  // rt[r]: the marginalization scale is eq_one(rt[r], c)
  LT<T>* eg;        // eq table (m entries in, m/2 out)
  CT<T>* sums;      // SUM_COPIES x 3 elements, zero on entry
  uint32_t* count;  // zero on entry
  LT<T>* qh;        // q_{r+1}
  uint32_t* flag;   // mailbox flag (null: device transcript)
  uint32_t val;
};
// Gruen rounds (as k_fold + k_marg + k_mult_round + k_sum_post), NT = 2
// tables (a, b) or 4 (a, b, c, e); 2 NT + 1 groups: group 2tb+h folds half
// h of table tb (group 2NT forms the scale first), then group 2NT the new
// eq entry and groups 0..2 the products P_s, then groups 0..2 add eg' P_s.
template <class T, int NT> __global__ void __launch_bounds__(32 * (2 * NT + 1)) k_tail_mult(TailM<T> A) {
  using C = CT<T>;
  using L = LT<T>;
  constexpr int NG = 2 * NT + 1;
  __shared__ __attribute__((aligned(16))) L sc[E], srt[E], sscale[E], seg[E];
  __shared__ __attribute__((aligned(16))) L sf[2 * NT][E];  // folded: [2tb] entry i, [2tb+1] entry i+mh
  __shared__ __attribute__((aligned(16))) L sp[3][E];       // P_s
  __shared__ __attribute__((aligned(16))) L sx[NG][E];      // per-group scratch
  __shared__ C ssum[3][E];
  __shared__ bool last;
  const int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  const uint64_t m = A.m, mh = m / 2;
  if (g == 0) {
    stage<T>(sc, A.rc, t);
    stage<T>(srt, A.rt, t);
  }
  __syncthreads();
  if (g == 2 * NT) {  // eq_one(rt[r], c) = 2 rt c - rt - c + 1
    Acc<T> pa;
    pa.clear();
    acc_mul<T>(pa, srt, sc, t);
    C o[4], a[4], b[4];
    acc_final<T>(pa, t, o);
    ld4<T>(srt + 4 * t, a);
    ld4<T>(sc + 4 * t, b);
    for (int c = 0; c < 4; c++) o[c] = o[c] + o[c] - a[c] - b[c];
    if (t == 0) o[0] += 1;
    st4<T>(sscale + 4 * t, o);
  }
  Acc<T> acc;  // groups 0..2: sum s = g
  acc.clear();
  for (uint64_t i = blockIdx.x; i < mh; i += gridDim.x) {
    if (g < 2 * NT) {  // stage 1: the folds
      const int tb = g >> 1, h = g & 1;
      C x[4];
      fold_one<T>(sc, sx[g], A.src[tb], i + h * mh, i + h * mh + m, t, x);
      st4<T>(sf[g] + 4 * t, x);
      st4<T>(A.dst[tb] + (i + h * mh) * E + 4 * t, x);
    }
    __syncthreads();
    // stage 2 (groups in different waves where they can be, so their
    // products run side by side): the last group eg'[i] = scale (eg[i] +
    // eg[i + mh]), groups 0..2 P_s (s = g) from the folded a, b (c, e)
    if (g == 2 * NT) {
      C e0[4], e1[4];
      ld4<T>(A.eg + i * E + 4 * t, e0);
      ld4<T>(A.eg + (i + mh) * E + 4 * t, e1);
      for (int c = 0; c < 4; c++) e0[c] += e1[c];
      st4<T>(sx[g] + 4 * t, e0);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      Acc<T> pa;
      pa.clear();
      acc_mul<T>(pa, sscale, sx[g], t);
      C eo[4];
      acc_final<T>(pa, t, eo);
      st4<T>(seg + 4 * t, eo);
      st4<T>(A.eg + i * E + 4 * t, eo);
    } else if (g < 3) {
      const int s = g;
      C a0[4], a1[4], b0[4], b1[4], x[4], y[4];
      ld4<T>(sf[0] + 4 * t, a0);
      ld4<T>(sf[1] + 4 * t, a1);
      ld4<T>(sf[2] + 4 * t, b0);
      ld4<T>(sf[3] + 4 * t, b1);
      for (int c = 0; c < 4; c++) {
        x[c] = s == 0 ? a0[c] : s == 1 ? a1[c] : a1[c] - a0[c];
        y[c] = s == 0 ? b0[c] : s == 1 ? b1[c] : b1[c] - b0[c];
      }
      st4<T>(sx[g] + 4 * t, x);
      st4<T>(sp[s] + 4 * t, y);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      Acc<T> pa;
      pa.clear();
      acc_mul<T>(pa, sx[g], sp[s], t);
      C p[4];
      acc_final<T>(pa, t, p);
      if constexpr (NT == 4) {
        if (s < 2) {
          C c4[4], e4[4];
          ld4<T>(sf[4 + s] + 4 * t, c4);
          ld4<T>(sf[6 + s] + 4 * t, e4);
          for (int c = 0; c < 4; c++) p[c] -= c4[c] + e4[c];
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      st4<T>(sp[s] + 4 * t, p);
    }
    __syncthreads();
    if (g < 3) acc_mul<T>(acc, seg, sp[g], t);  // stage 3
    __syncthreads();
  }
  if (g < 3) {
    C o[4];
    acc_final<T>(acc, t, o);
    C* dst = A.sums + (size_t)(blockIdx.x % SUM_COPIES) * 3 * E + g * E + 4 * t;
    for (int c = 0; c < 4; c++) atom_add(dst + c, o[c]);
  }
  if (!tail_last(A.count, &last)) return;
  __threadfence();
  gather_sums<T>(A.sums, 3, ssum);
  if (threadIdx.x == 0) *A.count = 0;
  __syncthreads();
  post_q<T>(ssum, 0, nullptr, nullptr, A.qh, A.flag, A.val);
}

template <class T> struct TailL {
  const LT<T>* f;  // the round's tables: f + t*fs, g + t*gs (2m entries each)
  const LT<T>* gt;
  uint64_t fs, gs;
  LT<T>* fw;  // folded tables at stride H (may be the sources)
  LT<T>* gw;
  uint64_t H, ntab, m;
  const LT<T>* rc;  // the round's challenge
  LT<T>* dq;        // q_r in, q_{r+1} out
  LT<T>* dc;        // the claim lin_claim(q_r, c) out
  CT<T>* sums;      // SUM_COPIES x 3 elements (2 used), zero on entry
  uint32_t* count;  // zero on entry
  LT<T>* qh;        // q_{r+1}
  uint32_t* flag;
  uint32_t val;
};
// {0,1,inf} linear rounds (as fold2 + k_lin_claim + k_gram_lin +
// k_sum_post): a block takes (table, pair) items in turn; groups 0..3 fold
// f at i, i+mh and g at i, i+mh, then groups 0, 2 add f'g' and df' dg'.
template <class T> __global__ void __launch_bounds__(128) k_tail_lin(TailL<T> A) {
  using C = CT<T>;
  using L = LT<T>;
  __shared__ __attribute__((aligned(16))) L sc[E], sq[E];
  __shared__ __attribute__((aligned(16))) L sf[4][E];  // f'_i, f'_{i+mh}, g'_i, g'_{i+mh}
  __shared__ __attribute__((aligned(16))) L sx[4][E];
  __shared__ C ssum[3][E];
  __shared__ bool last;
  const int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  const uint64_t m = A.m, mh = m / 2, total = A.ntab * mh;
  if (g == 0) stage<T>(sc, A.rc, t);
  __syncthreads();
  Acc<T> acc;  // groups 0, 2: sum s = g / 2
  acc.clear();
  for (uint64_t w = blockIdx.x; w < total; w += gridDim.x) {
    const uint64_t tab = w / mh, i = w % mh;
    {  // stage 1: group g folds table (g < 2 ? f : g), entry i + (g & 1) mh
      const int h = g & 1;
      const L* src = g < 2 ? A.f + tab * A.fs * E : A.gt + tab * A.gs * E;
      L* dst = g < 2 ? A.fw + tab * A.H * E : A.gw + tab * A.H * E;
      C x[4];
      fold_one<T>(sc, sx[g], src, i + h * mh, i + h * mh + m, t, x);
      st4<T>(sf[g] + 4 * t, x);
      st4<T>(dst + (i + h * mh) * E + 4 * t, x);
    }
    __syncthreads();
    if (g == 0 || g == 2) {  // stage 2 (groups 0 and 2 sit in different waves)
      if (g == 0) {
        acc_mul<T>(acc, sf[0], sf[2], t);
      } else {
        C f0[4], f1[4], g0[4], g1[4], x[4], y[4];
        ld4<T>(sf[0] + 4 * t, f0);
        ld4<T>(sf[1] + 4 * t, f1);
        ld4<T>(sf[2] + 4 * t, g0);
        ld4<T>(sf[3] + 4 * t, g1);
        for (int c = 0; c < 4; c++) {
          x[c] = f1[c] - f0[c];
          y[c] = g1[c] - g0[c];
        }
        st4<T>(sx[2] + 4 * t, x);
        st4<T>(sx[3] + 4 * t, y);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        acc_mul<T>(acc, sx[2], sx[3], t);
      }
    }
    __syncthreads();
  }
  if (g == 0 || g == 2) {
    C o[4];
    acc_final<T>(acc, t, o);
    C* dst = A.sums + (size_t)(blockIdx.x % SUM_COPIES) * 3 * E + (g / 2) * E + 4 * t;
    for (int c = 0; c < 4; c++) atom_add(dst + c, o[c]);
  }
  if (!tail_last(A.count, &last)) return;
  __threadfence();
  gather_sums<T>(A.sums, 2, ssum);
  if (threadIdx.x == 0) *A.count = 0;
  // the next claim c' = q0 + c (q1 + c q2) from q_r, kept in sq and written to dc
  if (g == 0) stage<T>(sq, A.dq + 2 * E, t);
  for (int step = 1; step >= 0; step--) {
    __syncthreads();
    C o[4], a[4];
    if (g == 0) {
      Acc<T> pa;
      pa.clear();
      acc_mul<T>(pa, sc, sq, t);
      acc_final<T>(pa, t, o);
      ld4<T>(A.dq + step * E + 4 * t, a);
      for (int k = 0; k < 4; k++) o[k] += a[k];
    }
    __syncthreads();
    if (g == 0) {
      st4<T>(sq + 4 * t, o);
      if (!step) st4<T>(A.dc + 4 * t, o);
    }
  }
  __syncthreads();
  post_q<T>(ssum, 1, sq, A.dq, A.qh, A.flag, A.val);
}

// ============================================================================
// Sumcheck scalar updates kept on the device (one 32-lane group)
// ============================================================================

// out = t r + (1 - t)(1 - r) = 2 t r - t - r + 1   (eq in one variable)
// The two scalar updates below run on threads 0..63 of any block of at least
// 64 threads (the others only meet the barriers): as kernels of their own,
// or as the side task of a fold launch (one extra block).
template <class T> __device__ __forceinline__ void eq_one_block(const LT<T>* t, const LT<T>* r, LT<T>* out) {
  __shared__ __attribute__((aligned(16))) LT<T> st[E], sr[E];
  const int g = threadIdx.x >> 5, l = threadIdx.x & 31;
  if (g == 0) {
    stage<T>(st, t, l);
    stage<T>(sr, r, l);
  }
  __syncthreads();
  if (g == 0) {
    Acc<T> acc;
    acc.clear();
    acc_mul<T>(acc, st, sr, l);
    CT<T> o[4], a[4], b[4];
    acc_final<T>(acc, l, o);
    ld4<T>(st + 4 * l, a);
    ld4<T>(sr + 4 * l, b);
    for (int c = 0; c < 4; c++) o[c] = o[c] + o[c] - a[c] - b[c];
    if (l == 0) o[0] += 1;
    st4<T>(out + 4 * l, o);
  }
}
template <class T> __global__ void __launch_bounds__(64) k_eq_one(const LT<T>* t, const LT<T>* r, LT<T>* out) {
  eq_one_block<T>(t, r, out);
}

// Next claim c = q0 + r (q1 + r q2).
template <class T> __device__ __forceinline__ void lin_claim_block(const LT<T>* q, const LT<T>* r, LT<T>* c) {
  __shared__ __attribute__((aligned(16))) LT<T> sr[E], sx[E];
  const int g = threadIdx.x >> 5, l = threadIdx.x & 31;
  if (g == 0) {
    stage<T>(sr, r, l);
    stage<T>(sx, q + 2 * E, l);
  }
  for (int step = 1; step >= 0; step--) {
    __syncthreads();
    CT<T> o[4], a[4];
    if (g == 0) {
      Acc<T> acc;
      acc.clear();
      acc_mul<T>(acc, sr, sx, l);
      acc_final<T>(acc, l, o);
      ld4<T>(q + step * E + 4 * l, a);
      for (int k = 0; k < 4; k++) o[k] += a[k];
    }
    __syncthreads();
    if (g == 0) st4<T>(step ? sx + 4 * l : c + 4 * l, o);
  }
}
template <class T> __global__ void __launch_bounds__(64) k_lin_claim(const LT<T>* q, const LT<T>* r, LT<T>* c) {
  lin_claim_block<T>(q, r, c);
}

// One product A * B (elements in LDS) by a 256-thread block: group p sums
// the pair terms 4p..4p+3 of every coefficient, the eight raw sums meet in
// LDS, lanes 0..31 reduce; lane t < 32 returns coefficient t in out.  All
// threads call it.
template <class T> __device__ __forceinline__ void block_mul256(const LT<T>* A, const LT<T>* B, CT<T> out[4]) {
  using C = CT<T>;
  __shared__ C red[8][14][32];
  const int t = threadIdx.x & 31, p = threadIdx.x >> 5;
  Acc<T> acc;
  acc.clear();
  acc_mul_range<T>(acc, A, B, t, 4 * p, 4 * p + 4);
#pragma unroll
  for (int c = 0; c < 7; c++) {
    red[p][c][t] = acc.c[c];
    red[p][7 + c][t] = acc.h[c];
  }
  __syncthreads();
  if (p == 0) {
#pragma unroll
    for (int q = 1; q < 8; q++)
#pragma unroll
      for (int c = 0; c < 7; c++) {
        acc.c[c] += red[q][c][t];
        acc.h[c] += red[q][7 + c][t];
      }
    acc_final<T>(acc, t, out);
  }
  __syncthreads();
}

// eq_one_block / lin_claim_block for 256-thread blocks (fold side tasks).
template <class T> __device__ __forceinline__ void eq_one_256(const LT<T>* t, const LT<T>* r, LT<T>* out) {
  __shared__ __attribute__((aligned(16))) LT<T> st[E], sr[E];
  const int l = threadIdx.x;
  if (l < 32) {
    stage<T>(st, t, l);
    stage<T>(sr, r, l);
  }
  __syncthreads();
  CT<T> o[4], a[4], b[4];
  block_mul256<T>(st, sr, o);
  if (l < 32) {
    ld4<T>(st + 4 * l, a);
    ld4<T>(sr + 4 * l, b);
    for (int c = 0; c < 4; c++) o[c] = o[c] + o[c] - a[c] - b[c];
    if (l == 0) o[0] += 1;
    st4<T>(out + 4 * l, o);
  }
}
template <class T> __device__ __forceinline__ void lin_claim_256(const LT<T>* q, const LT<T>* r, LT<T>* c) {
  __shared__ __attribute__((aligned(16))) LT<T> sr[E], sx[E];
  const int l = threadIdx.x;
  if (l < 32) {
    stage<T>(sr, r, l);
    stage<T>(sx, q + 2 * E, l);
  }
  __syncthreads();
  for (int step = 1; step >= 0; step--) {
    CT<T> o[4], a[4];
    block_mul256<T>(sr, sx, o);
    if (l < 32) {
      ld4<T>(q + step * E + 4 * l, a);
      for (int k = 0; k < 4; k++) o[k] += a[k];
      st4<T>(step ? sx + 4 * l : c + 4 * l, o);
    }
    __syncthreads();
  }
}

// ============================================================================
// Teichmuller lift of Fiat-Shamir challenges
// ============================================================================

// Element of C-typed LDS as 4 values.
template <class C> __device__ __forceinline__ void ldc4(const C* p, C o[4]) {
  if constexpr (sizeof(C) == 4) {
    uint4 v = *reinterpret_cast<const uint4*>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  } else {
    const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
    ulonglong2 a = q[0], b = q[1];
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
  }
}

// Cross-lane move inside a row of 16 lanes (DPP; CTRL is the dpp_ctrl code).
template <int CTRL, class C> __device__ __forceinline__ C dpp(C v) {
  if constexpr (sizeof(C) == 4) {
    return (C)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
  } else {
    int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
    return ((C)(uint32_t)hi << 32) | (C)(uint32_t)lo;
  }
}
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141; // lane i <-> 7-i within 8 lanes

// Teichmuller lift of `x[blockIdx.x]` in place.  The reference
// (teichmuller.rs:104-114) squares (ceil(log2 k)+1)*128 times; its result is
// the unique Teichmuller element congruent to x mod 2 (x^(2^128) already is
// that element, as 128 >= k-1).  With sigma the Frobenius automorphism of
// GR(2^k,128), the same element is (sigma^-(k-1) x)^(2^(k-1)): if
// z = sigma^-n(x) = t(1+2w) with t Teichmuller, z^(2^n) = t^(2^n) = Teich(x)
// once n+1 >= k.  sigma^-(k-1) is Z/2^k-linear; `frob_t` holds its 128x128
// matrix transposed ([j][o], computed on the host), so the lift is one matvec
// and `squarings` = k-1 squarings.
//
// Squaring with 256 threads: wave w owns result coefficients t = 8w..8w+7.
// The unordered pairs {i, j} with i + j = t mod 32 split into n0 = t/2 + 1
// unwrapped ones (i + j = t: i = q, j = t - q) and n1 = (32 - t)/2 wrapped
// ones (i + j = t + 32: i = t + 1 + q, j = 31 - q); lane (t, dl) takes three
// consecutive pairs of one kind (lanes dl < ceil(n0/3) the unwrapped ones),
// so it keeps a single accumulator and its kind is fixed.  Off-diagonal pairs
// count twice: they read the left factor from sx2 = 2 x.  The 8 lanes of a t
// sum their circ (all) and hi (wrapped) terms with a DPP transpose-reduction;
// the reduction Y^32 = -Y^3 - Y - alpha is applied by 128 threads reading the
// sums from LDS.
//
// square_times: `count` squarings of the element in sx (sx2 = 2 sx beside
// it), 256 threads, C the arithmetic width (the ring's, or a narrower one
// where only the low bits are needed).  Ends with sx, sx2 updated.  M24:
// the products on the 24-bit multiplier (C 32-bit; every product and sum is
// then exact mod 2^24 only, the limbs' upper bits are left undefined).
template <bool M24, class C> __device__ __forceinline__ void d_raw_sq(const C a[4], const C b[4], C p[7]) {
  if constexpr (M24) {
    static_assert(sizeof(C) == 4, "24-bit products on 32-bit limbs");
    auto m = [](C x, C y) -> C { return (C)__umul24((unsigned)x, (unsigned)y); };
    p[0] = m(a[0], b[0]);
    p[1] = m(a[0], b[1]) + m(a[1], b[0]);
    p[2] = m(a[0], b[2]) + m(a[1], b[1]) + m(a[2], b[0]);
    p[3] = m(a[0], b[3]) + m(a[1], b[2]) + m(a[2], b[1]) + m(a[3], b[0]);
    p[4] = m(a[1], b[3]) + m(a[2], b[2]) + m(a[3], b[1]);
    p[5] = m(a[2], b[3]) + m(a[3], b[2]);
    p[6] = m(a[3], b[3]);
  } else {
    d_raw(a, b, p);
  }
}
template <class C, bool M24 = false>
__device__ __forceinline__ void square_times(C* sx, C* sx2, C (*sred)[32][4], int count) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, dl = lane & 7;
  const int t = 8 * (tid >> 6) + (lane >> 3);
  const bool b2 = dl & 4, b1 = dl & 2, b0 = dl & 1;
  const int n0 = (t >> 1) + 1, n1 = (32 - t) >> 1, a0 = (n0 + 2) / 3;
  const bool wrap = dl >= a0;
  const int q0 = 3 * (wrap ? dl - a0 : dl), nq = wrap ? n1 : n0;
  const C hm = wrap ? (C)~(C)0 : (C)0;
  for (int it = 0; it < count; it++) {
    C ac[7];
#pragma unroll
    for (int c = 0; c < 7; c++) ac[c] = 0;
    // the three pairs' factors are read at once (one LDS latency): unused
    // pairs read clamped indices and contribute zero
    C a[3][4], b[3][4];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int q = q0 + k;
      int i = wrap ? t + 1 + q : q, j = wrap ? 31 - q : t - q;
      i = i < 0 ? 0 : (i > 31 ? 31 : i);
      j = j < 0 ? 0 : (j > 31 ? 31 : j);
      ldc4((i != j ? sx2 : sx) + 4 * i, a[k]);
      ldc4(sx + 4 * j, b[k]);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const C m = q0 + k < nq ? (C)~(C)0 : (C)0;
      C bm[4], p[7];
#pragma unroll
      for (int c = 0; c < 4; c++) bm[c] = b[k][c] & m;
      d_raw_sq<M24>(a[k], bm, p);
#pragma unroll
      for (int c = 0; c < 7; c++) ac[c] += p[c];
    }
    C v[8];
    d_red(ac, v);
#pragma unroll
    for (int c = 0; c < 4; c++) v[4 + c] = v[c] & hm;
    // lane dl ends with the sum over the 8 lanes of v[dl]
    C k4[4], k2[2];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      C send = b2 ? v[c] : v[c + 4];
      k4[c] = (b2 ? v[c + 4] : v[c]) + dpp<DPP_HALF_MIRROR>(send);
    }
#pragma unroll
    for (int c = 0; c < 2; c++) {
      C send = b1 ? k4[c] : k4[c + 2];
      k2[c] = (b1 ? k4[c + 2] : k4[c]) + dpp<DPP_XOR2>(send);
    }
    C sum = (b0 ? k2[1] : k2[0]) + dpp<DPP_XOR1>(b0 ? k2[0] : k2[1]);
    sred[dl >> 2][t][dl & 3] = sum;
    __syncthreads();
    if (tid < E) {
      // branch-free: every term is read (clamped index) and masked
      const int u = tid >> 2, c = tid & 3;
      auto hc = [&](int s, int cc) -> C {  // hi of position 32+s after the 61/62 cascade
        const C w = s < 2 ? ~(C)0 : (C)0;
        return sred[1][s][cc] - (sred[1][s < 2 ? s + 29 : 0][cc] & w);
      };
      const C m1 = u >= 1 ? ~(C)0 : (C)0, m3 = u >= 3 ? ~(C)0 : (C)0;
      const C lo = sred[0][u][c] - sred[1][u][c] - (hc(u >= 1 ? u - 1 : 0, c) & m1) - (hc(u >= 3 ? u - 3 : 0, c) & m3);
      // -alpha * hi(u): c0 += h3, c1 += h3 - h0, c2 -= h1, c3 -= h2
      const C ma = c <= 1 ? ~(C)0 : (C)0, mb = c >= 1 ? ~(C)0 : (C)0;
      const C y = lo + (hc(u, 3) & ma) - (hc(u, (c + 3) & 3) & mb);
      sx[tid] = y;
      sx2[tid] = y + y;
    }
    __syncthreads();
  }
}

// Lifts src (device or host-pinned memory) into el (if not null) with a
// 256-thread block; the lifted value is left in sx (LDS, C-typed, 16-byte
// aligned) as well.
template <class T>
__device__ __forceinline__ void teich_lift_block(const LT<T>* src, LT<T>* el, const LT<T>* frob_t, int squarings,
                                                 CT<T>* sx) {
  using C = CT<T>;
  __shared__ C part[E];
  __shared__ __attribute__((aligned(16))) C sx2[E];
  __shared__ C sred[2][32][4];  // [0] reduced circ, [1] reduced hi (before the cascade)
  const int tid = threadIdx.x;
  if (tid < E) sx[tid] = (C)src[tid];
  __syncthreads();
  {
    const int o = tid & (E - 1), h = tid >> 7;
    // all 64 matrix entries in flight at once (one memory latency)
    C f[64];
#pragma unroll
    for (int j = 0; j < 64; j++) f[j] = (C)frob_t[(size_t)(64 * h + j) * E + o];
    C s = 0;
#pragma unroll
    for (int j = 0; j < 64; j++) s += f[j] * sx[64 * h + j];
    if (h) part[o] = s;
    __syncthreads();
    if (!h) {
      const C y = s + part[o];
      sx[o] = y;
      sx2[o] = y + y;
    }
    __syncthreads();
  }
  // z_i = y^(2^i) is needed exact mod 2^K only at the end, and u = v mod 2^s
  // gives u^2 = v^2 mod 2^(s+1): z_{n-j} needs only K - j bits.  So squarings
  // run on the narrowest arithmetic that keeps this: the first ones on the
  // 24-bit multiplier (exact mod 2^24) as long as the squarings after them
  // bring the precision to K (32 at the end of the 32-bit stage below).
  auto n24 = [](int n, int bits) { return n + 24 - bits < 0 ? 0 : (n + 24 - bits < n ? n + 24 - bits : n); };
  if constexpr (sizeof(C) == 8) {
    // 64-bit limbs: the first squarings run mod 2^32 (z_31 exact mod 2^32)
    // and the last 32, from z_31's low words, leave z_63 exact mod 2^64
    __shared__ __attribute__((aligned(16))) uint32_t lx[E], lx2[E];
    __shared__ uint32_t lred[2][32][4];
    const int lo = squarings > 32 ? squarings - 32 : 0;
    if (lo) {
      if (tid < E) {
        lx[tid] = (uint32_t)sx[tid];
        lx2[tid] = (uint32_t)sx2[tid];
      }
      __syncthreads();
      const int a = n24(lo, 32);
      square_times<uint32_t, true>(lx, lx2, lred, a);
      square_times<uint32_t>(lx, lx2, lred, lo - a);
      if (tid < E) {
        // sx2 = 2 sx in 64 bits (the 32-bit double dropped its carry)
        sx[tid] = (C)lx[tid];
        sx2[tid] = (C)lx[tid] + (C)lx[tid];
      }
      __syncthreads();
    }
    square_times<C>(sx, sx2, sred, squarings - lo);
  } else {
    const int a = n24(squarings, T::K);
    square_times<C, true>(sx, sx2, sred, a);
    square_times<C>(sx, sx2, sred, squarings - a);
  }
  if (el && tid < E) el[tid] = (LT<T>)sx[tid];
}

template <class T>
__global__ void __launch_bounds__(256) k_lift_from(const LT<T>* src, LT<T>* x, const LT<T>* frob_t, int squarings,
                                                   MbWait w) {
  __shared__ __attribute__((aligned(16))) CT<T> sx[E];
  __shared__ uint32_t posted[E * sizeof(LT<T>) / 4];
  if (w.flag) {
    mb_wait(w);
    mb_fetch(reinterpret_cast<const uint32_t*>(src), E * sizeof(LT<T>) / 4, posted);
    src = reinterpret_cast<const LT<T>*>(posted);
  }
  teich_lift_block<T>(src, x, frob_t, squarings, sx);
}

template <class T>
__global__ void __launch_bounds__(256) k_lift(LT<T>* x, const LT<T>* frob_t, int squarings) {
  __shared__ __attribute__((aligned(16))) CT<T> sx[E];
  teich_lift_block<T>(x + (uint64_t)blockIdx.x * E, x + (uint64_t)blockIdx.x * E, frob_t, squarings, sx);
}

// ============================================================================
// Packing (phi), unpacking (psi) and matvecs
// ============================================================================

// out[rot*n + e] = pack_r2(rot_rot(slots[e]))  for rot in [0,nrot)
// (embedding.rs:63-65 as the R0-linear map with basis pack_r2(e_s), pcs.rs:49-55;
//  rot_l from reduce.rs:352-358).  basis: 32 x 128 limbs.
template <class T>
__global__ void __launch_bounds__(GB) k_pack(const LT<T>* slots, const LT<T>* basis, LT<T>* out, uint64_t n, int nrot,
                                             int rot0) {
  __shared__ __attribute__((aligned(16))) LT<T> sb[32 * E];
  __shared__ CT<T> ssl[GPB][32];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  // 16 groups of 4 limbs per thread, loads issued in batches of 8
#pragma unroll
  for (int u0 = 0; u0 < 32 * E / 4 / GB; u0 += 8) {
    CT<T> v[8][4];
#pragma unroll
    for (int u = 0; u < 8; u++) ld4<T>(basis + 4 * (threadIdx.x + (u0 + u) * GB), v[u]);
#pragma unroll
    for (int u = 0; u < 8; u++) st4<T>(sb + 4 * (threadIdx.x + (u0 + u) * GB), v[u]);
  }
  uint64_t total = n * (uint64_t)nrot;
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < total; base += (uint64_t)gridDim.x * GPB) {
    uint64_t w = base + g;
    bool ok = w < total;
    uint64_t rot = ok ? w / n : 0, e = ok ? w % n : 0;
    __syncthreads();
    if (ok) ssl[g][t] = slots[e * 32 + ((t + rot + rot0) & 31)];
    __syncthreads();
    if (ok) {
      CT<T> acc[4] = {0, 0, 0, 0};
      for (int s = 0; s < 32; s++) {
        CT<T> bv[4];
        ld4<T>(sb + s * E + 4 * t, bv);
        CT<T> x = ssl[g][s];
        for (int c = 0; c < 4; c++) acc[c] += x * bv[c];
      }
      st4<T>(out + w * E + 4 * t, acc);
    }
  }
}

// psi (embedding.rs:247-261, flattened as general.rs:192-206):
// slots[2i] = g_i(0), slots[2i+1] = g_i(1), g_i = sum_k x_k t_i^k in R1.
// pw: powers t_i^k as [16][32][4] limbs.  Lanes t and t+16 split the 32 k's.
template <class T>
__global__ void __launch_bounds__(GB) k_psi(const LT<T>* x, const LT<T>* pw, LT<T>* slots, uint64_t n) {
  __shared__ __attribute__((aligned(16))) LT<T> sp[16 * 32 * 4];
  __shared__ __attribute__((aligned(16))) LT<T> sx[GPB][E];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  {
    CT<T> v[16 * 32 / GB][4];  // all loads in flight before the first store
#pragma unroll
    for (int u = 0; u < 16 * 32 / GB; u++) ld4<T>(pw + 4 * (threadIdx.x + u * GB), v[u]);
#pragma unroll
    for (int u = 0; u < 16 * 32 / GB; u++) st4<T>(sp + 4 * (threadIdx.x + u * GB), v[u]);
  }
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < n; base += (uint64_t)gridDim.x * GPB) {
    uint64_t e = base + g;
    bool ok = e < n;
    __syncthreads();
    if (ok) stage<T>(sx[g], x + e * E, t);
    __syncthreads();
    int pt = t & 15, half = t >> 4;
    CT<T> raw[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int kk = 0; kk < 16; kk++) {
      int k = 16 * half + kk;
      CT<T> a[4], b[4], p[7];
      ld4<T>(sx[g] + 4 * k, a);
      ld4<T>(sp + (pt * 32 + k) * 4, b);
      d_raw(a, b, p);
      for (int c = 0; c < 7; c++) raw[c] += p[c];
    }
    for (int c = 0; c < 7; c++) raw[c] += __shfl_xor(raw[c], 16, 32);
    CT<T> gg[4];
    d_red(raw, gg);
    if (ok && half == 0) {
      slots[e * 32 + 2 * pt] = (LT<T>)gg[0];
      slots[e * 32 + 2 * pt + 1] = (LT<T>)(gg[0] + gg[1] + gg[2] + gg[3]);
    }
  }
}

// R0 CSR matvec over nvec vectors: out[r*ostride + v*ovstride] = sum A[r,c] w[v*wstride + c]
template <class T>
__global__ void k_r0_matvec(const uint32_t* ptr, const uint32_t* col, const LT<T>* val, const LT<T>* w, uint64_t wstride,
                            LT<T>* out, uint64_t nrows, int nvec, uint64_t ostride, uint64_t ovstride) {
  uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nrows * (uint64_t)nvec) return;
  uint64_t r = id / nvec;
  int v = (int)(id % nvec);
  CT<T> acc = 0;
  for (uint32_t k = ptr[r]; k < ptr[r + 1]; k++) acc += (CT<T>)val[k] * (CT<T>)w[v * wstride + col[k]];
  out[r * ostride + v * ovstride] = (LT<T>)acc;
}

// The 32 per-instance R0 matvecs of a MIMD matrix set over its union
// support (reduce.rs:194-201): out[r][s] = sum_{cells g of row r}
// val[g][s] * w[col[g]][s], i.e. (A_s w_s)[r] in slot layout [row][32].
template <class T>
__global__ void k_mimd_slots(const uint64_t* ptr, const uint32_t* col, const LT<T>* val, const LT<T>* w, LT<T>* out,
                             uint64_t nrows) {
  for (uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; id < nrows * 32;
       id += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = id >> 5, s = id & 31;
    CT<T> acc = 0;
    for (uint64_t g = ptr[r]; g < ptr[r + 1]; g++) acc += (CT<T>)val[g * 32 + s] * (CT<T>)w[(uint64_t)col[g] * 32 + s];
    out[id] = (LT<T>)acc;
  }
}

// out[o] (+)= sum_{k in ptr[o]..ptr[o+1]} val[k] * X[idx[k]]   (R0 scalar x R2,
// the mul_by_gr1 matvecs of matrix.rs:93-105 and simd.rs:570-580)
// Outputs with heavy_min or more entries keep only the accumulate baseline
// here; k_spmv_heavy / k_spmv_heavy_sum add their sums.
template <class T>
__global__ void __launch_bounds__(GB) k_spmv_gr1(const uint32_t* ptr, const uint32_t* idx, const LT<T>* val,
                                                 const LT<T>* X, LT<T>* out, uint64_t nout, int accumulate,
                                                 uint32_t heavy_min) {
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  for (uint64_t o = (uint64_t)blockIdx.x * GPB + g; o < nout; o += (uint64_t)gridDim.x * GPB) {
    CT<T> acc[4] = {0, 0, 0, 0};
    if (accumulate) ld4<T>(out + o * E + 4 * t, acc);
    const uint32_t k0 = ptr[o], k1 = ptr[o + 1];
    if (k1 - k0 < heavy_min) {
      for (uint32_t k = k0; k < k1; k++) {
        CT<T> x[4];
        ld4<T>(X + (uint64_t)idx[k] * E + 4 * t, x);
        CT<T> v = val[k];
        for (int c = 0; c < 4; c++) acc[c] += v * x[c];
      }
    }
    st4<T>(out + o * E + 4 * t, acc);
  }
}

// Outputs with many entries (a column every constraint uses, say): block
// (c, h) sums slice c of the entries of output heavy[h] (its groups stride
// through the slice) into part[h][c]; k_spmv_heavy_sum adds the slices.
template <class T>
__global__ void __launch_bounds__(GB) k_spmv_heavy(const uint32_t* ptr, const uint32_t* idx, const LT<T>* val,
                                                   const LT<T>* X, const uint32_t* heavy, LT<T>* part) {
  using C = CT<T>;
  __shared__ C red[GPB][E];
  const int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  const uint32_t o = heavy[blockIdx.y];
  const uint32_t k0 = ptr[o], k1 = ptr[o + 1];
  const uint32_t per = (k1 - k0 + gridDim.x - 1) / gridDim.x;
  const uint32_t a = min(k1, k0 + blockIdx.x * per), b = min(k1, a + per);
  C acc[4] = {0, 0, 0, 0};
  for (uint32_t k = a + g; k < b; k += GPB) {
    C x[4];
    ld4<T>(X + (uint64_t)idx[k] * E + 4 * t, x);
    const C v = val[k];
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] += v * x[c];
  }
#pragma unroll
  for (int c = 0; c < 4; c++) red[g][4 * t + c] = acc[c];
  __syncthreads();
  for (int l = threadIdx.x; l < E; l += GB) {
    C s = 0;
#pragma unroll
    for (int q = 0; q < GPB; q++) s += red[q][l];
    part[((uint64_t)blockIdx.y * gridDim.x + blockIdx.x) * E + l] = (LT<T>)s;
  }
}

template <class T>
__global__ void __launch_bounds__(E) k_spmv_heavy_sum(const LT<T>* part, uint32_t nch, const uint32_t* heavy,
                                                      LT<T>* out) {
  const uint64_t o = heavy[blockIdx.x];
  const int l = threadIdx.x;
  CT<T> s = (CT<T>)out[o * E + l];
  for (uint32_t c = 0; c < nch; c++) s += (CT<T>)part[((uint64_t)blockIdx.x * nch + c) * E + l];
  out[o * E + l] = (LT<T>)s;
}

// m-bar (pcs.rs:554-561): mbar[j] = sum_{i<rows} e0s[i] * M[i*k2 + j]  (M low-degree if NA=16)
template <class T, int NA>
__global__ void __launch_bounds__(GB) k_mbar(const LT<T>* e0s, const LT<T>* M, uint64_t rows, uint64_t k2,
                                             LT<T>* mbar, uint64_t mstride) {
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][2][E];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < k2; base += (uint64_t)gridDim.x * GPB) {
    uint64_t j = base + g;
    bool ok = j < k2;
    Acc<T> acc;
    acc.clear();
    for (uint64_t i = 0; i < rows; i++) {
      if (ok) {
        stage<T>(sh[g][0], M + (i * k2 + j) * E, t);
        stage<T>(sh[g][1], e0s + i * E, t);
      }
      __syncthreads();
      if (ok) acc_mul<T, NA>(acc, sh[g][0], sh[g][1], t);
      __syncthreads();
    }
    CT<T> r[4];
    acc_final<T>(acc, t, r);
    if (ok) st4<T>(mbar + j * E + 4 * t, r);
    (void)mstride;
  }
}

// Rotated m-bars (pcs.rs:984-1001): mbar[l][j] = sum_i e0[i] * rho_l[i*k2+j]
template <class T>
__global__ void __launch_bounds__(GB) k_rot_mbar(const LT<T>* e0, const LT<T>* rho, uint64_t nb, uint64_t k1,
                                                 uint64_t k2, int nl, LT<T>* mbar) {
  __shared__ __attribute__((aligned(16))) LT<T> sh[GPB][2][E];
  int g = threadIdx.x >> 5, t = threadIdx.x & 31;
  uint64_t total = k2 * (uint64_t)nl;
  for (uint64_t base = (uint64_t)blockIdx.x * GPB; base < total; base += (uint64_t)gridDim.x * GPB) {
    uint64_t w = base + g;
    bool ok = w < total;
    uint64_t l = ok ? w / k2 : 0, j = ok ? w % k2 : 0;
    Acc<T> acc;
    acc.clear();
    for (uint64_t i = 0; i < k1; i++) {
      if (ok) {
        stage<T>(sh[g][0], rho + (l * nb + i * k2 + j) * E, t);
        stage<T>(sh[g][1], e0 + i * E, t);
      }
      __syncthreads();
      if (ok) acc_mul<T, 16>(acc, sh[g][0], sh[g][1], t);
      __syncthreads();
    }
    CT<T> r[4];
    acc_final<T>(acc, t, r);
    if (ok) st4<T>(mbar + w * E + 4 * t, r);
  }
}

// ============================================================================
// RAA encoder (raa.rs:67-153), all rows of a commitment at once, W limbs per
// entry.  Each pass is gather-and-scale then an inclusive prefix sum over the
// codeword index, done as a segmented scan across many blocks:
//   gather  v[i] = src[perm[i] >> qs] * d[i]  (q = 2^qs) and per-segment sums
//   segscan exclusive scan of the segment sums, one block per row
//   scan    each block scans its segment (read once into registers) from its
//           segment's offset
// Encode = pass(msg -> t, perm1, d1, q) then pass(t -> z, perm2, d2, 1).
// Each thread moves V limbs of an entry with one access of up to 16 bytes
// (TPE threads per entry), SUB entries lie side by side and a thread takes
// PER consecutive entries: a block's segment is SEG = SUB * PER entries.
// ============================================================================
constexpr int RAA_TB = 256;  // threads per block
template <class T, int W> struct Raa {
  static constexpr int LB = (int)sizeof(LT<T>);
  static constexpr int V = (16 / LB) < (W / 8) ? 16 / LB : W / 8;  // limbs per access
  static constexpr int TPE = W / V;                                 // threads per entry
  static constexpr int SUB = RAA_TB / TPE;                          // entries side by side
  static constexpr int PB = 128 / (V * (int)sizeof(CT<T>));
  static constexpr int PER = PB > 4 ? PB : 4;  // entries per thread (about 32 registers of values)
  static constexpr int SEG = SUB * PER;                             // entries per block
};
template <class L, int V> struct alignas(V * sizeof(L)) LimbVec {
  L v[V];
};
// V limbs at p (one aligned access) widened to C, and back (truncated to L)
template <class L, class C, int V> __device__ __forceinline__ void ld_limbs(const L* p, C out[V]) {
  const LimbVec<L, V> x = *reinterpret_cast<const LimbVec<L, V>*>(p);
#pragma unroll
  for (int u = 0; u < V; u++) out[u] = (C)x.v[u];
}
template <class L, class C, int V> __device__ __forceinline__ void st_limbs(L* p, const C in[V]) {
  LimbVec<L, V> x;
#pragma unroll
  for (int u = 0; u < V; u++) x.v[u] = (L)in[u];
  *reinterpret_cast<LimbVec<L, V>*>(p) = x;
}

template <class T, int W>
__global__ void __launch_bounds__(RAA_TB) k_raa_gather(const LT<T>* src, uint64_t src_row, const uint32_t* perm,
                                                       const LT<T>* d, int qs, uint64_t n, LT<T>* dst,
                                                       CT<T>* segsum, uint64_t nseg) {
  using R = Raa<T, W>;
  using L = LT<T>;
  using C = CT<T>;
  constexpr int V = R::V, PER = R::PER;
  __shared__ C red[R::SUB][W];
  const uint64_t row = blockIdx.y, seg = blockIdx.x;
  const int tq = threadIdx.x % R::TPE, sub = threadIdx.x / R::TPE;
  const L* sr = src + row * src_row * W + tq * V;
  L* dr = dst + row * n * W + tq * V;
  const uint64_t b = seg * R::SEG + (uint64_t)sub * PER;
  C acc[V];
#pragma unroll
  for (int u = 0; u < V; u++) acc[u] = 0;
  if (b + PER <= n) {
    // all PER indices, scalars and source entries in flight at once
    uint32_t pi[PER];
    C dv[PER], x[PER][V];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      pi[k] = perm[b + k] >> qs;
      dv[k] = (C)d[b + k];
    }
#pragma unroll
    for (int k = 0; k < PER; k++) ld_limbs<L, C, V>(sr + (uint64_t)pi[k] * W, x[k]);
#pragma unroll
    for (int k = 0; k < PER; k++) {
#pragma unroll
      for (int u = 0; u < V; u++) {
        x[k][u] *= dv[k];
        acc[u] += x[k][u];
      }
      st_limbs<L, C, V>(dr + (b + k) * W, x[k]);
    }
  } else {
    for (int k = 0; k < PER && b + k < n; k++) {
      C x[V];
      ld_limbs<L, C, V>(sr + (uint64_t)(perm[b + k] >> qs) * W, x);
      const C dv = (C)d[b + k];
#pragma unroll
      for (int u = 0; u < V; u++) {
        x[u] *= dv;
        acc[u] += x[u];
      }
      st_limbs<L, C, V>(dr + (b + k) * W, x);
    }
  }
#pragma unroll
  for (int u = 0; u < V; u++) red[sub][tq * V + u] = acc[u];
  __syncthreads();
  for (int s = threadIdx.x; s < W; s += RAA_TB) {
    C t = 0;
#pragma unroll 8
    for (int k = 0; k < R::SUB; k++) t += red[k][s];
    segsum[(row * nseg + seg) * W + s] = t;
  }
}
// Exclusive scan of one row's segment sums (grid: rows): thread (sub, q)
// takes limbs [q * LPT, (q + 1) * LPT) of a contiguous range of segments,
// eight segments' loads in flight at a time; the 32 ranges' totals are
// scanned in LDS.
template <class T, int W>
__global__ void __launch_bounds__(1024) k_raa_segscan(CT<T>* segsum, uint64_t nseg) {
  using C = CT<T>;
  constexpr int LPT = W / 32, SUBS = 32;
  __shared__ C part[SUBS][W];
  const int q = threadIdx.x % 32, sub = threadIdx.x / 32;
  const uint64_t per = (nseg + SUBS - 1) / SUBS, g0 = (uint64_t)sub * per;
  const uint64_t g1 = g0 + per < nseg ? g0 + per : nseg;
  C* base = segsum + (uint64_t)blockIdx.x * nseg * W + q * LPT;
  C tot[LPT];
#pragma unroll
  for (int u = 0; u < LPT; u++) tot[u] = 0;
  for (uint64_t g = g0; g < g1; g += 8) {
    C x[8][LPT];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (g + k < g1) {
        ld_limbs<C, C, LPT>(base + (g + k) * W, x[k]);
      } else {
#pragma unroll
        for (int u = 0; u < LPT; u++) x[k][u] = 0;
      }
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
      for (int u = 0; u < LPT; u++) tot[u] += x[k][u];
  }
#pragma unroll
  for (int u = 0; u < LPT; u++) part[sub][q * LPT + u] = tot[u];
  __syncthreads();
  if (threadIdx.x < W) {
    C run = 0;
    for (int k = 0; k < SUBS; k++) {
      const C t = part[k][threadIdx.x];
      part[k][threadIdx.x] = run;
      run += t;
    }
  }
  __syncthreads();
  C run[LPT];
#pragma unroll
  for (int u = 0; u < LPT; u++) run[u] = part[sub][q * LPT + u];
  for (uint64_t g = g0; g < g1; g += 8) {
    C x[8][LPT];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (g + k < g1) {
        ld_limbs<C, C, LPT>(base + (g + k) * W, x[k]);
      } else {
#pragma unroll
        for (int u = 0; u < LPT; u++) x[k][u] = 0;
      }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (g + k < g1) st_limbs<C, C, LPT>(base + (g + k) * W, run);
#pragma unroll
      for (int u = 0; u < LPT; u++) run[u] += x[k][u];
    }
  }
}
template <class T, int W>
__global__ void __launch_bounds__(RAA_TB) k_raa_scan(LT<T>* buf, uint64_t n, const CT<T>* segsum, uint64_t nseg) {
  using R = Raa<T, W>;
  using L = LT<T>;
  using C = CT<T>;
  constexpr int V = R::V, PER = R::PER;
  __shared__ C red[R::SUB][W];
  const uint64_t row = blockIdx.y, seg = blockIdx.x;
  const int tq = threadIdx.x % R::TPE, sub = threadIdx.x / R::TPE;
  L* br = buf + row * n * W + tq * V;
  const uint64_t b = seg * R::SEG + (uint64_t)sub * PER;
  C v[PER][V], tot[V];
#pragma unroll
  for (int u = 0; u < V; u++) tot[u] = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    if (b + k < n) {
      ld_limbs<L, C, V>(br + (b + k) * W, v[k]);
    } else {
#pragma unroll
      for (int u = 0; u < V; u++) v[k][u] = 0;
    }
#pragma unroll
    for (int u = 0; u < V; u++) tot[u] += v[k][u];
  }
#pragma unroll
  for (int u = 0; u < V; u++) red[sub][tq * V + u] = tot[u];
  __syncthreads();
  // per limb: the segment's offset plus the totals of the entries before
  // this thread's within the segment
  for (int s = threadIdx.x; s < W; s += RAA_TB) {
    C run = segsum[(row * nseg + seg) * W + s];
    for (int k = 0; k < R::SUB; k++) {
      const C t = red[k][s];
      red[k][s] = run;
      run += t;
    }
  }
  __syncthreads();
  C run[V];
#pragma unroll
  for (int u = 0; u < V; u++) run[u] = red[sub][tq * V + u];
#pragma unroll
  for (int k = 0; k < PER; k++) {
#pragma unroll
    for (int u = 0; u < V; u++) run[u] += v[k][u];
    if (b + k < n) st_limbs<L, C, V>(br + (b + k) * W, run);
  }
}

// ============================================================================
// Merkle commitment over columns (merkle.rs:20-87, r2_pcs.rs:194-218)
// ============================================================================
// Column leaves, chunk-parallel.  Entries are es = 2^esl bytes (a multiple of
// 16), so a 64-byte block of a column is read as four 16-byte pieces with
// shifts and masks; bytes at or past blen in the block read as 0.
__device__ __forceinline__ void col_block(const uint8_t* base, uint64_t rowb, int esl, uint64_t o, uint32_t blen,
                                          uint32_t m[16]) {
  const uint64_t emask = (1ull << esl) - 1;
#pragma unroll
  for (int p = 0; p < 4; p++) {
    uint4 v = make_uint4(0, 0, 0, 0);
    if (16u * p < blen) {
      const uint64_t q = o + 16 * p;
      v = *reinterpret_cast<const uint4*>(base + (q >> esl) * rowb + (q & emask));
    }
    m[4 * p] = v.x; m[4 * p + 1] = v.y; m[4 * p + 2] = v.z; m[4 * p + 3] = v.w;
  }
}
// Chaining value of chunk c of a len-byte column (or, with root, the hash of
// a one-chunk column).
__device__ __forceinline__ void col_chunk(const uint8_t* base, uint64_t rowb, int esl, uint64_t len, uint64_t c,
                                          bool root, uint32_t out[8]) {
  const uint64_t cbeg = c * 1024, clen = len - cbeg < 1024 ? len - cbeg : 1024;
  const uint32_t nblk = (uint32_t)((clen + 63) / 64);
  uint32_t cv[8], m[16], o[16];
  b3::iv(cv);
  for (uint32_t b = 0; b + 1 < nblk; b++) {
    col_block(base, rowb, esl, cbeg + 64ull * b, 64, m);
    b3::compress(cv, m, c, 64, b == 0 ? b3::CHUNK_START : 0, o);
    for (int i = 0; i < 8; i++) cv[i] = o[i];
  }
  const uint32_t lb = nblk - 1, blen = (uint32_t)(clen - 64ull * lb);
  col_block(base, rowb, esl, cbeg + 64ull * lb, blen, m);
  const uint32_t fl = (lb == 0 ? b3::CHUNK_START : 0) | b3::CHUNK_END | (root ? b3::ROOT : 0);
  b3::compress(cv, m, c, blen, fl, o);
  for (int i = 0; i < 8; i++) out[i] = o[i];
}
// Thread per (column j, chunk c), columns fastest (n = 2^ln columns of
// rows x es bytes, rows x n row-major).  One-chunk columns write their leaf;
// longer ones write chunk chaining values cvs[j][c] for k_col_merge.
__global__ void k_col_chunks(const uint8_t* enc, uint64_t rows, int ln, int esl, uint64_t nch, uint32_t* cvs,
                             uint32_t* leaves) {
  const uint64_t n = 1ull << ln, len = rows << esl, rowb = n << esl;
  for (uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; id < n * nch;
       id += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t j = id & (n - 1), c = id >> ln;
    uint32_t out[8];
    col_chunk(enc + (j << esl), rowb, esl, len, c, nch == 1, out);
    uint32_t* dst = nch == 1 ? leaves + j * 8 : cvs + (j * nch + c) * 8;
    for (int i = 0; i < 8; i++) dst[i] = out[i];
  }
}
__device__ __forceinline__ void b3_parent(const uint32_t* l, const uint32_t* r, uint32_t flags, uint32_t out[8]) {
  uint32_t key[8], m[16], o[16];
  b3::iv(key);
  for (int i = 0; i < 8; i++) { m[i] = l[i]; m[8 + i] = r[i]; }
  b3::compress(key, m, 0, 64, b3::PARENT | flags, o);
  for (int i = 0; i < 8; i++) out[i] = o[i];
}
// Leaf of each multi-chunk column from its chunk chaining values: BLAKE3's
// left-complete tree is pairwise merging with an odd last node carried up;
// the last parent is the root.  In place in the column's own CV slots.
__global__ void k_col_merge(uint32_t* cvs, uint64_t n, uint64_t nch, uint32_t* leaves) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t* cv = cvs + j * nch * 8;
  uint64_t cnt = nch;
  uint32_t o[8];
  while (cnt > 2) {
    for (uint64_t i = 0; i < cnt / 2; i++) {
      b3_parent(cv + 16 * i, cv + 16 * i + 8, 0, o);
      for (int k = 0; k < 8; k++) cv[8 * i + k] = o[k];
    }
    if (cnt & 1)
      for (int k = 0; k < 8; k++) cv[8 * (cnt / 2) + k] = cv[8 * (cnt - 1) + k];
    cnt = (cnt + 1) / 2;
  }
  b3_parent(cv, cv + 8, b3::ROOT, o);
  for (int k = 0; k < 8; k++) leaves[j * 8 + k] = o[k];
}

// nodes: level-order, leaves first; next level = BLAKE3(left || right)
__global__ void k_merkle_level(uint32_t* nodes, uint64_t start, uint64_t len) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len / 2) return;
  const uint32_t* l = nodes + (start + 2 * i) * 8;
  uint32_t out[8];
  b3::hash64(l, l + 8, out);
  uint32_t* o = nodes + (start + len + i) * 8;
  for (int k = 0; k < 8; k++) o[k] = out[k];
}

// The top levels of a Merkle tree in one block: level `len` <= MERKLE_TOP
// nodes at `start` up to the root, staged through LDS, every level written
// to `nodes`.
constexpr int MERKLE_TOP = 1024;
__global__ void __launch_bounds__(1024) k_merkle_top(uint32_t* nodes, uint64_t start, uint64_t len) {
  __shared__ uint32_t a[MERKLE_TOP][8], b[MERKLE_TOP / 2][8];
  for (uint64_t i = threadIdx.x; i < len; i += blockDim.x)
    for (int k = 0; k < 8; k++) a[i][k] = nodes[(start + i) * 8 + k];
  __syncthreads();
  while (len > 1) {
    const uint64_t half = len / 2;
    for (uint64_t i = threadIdx.x; i < half; i += blockDim.x) {
      uint32_t o[8];
      b3::hash64(a[2 * i], a[2 * i + 1], o);
      for (int k = 0; k < 8; k++) {
        b[i][k] = o[k];
        nodes[(start + len + i) * 8 + k] = o[k];
      }
    }
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < half; i += blockDim.x)
      for (int k = 0; k < 8; k++) a[i][k] = b[i][k];
    __syncthreads();
    start += len;
    len = half;
  }
}

// ---- one long message (Fiat-Shamir query derivation over m-bar) ----
// Chunk chaining values of a byte message (len bytes; bytes past len read as 0
// from a padded buffer).  Thread per 1024-byte chunk; for a one-chunk message
// the root node inputs are written instead (cv, block, blen, flags).
__global__ void k_b3_chunks(const uint8_t* msg, uint64_t len, uint64_t nchunks, uint32_t* cvs, uint32_t* rootnode) {
  uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  uint64_t cbeg = c * 1024;
  uint64_t clen = len - cbeg < 1024 ? len - cbeg : 1024;
  uint32_t nblk = clen == 0 ? 1 : (uint32_t)((clen + 63) / 64);
  uint32_t cv[8], m[16], o[16];
  b3::iv(cv);
  // 16-byte pieces (the buffer is 16-byte aligned); the piece holding the
  // message end is read byte by byte, with bytes past len as 0
  auto block = [&](uint64_t byte) {
#pragma unroll
    for (int p = 0; p < 4; p++) {
      const uint64_t q = byte + 16 * p;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (q + 16 <= len) {
        v = *reinterpret_cast<const uint4*>(msg + q);
      } else if (q < len) {
        uint32_t w[4] = {0, 0, 0, 0};
        for (int i = 0; i < 16 && q + i < len; i++) w[i / 4] |= (uint32_t)msg[q + i] << (8 * (i % 4));
        v = make_uint4(w[0], w[1], w[2], w[3]);
      }
      m[4 * p] = v.x; m[4 * p + 1] = v.y; m[4 * p + 2] = v.z; m[4 * p + 3] = v.w;
    }
  };
  for (uint32_t b = 0; b + 1 < nblk; b++) {
    block(cbeg + 64 * (uint64_t)b);
    b3::compress(cv, m, c, 64, b == 0 ? b3::CHUNK_START : 0, o);
    for (int i = 0; i < 8; i++) cv[i] = o[i];
  }
  uint32_t lb = nblk - 1;
  uint32_t blen = (uint32_t)(clen - 64 * (uint64_t)lb);
  block(cbeg + 64 * (uint64_t)lb);
  uint32_t fl = (lb == 0 ? b3::CHUNK_START : 0) | b3::CHUNK_END;
  if (nchunks == 1) {
    for (int i = 0; i < 8; i++) rootnode[i] = cv[i];
    for (int i = 0; i < 16; i++) rootnode[8 + i] = m[i];
    rootnode[24] = blen;
    rootnode[25] = fl;
    return;
  }
  b3::compress(cv, m, c, blen, fl, o);
  for (int i = 0; i < 8; i++) cvs[c * 8 + i] = o[i];
}
// The last merge levels of a long message in one block: cnt <= MERKLE_TOP
// chaining values down to the two children of the root node (kids).
__global__ void __launch_bounds__(1024) k_b3_merge_top(const uint32_t* in, uint64_t cnt, uint32_t* kids) {
  __shared__ uint32_t a[MERKLE_TOP][8], b[MERKLE_TOP / 2 + 1][8];
  for (uint64_t i = threadIdx.x; i < cnt; i += blockDim.x)
    for (int k = 0; k < 8; k++) a[i][k] = in[i * 8 + k];
  __syncthreads();
  while (cnt > 2) {
    const uint64_t half = cnt / 2;
    for (uint64_t i = threadIdx.x; i < half; i += blockDim.x) b3_parent(a[2 * i], a[2 * i + 1], 0, b[i]);
    if (threadIdx.x == 0 && (cnt & 1))
      for (int k = 0; k < 8; k++) b[half][k] = a[cnt - 1][k];
    __syncthreads();
    cnt = (cnt + 1) / 2;
    for (uint64_t i = threadIdx.x; i < cnt; i += blockDim.x)
      for (int k = 0; k < 8; k++) a[i][k] = b[i][k];
    __syncthreads();
  }
  if (threadIdx.x < 16) kids[threadIdx.x] = a[threadIdx.x / 8][threadIdx.x % 8];
}

// One merge level of the left-complete BLAKE3 tree: pairs become parents, an
// odd last node is carried up unchanged.
__global__ void k_b3_merge(const uint32_t* in, uint64_t cnt, uint32_t* out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t half = cnt / 2;
  if (i < half) {
    uint32_t key[8], m[16], o[16];
    b3::iv(key);
    for (int k = 0; k < 8; k++) { m[k] = in[(2 * i) * 8 + k]; m[8 + k] = in[(2 * i + 1) * 8 + k]; }
    b3::compress(key, m, 0, 64, b3::PARENT, o);
    for (int k = 0; k < 8; k++) out[i * 8 + k] = o[k];
  } else if (i == half && (cnt & 1)) {
    for (int k = 0; k < 8; k++) out[i * 8 + k] = in[(cnt - 1) * 8 + k];
  }
}

// Serialized Merkle paths (pcs.rs:390-397): per query a u64 path length then
// the sibling hashes from the leaf up.  Thread per (query, level).
__global__ void k_merkle_paths(const uint32_t* nodes, uint64_t n, const uint32_t* idx, uint64_t nq, int plen,
                               uint8_t* out) {
  uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nq * plen) return;
  uint64_t q = id / plen;
  int lvl = (int)(id % plen);
  uint8_t* rec = out + q * (8 + (uint64_t)plen * 32);
  if (lvl == 0) *reinterpret_cast<uint64_t*>(rec) = (uint64_t)plen;
  uint64_t start = 0, len = n;
  for (int l = 0; l < lvl; l++) { start += len; len >>= 1; }
  uint64_t node = start + ((idx[q] >> lvl) ^ 1);
  uint32_t* dst = reinterpret_cast<uint32_t*>(rec + 8 + (uint64_t)lvl * 32);
  for (int k = 0; k < 8; k++) dst[k] = nodes[node * 8 + k];
}

// Query indices (pcs.rs:1130-1156; r2_pcs.rs:220-246) when the codeword
// length is a power of two (mask + 1): every 4-byte XOF word is kept, so
// index q is word q % 16 of the root node's output block q / 16, masked.
// The root node is the last parent (parent != 0: its 16 message words at
// src) or the single chunk's final compression (src: cv, block, blen,
// flags).  Written to idx (device) and to the proof buffer's index section.
__global__ void k_queries(const uint32_t* src, int parent, uint64_t t, uint32_t mask, uint32_t* idx, uint32_t* pidx) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= t) return;
  uint32_t cv[8], m[16], o[16], blen, flags;
  if (parent) {
    b3::iv(cv);
    for (int i = 0; i < 16; i++) m[i] = src[i];
    blen = 64;
    flags = b3::PARENT;
  } else {
    for (int i = 0; i < 8; i++) cv[i] = src[i];
    for (int i = 0; i < 16; i++) m[i] = src[8 + i];
    blen = src[24];
    flags = src[25];
  }
  b3::compress(cv, m, q / 16, blen, flags | b3::ROOT, o);
  const uint32_t v = o[q % 16] & mask;
  idx[q] = v;
  pidx[q] = v;
}

// Proof-section download: `bytes` (a multiple of 4) from device memory to
// the pinned proof buffer through its device address; 16-byte stores when
// the destination allows them.  Runs on the download stream (at most 256
// blocks) beside the prover's kernels.
__device__ __forceinline__ void copy_bytes(const uint8_t* src, uint8_t* dst, uint64_t bytes) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nth = (uint64_t)gridDim.x * blockDim.x;
  const uintptr_t al = (uintptr_t)src | (uintptr_t)dst;
  uint64_t done = 0;
  if ((al & 15) == 0) {
    const uint64_t n16 = bytes / 16;
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (uint64_t i = tid; i < n16; i += nth) d4[i] = s4[i];
    done = n16 * 16;
  }
  if ((al & 3) == 0) {
    const uint64_t nw = (bytes - done) / 4;
    const uint32_t* s1 = reinterpret_cast<const uint32_t*>(src + done);
    uint32_t* d1 = reinterpret_cast<uint32_t*>(dst + done);
    for (uint64_t i = tid; i < nw; i += nth) d1[i] = s1[i];
    done += nw * 4;
  }
  for (uint64_t i = done + tid; i < bytes; i += nth) dst[i] = src[i];
}
__global__ void __launch_bounds__(256) k_copy_out(const uint8_t* src, uint8_t* dst, uint64_t bytes) {
  copy_bytes(src, dst, bytes);
}
// Small uploads and root downloads between device memory and mapped pinned
// host memory (a kernel of our own rather than the runtime's copy path).
__global__ void __launch_bounds__(256) k_copy_small(const uint8_t* src, uint8_t* dst, uint64_t bytes) {
  copy_bytes(src, dst, bytes);
}

// slots[j][s] = w[s][j] for the 32 witnesses of m limbs each (the slot
// layout of simd_to_packed, reduce.rs:33-60): a block moves 64 columns
// through LDS so both the reads and the writes are contiguous.
template <class T>
__global__ void __launch_bounds__(256) k_slots(const LT<T>* w, uint64_t m, LT<T>* slots) {
  __shared__ LT<T> t[32][65];
  for (uint64_t j0 = (uint64_t)blockIdx.x * 64; j0 < m; j0 += (uint64_t)gridDim.x * 64) {
    for (int x = threadIdx.x; x < 32 * 64; x += 256) {
      const int s = x >> 6, jj = x & 63;
      const uint64_t j = j0 + jj;
      t[s][jj] = j < m ? w[s * m + j] : (LT<T>)0;
    }
    __syncthreads();
    for (int x = threadIdx.x; x < 32 * 64; x += 256) {
      const int jj = x >> 5, s = x & 31;
      const uint64_t j = j0 + jj;
      if (j < m) slots[j * 32 + s] = t[s][jj];
    }
    __syncthreads();
  }
}

// Serialized rotated m-bars (pcs.rs:426-471): per rotation a u64 length
// `cnt`, then the rotation's `per` 8-byte words of src.
__global__ void k_len_prefixed(const uint64_t* src, uint64_t nrec, uint64_t per, uint64_t cnt, uint64_t* dst) {
  const uint64_t rec = per + 1, n = nrec * rec;
  for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t l = x / rec, o = x % rec;
    dst[x] = o == 0 ? cnt : src[l * per + o - 1];
  }
}

// Serialized queried columns (pcs.rs:379-388): per query a u64 entry count,
// then enc[i][idx[q]] for rows i (W limbs each).  Thread per limb.
template <class T>
__global__ void k_gather_cols(const LT<T>* enc, const uint32_t* idx, uint64_t nq, uint64_t rows, uint64_t n, int W,
                              uint8_t* out) {
  uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t per = rows * W;
  if (id >= nq * per) return;
  uint64_t q = id / per, r = (id % per) / W, s = id % W;
  uint8_t* rec = out + q * (8 + per * sizeof(LT<T>));
  if (r == 0 && s == 0) *reinterpret_cast<uint64_t*>(rec) = rows;
  reinterpret_cast<LT<T>*>(rec + 8)[r * W + s] = enc[(r * n + idx[q]) * W + s];
}

// small element-wise helpers (thread per limb)
template <class T>
__global__ void k_sub(const LT<T>* a, const LT<T>* b, LT<T>* o, uint64_t nl) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nl) o[i] = (LT<T>)(a[i] - b[i]);
}
template <class T>
__global__ void k_hadamard(const LT<T>* a, const LT<T>* b, LT<T>* o, uint64_t nl) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nl) o[i] = (LT<T>)((CT<T>)a[i] * (CT<T>)b[i]);
}

}  // namespace nenya
