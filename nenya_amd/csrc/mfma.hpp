// Products by a fixed element on the gfx950 matrix cores.
//
// For fixed r in GR(2^k,128) the map x -> r*x is Z/2^k-linear: r*x = M_r x
// with M_r the 128x128 matrix whose column j is r*e_j (e_{4i+c} = X^c Y^i,
// the limb order of an element).  Applying r to many elements is therefore a
// GEMM over Z/2^k, which int8 MFMA computes exactly through balanced digits:
// u = sum_p 256^p u_p (mod 2^k) with u_p in [-128, 127]; the bytes of
// (u + 0x80..80) ^ 0x80..80 are those digits as int8.  With ND = k/8,
//   M x = sum_{s<ND} 256^s sum_{p+q=s} M_p x_q   (mod 2^k),
// every M_p x_q is an exact small-integer product (|.| <= 128 * 2^14), and
// the i32 accumulators wrap mod 2^32, so Horner in 256 over s leaves M x mod
// 2^32 in one accumulator set: acc = ((Y_3 * 256 + Y_2) * 256 + Y_1) * 256 + Y_0.
//
// MFMA used: v_mfma_i32_32x32x32_i8.  Lane l holds A[row l%32][k] and
// B[k][col l%32] for the 16 k of its half h = l/32 in the bytes of its
// 16-byte operands, and C[8(i/4) + 4h + i%4][l%32] in accumulator i.  Rows
// are limbs of the product, columns are elements, k runs over the limbs of
// the input.  Any assignment of the 32 k of a block to (h, byte) gives the
// same product as long as A and B share it; byte b of half h carries
// k = kappa(h, b) = 8(b/4) + 4h + b%4, which makes a lane's input limbs the
// same as its output rows (the fold adds lo at exactly the limbs it loaded).
#pragma once
#include "kernels.hpp"

namespace nenya {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

template <class T> constexpr bool kMx = T::K <= 32;  // Horner-accumulated digit GEMMs (mod 2^32)
template <class T> constexpr int kND = T::K / 8;     // balanced base-256 digits per limb
// 64-bit limbs: fixed-multiplier products (folds, scalings) on the matrix
// cores with exact digit classes (k_fold_mx64, k_mulfix_mx64)
template <class T> constexpr bool kMx64 = T::K == 64;
template <class T> constexpr bool kFrag = kMx<T> || kMx64<T>;  // challenges carry MFMA fragments

// Balanced digits of v (mod 2^(8 ND)): byte p of the result is digit p.
template <int ND> __device__ __forceinline__ uint32_t bal_digits(uint32_t v) {
  if constexpr (ND == 1) return v & 0xffu;
  else if constexpr (ND == 2) return ((v + 0x8080u) ^ 0x8080u) & 0xffffu;
  else return (v + 0x80808080u) ^ 0x80808080u;
}

// 4 x 4 byte transpose by byte permutes: w[p] holds byte p of z0, z1, z2,
// z3 (in that order), for p < NB.
template <int NB>
__device__ __forceinline__ void bytes_t4(uint32_t z0, uint32_t z1, uint32_t z2, uint32_t z3, uint32_t w[NB]) {
  // perm(hi, lo, sel): byte i of the result is byte sel_i of hi:lo (lo = 0..3)
  const uint32_t t0 = __builtin_amdgcn_perm(z1, z0, 0x05010400u);  // z0.0 z1.0 z0.1 z1.1
  const uint32_t t1 = __builtin_amdgcn_perm(z3, z2, 0x05010400u);  // z2.0 z3.0 z2.1 z3.1
  w[0] = __builtin_amdgcn_perm(t1, t0, 0x05040100u);
  if constexpr (NB > 1) w[1] = __builtin_amdgcn_perm(t1, t0, 0x07060302u);
  if constexpr (NB > 2) {
    const uint32_t t2 = __builtin_amdgcn_perm(z1, z0, 0x07030602u);  // z0.2 z1.2 z0.3 z1.3
    const uint32_t t3 = __builtin_amdgcn_perm(z3, z2, 0x07030602u);
    w[2] = __builtin_amdgcn_perm(t3, t2, 0x05040100u);
    if constexpr (NB > 3) w[3] = __builtin_amdgcn_perm(t3, t2, 0x07060302u);
  }
}

// 16 values (k = 16 consecutive limbs) -> ND digit planes of 16 int8 each.
template <int ND> __device__ __forceinline__ void planes16(const uint32_t v[16], v4i out[ND]) {
  uint32_t z[16];
#pragma unroll
  for (int b = 0; b < 16; b++) z[b] = bal_digits<ND>(v[b]);
#pragma unroll
  for (int w = 0; w < 4; w++) {
    uint32_t x[ND];
    bytes_t4<ND>(z[4 * w], z[4 * w + 1], z[4 * w + 2], z[4 * w + 3], x);
#pragma unroll
    for (int p = 0; p < ND; p++) out[p][w] = (int)x[p];
  }
}

// Workgroup barrier that orders LDS only: waits for this wave's LDS
// operations (lgkmcnt 0; vmcnt and expcnt left at their maxima), so global
// loads issued before it stay in flight.  __syncthreads() would also wait
// for them (its release fence drains vmcnt).
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Copies one matrix's fragments (ND * 16 KB) into LDS with a 256-thread
// block: every load is issued before the first LDS write, so the copy costs
// one memory latency, not ND * 4 of them.
template <int ND> __device__ __forceinline__ void stage_frags(const v4i* gA, v4i* sA) {
  v4i t[ND * 4];
#pragma unroll
  for (int u = 0; u < ND * 4; u++) t[u] = gA[threadIdx.x + 256 * u];
#pragma unroll
  for (int u = 0; u < ND * 4; u++) sA[threadIdx.x + 256 * u] = t[u];
}

// Fragment index of M_r digit plane p, k-block kb, row-block rb (16 bytes per lane).
__device__ __forceinline__ int frag_at(int p, int kb, int rb) { return (p * 4 + kb) * 4 + rb; }

// v[i] = Y^i v[0] for i < 32 by Y^32 = -Y^3 - Y - alpha (all threads of the
// block call this).  Doubling: level k makes v[k..2k) = Y^k v[0..k), five
// levels with one barrier each.  Multiplying by Y^k (k <= 16) shifts by k
// and sends each of the k top elements h = o[s + 32 - k] (s < k) to
// -alpha h at Y^s, -h at Y^(s+1) and -h at Y^(s+3), all below Y^32.
// Limb l of Y^k o (k <= 29: the wrapped terms o_t Y^(t+k), t + k >= 32, go
// to Y^s, Y^(s+1), Y^(s+3) with s = t + k - 32 < k, all below Y^32).
template <class C> __device__ __forceinline__ C mul_yk_limb(const C* o, int k, int l) {
  const int t = l >> 2, c = l & 3;
  C x = t >= k ? o[l - 4 * k] : (C)0;
  if (t < k) {  // -alpha * h = (h3, h3 - h0, -h1, -h2)
    const C* h = o + 4 * (t + 32 - k);
    x += (c == 0) ? h[3] : (c == 1) ? (C)(h[3] - h[0]) : (c == 2) ? (C)(0 - h[1]) : (C)(0 - h[2]);
  }
  if (t >= 1 && t <= k) x -= o[4 * (t - 1 + 32 - k) + c];
  if (t >= 3 && t <= k + 2) x -= o[4 * (t - 3 + 32 - k) + c];
  return x;
}
template <class C> __device__ __forceinline__ void mulmat_chain(C (*v)[E]) {
  for (int k = 1; k < 32; k <<= 1) {
    for (int id = threadIdx.x; id < k * E; id += blockDim.x) {
      const int i = k + id / E, l = id % E;
      v[i][l] = mul_yk_limb<C>(v[i - k], k, l);
    }
    __syncthreads();
  }
}

// A-fragments of M_r from v[i] = Y^i r: column 4i+c of M_r is X^c v_i with
// X*(a0,a1,a2,a3) = (-a3, a0-a3, a1, a2) (X^4 = -X - 1).  Item (row, kb, h)
// writes the 16 bytes of lane (row%32)+32h in fragment (p, kb, row/32): word
// j holds columns 4i..4i+3 with i = 8kb + 2j + h (k = 32kb + kappa(h, 4j+c)),
// read from vj[j] = v[8kb + 2j + h].
template <class T>
__device__ __forceinline__ void mulmat_frag_item(const CT<T>* const (&vj)[4], int row, int kb, int h, uint32_t* frag) {
  using C = CT<T>;
  constexpr int ND = kND<T>;
  const int t = row >> 2, cp = row & 3;
  uint32_t words[ND][4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const C* vi = vj[j];
    C a[4] = {vi[4 * t], vi[4 * t + 1], vi[4 * t + 2], vi[4 * t + 3]};
    uint32_t z[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      z[c] = bal_digits<ND>((uint32_t)a[cp]);
      C n0 = 0 - a[3], n1 = a[0] - a[3], n2 = a[1], n3 = a[2];
      a[0] = n0; a[1] = n1; a[2] = n2; a[3] = n3;
    }
#pragma unroll
    for (int p = 0; p < ND; p++) {
      uint32_t wd = 0;
#pragma unroll
      for (int c = 0; c < 4; c++) wd |= ((z[c] >> (8 * p)) & 0xffu) << (8 * c);
      words[p][j] = wd;
    }
  }
  const int rb = row >> 5, lane = (row & 31) + 32 * h;
#pragma unroll
  for (int p = 0; p < ND; p++)
    reinterpret_cast<uint4*>(frag)[frag_at(p, kb, rb) * 64 + lane] =
        make_uint4(words[p][0], words[p][1], words[p][2], words[p][3]);
}
template <class T> __device__ __forceinline__ void mulmat_frag(const CT<T> (*v)[E], uint32_t* frag) {
  for (int id = threadIdx.x; id < E * 8; id += blockDim.x) {
    const int row = id & (E - 1), kb = (id >> 7) & 3, h = id >> 9;
    const CT<T>* const vj[4] = {v[8 * kb + h], v[8 * kb + 2 + h], v[8 * kb + 4 + h], v[8 * kb + 6 + h]};
    mulmat_frag_item<T>(vj, row, kb, h, frag);
  }
}

// M_r fragments for r already in device memory; block b handles r + b
// (fragments at frag + b * ND * 4096 words).
template <class T>
__global__ void __launch_bounds__(256) k_mulmat(const LT<T>* r, uint32_t* frag) {
  using C = CT<T>;
  __shared__ C v[32][E];
  r += (uint64_t)blockIdx.x * E;
  frag += (uint64_t)blockIdx.x * kND<T> * 4096;
  if (threadIdx.x < E) v[0][threadIdx.x] = (C)r[threadIdx.x];
  __syncthreads();
  mulmat_chain<C>(v);
  mulmat_frag<T>(v, frag);
}

// Lift one challenge (natural sample src, may be host-pinned; with w.flag
// set, posted by the host into the mailbox) into x and emit the M_r
// fragments of the result.  With one block, that block makes all of them.
// With LIFT_MM_BLOCKS blocks every block lifts the same sample (the lift is
// one block's dependent chain, so the copies run side by side and add no
// latency), block 0 stores it, and block b makes the fragments of k-block
// b/2, half b%2 from its four vectors Y^i r (i = 8kb + h + 2j), reached in
// two single-vector steps (Y^(8kb+h) r, then its Y^2, Y^4 and Y^6 multiples)
// instead of the five levels of the 31-vector chain.  src
// must not alias x when several blocks run (a block may still read src
// after block 0 has stored x).
constexpr int LIFT_MM_BLOCKS = 8;
template <class T>
__global__ void __launch_bounds__(256) k_lift_mulmat(const LT<T>* src, LT<T>* x, const LT<T>* frob_t, int squarings,
                                                     uint32_t* frag, MbWait w) {
  using C = CT<T>;
  __shared__ __attribute__((aligned(16))) C v[32][E];
  __shared__ uint32_t posted[E * sizeof(LT<T>) / 4];
  if (w.flag) {
    mb_wait(w);
    mb_fetch(reinterpret_cast<const uint32_t*>(src), E * sizeof(LT<T>) / 4, posted);
    src = reinterpret_cast<const LT<T>*>(posted);
  }
  teich_lift_block<T>(src, blockIdx.x == 0 ? x : nullptr, frob_t, squarings, v[0]);
  __syncthreads();
  if (gridDim.x == 1) {
    mulmat_chain<C>(v);
    mulmat_frag<T>(v, frag);
    return;
  }
  const int kb = (int)blockIdx.x >> 1, h = (int)blockIdx.x & 1, tid = threadIdx.x, i0 = 8 * kb + h;
  // v[1] = Y^i0 r (i0 <= 25), then v[2 + j] = Y^(2j + 2) v[1]
  if (i0 > 0) {
    if (tid < E) v[1][tid] = mul_yk_limb<C>(v[0], i0, tid);
    __syncthreads();
  }
  const C* u = v[i0 > 0 ? 1 : 0];
  for (int id = tid; id < 3 * E; id += blockDim.x) v[2 + id / E][id % E] = mul_yk_limb<C>(u, 2 * (id / E) + 2, id % E);
  __syncthreads();
  if (tid < E) {
    const C* const vj[4] = {u, v[2], v[3], v[4]};
    mulmat_frag_item<T>(vj, tid, kb, h, frag);
  }
}

// The 16 limbs of half h in k-block kb (kappa order): groups of 4 at 8j + 4h.
template <class T> __device__ __forceinline__ void ld16k(const LT<T>* el, int kb, int h, uint32_t v[16]) {
#pragma unroll
  for (int j = 0; j < 4; j++) {
    CT<T> x[4];
    ld4<T>(el + 32 * kb + 8 * j + 4 * h, x);
#pragma unroll
    for (int c = 0; c < 4; c++) v[4 * j + c] = (uint32_t)x[c];
  }
}

// Up to two groups of tables folded by one launch (same challenge, same m):
// group g has ntab[g] tables, table t at src[g] + t*sstride[g] elements,
// output at dst[g] + t*dstride[g].
// side: a scalar update run by one extra block of the launch, next to the
// folds (1: so = eq_one(sa, sb), 2: so = lin_claim(q = sa, r = sb)).
template <class T> struct FoldJob {
  const LT<T>* src[2];
  LT<T>* dst[2];
  uint64_t ntab[2], sstride[2], dstride[2];
  uint64_t m;
  int side = 0;
  const LT<T>* sa = nullptr;
  const LT<T>* sb = nullptr;
  LT<T>* so = nullptr;
};

// Element e of the job: lo pointer (hi = lo + m) and output pointer.
template <class T>
__device__ __forceinline__ bool fold_locate(const FoldJob<T>& J, uint64_t e, const LT<T>*& lo, LT<T>*& out) {
  // selects rather than J.x[g]: a dynamic index into the argument struct
  // would move it to scratch memory
  const uint64_t n0 = J.m * J.ntab[0];
  const bool g1 = e >= n0;
  const uint64_t x = g1 ? e - n0 : e;
  if (x >= J.m * (g1 ? J.ntab[1] : J.ntab[0])) {
    lo = J.src[0];
    out = nullptr;
    return false;
  }
  const uint64_t tab = x / J.m, i = x % J.m;
  lo = (g1 ? J.src[1] : J.src[0]) + (tab * (g1 ? J.sstride[1] : J.sstride[0]) + i) * E;
  out = (g1 ? J.dst[1] : J.dst[0]) + (tab * (g1 ? J.dstride[1] : J.dstride[0]) + i) * E;
  return true;
}

// One 32-element fold tile: loads lo/hi, forms the digit planes of hi - lo,
// then row-blocks rb0 .. rb0+NRB-1 of lo + M_r (hi - lo).  A points at the
// fragments (LDS or global).  SPLIT: the block's four waves share the tile,
// one row-block each; every wave first loads the fragments it needs into
// registers (global source: one round of latency, not one per MFMA).
template <class T, int NRB, bool SPLIT>
__device__ __forceinline__ void fold_tile(const v4i* A, const FoldJob<T>& J, uint64_t tile, int rb0) {
  constexpr bool PRE = SPLIT;
  using C = CT<T>;
  using L = LT<T>;
  constexpr int ND = kND<T>;
  const int l = threadIdx.x & 63, col = l & 31, h = l >> 5;
  const L* lo;
  L* out;
  const bool ok = fold_locate<T>(J, tile * 32 + col, lo, out);
  const L* hi = lo + J.m * E;
  v4i Apre[PRE ? ND : 1][PRE ? 4 : 1];
  if constexpr (PRE) {
#pragma unroll
    for (int p = 0; p < ND; p++)
#pragma unroll
      for (int kb = 0; kb < 4; kb++) Apre[p][kb] = A[frag_at(p, kb, rb0) * 64 + l];
  }
  uint32_t a[4][16], d[4][16];
  if (ok) {
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      ld16k<T>(lo, kb, h, a[kb]);
      ld16k<T>(hi, kb, h, d[kb]);
    }
  } else {
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
#pragma unroll
      for (int b = 0; b < 16; b++) a[kb][b] = d[kb][b] = 0;
  }
  v4i B[4][ND];
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
#pragma unroll
    for (int b = 0; b < 16; b++) d[kb][b] -= a[kb][b];
    planes16<ND>(d[kb], B[kb]);
  }
  if constexpr (SPLIT) {
    // Each wave has read all 128 limbs of lo and hi but writes only its own
    // row-block of lo.  The fold may be in place (dst == src), so no wave
    // stores before every wave of the block holds its operands.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  constexpr int NU = NRB >= 2 ? 2 : 1;  // row-blocks per accumulator set
#pragma unroll
  for (int rp = 0; rp < NRB; rp += NU) {
    v16i acc[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) acc[u] = (v16i)0;
#pragma unroll
    for (int s = ND - 1; s >= 0; s--) {
      if (s != ND - 1) {
#pragma unroll
        for (int u = 0; u < NU; u++) acc[u] <<= 8;
      }
#pragma unroll
      for (int p = 0; p <= s; p++)
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
#pragma unroll
          for (int u = 0; u < NU; u++) {
            const v4i af = PRE ? Apre[PRE ? p : 0][PRE ? kb : 0] : A[frag_at(p, kb, rb0 + rp + u) * 64 + l];
            acc[u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, B[kb][s - p], acc[u], 0, 0, 0);
          }
    }
    if (ok) {
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const int rb = rb0 + rp + u;
        // lo at this row-block; rb0 may be a runtime value, so select the
        // row statically (a dynamic register index would go to scratch)
        uint32_t arow[16];
#pragma unroll
        for (int b = 0; b < 16; b++) arow[b] = a[0][b];
#pragma unroll
        for (int kb = 1; kb < 4; kb++)
#pragma unroll
          for (int b = 0; b < 16; b++) arow[b] = (rb == kb) ? a[kb][b] : arow[b];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          C x[4];
#pragma unroll
          for (int c = 0; c < 4; c++) x[c] = (C)(arow[4 * j + c] + (uint32_t)acc[u][4 * j + c]);
          st4<T>(out + 32 * rb + 8 * j + 4 * h, x);
        }
      }
    }
  }
}

// Sumcheck fold on the matrix cores (same result as k_fold), split mode
// (few tiles):
//   dst[tab][i] = lo + r (hi - lo),  lo = src[tab][i], hi = src[tab][i+m].
// Block b takes tile b and wave w its row-block w, with its fragments loaded
// from global memory (L2-resident) up front; a barrier between the loads and
// the stores keeps in-place folds exact.  With a side task the last block
// runs it instead of folding.  Many tiles: k_fold_mx_bulk.
template <class T>
__global__ void __launch_bounds__(256) k_fold_mx(const uint32_t* frag, FoldJob<T> J) {
  const v4i* gA = reinterpret_cast<const v4i*>(frag);
  const unsigned nfold = gridDim.x - (J.side ? 1 : 0);
  if (J.side && blockIdx.x == nfold) {
    if (J.side == 1) eq_one_256<T>(J.sa, J.sb, J.so);
    else lin_claim_256<T>(J.sa, J.sb, J.so);
    return;
  }
  fold_tile<T, 1, true>(gA, J, blockIdx.x, threadIdx.x >> 6);
}

// Sumcheck fold on the matrix cores with loads in flight (many tiles):
// persistent blocks of 4 waves, wave w owns output row-block w
// and keeps its fragments of M_r (ND planes x 4 k-blocks) in registers for
// the block's lifetime.  A tile's 32 elements are shared by the block: wave
// w loads k-block w of lo and hi (one 16-limb access pair per lane), forms
// the digit planes of hi - lo for that k-block into one of two LDS buffers,
// and keeps lo's k-block w, which is exactly the lo part of its output rows.
// After one barrier every wave reads the tile's four k-blocks of planes and
// runs its row-block's MFMAs.  The raw loads of the block's next two tiles
// are issued unconditionally (past-the-end elements read element 0 and are
// zeroed) and stay in flight across the barrier (lds_barrier orders LDS
// only).  In-place folds are safe: wave w overwrites only the lo limbs it
// read itself.  With m a multiple of 32 a tile lies in one table (one
// division per tile).  Two blocks per CU.  Tiles of block b: b, b + nfold,
// ...
template <class T>
__global__ void __launch_bounds__(256, 2) k_fold_mx_bulk(const uint32_t* frag, FoldJob<T> J) {
  using C = CT<T>;
  using L = LT<T>;
  constexpr int ND = kND<T>;
  __shared__ v4i sD[2][4 * ND * 64];  // [buffer][k-block][plane][lane]
  const int l64 = threadIdx.x & 63, col = l64 & 31, h = l64 >> 5, w = threadIdx.x >> 6;
  const uint64_t total = J.m * (J.ntab[0] + J.ntab[1]), ntiles = (total + 31) / 32;
  const unsigned nfold = gridDim.x - (J.side ? 1 : 0);
  if (J.side && blockIdx.x == nfold) {
    if (J.side == 1) eq_one_256<T>(J.sa, J.sb, J.so);
    else lin_claim_256<T>(J.sa, J.sb, J.so);
    return;
  }
  const v4i* gA = reinterpret_cast<const v4i*>(frag);
  v4i A[ND][4];
#pragma unroll
  for (int p = 0; p < ND; p++)
#pragma unroll
    for (int kb = 0; kb < 4; kb++) A[p][kb] = gA[frag_at(p, kb, w) * 64 + l64];
  const bool whole = J.m % 32 == 0;
  // element pointers of a tile: lo (hi = lo + m) and out; ok false past the end
  auto locate = [&](uint64_t tile, const L*& lo, L*& out, bool& ok) {
    const uint64_t x0 = tile * 32;
    ok = tile < ntiles && x0 + col < total;
    if (whole && tile < ntiles) {
      const uint64_t n0 = J.m * J.ntab[0];
      const bool g1 = x0 >= n0;
      const uint64_t x = g1 ? x0 - n0 : x0, tab = x / J.m, i = x - tab * J.m + col;
      lo = (g1 ? J.src[1] : J.src[0]) + (tab * (g1 ? J.sstride[1] : J.sstride[0]) + i) * E;
      out = (g1 ? J.dst[1] : J.dst[0]) + (tab * (g1 ? J.dstride[1] : J.dstride[0]) + i) * E;
    } else {
      ok = fold_locate<T>(J, tile < ntiles ? x0 + col : total, lo, out) && ok;
    }
    if (!ok) lo = J.src[0];
  };
  // raw limbs of k-block w of lo and hi
  auto load = [&](uint64_t tile, uint32_t (*R)[16], L*& out, bool& ok) {
    const L* lo;
    locate(tile, lo, out, ok);
    ld16k<T>(lo, w, h, R[0]);
    ld16k<T>(lo + (ok ? J.m : 0) * E, w, h, R[1]);
  };
  auto stage = [&](uint32_t (*R)[16], bool ok, int buf) {
    uint32_t d[16];
#pragma unroll
    for (int b = 0; b < 16; b++) {
      if (!ok) R[0][b] = R[1][b] = 0;
      d[b] = R[1][b] - R[0][b];
    }
    v4i P[ND];
    planes16<ND>(d, P);
#pragma unroll
    for (int p = 0; p < ND; p++) sD[buf][(w * ND + p) * 64 + l64] = P[p];
  };
  // the B operands are read from LDS next to their MFMAs (registers go to
  // the loads in flight instead)
  auto compute = [&](const uint32_t* lo, L* out, bool ok, int buf) {
    v16i acc = (v16i)0;
#pragma unroll
    for (int sd = ND - 1; sd >= 0; sd--) {
      if (sd != ND - 1) acc <<= 8;
#pragma unroll
      for (int p = 0; p <= sd; p++)
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
          acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[p][kb], sD[buf][(kb * ND + sd - p) * 64 + l64], acc, 0, 0, 0);
    }
    if (ok) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        C x[4];
#pragma unroll
        for (int c = 0; c < 4; c++) x[c] = (C)(lo[4 * j + c] + (uint32_t)acc[4 * j + c]);
        st4<T>(out + 32 * w + 8 * j + 4 * h, x);
      }
    }
  };
  // two tiles' raw loads in flight: the register set of tile t is staged,
  // then reloaded with tile t + 2 nfold (a third set spills registers and
  // ran slower)
  constexpr int NP = 2;
  uint32_t R[NP][2][16], lo[16];
  L* o[NP];
  bool k[NP];
  uint64_t tile = blockIdx.x;
  int buf = 0;
#pragma unroll
  for (int u = 0; u < NP; u++) load(tile + u * (uint64_t)nfold, R[u], o[u], k[u]);
  while (true) {
#pragma unroll
    for (int u = 0; u < NP; u++) {
      if (tile >= ntiles) break;
      stage(R[u], k[u], buf);
#pragma unroll
      for (int b = 0; b < 16; b++) lo[b] = R[u][0][b];
      L* oc = o[u];
      const bool kc = k[u];
      load(tile + NP * (uint64_t)nfold, R[u], o[u], k[u]);
      lds_barrier();  // the tile's planes are in LDS
      compute(lo, oc, kc, buf);
      buf ^= 1;
      tile += nfold;
    }
    if (tile >= ntiles) break;
  }
}

// y = M x for one 32-element tile (x[e] at x + e*E, e < n), all four
// row-blocks, written to out + e*ostride*E (A: fragments in LDS).
template <class T>
__device__ __forceinline__ void mul_tile(const v4i* A, const LT<T>* x, uint64_t n, LT<T>* out, uint64_t ostride,
                                         uint64_t tile) {
  using C = CT<T>;
  constexpr int ND = kND<T>;
  const int l = threadIdx.x & 63, col = l & 31, h = l >> 5;
  const uint64_t e = tile * 32 + col;
  const bool ok = e < n;
  v4i B[4][ND];
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    uint32_t d[16];
    if (ok) {
      ld16k<T>(x + e * E, kb, h, d);
    } else {
#pragma unroll
      for (int b = 0; b < 16; b++) d[b] = 0;
    }
    planes16<ND>(d, B[kb]);
  }
#pragma unroll
  for (int rp = 0; rp < 4; rp += 2) {
    v16i acc[2] = {(v16i)0, (v16i)0};
#pragma unroll
    for (int s = ND - 1; s >= 0; s--) {
      if (s != ND - 1) {
        acc[0] <<= 8;
        acc[1] <<= 8;
      }
#pragma unroll
      for (int p = 0; p <= s; p++)
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
#pragma unroll
          for (int u = 0; u < 2; u++)
            acc[u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[frag_at(p, kb, rp + u) * 64 + l], B[kb][s - p], acc[u], 0,
                                                           0, 0);
    }
    if (ok) {
#pragma unroll
      for (int u = 0; u < 2; u++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          C y[4];
#pragma unroll
          for (int c = 0; c < 4; c++) y[c] = (C)(uint32_t)acc[u][4 * j + c];
          st4<T>(out + e * ostride * E + 32 * (rp + u) + 8 * j + 4 * h, y);
        }
    }
  }
}

// out[e*nmat + i] = M_i x[e] (or out[i*n + e] when imajor) for the nmat
// fixed matrices of frag (matrix i at fragment offset i*ND*16); block
// (bx, i) keeps matrix i in LDS.
template <class T>
__global__ void __launch_bounds__(256) k_mulfix_mx(const uint32_t* frag, const LT<T>* x, uint64_t n, LT<T>* out,
                                                   int imajor) {
  constexpr int ND = kND<T>;
  __shared__ v4i sA[ND * 16 * 64];
  const int i = blockIdx.y, nmat = gridDim.y;
  const v4i* gA = reinterpret_cast<const v4i*>(frag) + (size_t)i * ND * 16 * 64;
  stage_frags<ND>(gA, sA);
  __syncthreads();
  const uint64_t ntiles = (n + 31) / 32;
  for (uint64_t tile = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); tile < ntiles; tile += (uint64_t)gridDim.x * 4) {
    asm volatile("" ::: "memory");
    if (imajor) mul_tile<T>(sA, x, n, out + (size_t)i * n * E, 1, tile);
    else mul_tile<T>(sA, x, n, out + (size_t)i * E, (uint64_t)nmat, tile);
  }
}

// build_layer_rows (general.rs:233-271) from the products G[w][p][i] =
// eqs[w][p] * l_beta_i: the packing basis is b_{2i+h} = l_beta_i u_h with
// u_0 = 1 - alpha, u_1 = alpha (R1 scalars acting per coefficient), so an
// entry (row 32p+s of matrix w, value v) adds v * G[w][p][s/2] * u_{s%2} to
// rows[l][q].  One 32-lane group per output (l, q); entries CSR by output,
// code = p << 7 | w << 5 | s.  The group reads up to 32 entries' codes and
// values at once (lane i: entry k0 + i), then gathers their G elements four
// at a time: the four loads are issued before any is used (indices past the
// output's last entry repeat that entry with value 0, so no load is
// conditional), keeping 2 KiB per group in flight instead of one element
// behind a dependent code read.
template <class T>
__global__ void __launch_bounds__(256) k_rows_gather(const uint64_t* ptr, const uint32_t* code, const LT<T>* val,
                                                     const LT<T>* G, uint64_t nb, LT<T>* rows, uint64_t nout) {
  using C = CT<T>;
  constexpr int U = 4;
  const int t = threadIdx.x & 31;
  for (uint64_t o = (uint64_t)blockIdx.x * 8 + (threadIdx.x >> 5); o < nout; o += (uint64_t)gridDim.x * 8) {
    C acc[4] = {0, 0, 0, 0};
    const uint64_t k0 = ptr[o], k1 = ptr[o + 1];
    for (uint64_t kb = k0; kb < k1; kb += 32) {
      const uint32_t ne = k1 - kb < 32 ? (uint32_t)(k1 - kb) : 32u;
      const uint64_t ke = kb + ((uint32_t)t < ne ? (uint32_t)t : ne - 1);
      const uint32_t cdl = code[ke];
      const C vl = (uint32_t)t < ne ? (C)val[ke] : (C)0;
      for (uint32_t j = 0; j < ne; j += U) {
        C g[U][4];
        uint32_t sj[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          // lanes j + u >= ne hold the last entry's code (value 0 below)
          const uint32_t cd = __shfl(cdl, (int)(j + u < ne ? j + u : ne - 1), 32);
          const uint64_t p = cd >> 7, w = (cd >> 5) & 3;
          sj[u] = cd & 31;
          ld4<T>(G + ((w * nb + p) * 16 + (sj[u] >> 1)) * E + 4 * t, g[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          // j + u <= 31: lanes at or past ne carry value 0
          const C v = __shfl(vl, (int)(j + u), 32);
          C f[4];
          if (sj[u] & 1) {  // alpha g
            f[0] = 0 - g[u][3]; f[1] = g[u][0] - g[u][3]; f[2] = g[u][1]; f[3] = g[u][2];
          } else {          // (1 - alpha) g
            f[0] = g[u][0] + g[u][3]; f[1] = g[u][1] - g[u][0] + g[u][3]; f[2] = g[u][2] - g[u][1];
            f[3] = g[u][3] - g[u][2];
          }
#pragma unroll
          for (int c = 0; c < 4; c++) acc[c] += v * f[c];
        }
      }
    }
    st4<T>(rows + o * E + 4 * t, acc);
  }
}

// ---------------------------------------------------------------------------
// Layered matvec on the matrix cores (k <= 32).
//
// layered_matvec (general.rs:211-230) sums D_hat(l,p,q) * rho_l[q] over the
// diagonal cells of block row p.  pack_r2 is Z/2^k-linear with basis
// b_s = pack_r2(e_s), so with D_hat = sum_s D[s] b_s and
// rho_l[q] = sum_s' w[32q + (s'+l) mod 32] b_s' (l = (c - r) mod 32 of the
// entry, D[s] = A[32p+s][c]):
//   sum_cells D_hat rho = sum_{s,s'} C_p[s][s'] (b_s b_s'),
//   C_p[s][s'] = sum_{entries (32p+s, c, v)} v * w[32(c/32) + (s'+c-s) mod 32].
// k_cmat forms C_p from the row CSR of the matrix (stored k-block-major:
// row s of C_p at cmat + (s * nb + p) * 32); k_cgemm multiplies by the fixed
// 128 x 1024 matrix of the b_s b_s' (fragments built on the host).
// ---------------------------------------------------------------------------
// A 32-lane group per block row p: the entries of its 32 rows are contiguous
// in the CSR, so the group reads them 32 at a time (lane i: entry k0 + i,
// its column, value and row s found among the 33 row pointers), then every
// lane s' adds each entry's term into C_p[s][s'] (LDS, column s' owned by
// lane s'); the w reads of a chunk are independent, so they overlap.
template <class T>
__global__ void __launch_bounds__(256) k_cmat(const uint32_t* ptr, const uint32_t* col, const LT<T>* val,
                                              const LT<T>* w, LT<T>* cmat, uint64_t nb) {
  using C = CT<T>;
  __shared__ C cm[8][32][32];
  const uint32_t lane = threadIdx.x & 31, g = threadIdx.x >> 5;
  for (uint64_t pb = (uint64_t)blockIdx.x * 8; pb < nb; pb += (uint64_t)gridDim.x * 8) {
    const uint64_t p = pb + g;
    const bool ok = p < nb;
    // row pointers 32p .. 32p+32 (lane i holds ptr[32p + i]; lane 0 also the end)
    const uint32_t rp = ok ? ptr[32 * p + lane] : 0, rend = ok ? ptr[32 * p + 32] : 0;
    const uint32_t k0 = __shfl(rp, 0, 32);
#pragma unroll
    for (int s = 0; s < 32; s++) cm[g][s][lane] = 0;
    for (uint32_t kb = k0; kb < rend; kb += 32) {
      const uint32_t k = kb + lane;
      const bool ek = k < rend;
      const uint32_t c = ek ? col[k] : 0;
      const C v = ek ? (C)val[k] : (C)0;
      // row of entry k: the number of row starts rp[1..31] at or below k
      uint32_t s = 0;
#pragma unroll
      for (int i = 1; i < 32; i++) s += (__shfl(rp, i, 32) <= k) ? 1u : 0u;
      const uint32_t ne = rend - kb < 32 ? rend - kb : 32;
      C term[32];
#pragma unroll
      for (int j = 0; j < 32; j++) {
        const uint32_t cj = __shfl(c, j, 32), sj = __shfl(s, j, 32);
        term[j] = (uint32_t)j < ne ? (C)w[(cj & ~31u) + ((lane + cj - sj) & 31u)] : (C)0;
      }
#pragma unroll
      for (int j = 0; j < 32; j++) {
        const C vj = __shfl(v, j, 32);
        const uint32_t sj = __shfl(s, j, 32);
        if ((uint32_t)j < ne) cm[g][sj][lane] += vj * term[j];
      }
    }
    if (ok)
#pragma unroll
      for (int s = 0; s < 32; s++) cmat[((uint64_t)s * nb + p) * 32 + lane] = (LT<T>)cm[g][s][lane];
  }
}

// MIMD rows: C_r[s][s'] = sum_{cells g of row r} val[g][s] * w[col[g]][s']
// (cells of the union support with their 32 instance values, w the witness
// slots [m][32]); one 32-lane group per row, lane s'.  With
// Ahat[r,c] = sum_s a_s b_s and w_phi[c] = sum_s' w_s' b_s', the row's
// sum_c Ahat[r,c] w_phi[c] is sum_{s,s'} C_r[s][s'] b_s b_s' (k_cgemm; C_r
// stored k-block-major as in k_cmat).
template <class T>
__global__ void __launch_bounds__(256) k_cmat_mimd(const uint64_t* ptr, const uint32_t* col, const LT<T>* val,
                                                   const LT<T>* w, LT<T>* cmat, uint64_t n) {
  using C = CT<T>;
  const uint32_t lane = threadIdx.x & 31;
  for (uint64_t r = (uint64_t)blockIdx.x * 8 + (threadIdx.x >> 5); r < n; r += (uint64_t)gridDim.x * 8) {
    C acc[32];
#pragma unroll
    for (int s = 0; s < 32; s++) acc[s] = 0;
    for (uint64_t g = ptr[r]; g < ptr[r + 1]; g++) {
      const C wv = (C)w[(uint64_t)col[g] * 32 + lane];
      const C vl = (C)val[g * 32 + lane];
#pragma unroll
      for (int s = 0; s < 32; s++) acc[s] += __shfl(vl, s, 32) * wv;
    }
#pragma unroll
    for (int s = 0; s < 32; s++) cmat[((uint64_t)s * n + r) * 32 + lane] = (LT<T>)acc[s];
  }
}

// out[p] = sum_k B2[.][k] C_p[k] over k = 32s + s' (1024).  Block: 4 waves on
// the same 32 block rows, wave w computing output rows 32w..32w+31; one
// accumulator per digit class (K streams, so no Horner), combined at the end.
template <class T>
__global__ void __launch_bounds__(256) k_cgemm(const uint32_t* frag, const LT<T>* cmat, LT<T>* out, uint64_t nb) {
  using C = CT<T>;
  constexpr int ND = kND<T>;
  const v4i* A = reinterpret_cast<const v4i*>(frag);
  const int l = threadIdx.x & 63, col = l & 31, h = l >> 5, rb = threadIdx.x >> 6;
  const uint64_t ntiles = (nb + 31) / 32;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t p = tile * 32 + col;
    const bool ok = p < nb;
    // k-block kb of C_p at cmat + (kb * nb + p) * 32: a wave's loads are contiguous
    const LT<T>* cp = cmat + (ok ? p : 0) * 32;
    v16i acc[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) acc[d] = (v16i)0;
    // k-block kb+1's operands are loaded during kb's matrix work
    uint32_t v[16];
    v4i a[ND];
    auto load_kb = [&](int kb) {
      if (ok && kb < 32) {
        ld16k<T>(cp + (uint64_t)kb * nb * 32, 0, h, v);
      } else {
#pragma unroll
        for (int b = 0; b < 16; b++) v[b] = 0;
      }
#pragma unroll
      for (int pd = 0; pd < ND; pd++) a[pd] = A[((pd * 32 + (kb & 31)) * 4 + rb) * 64 + l];
    };
    load_kb(0);
    for (int kb = 0; kb < 32; kb++) {
      v4i B[ND], ac[ND];
      planes16<ND>(v, B);
#pragma unroll
      for (int pd = 0; pd < ND; pd++) ac[pd] = a[pd];
      load_kb(kb + 1);
#pragma unroll
      for (int pd = 0; pd < ND; pd++)
#pragma unroll
        for (int q = 0; pd + q < ND; q++)
          acc[pd + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ac[pd], B[q], acc[pd + q], 0, 0, 0);
    }
    if (ok) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        C x[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          uint32_t y = 0;
#pragma unroll
          for (int d = 0; d < ND; d++) y += (uint32_t)acc[d][4 * j + c] << (8 * d);
          x[c] = (C)y;
        }
        st4<T>(out + p * E + 32 * rb + 8 * j + 4 * h, x);
      }
    }
  }
}

// Split-K form of k_cgemm: block (bx, z) keeps the fragments of k-blocks
// [8z, 8z + 8) (all four row-blocks, ND planes, ND * 32 KiB) in LDS for its
// lifetime, so no fragment load waits on L2 inside the loop; each wave takes
// 32-row tiles with all four row-blocks (a k-block's digit planes feed four
// row-blocks), Horner over the digit classes per k-block (i32 wrap =
// mod 2^k), and writes the slice's partial sums to part + z * nb * E
// (k_sum_split adds the four slices).  One block per CU (LDS), one wave per
// SIMD: the C loads run four k-blocks ahead to cover HBM latency.
template <class T>
__global__ void __launch_bounds__(256) k_cgemm_sk(const uint32_t* frag, const LT<T>* cmat, LT<T>* part, uint64_t nb) {
  using C = CT<T>;
  constexpr int ND = kND<T>, KS = 8;
  __shared__ v4i sA[ND * KS * 4 * 64];  // [pd][kbl][rb][lane]
  const v4i* A = reinterpret_cast<const v4i*>(frag);
  const int z = blockIdx.y;
  {
    v4i t[ND * KS * 4 * 64 / 256];
#pragma unroll
    for (int u = 0; u < ND * KS * 4 * 64 / 256; u++) {
      const int i = threadIdx.x + 256 * u;
      const int l = i & 63, rb = (i >> 6) & 3, kbl = (i >> 8) % KS, pd = (i >> 8) / KS;
      t[u] = A[((pd * 32 + z * KS + kbl) * 4 + rb) * 64 + l];
    }
#pragma unroll
    for (int u = 0; u < ND * KS * 4 * 64 / 256; u++) sA[threadIdx.x + 256 * u] = t[u];
  }
  __syncthreads();
  const int l = threadIdx.x & 63, col = l & 31, h = l >> 5;
  const uint64_t ntiles = (nb + 31) / 32, stride = (uint64_t)gridDim.x * 4;
  // k-blocks of C four steps ahead, across tile boundaries: at step
  // kbl = 4 half + q the wave computes ring slot q and refills it with
  // k-block kbl + 4 (of the next tile in the second half)
  uint32_t ring[4][16];
  auto load = [&](uint64_t tile, int kbl, uint32_t (&v)[16]) {
    const uint64_t p = tile * 32 + col;
    if (tile < ntiles && p < nb) {
      ld16k<T>(cmat + p * 32 + (uint64_t)(z * KS + kbl) * nb * 32, 0, h, v);
    } else {
#pragma unroll
      for (int b = 0; b < 16; b++) v[b] = 0;
    }
  };
  uint64_t tile = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
#pragma unroll
  for (int q = 0; q < 4; q++) load(tile, q, ring[q]);
  for (; tile < ntiles; tile += stride) {
    const uint64_t p = tile * 32 + col;
    const bool ok = p < nb;
    v16i acc[4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) acc[rb] = (v16i)0;
#pragma unroll 1
    for (int half = 0; half < 2; half++) {
#pragma unroll
     for (int q = 0; q < 4; q++) {
      const int kbl = 4 * half + q;
      v4i B[ND];
      planes16<ND>(ring[q], B);
      if (half == 0) load(tile, q + 4, ring[q]);
      else load(tile + stride, q, ring[q]);
#pragma unroll
      for (int rb = 0; rb < 4; rb += 2) {
        // keeps each row-block pair's fragment reads in place (hoisted, the
        // unrolled loop's reads would pin hundreds of registers); the pair's
        // two MFMA chains interleave, so neither waits on its own results
        asm volatile("" ::: "memory");
        v16i t0 = (v16i)0, t1 = (v16i)0;
#pragma unroll
        for (int sd = ND - 1; sd >= 0; sd--) {
          if (sd != ND - 1) {
            t0 <<= 8;
            t1 <<= 8;
          }
#pragma unroll
          for (int pd = 0; pd <= sd; pd++) {
            t0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(sA[((pd * KS + kbl) * 4 + rb) * 64 + l], B[sd - pd], t0, 0, 0, 0);
            t1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(sA[((pd * KS + kbl) * 4 + rb + 1) * 64 + l], B[sd - pd], t1, 0, 0,
                                                       0);
          }
        }
        acc[rb] += t0;
        acc[rb + 1] += t1;
      }
     }
    }
    if (ok) {
      LT<T>* o = part + ((uint64_t)z * nb + p) * E;
#pragma unroll
      for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          C x[4];
#pragma unroll
          for (int c = 0; c < 4; c++) x[c] = (C)(uint32_t)acc[rb][4 * j + c];
          st4<T>(o + 32 * rb + 8 * j + 4 * h, x);
        }
    }
  }
}

// ---------------------------------------------------------------------------
// Fixed-multiplier products with 64-bit limbs on the matrix cores.
//
// The Horner accumulation above relies on the i32 accumulators wrapping mod
// 2^32.  With k = 64 every digit class
//   Y_s = sum_{p+q=s} M_q x_p     (s < 8, balanced base-256 digits)
// is kept exact instead: |M_q x_p| <= 128 * 2^14 = 2^21 per 128-term row and
// at most 8 pairs per class, so |Y_s| <= 2^24 fits an i32 accumulator, and
// M x = sum_s 256^s Y_s (mod 2^64) is formed in 64-bit Horner on the vector
// ALUs.  36 digit-pair MFMAs per k-block and row-block (10 at k = 32).  The
// fragments of M (8 digit planes, 128 KB) stay in LDS for the block's
// lifetime; each fragment is read once per row-block and applied to the
// 8 - q classes it feeds (eight class accumulators, 128 AGPRs).
__device__ __forceinline__ uint64_t bal_digits64(uint64_t v) {
  return (v + 0x8080808080808080ull) ^ 0x8080808080808080ull;
}
// 16 limbs -> 8 digit planes of 16 int8 each
__device__ __forceinline__ void planes16_64(const uint64_t v[16], v4i out[8]) {
  uint64_t z[16];
#pragma unroll
  for (int b = 0; b < 16; b++) z[b] = bal_digits64(v[b]);
#pragma unroll
  for (int w = 0; w < 4; w++) {
    uint32_t lo[4], hi[4];
    bytes_t4<4>((uint32_t)z[4 * w], (uint32_t)z[4 * w + 1], (uint32_t)z[4 * w + 2], (uint32_t)z[4 * w + 3], lo);
    bytes_t4<4>((uint32_t)(z[4 * w] >> 32), (uint32_t)(z[4 * w + 1] >> 32), (uint32_t)(z[4 * w + 2] >> 32),
                (uint32_t)(z[4 * w + 3] >> 32), hi);
#pragma unroll
    for (int p = 0; p < 4; p++) {
      out[p][w] = (int)lo[p];
      out[4 + p][w] = (int)hi[p];
    }
  }
}
// The 16 limbs of half h in k-block kb (kappa order), 64-bit.
__device__ __forceinline__ void ld16k64(const uint64_t* el, int kb, int h, uint64_t v[16]) {
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const ulonglong2* q = reinterpret_cast<const ulonglong2*>(el + 32 * kb + 8 * j + 4 * h);
    const ulonglong2 a = q[0], b = q[1];
    v[4 * j] = a.x;
    v[4 * j + 1] = a.y;
    v[4 * j + 2] = b.x;
    v[4 * j + 3] = b.y;
  }
}
// Fragments of M_r (8 digit planes) from v[i] = Y^i r, as mulmat_frag;
// items [id0, id1) of the E * 8 (row, k-block, half) items.
// One fragment item (row, kb, h) of mulmat_frag64, from vj[j] = v[8kb + 2j + h].
__device__ __forceinline__ void mulmat_frag64_item(const uint64_t* const (&vj)[4], int row, int kb, int h,
                                                   uint32_t* frag) {
  const int t = row >> 2, cp = row & 3;
  uint32_t words[8][4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const uint64_t* vi = vj[j];
    uint64_t a[4] = {vi[4 * t], vi[4 * t + 1], vi[4 * t + 2], vi[4 * t + 3]};
    uint64_t z[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      z[c] = bal_digits64(a[cp]);
      const uint64_t n0 = 0 - a[3], n1 = a[0] - a[3], n2 = a[1], n3 = a[2];
      a[0] = n0; a[1] = n1; a[2] = n2; a[3] = n3;
    }
#pragma unroll
    for (int p = 0; p < 8; p++) {
      uint32_t wd = 0;
#pragma unroll
      for (int c = 0; c < 4; c++) wd |= (uint32_t)((z[c] >> (8 * p)) & 0xffu) << (8 * c);
      words[p][j] = wd;
    }
  }
  const int rb = row >> 5, lane = (row & 31) + 32 * h;
#pragma unroll
  for (int p = 0; p < 8; p++)
    reinterpret_cast<uint4*>(frag)[frag_at(p, kb, rb) * 64 + lane] =
        make_uint4(words[p][0], words[p][1], words[p][2], words[p][3]);
}
__device__ __forceinline__ void mulmat_frag64(const uint64_t (*v)[E], uint32_t* frag, int id0, int id1) {
  for (int id = id0 + threadIdx.x; id < id1; id += blockDim.x) {
    const int row = id & (E - 1), kb = (id >> 7) & 3, h = id >> 9;
    const uint64_t* const vj[4] = {v[8 * kb + h], v[8 * kb + 2 + h], v[8 * kb + 4 + h], v[8 * kb + 6 + h]};
    mulmat_frag64_item(vj, row, kb, h, frag);
  }
}
// M_r fragments for r in device memory; blocks (b, y) handle r + b
// (fragments at frag + b * 8 * 4096 words), block y the y-th share of the
// fragment items (each block forms the Y^i r chain itself: it is cheap, the
// digit planes are not).
__global__ void __launch_bounds__(256) k_mulmat64(const uint64_t* r, uint32_t* frag) {
  __shared__ uint64_t v[32][E];
  r += (uint64_t)blockIdx.x * E;
  frag += (uint64_t)blockIdx.x * 8 * 4096;
  if (threadIdx.x < E) v[0][threadIdx.x] = r[threadIdx.x];
  __syncthreads();
  mulmat_chain<uint64_t>(v);
  const int per = E * 8 / gridDim.y;
  mulmat_frag64(v, frag, blockIdx.y * per, (blockIdx.y + 1) * per);
}

// The 64-bit counterpart of k_lift_mulmat with LIFT_MM_BLOCKS blocks: every
// block lifts the sample (posted by the host when w.flag is set), block 0
// stores it in x, and block b makes the fragments of k-block b/2, half b%2
// (8 digit planes) from Y^(8kb+h) r and its Y^2, Y^4, Y^6 multiples.  One
// launch instead of the lift and k_mulmat64; src must not alias x.
__global__ void __launch_bounds__(256) k_lift_mulmat64(const uint64_t* src, uint64_t* x, const uint64_t* frob_t,
                                                       int squarings, uint32_t* frag, MbWait w) {
  __shared__ __attribute__((aligned(16))) uint64_t v[5][E];
  __shared__ uint32_t posted[E * 2];
  if (w.flag) {
    mb_wait(w);
    mb_fetch(reinterpret_cast<const uint32_t*>(src), E * 2, posted);
    src = reinterpret_cast<const uint64_t*>(posted);
  }
  teich_lift_block<W<64>>(src, blockIdx.x == 0 ? x : nullptr, frob_t, squarings, v[0]);
  __syncthreads();
  const int kb = (int)blockIdx.x >> 1, h = (int)blockIdx.x & 1, tid = threadIdx.x, i0 = 8 * kb + h;
  if (i0 > 0) {
    if (tid < E) v[1][tid] = mul_yk_limb<uint64_t>(v[0], i0, tid);
    __syncthreads();
  }
  const uint64_t* u = v[i0 > 0 ? 1 : 0];
  for (int id = tid; id < 3 * E; id += blockDim.x)
    v[2 + id / E][id % E] = mul_yk_limb<uint64_t>(u, 2 * (id / E) + 2, id % E);
  __syncthreads();
  if (tid < E) {
    const uint64_t* const vj[4] = {u, v[2], v[3], v[4]};
    mulmat_frag64_item(vj, tid, kb, h, frag);
  }
}
// Row-block rb of M x for one 32-element tile: R[i] (lane (col, h), rows
// 8(i/4) + 4h + i%4 of the row-block) from the digit planes B of x.
__device__ __forceinline__ void mul_rb64(const v4i* A, const v4i B[4][8], int rb, uint64_t R[16]) {
  const int l = threadIdx.x & 63;
  v16i acc[8];
#pragma unroll
  for (int s = 0; s < 8; s++) acc[s] = (v16i)0;
#pragma unroll
  for (int q = 0; q < 8; q++)
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      const v4i af = A[frag_at(q, kb, rb) * 64 + l];
#pragma unroll
      for (int p = 0; p + q < 8; p++) acc[q + p] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, B[kb][p], acc[q + p], 0, 0, 0);
    }
#pragma unroll
  for (int i = 0; i < 16; i++) {
    uint64_t x = (uint64_t)(int64_t)acc[7][i];
#pragma unroll
    for (int s = 6; s >= 0; s--) x = (x << 8) + (uint64_t)(int64_t)acc[s][i];
    R[i] = x;
  }
}
// Digit planes of a tile's 32 elements (x[e] at x + e * E, or their
// differences hi - lo when lo is given); zero for elements past the end.
__device__ __forceinline__ void tile_planes64(const uint64_t* x, const uint64_t* lo, bool ok, v4i B[4][8]) {
  const int h = (threadIdx.x & 63) >> 5;
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    uint64_t d[16];
    if (ok) {
      ld16k64(x, kb, h, d);
      if (lo) {
        uint64_t a[16];
        ld16k64(lo, kb, h, a);
#pragma unroll
        for (int b = 0; b < 16; b++) d[b] -= a[b];
      }
    } else {
#pragma unroll
      for (int b = 0; b < 16; b++) d[b] = 0;
    }
    planes16_64(d, B[kb]);
  }
}

// Sumcheck fold with 64-bit limbs on the matrix cores (same result as
// k_fold): dst[tab][i] = lo + r (hi - lo).  Persistent blocks of 4 waves,
// a wave per 32-element tile; the fold may be in place (a wave reloads the
// lo limbs of a row-block before it stores that row-block, and no other
// wave touches its tile).  With a side task the last block runs it.
template <class T>
__global__ void __launch_bounds__(256) k_fold_mx64(const uint32_t* frag, FoldJob<T> J) {
  static_assert(T::K == 64, "64-bit limbs");
  __shared__ v4i sA[8 * 16 * 64];
  const uint64_t total = J.m * (J.ntab[0] + J.ntab[1]), ntiles = (total + 31) / 32;
  const unsigned nfold = gridDim.x - (J.side ? 1 : 0);
  if (J.side && blockIdx.x == nfold) {
    if (J.side == 1) eq_one_256<T>(J.sa, J.sb, J.so);
    else lin_claim_256<T>(J.sa, J.sb, J.so);
    return;
  }
  stage_frags<8>(reinterpret_cast<const v4i*>(frag), sA);
  __syncthreads();
  const int l = threadIdx.x & 63, col = l & 31, h = l >> 5;
  for (uint64_t tile = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); tile < ntiles; tile += (uint64_t)nfold * 4) {
    asm volatile("" ::: "memory");
    const uint64_t* lo;
    uint64_t* out;
    const bool ok = fold_locate<T>(J, tile * 32 + col, lo, out);
    v4i B[4][8];
    tile_planes64(lo + J.m * E, lo, ok, B);
#pragma unroll 1
    for (int rb = 0; rb < 4; rb++) {
      uint64_t R[16];
      mul_rb64(sA, B, rb, R);
      if (ok) {
        uint64_t a[16];
        ld16k64(lo, rb, h, a);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          ulonglong2* o = reinterpret_cast<ulonglong2*>(out + 32 * rb + 8 * j + 4 * h);
          o[0] = make_ulonglong2(a[4 * j] + R[4 * j], a[4 * j + 1] + R[4 * j + 1]);
          o[1] = make_ulonglong2(a[4 * j + 2] + R[4 * j + 2], a[4 * j + 3] + R[4 * j + 3]);
        }
      }
    }
  }
}

// out[e * nmat + i] = M_i x[e] (or out[i * n + e] when imajor) for the nmat
// fixed matrices of frag (matrix i at fragment offset i * 8 * 4096 words);
// block (bx, i) keeps M_i in LDS.
__global__ void __launch_bounds__(256) k_mulfix_mx64(const uint32_t* frag, const uint64_t* x, uint64_t n,
                                                     uint64_t* out, int imajor) {
  __shared__ v4i sA[8 * 16 * 64];
  const int i = blockIdx.y, nmat = gridDim.y;
  stage_frags<8>(reinterpret_cast<const v4i*>(frag) + (size_t)i * 8 * 16 * 64, sA);
  __syncthreads();
  const int l = threadIdx.x & 63, col = l & 31, h = l >> 5;
  const uint64_t ntiles = (n + 31) / 32;
  for (uint64_t tile = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); tile < ntiles; tile += (uint64_t)gridDim.x * 4) {
    asm volatile("" ::: "memory");
    const uint64_t e = tile * 32 + col;
    const bool ok = e < n;
    v4i B[4][8];
    tile_planes64(x + (ok ? e : 0) * E, nullptr, ok, B);
#pragma unroll 1
    for (int rb = 0; rb < 4; rb++) {
      uint64_t R[16];
      mul_rb64(sA, B, rb, R);
      if (ok) {
        uint64_t* o = out + (imajor ? (uint64_t)i * n + e : e * nmat + i) * E + 32 * rb + 4 * h;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          ulonglong2* q = reinterpret_cast<ulonglong2*>(o + 8 * j);
          q[0] = make_ulonglong2(R[4 * j], R[4 * j + 1]);
          q[1] = make_ulonglong2(R[4 * j + 2], R[4 * j + 3]);
        }
      }
    }
  }
}

// A-fragments (8 digit planes) of the basis products B2[k][row] = (b_s
// b_s')[row], k = 32 s + s' (1024 columns), with 64-bit limbs: thread
// (kb, rb, lane, word) forms the digits of its four k values, as mx_frags
// lays them out ([plane][kb][rb][lane][16 bytes], byte c of word w at
// k = 32 kb + 8 w + 4 h + c, h = lane / 32).
__global__ void __launch_bounds__(256) k_b2frag64(const uint64_t* b2, uint32_t* frag) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;  // 32 * 4 * 64 * 4 items
  if (id >= 32 * 4 * 64 * 4) return;
  const int wd = id & 3, lane = (id >> 2) & 63, rb = (id >> 8) & 3, kb = id >> 10;
  const int h = lane >> 5, row = 32 * rb + (lane & 31);
  uint64_t z[4];
#pragma unroll
  for (int c = 0; c < 4; c++) z[c] = bal_digits64(b2[(size_t)(32 * kb + 8 * wd + 4 * h + c) * E + row]);
#pragma unroll
  for (int p = 0; p < 8; p++) {
    uint32_t word = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) word |= (uint32_t)((z[c] >> (8 * p)) & 0xffu) << (8 * c);
    frag[((((size_t)p * 32 + kb) * 4 + rb) * 64 + lane) * 4 + wd] = word;
  }
}

// The layered matvec's GEMM with 64-bit limbs on the matrix cores:
//   out[p] = sum_{k < 1024} B2[k] C_p[k]   (k = 32 s + s', C_p as k_cmat
// stores it, k-block s at cmat + (s nb + p) 32).  Every digit class Y_s
// (s < 8) is exact in i32 over the whole sum: |Y_s| <= 8 * 1024 * 2^14 =
// 2^27.  Block: 4 waves on 32 block rows, wave w output rows 32w..32w+31
// (one row-block); per k-block 36 digit-pair MFMAs, C and the fragments of
// k-block kb+1 loaded during kb's matrix work; then a 64-bit Horner over
// the classes.  Two blocks per CU: a few registers spill, but the second wave
// per SIMD hides more load latency than the spills cost.
__global__ void __launch_bounds__(256, 2) k_cgemm64(const uint32_t* frag, const uint64_t* cmat, uint64_t* out,
                                                 uint64_t nb) {
  const v4i* A = reinterpret_cast<const v4i*>(frag);
  const int l = threadIdx.x & 63, col = l & 31, h = l >> 5, rb = threadIdx.x >> 6;
  const uint64_t ntiles = (nb + 31) / 32;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t p = tile * 32 + col;
    const bool ok = p < nb;
    const uint64_t* cp = cmat + (ok ? p : 0) * 32;
    v16i acc[8];
#pragma unroll
    for (int s = 0; s < 8; s++) acc[s] = (v16i)0;
    uint64_t v[16];
    v4i a[8];
    auto load_kb = [&](int kb) {
      if (ok && kb < 32) {
        ld16k64(cp + (uint64_t)kb * nb * 32, 0, h, v);
      } else {
#pragma unroll
        for (int b = 0; b < 16; b++) v[b] = 0;
      }
      const int kk = kb & 31;
#pragma unroll
      for (int q = 0; q < 8; q++) a[q] = A[((q * 32 + kk) * 4 + rb) * 64 + l];
    };
    load_kb(0);
    for (int kb = 0; kb < 32; kb++) {
      v4i B[8], ac[8];
      planes16_64(v, B);
#pragma unroll
      for (int q = 0; q < 8; q++) ac[q] = a[q];
      load_kb(kb + 1);
#pragma unroll
      for (int q = 0; q < 8; q++)
#pragma unroll
        for (int pd = 0; pd + q < 8; pd++)
          acc[q + pd] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ac[q], B[pd], acc[q + pd], 0, 0, 0);
    }
    if (ok) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        uint64_t x[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          uint64_t y = (uint64_t)(int64_t)acc[7][4 * j + c];
#pragma unroll
          for (int s = 6; s >= 0; s--) y = (y << 8) + (uint64_t)(int64_t)acc[s][4 * j + c];
          x[c] = y;
        }
        ulonglong2* o = reinterpret_cast<ulonglong2*>(out + p * E + 32 * rb + 8 * j + 4 * h);
        o[0] = make_ulonglong2(x[0], x[1]);
        o[1] = make_ulonglong2(x[2], x[3]);
      }
    }
  }
}

// A-fragments (8 digit planes) of pack_r2 o rot_L for L < 32 with 64-bit
// limbs: P_L[row][k] = b_{(k - L) mod 32}[row] (one k-block; layout as
// k_pack_mx reads it: [L][plane][rb][lane][16 bytes]).
__global__ void __launch_bounds__(256) k_packfrag64(const uint64_t* basis, uint32_t* frag) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;  // 32 * 4 * 64 * 4 items
  if (id >= 32 * 4 * 64 * 4) return;
  const int wd = id & 3, lane = (id >> 2) & 63, rb = (id >> 8) & 3, L = id >> 10;
  const int h = lane >> 5, row = 32 * rb + (lane & 31);
  uint64_t z[4];
#pragma unroll
  for (int c = 0; c < 4; c++) z[c] = bal_digits64(basis[(size_t)((8 * wd + 4 * h + c - L) & 31) * E + row]);
#pragma unroll
  for (int p = 0; p < 8; p++) {
    uint32_t word = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) word |= (uint32_t)((z[c] >> (8 * p)) & 0xffu) << (8 * c);
    frag[((((size_t)L * 8 + p) * 4 + rb) * 64 + lane) * 4 + wd] = word;
  }
}

// pack_r2 on the matrix cores with 64-bit limbs (as k_pack_mx):
// out[r * n + e] = P_L slots[e], L = (r + rot0) mod 32; the 32-long sum of
// one k-block keeps every digit class exact in i32 (|Y_s| <= 8 * 32 * 2^14),
// 36 MFMAs per row-block, then a 64-bit Horner.  Block: 32 elements, wave w
// output rows 32w..32w+31; the slots are read once for all nrot rotations.
__global__ void __launch_bounds__(256) k_pack_mx64(const uint32_t* frag, const uint64_t* slots, uint64_t n, int nrot,
                                                   int rot0, uint64_t* out) {
  const int l64 = threadIdx.x & 63, col = l64 & 31, h = l64 >> 5, rb = threadIdx.x >> 6;
  const v4i* A = reinterpret_cast<const v4i*>(frag);
  const uint64_t ntiles = (n + 31) / 32;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t e = tile * 32 + col;
    const bool ok = e < n;
    uint64_t v[16];
    if (ok) {
      ld16k64(slots + e * 32, 0, h, v);
    } else {
#pragma unroll
      for (int b = 0; b < 16; b++) v[b] = 0;
    }
    v4i B[8];
    planes16_64(v, B);
    for (int r = 0; r < nrot; r++) {
      const int L = (r + rot0) & 31;
      v16i acc[8];
#pragma unroll
      for (int d = 0; d < 8; d++) acc[d] = (v16i)0;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const v4i a = A[(((size_t)L * 8 + q) * 4 + rb) * 64 + l64];
#pragma unroll
        for (int pd = 0; pd + q < 8; pd++)
          acc[q + pd] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, B[pd], acc[q + pd], 0, 0, 0);
      }
      if (ok) {
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
          uint64_t y4[4];
#pragma unroll
          for (int c = 0; c < 4; c++) {
            uint64_t y = (uint64_t)(int64_t)acc[7][4 * jj + c];
#pragma unroll
            for (int d = 6; d >= 0; d--) y = (y << 8) + (uint64_t)(int64_t)acc[d][4 * jj + c];
            y4[c] = y;
          }
          ulonglong2* o = reinterpret_cast<ulonglong2*>(out + ((uint64_t)r * n + e) * E + 32 * rb + 8 * jj + 4 * h);
          o[0] = make_ulonglong2(y4[0], y4[1]);
          o[1] = make_ulonglong2(y4[2], y4[3]);
        }
      }
    }
  }
}

// m-bar with 64-bit limbs on the matrix cores (as k_mbar_mx):
// out[l][j] = sum_{i in rows of split s} M_{e_i} X[l * lstride + i * k2 + j],
// frag: the fragments of every M_{e_i} (k_mulmat64, 8 planes, 128 KB each).
// Block: one tile of 32 consecutive outputs of one layer, wave w output
// rows 32w..32w+31, 36 digit-pair MFMAs per (row, k-block).  The tile's X
// row is staged once per block as digit planes in LDS (thread (kb, lane)
// forms the 16 limbs of its k-block; two buffers, one barrier per row), its
// global loads issued one row ahead; each wave's fragments are loaded one
// k-block ahead.  The digit classes add up exactly in i32 over 64 rows
// (|Y_s| <= 8 * 128 * 64 * 2^14 = 2^30) and are folded into 64-bit totals
// by a Horner step every 64 rows.  Block b takes split b % nsplit (rows
// [s*rows/nsplit, (s+1)*rows/nsplit)) and writes out + s * nl * k2 elements;
// workgroups go to the eight XCDs round-robin, so with nsplit a multiple of
// 8 an XCD reads the fragments of one eighth of the rows.  Two blocks per
// CU (a few registers spill; the second wave per SIMD hides load latency).
__global__ void __launch_bounds__(256, 2) k_mbar_mx64(const uint32_t* frag, const uint64_t* X, uint64_t rows, uint64_t k2,
                                                   uint64_t lstride, uint64_t nl, uint64_t nsplit, uint64_t* out) {
  __shared__ v4i sB[2][4 * 8 * 64];  // [buffer][k-block][plane][lane], 2 x 32 KB
  const v4i* A = reinterpret_cast<const v4i*>(frag);
  const int l64 = threadIdx.x & 63, col = l64 & 31, h = l64 >> 5, w = threadIdx.x >> 6;
  const uint64_t tpl = (k2 + 31) / 32, ntiles = tpl * nl;
  const uint64_t sp = blockIdx.x % nsplit, nq = gridDim.x / nsplit;
  const uint64_t i0 = rows * sp / nsplit, i1 = rows * (sp + 1) / nsplit;
  out += sp * nl * k2 * E;
  for (uint64_t tile = blockIdx.x / nsplit; tile < ntiles; tile += nq) {
    const uint64_t l = tile / tpl, j = (tile % tpl) * 32 + col;
    const bool ok = j < k2;
    const uint64_t* xl = X + (l * lstride + (ok ? j : 0)) * E;
    uint64_t tot[16];
#pragma unroll
    for (int c = 0; c < 16; c++) tot[c] = 0;
    uint64_t v[16];
    // this thread's share of row i: k-block w of its lane's column
    auto ldrow = [&](uint64_t i, uint64_t ie) {
      if (ok && i < ie) {
        ld16k64(xl + i * k2 * E, w, h, v);
      } else {
#pragma unroll
        for (int b = 0; b < 16; b++) v[b] = 0;
      }
    };
    auto strow = [&](int buf) {
      v4i P[8];
      planes16_64(v, P);
#pragma unroll
      for (int p = 0; p < 8; p++) sB[buf][(w * 8 + p) * 64 + l64] = P[p];
    };
    for (uint64_t ib = i0; ib < i1; ib += 64) {
      const uint64_t ie = i1 < ib + 64 ? i1 : ib + 64, nt = 4 * (ie - ib);
      v16i acc[8];
#pragma unroll
      for (int s = 0; s < 8; s++) acc[s] = (v16i)0;
      ldrow(ib, ie);
      lds_barrier();  // the buffers are no longer read
      strow(0);
      ldrow(ib + 1, ie);
      lds_barrier();
      v4i a[8];
#pragma unroll
      for (int q = 0; q < 8; q++) a[q] = A[ib * 8192 + frag_at(q, 0, w) * 64 + l64];
      for (uint64_t t = 0; t < nt; t++) {
        const uint64_t i = ib + (t >> 2), tn = t + 1;
        const int kb = (int)(t & 3), buf = (int)((i - ib) & 1);
        v4i an[8];
        const uint64_t in = tn < nt ? ib + (tn >> 2) : ib;
#pragma unroll
        for (int q = 0; q < 8; q++) an[q] = A[in * 8192 + frag_at(q, (int)(tn & 3), w) * 64 + l64];
        const v4i* Bs = &sB[buf][kb * 8 * 64 + l64];
#pragma unroll
        for (int p = 0; p < 8; p++) {
          const v4i B = Bs[p * 64];
#pragma unroll
          for (int q = 0; p + q < 8; q++)
            acc[q + p] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[q], B, acc[q + p], 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 8; q++) a[q] = an[q];
        if (kb == 3) {  // row i done: stage row i + 1, load row i + 2
          if (i + 1 < ie) strow(buf ^ 1);
          ldrow(i + 2, ie);
          lds_barrier();
        }
      }
#pragma unroll
      for (int c = 0; c < 16; c++) {
        uint64_t y = (uint64_t)(int64_t)acc[7][c];
#pragma unroll
        for (int s = 6; s >= 0; s--) y = (y << 8) + (uint64_t)(int64_t)acc[s][c];
        tot[c] += y;
      }
    }
    if (ok) {
#pragma unroll
      for (int jj = 0; jj < 4; jj++) {
        ulonglong2* o = reinterpret_cast<ulonglong2*>(out + (l * k2 + j) * E + 32 * w + 8 * jj + 4 * h);
        o[0] = make_ulonglong2(tot[4 * jj], tot[4 * jj + 1]);
        o[1] = make_ulonglong2(tot[4 * jj + 2], tot[4 * jj + 3]);
      }
    }
  }
}

// Linear-check round sums with 64-bit limbs (as k_gram_lin): the two Gram
// matrices over (f_lo, g_lo) and (f_hi - f_lo, g_hi - g_lo) with exact digit
// classes.  Per 32-element chunk and 32 x 32 block, class s of the digit
// products, sum_{p+q=s} X_p^T Y_q, is exact in i32 (|.| <= 8 * 32 * 2^14),
// and 256^s (class s) is added to the wave's running 64-bit totals.  512 threads (8 waves, 2 per SIMD, the digit
// planes of the chunk's four 128 x 32 limb matrices fill 128 KB of LDS);
// wave w takes the (gram, bx, by) blocks w, w + 8, w + 16, w + 24.  The
// contraction into 128 limbs is k_gram_lin's, in 64-bit.
// (block blk of nblk; sums != null as in gram_lin_body)
__device__ __forceinline__ void gram_lin64_body(const uint64_t* f, uint64_t fstride, const uint64_t* gt,
                                                uint64_t gstride, uint64_t ntab, uint64_t m, uint64_t hoff, int xrows,
                                                uint64_t* partials, int pstride, int poff, uint64_t* sums, unsigned blk,
                                                unsigned nblk) {
  __shared__ __attribute__((aligned(16))) uint32_t lds[32768];
  __shared__ unsigned long long raw[2][63][8];
  __shared__ uint64_t q[2][63][4];
  uint32_t (*sF)[8][4][64][4] = reinterpret_cast<uint32_t (*)[8][4][64][4]>(lds);
  const int tid = threadIdx.x, j = tid & 127, eg0 = tid >> 7;
  const int w = tid >> 6, l = tid & 63;
  const int xb = xrows / 32, npairs = (hoff ? 2 : 1) * xb * 4;  // gram 1 only with a hi half
  const uint64_t total = m * ntab, nchunks = (total + 31) / 32;
  for (int i = tid; i < 2 * 63 * 8; i += 512) (&raw[0][0][0])[i] = 0;
  uint64_t tot[4][16];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int i = 0; i < 16; i++) tot[u][i] = 0;
  const int rb = j >> 5;
  // with m a multiple of 32 a chunk lies in one table: one base pointer per
  // array (uniform table and index), unconditional loads at constant offsets
  const bool whole = m % 32 == 0 && total % 32 == 0;
  for (uint64_t ch = blk; ch < nchunks; ch += nblk) {
    // thread (j, eg0): limb j of the element groups eg = eg0 and eg0 + 4
    // (elements 4eg .. 4eg+3, k = kappa(eg & 1, 4(eg >> 1) + c))
    uint64_t R[2][4][4];  // [half][element][f lo, f hi, g lo, g hi]
    if (whole) {
      const uint64_t tab = ch * 32 / m, i0 = ch * 32 - tab * m + 4 * eg0;
      const uint64_t* pf = f + (tab * fstride + i0) * E + j;
      const uint64_t* pg = gt + (tab * gstride + i0) * E + j;
#pragma unroll
      for (int half = 0; half < 2; half++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const uint64_t o = (uint64_t)(16 * half + c) * E;
          R[half][c][0] = pf[o];
          R[half][c][1] = pf[hoff * E + o];
          R[half][c][2] = pg[o];
          R[half][c][3] = pg[hoff * E + o];
        }
    } else {
#pragma unroll
      for (int half = 0; half < 2; half++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const uint64_t e = ch * 32 + 4 * (eg0 + 4 * half) + c;
          R[half][c][0] = R[half][c][1] = R[half][c][2] = R[half][c][3] = 0;
          if (e < total) {
            const uint64_t tab = e / m, i = e % m;
            const uint64_t* F = f + (tab * fstride + i) * E;
            const uint64_t* G = gt + (tab * gstride + i) * E;
            R[half][c][0] = F[j];
            R[half][c][1] = F[hoff * E + j];
            R[half][c][2] = G[j];
            R[half][c][3] = G[hoff * E + j];
          }
        }
    }
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int eg = eg0 + 4 * half;
      const int lane = (j & 31) + 32 * (eg & 1), wd = eg >> 1;
      uint64_t z[4][4];  // [x0, y0, x1, y1][element]
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const uint64_t f0 = R[half][c][0], f1 = R[half][c][1], g0 = R[half][c][2], g1 = R[half][c][3];
        z[0][c] = bal_digits64(f0);
        z[1][c] = bal_digits64(g0);
        z[2][c] = bal_digits64(f1 - f0);
        z[3][c] = bal_digits64(g1 - g0);
      }
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        uint32_t lo[4], hi[4];
        bytes_t4<4>((uint32_t)z[mt][0], (uint32_t)z[mt][1], (uint32_t)z[mt][2], (uint32_t)z[mt][3], lo);
        bytes_t4<4>((uint32_t)(z[mt][0] >> 32), (uint32_t)(z[mt][1] >> 32), (uint32_t)(z[mt][2] >> 32),
                    (uint32_t)(z[mt][3] >> 32), hi);
#pragma unroll
        for (int p = 0; p < 4; p++) {
          sF[mt][p][rb][lane][wd] = lo[p];
          sF[mt][4 + p][rb][lane][wd] = hi[p];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int pr = w + 8 * u;
      if (pr >= npairs) continue;
      const int pg = pr / (xb * 4), pbx = (pr % (xb * 4)) >> 2, pby = pr & 3;
      const int xm = 2 * pg, ym = 2 * pg + 1;
#pragma unroll
      for (int sd = 7; sd >= 0; sd--) {
        v16i tmp = (v16i)0;
#pragma unroll
        for (int pd = 0; pd <= sd; pd++) {
          const v4i a = *reinterpret_cast<const v4i*>(sF[xm][pd][pbx][l]);
          const v4i b = *reinterpret_cast<const v4i*>(sF[ym][sd - pd][pby][l]);
          tmp = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, tmp, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 16; i++) tot[u][i] += (uint64_t)(int64_t)tmp[i] << (8 * sd);
      }
    }
    __syncthreads();
  }
  // accumulator entry i of lane l in block (bx, by): row 32bx + 8(i>>2) +
  // 4(l>>5) + (i&3), column 32by + (l&31)
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int pr = w + 8 * u;
    if (pr >= npairs) continue;
    const int pg = pr / (xb * 4), pbx = (pr % (xb * 4)) >> 2, pby = pr & 3;
    const int mp0 = 8 * (pbx + pby) + (l >> 5) + ((l & 31) >> 2), cp = l & 3;
#pragma unroll
    for (int i = 0; i < 16; i++)
      atomicAdd(&raw[pg][mp0 + 2 * (i >> 2)][(i & 3) + cp], (unsigned long long)tot[u][i]);
  }
  __syncthreads();
  if (tid < 2 * 63) {
    const int g = tid / 63, p = tid % 63;
    const unsigned long long* r = raw[g][p];
    q[g][p][0] = r[0] - r[4];
    q[g][p][1] = r[1] - r[4] - r[5];
    q[g][p][2] = r[2] - r[5] - r[6];
    q[g][p][3] = r[3] - r[6];
  }
  __syncthreads();
  if (tid < 2 * E) {
    const int g = tid >> 7, u = (tid & 127) >> 2, c = tid & 3;
    auto hc = [&](int s2, int cc) -> uint64_t { return q[g][32 + s2][cc] - (s2 < 2 ? q[g][61 + s2][cc] : 0ull); };
    uint64_t v = q[g][u][c];
    if (u >= 1) v -= hc(u - 1, c);
    if (u >= 3) v -= hc(u - 3, c);
    if (u <= 30) {
      if (c == 0) v += hc(u, 3);
      else if (c == 1) v += hc(u, 3) - hc(u, 0);
      else v -= hc(u, c - 1);
    }
    if (g == 0 || hoff) {
      if (sums) atom_add(sums + (size_t)(poff + g) * E + (tid & 127), v);
      else partials[((size_t)blk * pstride + poff + g) * E + (tid & 127)] = v;
    }
  }
}
__global__ void __launch_bounds__(512) k_gram_lin64(const uint64_t* f, uint64_t fstride, const uint64_t* gt,
                                                    uint64_t gstride, uint64_t ntab, uint64_t m, uint64_t hoff,
                                                    int xrows, uint64_t* partials, int pstride, int poff) {
  gram_lin64_body(f, fstride, gt, gstride, ntab, m, hoff, xrows, partials, pstride, poff, nullptr, blockIdx.x,
                  gridDim.x);
}

// m-bars on the matrix cores (pcs.rs:554-561, 984-1001):
//   out[l][j] = sum_{i < rows} e_i X[l * lstride + i * k2 + j]
// with the M_{e_i} fragments (frag + i * ND * 4096 words, from k_mulmat).
// A block takes four tiles of 32 consecutive outputs (tile = l * tiles per
// layer + j / 32), wave w tile 4b + w, all four output row-blocks; for each
// row i the block stages M_{e_i} (ND * 16 KB) in LDS once for its four
// tiles, each wave forms the digit planes of its 32 X elements (loaded
// during the previous row's matrix work) and accumulates one digit class per
// accumulator, adding sum_d acc_d << 8d to its running sums (i32 wrap =
// mod 2^32, so the sum over i is exact).
// Block b takes split s = b % nsplit, rows [s*rows/nsplit, (s+1)*rows/nsplit),
// and quads b / nsplit, b / nsplit + gridDim.x / nsplit, ...; it writes its
// partial sums to out + s * nl * k2 elements (summed afterwards by
// k_sum_split when nsplit > 1).  Workgroups are dealt to the eight XCDs
// round-robin, so with nsplit a multiple of 8 an XCD's blocks share one
// eighth of the rows and their fragments stay in that XCD's L2.
template <class T>
__global__ void __launch_bounds__(256, 2) k_mbar_mx(const uint32_t* frag, const LT<T>* X, uint64_t rows, uint64_t k2,
                                                 uint64_t lstride, uint64_t nl, uint64_t nsplit, LT<T>* out) {
  using C = CT<T>;
  constexpr int ND = kND<T>;
  __shared__ v4i sA[ND * 16 * 64];
  const int l64 = threadIdx.x & 63, col = l64 & 31, h = l64 >> 5, w = threadIdx.x >> 6;
  const uint64_t tpl = (k2 + 31) / 32, ntiles = tpl * nl;
  const uint64_t sp = blockIdx.x % nsplit, nq = gridDim.x / nsplit;
  const uint64_t i0 = rows * sp / nsplit, i1 = rows * (sp + 1) / nsplit;
  out += sp * nl * k2 * E;
  for (uint64_t tb = (uint64_t)(blockIdx.x / nsplit) * 4; tb < ntiles; tb += nq * 4) {
    const uint64_t tile = tb + w;
    const uint64_t l = tile < ntiles ? tile / tpl : 0, j = (tile < ntiles ? (tile % tpl) * 32 : k2) + col;
    const bool ok = tile < ntiles && j < k2;
    const LT<T>* xl = X + (l * lstride + (ok ? j : 0)) * E;
    v16i tot[4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) tot[rb] = (v16i)0;
    for (uint64_t i = i0; i < i1; i++) {
      lds_barrier();  // the previous row's fragments are no longer read
      const v4i* gA = reinterpret_cast<const v4i*>(frag + i * ND * 4096);
      stage_frags<ND>(gA, sA);
      v4i B[4][ND];
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        uint32_t v[16];
        if (ok) {
          ld16k<T>(xl + i * k2 * E, kb, h, v);
        } else {
#pragma unroll
          for (int b = 0; b < 16; b++) v[b] = 0;
        }
        planes16<ND>(v, B[kb]);
      }
      lds_barrier();  // the fragments are in LDS
#pragma unroll
      for (int rb = 0; rb < 4; rb++) {
        // keeps the fragment reads next to their matrix ops
        asm volatile("" ::: "memory");
        // one accumulator per digit class: each fragment is read from LDS
        // once and used for every class it feeds
        v16i acc[ND];
#pragma unroll
        for (int d = 0; d < ND; d++) acc[d] = (v16i)0;
#pragma unroll
        for (int p = 0; p < ND; p++)
#pragma unroll
          for (int kb = 0; kb < 4; kb++) {
            const v4i af = sA[frag_at(p, kb, rb) * 64 + l64];
#pragma unroll
            for (int q = 0; p + q < ND; q++)
              acc[p + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, B[kb][q], acc[p + q], 0, 0, 0);
          }
#pragma unroll
        for (int d = 0; d < ND; d++) tot[rb] += acc[d] << (8 * d);
      }
    }
    if (ok) {
#pragma unroll
      for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
          C y4[4];
#pragma unroll
          for (int c = 0; c < 4; c++) y4[c] = (C)(uint32_t)tot[rb][4 * jj + c];
          st4<T>(out + (l * k2 + j) * E + 32 * rb + 8 * jj + 4 * h, y4);
        }
    }
  }
}

// pack_r2 on the matrix cores: out[r * n + e] = pack_r2(rot_{r + rot0}(slots[e]))
// = P_L slots[e] with L = (r + rot0) mod 32 (frag: the 32 P_L, ND * 1024
// words each).  Block: 32 elements, wave w output rows 32w..32w+31; the
// slots are read once for all nrot rotations.
template <class T>
__global__ void __launch_bounds__(256) k_pack_mx(const uint32_t* frag, const LT<T>* slots, uint64_t n, int nrot,
                                                 int rot0, LT<T>* out) {
  using C = CT<T>;
  constexpr int ND = kND<T>;
  const int l64 = threadIdx.x & 63, col = l64 & 31, h = l64 >> 5, rb = threadIdx.x >> 6;
  const v4i* A = reinterpret_cast<const v4i*>(frag);
  const uint64_t ntiles = (n + 31) / 32;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t e = tile * 32 + col;
    const bool ok = e < n;
    uint32_t v[16];
    if (ok) {
      ld16k<T>(slots + e * 32, 0, h, v);
    } else {
#pragma unroll
      for (int b = 0; b < 16; b++) v[b] = 0;
    }
    v4i B[ND];
    planes16<ND>(v, B);
    for (int r = 0; r < nrot; r++) {
      const int L = (r + rot0) & 31;
      v16i acc[ND];
#pragma unroll
      for (int d = 0; d < ND; d++) acc[d] = (v16i)0;
#pragma unroll
      for (int pd = 0; pd < ND; pd++) {
        const v4i a = A[((size_t)L * ND * 4 + pd * 4 + rb) * 64 + l64];
#pragma unroll
        for (int q = 0; pd + q < ND; q++) acc[pd + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, B[q], acc[pd + q], 0, 0, 0);
      }
      if (ok) {
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
          C y4[4];
#pragma unroll
          for (int c = 0; c < 4; c++) {
            uint32_t y = 0;
#pragma unroll
            for (int d = 0; d < ND; d++) y += (uint32_t)acc[d][4 * jj + c] << (8 * d);
            y4[c] = (C)y;
          }
          st4<T>(out + ((uint64_t)r * n + e) * E + 32 * rb + 8 * jj + 4 * h, y4);
        }
      }
    }
  }
}

// out[e] = sum_{s < nsplit} part[s * cnt + e] (limbwise, cnt elements)
template <class T>
__global__ void k_sum_split(const LT<T>* part, uint64_t nsplit, uint64_t cnt, LT<T>* out) {
  const uint64_t nl = cnt * E;
  for (uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; id < nl; id += (uint64_t)gridDim.x * blockDim.x) {
    CT<T> a = 0;
    for (uint64_t s = 0; s < nsplit; s++) a += (CT<T>)part[s * nl + id];
    out[id] = (LT<T>)a;
  }
}

// ---------------------------------------------------------------------------
// Sums of products on the matrix cores (k <= 32).
// sum_e x_e y_e = contract(G) with G = sum_e x_e y_e^T, the 128 x 128 Gram
// matrix of the limb vectors over Z/2^32 (wrap-around), and contract the
// product formula applied once: limb j = 4i+c of x and j' = 4i'+c' of y
// meet at beta^(i+i') alpha^(c+c').
// ---------------------------------------------------------------------------

// The end of a linear-check Gram block: the wave accumulators (NACC
// (gram, bx, by) blocks per wave) are contracted into 128 limbs per Gram
// matrix (bins raw[i+i'][c+c'] by LDS atomics, alpha^4 = -alpha - 1, then
// Y^32 = -Y^3 - Y - alpha) and written as the block's partials (or added
// atomically into sums).  raw must be zero on entry.
template <class T, int NACC>
__device__ __forceinline__ void gram_finish(const v16i* tot, const bool* pon, const int* pg, const int* pbx,
                                            const int* pby, uint32_t (*raw)[63][8], uint32_t (*q)[63][4], uint64_t hoff,
                                            LT<T>* partials, int pstride, int poff, CT<T>* sums, unsigned blk) {
  using L = LT<T>;
  const int tid = threadIdx.x, l = tid & 63;
  __syncthreads();
  // accumulator entry i of lane l in block (bx, by): row 32bx + 8(i>>2) +
  // 4(l>>5) + (i&3), column 32by + (l&31)
#pragma unroll
  for (int u = 0; u < NACC; u++) {
    if (!pon[u]) continue;
    const int mp0 = 8 * (pbx[u] + pby[u]) + (l >> 5) + ((l & 31) >> 2), cp = l & 3;
#pragma unroll
    for (int i = 0; i < 16; i++) atomicAdd(&raw[pg[u]][mp0 + 2 * (i >> 2)][(i & 3) + cp], (uint32_t)tot[u][i]);
  }
  __syncthreads();
  if (tid < 2 * 63) {
    const int g = tid / 63, p = tid % 63;
    const uint32_t* r = raw[g][p];
    q[g][p][0] = r[0] - r[4];
    q[g][p][1] = r[1] - r[4] - r[5];
    q[g][p][2] = r[2] - r[5] - r[6];
    q[g][p][3] = r[3] - r[6];
  }
  __syncthreads();
  if (tid < 2 * E) {
    const int g = tid >> 7, u = (tid & 127) >> 2, c = tid & 3;
    // position 32 + s after positions 61 and 62 have folded into 32 and 33
    auto hc = [&](int s, int cc) -> uint32_t { return q[g][32 + s][cc] - (s < 2 ? q[g][61 + s][cc] : 0u); };
    uint32_t v = q[g][u][c];
    if (u >= 1) v -= hc(u - 1, c);  // -Y from position u + 31
    if (u >= 3) v -= hc(u - 3, c);  // -Y^3 from position u + 29
    if (u <= 30) {                  // -alpha from position u + 32
      if (c == 0) v += hc(u, 3);
      else if (c == 1) v += hc(u, 3) - hc(u, 0);
      else v -= hc(u, c - 1);
    }
    if (g == 0 || hoff) {
      if (sums) atom_add(sums + (size_t)(poff + g) * E + (tid & 127), v);
      else partials[((size_t)blk * pstride + poff + g) * E + (tid & 127)] = (L)v;
    }
  }
}

// Linear-check round (as k_lin_round): gram 0 over (f_lo, g_lo), gram 1
// over (f_hi - f_lo, g_hi - g_lo), both from one read of the tables (hi at
// hoff elements past lo; hoff = 0 makes gram 1 zero and gram 0 a plain sum
// of products, the inner products of Engine::dots).  Partials of block b:
// gram g at partials[b * pstride + poff + g] (gram 1 only when hoff != 0).  A
// block (16 waves) walks chunks of 32 elements: its threads load the four
// 128 x 32 limb matrices (thread (j, eg): limb j of elements 4eg..4eg+3,
// which are k = kappa(eg&1, 4(eg>>1) + c)) and store their digit planes in
// LDS; then each wave takes one or two (gram, 32x32 block) pairs.  Per
// chunk a Horner pass over the digit classes fills a temporary accumulator
// that is added to the wave's running total (i32 wrap = mod 2^32).  x rows
// at limbs >= xrows are zero (low-degree f) and skipped.
// Contraction: limb j = 4i+c of x and j' = 4i'+c' of y meet at
// beta^(i+i') alpha^(c+c'), so each wave adds its accumulator entries into
// the 63 x 7 bins raw[i+i'][c+c'] with LDS atomics (both Gram matrices at
// once); then alpha^4 = -alpha - 1 and Y^32 = -Y^3 - Y - alpha reduce the
// bins to 128 limbs per matrix, written as partials [block][gram].
// (block blk of nblk; sums != null: gram g is added atomically into
// sums[poff + g] instead of written as a partial)
template <class T>
__device__ __forceinline__ void gram_lin_body(const LT<T>* f, uint64_t fstride, const LT<T>* gt, uint64_t gstride,
                                              uint64_t ntab, uint64_t m, uint64_t hoff, int xrows, LT<T>* partials,
                                              int pstride, int poff, CT<T>* sums, unsigned blk, unsigned nblk) {
  using L = LT<T>;
  constexpr int ND = kND<T>;
  // digit planes [x0|y0|x1|y1][plane][rb][lane][word] in the loop, a Gram matrix after
  __shared__ __attribute__((aligned(16))) uint32_t lds[16384];
  __shared__ uint32_t raw[2][63][8], q[2][63][4];
  uint32_t (*sF)[ND][4][64][4] = reinterpret_cast<uint32_t (*)[ND][4][64][4]>(lds);
  const int tid = threadIdx.x, j = tid & 127, eg = tid >> 7;
  const int w = tid >> 6, l = tid & 63;
  // (gram, bx, by) pairs: xb = xrows/32 row-blocks of x per gram
  const int xb = xrows / 32, npairs = (hoff ? 2 : 1) * xb * 4;  // gram 1 only with a hi half
  int pg[2], pbx[2], pby[2];
  bool pon[2];
#pragma unroll
  for (int u = 0; u < 2; u++) {
    const int pr = w + 16 * u;
    pon[u] = pr < npairs;
    pg[u] = pr / (xb * 4);
    pbx[u] = (pr % (xb * 4)) >> 2;
    pby[u] = pr & 3;
  }
  const uint64_t total = m * ntab, nchunks = (total + 31) / 32;
  if (tid < 2 * 63 * 8) (&raw[0][0][0])[tid] = 0;
  v16i tot[2] = {(v16i)0, (v16i)0};
  const int lane = (j & 31) + 32 * (eg & 1), rb = j >> 5, wd = eg >> 1;
  // raw limbs of the chunk: loads for chunk ch+stride are issued before the
  // MFMA work of chunk ch, so HBM latency hides behind the matrix cores
  uint32_t rf[4][2], rg[4][2];  // [element][lo, hi]
  auto load = [&](uint64_t ch) {
    // one division per chunk (block-uniform): with m >= 32 a chunk spans at
    // most two tables
    const uint64_t base = ch * 32, tab0 = base / m, i0 = base - tab0 * m;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const uint64_t e = base + 4 * eg + c;
      rf[c][0] = rf[c][1] = rg[c][0] = rg[c][1] = 0;
      if (ch < nchunks && e < total) {
        uint64_t tab, i;
        if (m >= 32) {
          const uint64_t ii = i0 + 4 * eg + c;
          const bool wr = ii >= m;
          tab = tab0 + (wr ? 1 : 0);
          i = wr ? ii - m : ii;
        } else {
          tab = e / m;
          i = e % m;
        }
        const L* F = f + (tab * fstride + i) * E;
        const L* G = gt + (tab * gstride + i) * E;
        rf[c][0] = (uint32_t)F[j];
        rf[c][1] = (uint32_t)F[hoff * E + j];
        rg[c][0] = (uint32_t)G[j];
        rg[c][1] = (uint32_t)G[hoff * E + j];
      }
    }
  };
  load(blk);
  for (uint64_t ch = blk; ch < nchunks; ch += nblk) {
    uint32_t z[4][4];  // [x0, y0, x1, y1][element]
#pragma unroll
    for (int c = 0; c < 4; c++) {
      z[0][c] = bal_digits<ND>(rf[c][0]);
      z[1][c] = bal_digits<ND>(rg[c][0]);
      z[2][c] = bal_digits<ND>(rf[c][1] - rf[c][0]);
      z[3][c] = bal_digits<ND>(rg[c][1] - rg[c][0]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      uint32_t wv[ND];
      bytes_t4<ND>(z[mt][0], z[mt][1], z[mt][2], z[mt][3], wv);
#pragma unroll
      for (int p = 0; p < ND; p++) sF[mt][p][rb][lane][wd] = wv[p];
    }
    __syncthreads();
    load(ch + nblk);
#pragma unroll
    for (int u = 0; u < 2; u++) {
      if (!pon[u]) continue;
      const int xm = 2 * pg[u], ym = 2 * pg[u] + 1;
      v4i a[ND], b[ND];
#pragma unroll
      for (int p = 0; p < ND; p++) {
        a[p] = *reinterpret_cast<const v4i*>(sF[xm][p][pbx[u]][l]);
        b[p] = *reinterpret_cast<const v4i*>(sF[ym][p][pby[u]][l]);
      }
      v16i tmp = (v16i)0;
#pragma unroll
      for (int sd = ND - 1; sd >= 0; sd--) {
        if (sd != ND - 1) tmp <<= 8;
#pragma unroll
        for (int pd = 0; pd <= sd; pd++) tmp = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[pd], b[sd - pd], tmp, 0, 0, 0);
      }
      tot[u] += tmp;
    }
    __syncthreads();
  }
  gram_finish<T, 2>(tot, pon, pg, pbx, pby, raw, q, hoff, partials, pstride, poff, sums, blk);
}
// The same sums with 512 threads, two chunks of loads in flight and two LDS
// buffers: thread (j, eg2) loads limb j of elements 8eg2..8eg2+7 of a chunk
// (f and g at lo and hi: 32 values), so the raw limbs of the next two
// chunks are in registers (128 KB per CU); while the waves' MFMAs read one
// chunk's digit planes from one LDS buffer, its threads write the next
// chunk's into the other (one barrier per chunk).  Wave w accumulates the
// (gram, bx, by) pairs w, w+8, w+16, w+24.  Chunks of block b: b, b + nblk,
// ...
template <class T>
__global__ void __launch_bounds__(512) k_gram_lin2(const LT<T>* f, uint64_t fstride, const LT<T>* gt, uint64_t gstride,
                                                   uint64_t ntab, uint64_t m, uint64_t hoff, int xrows, LT<T>* partials,
                                                   int pstride, int poff) {
  using L = LT<T>;
  constexpr int ND = kND<T>;
  constexpr int BUFW = 4 * ND * 4 * 64 * 4;  // words of one chunk's planes
  __shared__ __attribute__((aligned(16))) uint32_t lds[2 * BUFW];
  __shared__ uint32_t raw[2][63][8], q[2][63][4];
  const int tid = threadIdx.x, j = tid & 127, eg2 = tid >> 7, rb = j >> 5;
  const int w = tid >> 6, l = tid & 63;
  const int xb = xrows / 32, npairs = (hoff ? 2 : 1) * xb * 4;
  int pg[4], pbx[4], pby[4];
  bool pon[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int pr = w + 8 * u;
    pon[u] = pr < npairs;
    pg[u] = pr / (xb * 4);
    pbx[u] = (pr % (xb * 4)) >> 2;
    pby[u] = pr & 3;
  }
  const uint64_t total = m * ntab, nchunks = (total + 31) / 32, nblk = gridDim.x;
  for (int i = tid; i < 2 * 63 * 8; i += 512) (&raw[0][0][0])[i] = 0;
  v16i tot[4];
#pragma unroll
  for (int u = 0; u < 4; u++) tot[u] = (v16i)0;
  uint32_t R0[8][4], R1[8][4];  // [element][f lo, f hi, g lo, g hi]
  // Loads are unconditional, so they stay in flight across the stage and
  // MFMA work.  With m a multiple of 32 a whole chunk lies in one table and
  // its elements are consecutive: one base pointer per array (uniform table
  // and index), constant offsets per element.  Otherwise (short tables, the
  // tail chunk) each element finds its table and index, elements past the
  // end read element 0 and are zeroed.
  const bool whole = m % 32 == 0 && total % 32 == 0;
  auto load = [&](uint64_t ch, uint32_t (*R)[4]) {
    const uint64_t base = ch * 32;
    if (whole && ch < nchunks) {
      const uint64_t tab = base / m, i0 = base - tab * m + 8 * eg2;
      const L* pf = f + (tab * fstride + i0) * E + j;
      const L* pg = gt + (tab * gstride + i0) * E + j;
      const L* pfh = pf + hoff * E;
      const L* pgh = pg + hoff * E;
#pragma unroll
      for (int c = 0; c < 8; c++) {
        R[c][0] = (uint32_t)pf[c * E];
        R[c][1] = (uint32_t)pfh[c * E];
        R[c][2] = (uint32_t)pg[c * E];
        R[c][3] = (uint32_t)pgh[c * E];
      }
    } else {
      uint64_t fo[8], go[8];
      bool ok[8];
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const uint64_t e = base + 8 * eg2 + c, tab = e / m, i = e % m;
        ok[c] = ch < nchunks && e < total;
        fo[c] = ok[c] ? (tab * fstride + i) * E : 0;
        go[c] = ok[c] ? (tab * gstride + i) * E : 0;
      }
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const uint32_t x0 = (uint32_t)f[fo[c] + j], x1 = (uint32_t)f[fo[c] + hoff * E + j];
        const uint32_t y0 = (uint32_t)gt[go[c] + j], y1 = (uint32_t)gt[go[c] + hoff * E + j];
        R[c][0] = ok[c] ? x0 : 0u;
        R[c][1] = ok[c] ? x1 : 0u;
        R[c][2] = ok[c] ? y0 : 0u;
        R[c][3] = ok[c] ? y1 : 0u;
      }
    }
  };
  // digit planes of a chunk into buffer sF ([x0|y0|x1|y1][plane][rb][lane][word]):
  // elements 4k..4k+3 of this thread's eight are k = 8 eg2 + 4k + c of the
  // operands (lane (j%32) + 32k, word eg2)
  auto stage = [&](const uint32_t (*R)[4], uint32_t* buf) {
    uint32_t (*sF)[ND][4][64][4] = reinterpret_cast<uint32_t (*)[ND][4][64][4]>(buf);
#pragma unroll
    for (int k = 0; k < 2; k++) {
      uint32_t z[4][4];  // [x0, y0, x1, y1][element]
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const uint32_t* r = R[4 * k + c];
        z[0][c] = bal_digits<ND>(r[0]);
        z[1][c] = bal_digits<ND>(r[2]);
        z[2][c] = bal_digits<ND>(r[1] - r[0]);
        z[3][c] = bal_digits<ND>(r[3] - r[2]);
      }
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        uint32_t wv[ND];
        bytes_t4<ND>(z[mt][0], z[mt][1], z[mt][2], z[mt][3], wv);
#pragma unroll
        for (int p = 0; p < ND; p++) sF[mt][p][rb][(j & 31) + 32 * k][eg2] = wv[p];
      }
    }
  };
  // each operand plane is read from LDS once per pair (2 ND reads for
  // ND (ND + 1) / 2 MFMAs)
  auto grams = [&](const uint32_t* buf) {
    const uint32_t (*sF)[ND][4][64][4] = reinterpret_cast<const uint32_t (*)[ND][4][64][4]>(buf);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (!pon[u]) continue;
      const int xm = 2 * pg[u], ym = 2 * pg[u] + 1;
      v4i a[ND], b[ND];
#pragma unroll
      for (int p = 0; p < ND; p++) {
        a[p] = *reinterpret_cast<const v4i*>(sF[xm][p][pbx[u]][l]);
        b[p] = *reinterpret_cast<const v4i*>(sF[ym][p][pby[u]][l]);
      }
      v16i tmp = (v16i)0;
#pragma unroll
      for (int sd = ND - 1; sd >= 0; sd--) {
        if (sd != ND - 1) tmp <<= 8;
#pragma unroll
        for (int pd = 0; pd <= sd; pd++) tmp = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[pd], b[sd - pd], tmp, 0, 0, 0);
      }
      tot[u] += tmp;
    }
  };
  uint64_t ch = blockIdx.x;
  int cur = 0;
  load(ch, R0);
  load(ch + nblk, R1);
  stage(R0, lds);
  load(ch + 2 * nblk, R0);
  __syncthreads();
  // chunk ch's planes in buffer cur; the register set that holds chunk
  // ch + nblk is staged into the other buffer and reloaded with ch + 3 nblk
  while (ch < nchunks) {
    grams(lds + cur * BUFW);
    stage(R1, lds + (cur ^ 1) * BUFW);
    load(ch + 3 * nblk, R1);
    __syncthreads();
    ch += nblk;
    cur ^= 1;
    if (ch >= nchunks) break;
    grams(lds + cur * BUFW);
    stage(R0, lds + (cur ^ 1) * BUFW);
    load(ch + 3 * nblk, R0);
    __syncthreads();
    ch += nblk;
    cur ^= 1;
  }
  gram_finish<T, 4>(tot, pon, pg, pbx, pby, raw, q, hoff, partials, pstride, poff, nullptr, blockIdx.x);
}
template <class T>
__global__ void __launch_bounds__(1024) k_gram_lin(const LT<T>* f, uint64_t fstride, const LT<T>* gt, uint64_t gstride,
                                                   uint64_t ntab, uint64_t m, uint64_t hoff, int xrows, LT<T>* partials,
                                                   int pstride, int poff) {
  gram_lin_body<T>(f, fstride, gt, gstride, ntab, m, hoff, xrows, partials, pstride, poff, nullptr, blockIdx.x,
                   gridDim.x);
}

}  // namespace nenya
