// Shared device helpers for the paces HIP library (gfx950, wave64).
//
// Rows are packed Fock states: W uint64 words per row, row-major, field 0 (exciton position)
// in the most significant bits of word 0 (big-endian bit concatenation, see
// paces/core/hamiltonian.py:218-248 for the reference's uint32 form).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "paces_hip.h"

typedef unsigned long long u64;

// A pointer in the global address space: loads through it are global, not flat, accesses (a
// flat load also counts against the LDS wait counter, so later LDS reads would wait for it).
#define PACES_GPTR(T) __attribute__((address_space(1))) T*
// A pointer into LDS: 32-bit address arithmetic.
#define PACES_LPTR(T) __attribute__((address_space(3))) T*

#define PACES_MAX_W 16

namespace paces {

constexpr int kBlock = 256;
constexpr int kScanTile = 2048;  // elements per workgroup of the device-wide scans

// ---- status plumbing --------------------------------------------------------------------------
inline int launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? PACES_OK : PACES_E_HIP;
}

inline unsigned grid_for(int64_t n, int per_block) {
  int64_t g = (n + per_block - 1) / per_block;
  return (unsigned)(g < 1 ? 1 : g);
}

// ---- rows -------------------------------------------------------------------------------------
template <int W>
struct Row {
  u64 w[W];
};

template <int W>
__device__ __forceinline__ Row<W> load_row(const u64* __restrict__ base, int64_t i) {
  Row<W> r;
  const u64* p = base + i * W;
#pragma unroll
  for (int k = 0; k < W; ++k) r.w[k] = p[k];
  return r;
}

template <int W>
__device__ __forceinline__ void store_row(u64* __restrict__ base, int64_t i, const Row<W>& r) {
  u64* p = base + i * W;
#pragma unroll
  for (int k = 0; k < W; ++k) p[k] = r.w[k];
}

// -1 if a < b, 0 if equal, +1 if a > b (lexicographic over words)
template <int W>
__device__ __forceinline__ int cmp_row(const Row<W>& a, const u64* __restrict__ b) {
#pragma unroll
  for (int k = 0; k < W; ++k) {
    u64 bk = b[k];
    if (a.w[k] != bk) return a.w[k] < bk ? -1 : 1;
  }
  return 0;
}

template <int W>
__device__ __forceinline__ bool eq_row(const Row<W>& a, const u64* __restrict__ b) {
  bool e = true;
#pragma unroll
  for (int k = 0; k < W; ++k) e &= (a.w[k] == b[k]);
  return e;
}

// lexicographic lower_bound of row q in sorted book[0, M): first index m with book[m] >= q
template <int W>
__device__ __forceinline__ int64_t lower_bound_row(const u64* __restrict__ book, int64_t lo,
                                                   int64_t hi, const Row<W>& q) {
  while (lo < hi) {
    int64_t m = (lo + hi) >> 1;
    if (cmp_row<W>(q, book + m * W) > 0)
      lo = m + 1;
    else
      hi = m;
  }
  return lo;
}

// ---- block-cooperative lookup ---------------------------------------------------------------------
// Every thread of a workgroup holds at most one query row (valid = false for none). The workgroup
// reduces the lexicographic min and max of its valid queries, finds their lower bounds lo <= hi in
// the sorted book with two searches, and stages book[lo, hi) in LDS when it fits; otherwise it
// stages evenly spaced fence rows of that range, so each query finishes with an LDS search plus a
// short search inside one fence window. Queries that are sorted or clustered (every call on the
// step) give a narrow range, so the book is read once, coalesced. Must be called by every thread
// of the workgroup (it synchronises).

constexpr int kLkWords = 3072;  // 24 KiB of staged rows per workgroup

// a < b over W words, word 0 most significant: the borrow out of the multiword a - b.
template <int W>
__device__ __forceinline__ bool lex_less(const u64 (&a)[W], const u64 (&b)[W]) {
  unsigned long long borrow = 0;
#pragma unroll
  for (int k = W - 1; k >= 0; --k) (void)__builtin_subcll(a[k], b[k], borrow, &borrow);
  return borrow != 0;
}

// ((a < b) != INV) ? x1 : x0, for W <= 4 by the borrow chain of a - b in VCC and one select
// (VALU only: the compiler's form of a multiword compare combines per-word lane masks with
// scalar instructions, which a binary search repeats at every probe).
template <int W, bool INV>
__device__ __forceinline__ uint32_t less_select(const u64 (&a)[W], const u64 (&b)[W], uint32_t x1,
                                                uint32_t x0) {
#define PACES_LO(x) (uint32_t)(x)
#define PACES_HI(x) (uint32_t)((x) >> 32)
  // v_cndmask_b32 d, s0, s1, vcc: d = vcc ? s1 : s0, and vcc = borrow = (a < b)
  const uint32_t on_borrow = INV ? x0 : x1, on_none = INV ? x1 : x0;
  uint32_t t, f;
  if constexpr (W == 1) {
    asm("v_sub_co_u32 %0, vcc, %2, %3\n\t"
        "v_subb_co_u32 %0, vcc, %4, %5, vcc\n\t"
        "v_cndmask_b32 %1, %6, %7, vcc"
        : "=&v"(t), "=v"(f)
        : "v"(PACES_LO(a[0])), "v"(PACES_LO(b[0])), "v"(PACES_HI(a[0])), "v"(PACES_HI(b[0])),
          "v"(on_none), "v"(on_borrow)
        : "vcc");
  } else if constexpr (W == 2) {
    asm("v_sub_co_u32 %0, vcc, %2, %3\n\t"
        "v_subb_co_u32 %0, vcc, %4, %5, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %6, %7, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %8, %9, vcc\n\t"
        "v_cndmask_b32 %1, %10, %11, vcc"
        : "=&v"(t), "=v"(f)
        : "v"(PACES_LO(a[1])), "v"(PACES_LO(b[1])), "v"(PACES_HI(a[1])), "v"(PACES_HI(b[1])),
          "v"(PACES_LO(a[0])), "v"(PACES_LO(b[0])), "v"(PACES_HI(a[0])), "v"(PACES_HI(b[0])),
          "v"(on_none), "v"(on_borrow)
        : "vcc");
  } else if constexpr (W == 3) {
    asm("v_sub_co_u32 %0, vcc, %2, %3\n\t"
        "v_subb_co_u32 %0, vcc, %4, %5, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %6, %7, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %8, %9, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %10, %11, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %12, %13, vcc\n\t"
        "v_cndmask_b32 %1, %14, %15, vcc"
        : "=&v"(t), "=v"(f)
        : "v"(PACES_LO(a[2])), "v"(PACES_LO(b[2])), "v"(PACES_HI(a[2])), "v"(PACES_HI(b[2])),
          "v"(PACES_LO(a[1])), "v"(PACES_LO(b[1])), "v"(PACES_HI(a[1])), "v"(PACES_HI(b[1])),
          "v"(PACES_LO(a[0])), "v"(PACES_LO(b[0])), "v"(PACES_HI(a[0])), "v"(PACES_HI(b[0])),
          "v"(on_none), "v"(on_borrow)
        : "vcc");
  } else if constexpr (W == 4) {
    asm("v_sub_co_u32 %0, vcc, %2, %3\n\t"
        "v_subb_co_u32 %0, vcc, %4, %5, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %6, %7, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %8, %9, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %10, %11, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %12, %13, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %14, %15, vcc\n\t"
        "v_subb_co_u32 %0, vcc, %16, %17, vcc\n\t"
        "v_cndmask_b32 %1, %18, %19, vcc"
        : "=&v"(t), "=v"(f)
        : "v"(PACES_LO(a[3])), "v"(PACES_LO(b[3])), "v"(PACES_HI(a[3])), "v"(PACES_HI(b[3])),
          "v"(PACES_LO(a[2])), "v"(PACES_LO(b[2])), "v"(PACES_HI(a[2])), "v"(PACES_HI(b[2])),
          "v"(PACES_LO(a[1])), "v"(PACES_LO(b[1])), "v"(PACES_HI(a[1])), "v"(PACES_HI(b[1])),
          "v"(PACES_LO(a[0])), "v"(PACES_LO(b[0])), "v"(PACES_HI(a[0])), "v"(PACES_HI(b[0])),
          "v"(on_none), "v"(on_borrow)
        : "vcc");
  } else {
    (void)t;
    f = lex_less<W>(a, b) ? on_borrow : on_none;
  }
#undef PACES_LO
#undef PACES_HI
  return f;
}

template <int W>
__device__ __forceinline__ void load_lds_row(PACES_LPTR(const unsigned char) rows, unsigned i,
                                             u64 (&w)[W]) {
  const PACES_LPTR(const u64) p = (PACES_LPTR(const u64))(rows + __umul24(i, W * 8));
#pragma unroll
  for (int k = 0; k < W; ++k) w[k] = p[k];
}

// Bound of q in n sorted unique rows staged in LDS (byte address rows): the number of rows below
// q, or with UPPER the number of rows at most q. Halving search: the window [base, base + len)
// holds the answer's neighbourhood, every probe (base + len/2) is in bounds, and the window
// lengths depend on n only, so every lane takes ceil(log2 n) steps plus one final compare
// (n is uniform across the workgroup).
template <int W, bool UPPER>
__device__ __forceinline__ int bound_lds(PACES_LPTR(const unsigned char) rows, int n,
                                         const Row<W>& q) {
  if (n <= 0) return 0;
  // the window's first row is tracked by its LDS address: one add forms a probe's address, and
  // one select (on the compare's borrow) moves the window to it when q lies above it (lower
  // bound) or not below it (upper bound)
  typedef PACES_LPTR(const unsigned char) lptr;
  lptr a = rows;
  for (int len = n; len > 1;) {
    const int half = len >> 1;
    uint32_t hb = (uint32_t)half * (W * 8u);  // uniform: kept in a scalar register, so the
    asm("" : "+s"(hb));                       // probe address is one 32-bit add (not a mad)
    const lptr p = a + hb;
    u64 w[W];
    load_lds_row<W>(p, 0u, w);
    const uint32_t pa = (uint32_t)(size_t)p, aa = (uint32_t)(size_t)a;  // LDS addresses: 32-bit
    a = (lptr)(size_t)(UPPER ? less_select<W, true>(q.w, w, pa, aa)
                             : less_select<W, false>(w, q.w, pa, aa));
    len -= half;
  }
  u64 w[W];
  load_lds_row<W>(a, 0u, w);
  const uint32_t pred = UPPER ? less_select<W, true>(q.w, w, 1u, 0u)
                              : less_select<W, false>(w, q.w, 1u, 0u);
  return (int)((uint32_t)(a - rows) / (W * 8) + pred);
}

// Copy the 16-byte blocks holding bytes [src, src + nbytes) global -> LDS at dst (16-byte
// aligned) by direct loads (global_load_lds_dwordx4: lane l's block lands at the wave's base +
// 16 l, no registers); returns the number of blocks. A block never crosses a page, so the bytes
// read beyond the range cannot fault. Every wave of the workgroup takes part; the caller waits
// (vmcnt 0) and synchronises the workgroup before reading dst.
__device__ __forceinline__ int dma_blocks_to_lds(const void* src, size_t nbytes, u64* dst,
                                                 int& lead_words) {
  const uintptr_t a0 = (uintptr_t)src;
  const uintptr_t b0 = a0 & ~(uintptr_t)15, b1 = (a0 + nbytes + 15) & ~(uintptr_t)15;
  const int nb = nbytes ? (int)((b1 - b0) >> 4) : 0;
  lead_words = (int)((a0 - b0) >> 3);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  PACES_GPTR(const unsigned char) g = (PACES_GPTR(const unsigned char))b0;
  for (int c0 = wave * 64; c0 < nb; c0 += (int)blockDim.x)
    if (c0 + lane < nb)
      __builtin_amdgcn_global_load_lds((PACES_GPTR(const void))(g + (size_t)(c0 + lane) * 16),
                                       (PACES_LPTR(void))(dst + 2 * c0), 16, 0, 0);
  return nb;
}

template <int W>
struct LookupSmem {
  __align__(16) u64 rows[kLkWords];
  u64 wmin[kBlock / 64][W];
  u64 wmax[kBlock / 64][W];
  int wany[kBlock / 64];
  long long lo, hi, stride, nst;
  int any;
};

template <int W>
__device__ __forceinline__ bool row_less(const Row<W>& a, const Row<W>& b) {
#pragma unroll
  for (int k = 0; k < W; ++k)
    if (a.w[k] != b.w[k]) return a.w[k] < b.w[k];
  return false;
}

template <int W>
__device__ __forceinline__ int64_t lower_bound_lds(const u64* rows, int64_t n, const Row<W>& q) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t m = (lo + hi) >> 1;
    if (cmp_row<W>(q, rows + m * W) > 0)
      lo = m + 1;
    else
      hi = m;
  }
  return lo;
}

template <int W>
__device__ int64_t block_lower_bound(const Row<W>& q, bool valid, const u64* __restrict__ book,
                                     int64_t M, LookupSmem<W>& sm) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  Row<W> mn = q, mx = q;
  int v = valid ? 1 : 0;
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    Row<W> omn, omx;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      omn.w[k] = __shfl_xor(mn.w[k], s, 64);
      omx.w[k] = __shfl_xor(mx.w[k], s, 64);
    }
    int ov = __shfl_xor(v, s, 64);
    if (ov && (!v || row_less<W>(omn, mn))) mn = omn;
    if (ov && (!v || row_less<W>(mx, omx))) mx = omx;
    v |= ov;
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < W; ++k) {
      sm.wmin[wid][k] = mn.w[k];
      sm.wmax[wid][k] = mx.w[k];
    }
    sm.wany[wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    // thread 0 searches the block minimum, thread 1 the block maximum
    Row<W> best;
    int have = 0;
    for (int q2 = 0; q2 < kBlock / 64; ++q2) {
      if (!sm.wany[q2]) continue;
      Row<W> c;
#pragma unroll
      for (int k = 0; k < W; ++k) c.w[k] = threadIdx.x == 0 ? sm.wmin[q2][k] : sm.wmax[q2][k];
      if (!have || (threadIdx.x == 0 ? row_less<W>(c, best) : row_less<W>(best, c))) best = c;
      have = 1;
    }
    if (have) {
      int64_t b = lower_bound_row<W>(book, 0, M, best);
      if (threadIdx.x == 0)
        sm.lo = b;
      else
        sm.hi = b;
    }
    if (threadIdx.x == 0) sm.any = have;
  }
  __syncthreads();
  if (!sm.any) return 0;
  const int64_t lo = sm.lo, hi = sm.hi;
  const int64_t span = hi - lo;
  if (span * W + 3 <= kLkWords) {  // the whole range fits: direct loads, branchless search
    int lead;
    dma_blocks_to_lds(book + lo * W, (size_t)span * W * 8, sm.rows, lead);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    int64_t res = 0;
    if (valid)
      res = lo + bound_lds<W, false>((PACES_LPTR(const unsigned char))(sm.rows + lead),
                                     (int)span, q);
    __syncthreads();
    return res;
  }
  constexpr int cap = kLkWords / W;
  int64_t stride = 1, nst = span;
  if (span > cap) {
    stride = (span + cap - 1) / cap;
    nst = (span + stride - 1) / stride;
  }
  // stage rows lo + f*stride, f < nst
  for (int64_t e = threadIdx.x; e < nst * W; e += blockDim.x) {
    int64_t f = e / W, k = e - f * W;
    sm.rows[e] = book[(lo + f * stride) * W + k];
  }
  __syncthreads();
  int64_t res = 0;
  if (valid) {
    int64_t fi = lower_bound_lds<W>(sm.rows, nst, q);
    if (stride == 1) {
      res = lo + fi;
    } else {
      int64_t a = (fi == 0) ? lo : lo + (fi - 1) * stride + 1;
      int64_t b = (fi == nst) ? hi : lo + fi * stride;
      res = lower_bound_row<W>(book, a, b, q);
    }
  }
  __syncthreads();  // the staging area is reused by the next call
  return res;
}

// Same as block_lower_bound for queries whose lower bounds are known to lie in [lo, hi] (the
// range comes from a partition pass, so no thread searches the whole book here).
template <int W>
__device__ int64_t block_lower_bound_range(const Row<W>& q, bool valid,
                                           const u64* __restrict__ book, int64_t lo, int64_t hi,
                                           LookupSmem<W>& sm) {
  const int64_t span = hi - lo;
  if (span * W + 3 <= kLkWords) {  // the whole range fits: direct loads, branchless search
    int lead;
    dma_blocks_to_lds(book + lo * W, (size_t)span * W * 8, sm.rows, lead);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    int64_t res = 0;
    if (valid)
      res = lo + bound_lds<W, false>((PACES_LPTR(const unsigned char))(sm.rows + lead),
                                     (int)span, q);
    __syncthreads();
    return res;
  }
  constexpr int cap = kLkWords / W;
  int64_t stride = 1, nst = span;
  if (span > cap) {
    stride = (span + cap - 1) / cap;
    nst = (span + stride - 1) / stride;
  }
  for (int64_t e = threadIdx.x; e < nst * W; e += blockDim.x) {
    int64_t f = e / W, k = e - f * W;
    sm.rows[e] = book[(lo + f * stride) * W + k];
  }
  __syncthreads();
  int64_t res = 0;
  if (valid) {
    int64_t fi = lower_bound_lds<W>(sm.rows, nst, q);
    if (stride == 1) {
      res = lo + fi;
    } else {
      int64_t a = (fi == 0) ? lo : lo + (fi - 1) * stride + 1;
      int64_t b = (fi == nst) ? hi : lo + fi * stride;
      res = lower_bound_row<W>(book, a, b, q);
    }
  }
  __syncthreads();
  return res;
}

// ---- bit fields ---------------------------------------------------------------------------------
// A field at bit offset `off` (from the MSB of word 0) and width `b` (1..64). It lies in word
// k = off/64 or, when o + b > 64 (o = off%64), straddles words k and k+1. Only 64-bit shifts by
// 0..63 are used: the 128-bit formulation gave unstable results for straddling fields.
template <int W>
__device__ __forceinline__ void field_words(const Row<W>& r, int k, u64& hi, u64& lo) {
  hi = 0;
  lo = 0;
#pragma unroll
  for (int kk = 0; kk < W; ++kk) {  // select without dynamic register indexing
    hi = (kk == k) ? r.w[kk] : hi;
    lo = (kk == k + 1) ? r.w[kk] : lo;
  }
}

template <int W>
__device__ __forceinline__ u64 get_field(const Row<W>& r, int off, int b) {
  // branch-free (selects only), so loads of later rows are not held behind this one's branches
  const int k = off >> 6, o = off & 63;
  u64 hi, lo;
  field_words<W>(r, k, hi, lo);
  const u64 mask = (b >= 64) ? ~0ull : ((1ull << (b & 63)) - 1ull);  // 0 for an empty field
  const int end = o + b;   // 0..127
  const int s = end - 64;  // when end > 64: the field's bits in word k+1, 1..63
  const u64 one = hi >> ((64 - end) & 63);
  const u64 two = (hi << (s & 63)) | (lo >> ((64 - s) & 63));
  return (end <= 64 ? one : two) & mask;
}

// add a signed delta to a field; the caller guarantees the result stays inside the field, so the
// carry or borrow never leaves the two words the field touches.
template <int W>
__device__ __forceinline__ void add_field(Row<W>& r, int off, int b, int64_t delta) {
  if (b == 0 || delta == 0) return;
  const int k = off >> 6, o = off & 63;
  u64 hi, lo;
  field_words<W>(r, k, hi, lo);
  const int end = o + b;
  const bool neg = delta < 0;
  const u64 d = neg ? (u64)(-delta) : (u64)delta;
  if (end <= 64) {
    const u64 a = d << (64 - end);
    hi = neg ? hi - a : hi + a;
  } else {
    const int s = end - 64;            // field LSB is bit 64-s of word k+1
    const u64 a_lo = d << (64 - s);    // part of d*2^(64-s) inside word k+1
    const u64 a_hi = d >> s;           // part that lands in word k
    if (neg) {
      const u64 borrow = lo < a_lo ? 1ull : 0ull;
      lo = lo - a_lo;
      hi = hi - a_hi - borrow;
    } else {
      const u64 nlo = lo + a_lo;
      const u64 carry = nlo < lo ? 1ull : 0ull;
      lo = nlo;
      hi = hi + a_hi + carry;
    }
  }
#pragma unroll
  for (int kk = 0; kk < W; ++kk) {
    if (kk == k) r.w[kk] = hi;
    if (kk == k + 1 && end > 64) r.w[kk] = lo;
  }
}

template <int W>
__device__ __forceinline__ void set_field(Row<W>& r, int off, int b, u64 v) {
  if (b == 0) return;
  u64 cur = get_field<W>(r, off, b);
  add_field<W>(r, off, b, (int64_t)v - (int64_t)cur);
}

// ---- wave / block reductions ----------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
  return v;
}

__device__ __forceinline__ u64 wave_max_u64(u64 v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    u64 o = __shfl_xor(v, s, 64);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
  return v;
}

// ---- launch shapes (N4): workgroup sizes 128..1024 x rows per thread 1, 2, 4 ----------------------
inline bool cfg_ok(const paces_launch_cfg_t* cfg) {
  if (!cfg) return true;
  const int b = cfg->block, r = cfg->rows_per_thread;
  return (b == 128 || b == 256 || b == 512 || b == 1024) && (r == 1 || r == 2 || r == 4);
}
inline void cfg_get(const paces_launch_cfg_t* cfg, int* b, int* r) {
  *b = cfg ? cfg->block : kBlock;
  *r = cfg ? cfg->rows_per_thread : 1;
}

#define PACES_CFG_SWITCH(B, RR, LAUNCH)                            \
  switch ((B) * 8 + (RR)) {                                       \
    case 128 * 8 + 1: LAUNCH(128, 1); break;                      \
    case 128 * 8 + 2: LAUNCH(128, 2); break;                      \
    case 128 * 8 + 4: LAUNCH(128, 4); break;                      \
    case 256 * 8 + 1: LAUNCH(256, 1); break;                      \
    case 256 * 8 + 2: LAUNCH(256, 2); break;                      \
    case 256 * 8 + 4: LAUNCH(256, 4); break;                      \
    case 512 * 8 + 1: LAUNCH(512, 1); break;                      \
    case 512 * 8 + 2: LAUNCH(512, 2); break;                      \
    case 512 * 8 + 4: LAUNCH(512, 4); break;                      \
    case 1024 * 8 + 1: LAUNCH(1024, 1); break;                    \
    case 1024 * 8 + 2: LAUNCH(1024, 2); break;                    \
    case 1024 * 8 + 4: LAUNCH(1024, 4); break;                    \
    default: return PACES_E_INVALID;                              \
  }

// ---- sparse H.x of several rows per thread --------------------------------------------------------
// H.x for the RPT rows ii[] of one thread (diag*x plus each term's CSR row), in load phases so a
// thread keeps many independent loads in flight instead of walking each row's dependent chain
// (bounds, entries, gathers) in turn:
//   1. own x, diagonal and every term's row bounds of every row;
//   2. the first kHxEnt entries (value, column) of every (row, term);
//   3. the gathers of x at those columns.
// Every load is unconditional and in bounds (rows past the tile read row `spare`, an absent
// entry reads the term's indptr, an absent gather row `spare`; the results are unused), so the
// code has no branches and each use waits only for its own load. The sums are formed in the
// reference order: y = d*x; per term t = 0, t = t + a_k x_k in column order, y = y + t. Entries
// past the first kHxEnt of a row's term are taken in a loop. NT = the number of off-diagonal terms
// (a template argument, so the per-term arrays stay in registers). CSR: indptr[t], indices[t],
// data[t] (F values; data may be typed or void*).
constexpr int kHxEnt = 2;

// Entry values read from the CSR's data arrays. (A value loader gives each entry's raw load --
// issued in the entry phase, in bounds: an absent entry reads the term's indptr -- and its value.)
template <typename F>
struct PlainVals {
  typedef F raw_t;
  template <typename CSR>
  __device__ __forceinline__ F load(const CSR& H, int t, int32_t k, bool ok) const {
    const F* pa = ok ? ((const F*)H.data[t]) + k : (const F*)H.indptr[t];
    return *pa;
  }
  __device__ __forceinline__ F value(int, F raw) const { return raw; }
};

template <typename F, typename V, int RPT, int NT, typename CSR, typename VL = PlainVals<F>>
__device__ __forceinline__ void hx_rows_phased(const int64_t (&ii)[RPT], const bool (&vv)[RPT],
                                               int64_t spare, const F* __restrict__ diag,
                                               const CSR& H, const V* __restrict__ x,
                                               V (&h)[RPT], V (&xi)[RPT], const VL& vl = VL()) {
  F d[RPT];
  int32_t b[RPT][NT], e[RPT][NT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int64_t i = vv[q] ? ii[q] : spare;
    xi[q] = x[i];
    d[q] = diag[i];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      b[q][t] = H.indptr[t][i];
      e[q][t] = H.indptr[t][i + 1];
    }
  }
  typename VL::raw_t a[RPT][NT][kHxEnt];
  uint32_t ix[RPT][NT][kHxEnt];  // unsigned: no sign extension (and wait) before the gathers
  bool ok[RPT][NT][kHxEnt];
#pragma unroll
  for (int q = 0; q < RPT; ++q)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int m = 0; m < kHxEnt; ++m) {
        const int32_t k = b[q][t] + m;
        ok[q][t][m] = vv[q] & (k < e[q][t]);
        const int32_t* pi = ok[q][t][m] ? H.indices[t] + k : H.indptr[t];
        a[q][t][m] = vl.load(H, t, k, ok[q][t][m]);
        ix[q][t][m] = (uint32_t)*pi;
      }
  V xj[RPT][NT][kHxEnt];
#pragma unroll
  for (int q = 0; q < RPT; ++q)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int m = 0; m < kHxEnt; ++m)
        xj[q][t][m] = x[ok[q][t][m] ? (int64_t)ix[q][t][m] : spare];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    F yr = d[q] * xi[q].x, yi = d[q] * xi[q].y;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      F tr = 0, ti = 0;
#pragma unroll
      for (int m = 0; m < kHxEnt; ++m) {
        if (ok[q][t][m]) {
          const F av = vl.value(t, a[q][t][m]);
          tr = tr + av * xj[q][t][m].x;
          ti = ti + av * xj[q][t][m].y;
        }
      }
      if (vv[q]) {
        for (int32_t k = b[q][t] + kHxEnt; k < e[q][t]; ++k) {
          const F av = vl.value(t, vl.load(H, t, k, true));
          const V xv = x[H.indices[t][k]];
          tr = tr + av * xv.x;
          ti = ti + av * xv.y;
        }
      }
      yr = yr + tr;
      yi = yi + ti;
    }
    h[q].x = yr;
    h[q].y = yi;
  }
}

}  // namespace paces

// Dispatch a templated launcher over the supported row widths.
#define PACES_DISPATCH_W(Wv, FN, ...)                 \
  switch (Wv) {                                       \
    case 1: return FN<1>(__VA_ARGS__);                \
    case 2: return FN<2>(__VA_ARGS__);                \
    case 3: return FN<3>(__VA_ARGS__);                \
    case 4: return FN<4>(__VA_ARGS__);                \
    case 5: return FN<5>(__VA_ARGS__);                \
    case 6: return FN<6>(__VA_ARGS__);                \
    case 7: return FN<7>(__VA_ARGS__);                \
    case 8: return FN<8>(__VA_ARGS__);                \
    case 9: return FN<9>(__VA_ARGS__);                \
    case 10: return FN<10>(__VA_ARGS__);              \
    case 11: return FN<11>(__VA_ARGS__);              \
    case 12: return FN<12>(__VA_ARGS__);              \
    case 13: return FN<13>(__VA_ARGS__);              \
    case 14: return FN<14>(__VA_ARGS__);              \
    case 15: return FN<15>(__VA_ARGS__);              \
    case 16: return FN<16>(__VA_ARGS__);              \
    default: return PACES_E_UNSUPPORTED;              \
  }
