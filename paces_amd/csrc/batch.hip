// N1: many independent trajectories advanced together. All trajectories share one packed basis
// whose field 0 is the trajectory id, so the basis of trajectory t is the contiguous segment
// [seg_off[t], seg_off[t+1]) of the sorted rows, and H is block diagonal. The union, the
// reassignment, the diagonal and the products run on the whole batch unchanged; this file holds
// the per-trajectory parts: segments, the top-k select (_determine_select_whoami,
// paces/core/time_evolution.py:833-910, per trajectory), the observables and the Taylor stopping
// test (time_evolution.py:482-512, per trajectory). Workgroups are aligned to segments: block b
// covers tile b - boff[t] of trajectory t, with boff[t] = sum_{u<t} ceil(R_u / kTile).
#include <stddef.h>

#include "common.hpp"

namespace paces {

int scan_u8(const uint8_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
int scan_i32(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws_bytes(int64_t n);

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr int kTile = kBlock;  // rows per segment-aligned workgroup

// ---- segments -------------------------------------------------------------------------------------

// seg_off[t] = first row whose trajectory id is >= t (t = 0..ntraj); rows are sorted
template <int W>
__global__ void k_seg_offsets(const u64* __restrict__ rows, int64_t R, int trajbits, int ntraj,
                              int32_t* __restrict__ seg_off) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntraj) return;
  if (t == ntraj) {
    seg_off[t] = (int32_t)R;
    return;
  }
  Row<W> q;
#pragma unroll
  for (int k = 0; k < W; ++k) q.w[k] = 0;
  q.w[0] = trajbits ? ((u64)t << (64 - trajbits)) : 0;
  seg_off[t] = (int32_t)lower_bound_row<W>(rows, 0, R, q);
}

// boff[t]: first workgroup of trajectory t; boff[ntraj] = number of workgroups
// boff[t] = the tiles of `tile` rows before trajectory t's (boff[ntraj]: all of them); one
// workgroup of kSegThreads threads, each taking a run of trajectories, with a scan across them
constexpr int kSegThreads = 256;

__device__ __forceinline__ void seg_blocks_body(const int32_t* __restrict__ seg_off, int ntraj,
                                                int tile, int64_t* __restrict__ boff) {
  __shared__ int64_t sc[kSegThreads];
  const int per = (ntraj + kSegThreads - 1) / kSegThreads;
  const int t0 = (int)threadIdx.x * per;
  const int t1 = t0 + per < ntraj ? t0 + per : ntraj;
  int64_t loc = 0;
  for (int t = t0; t < t1; ++t) loc += (seg_off[t + 1] - seg_off[t] + tile - 1) / tile;
  sc[threadIdx.x] = loc;
  __syncthreads();
  for (int o = 1; o < kSegThreads; o <<= 1) {  // inclusive scan
    const int64_t v = (int)threadIdx.x >= o ? sc[threadIdx.x - o] : 0;
    __syncthreads();
    sc[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t b = sc[threadIdx.x] - loc;
  for (int t = t0; t < t1; ++t) {
    boff[t] = b;
    b += (seg_off[t + 1] - seg_off[t] + tile - 1) / tile;
  }
  if (threadIdx.x == kSegThreads - 1) boff[ntraj] = sc[kSegThreads - 1];
}

__global__ void __launch_bounds__(kSegThreads) k_seg_blocks(const int32_t* __restrict__ seg_off,
                                                            int ntraj, int tile,
                                                            int64_t* __restrict__ boff) {
  seg_blocks_body(seg_off, ntraj, tile, boff);
}

struct SegTile {
  int t;        // trajectory, -1 if the workgroup is past the last tile
  int64_t r0;   // first row of the tile
  int64_t r1;   // one past the last row
  int64_t b;    // the tile's index (the workgroup's, unless remapped)
};

// kXcdChunk > 0: workgroup b runs tile (k / C) * 8C + (b % 8) * C + k % C, k = b / 8, C =
// kXcdChunk: each XCD (workgroup b runs on XCD b % 8) takes runs of C consecutive tiles, so that
// tiles whose rows neighbour each other share an L2 (the grid is padded to a multiple of 8C)
#ifndef PACES_XCD_CHUNK
#define PACES_XCD_CHUNK 64
#endif
constexpr int kXcdChunk = PACES_XCD_CHUNK;

__device__ __forceinline__ int64_t xcd_tile(int64_t b) {
  if (kXcdChunk == 0) return b;
  const int64_t k = b >> 3;
  return (k / kXcdChunk) * (8 * kXcdChunk) + (b & 7) * kXcdChunk + k % kXcdChunk;
}

static unsigned xcd_grid(unsigned g) {
  if (kXcdChunk == 0) return g;
  const unsigned q = 8u * kXcdChunk;
  return (g + q - 1) / q * q;
}

template <bool XCD = false>
__device__ __forceinline__ SegTile seg_tile(const int32_t* __restrict__ seg_off,
                                            const int64_t* __restrict__ boff, int ntraj, int tile) {
  SegTile st;
  const int64_t b = XCD ? xcd_tile(blockIdx.x) : (int64_t)blockIdx.x;
  st.b = b;
  if (b >= boff[ntraj]) {
    st.t = -1;
    st.r0 = st.r1 = 0;
    return st;
  }
  int lo = 0, hi = ntraj;  // last t with boff[t] <= b (empty trajectories share boff values)
  while (hi - lo > 1) {
    int m = (lo + hi) >> 1;
    if (boff[m] <= b)
      lo = m;
    else
      hi = m;
  }
  while (lo + 1 < ntraj && boff[lo + 1] <= b) ++lo;
  st.t = lo;
  st.r0 = seg_off[lo] + (b - boff[lo]) * tile;
  int64_t e = st.r0 + tile;
  st.r1 = e < seg_off[lo + 1] ? e : seg_off[lo + 1];
  return st;
}

static unsigned seg_grid(int64_t R, int ntraj, int tile) {
  return (unsigned)((R + tile - 1) / tile + ntraj + 1);
}

// ---- per-trajectory top-k select --------------------------------------------------------------------

struct SegSel {
  u64 prefix;
  u64 mask;
  long long krem;
  int all;
  int pad;
};

__global__ void k_bsel_keys(const double2* __restrict__ psi, int64_t R, u64* __restrict__ keys) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  double2 v = psi[i];
  double k = v.x * v.x + v.y * v.y;
  keys[i] = (u64)__double_as_longlong(k);
}

// Radix select over the 64-bit keys in digits of up to 11 bits, most significant first: five
// 11-bit passes and a last 9-bit one (6 passes), with 2048-bin per-trajectory histograms.
constexpr int kSelBits = 11;
constexpr int kSelBins = 1 << kSelBits;
constexpr int kSelPasses = 6;
__host__ __device__ constexpr int sel_shift(int pass) { return pass < 5 ? 53 - 11 * pass : 0; }
__host__ __device__ constexpr int sel_width(int pass) { return pass < 5 ? 11 : 9; }

__global__ void k_bsel_init(const int32_t* __restrict__ seg_off, const int64_t* __restrict__ kpt,
                            int ntraj, SegSel* __restrict__ st, unsigned* __restrict__ hist,
                            int32_t* __restrict__ ccount, int32_t* __restrict__ out_count) {
  const int t = blockIdx.x;
  if (t >= ntraj) return;
  for (int d = threadIdx.x; d < kSelBins; d += blockDim.x) hist[(int64_t)t * kSelBins + d] = 0;
  if (t == 0 && threadIdx.x == 0) *out_count = 0;
  if (threadIdx.x == 0) {
    ccount[t] = 0;
    const long long Rt = seg_off[t + 1] - seg_off[t];
    SegSel s;
    s.prefix = 0;
    s.mask = 0;
    s.all = (kpt[t] >= Rt) ? 1 : 0;
    s.krem = s.all ? 0 : kpt[t];
    s.pad = 0;
    st[t] = s;
  }
}

// histogram passes take kSelTile rows per workgroup (8 per thread, strided for coalescing), so
// each workgroup's LDS histogram is flushed to the trajectory's global one 8x less often
constexpr int kSelRpt = 8;
constexpr int kSelTile = kBlock * kSelRpt;

__global__ void __launch_bounds__(kBlock) k_bsel_hist(const u64* __restrict__ keys,
                                                      const int32_t* __restrict__ seg_off,
                                                      const int64_t* __restrict__ boff, int ntraj,
                                                      const SegSel* __restrict__ st, int shift,
                                                      int width, unsigned* __restrict__ hist) {
  __shared__ unsigned lh[kSelBins];
  const SegTile T = seg_tile(seg_off, boff, ntraj, kSelTile);
  if (T.t < 0 || st[T.t].all || st[T.t].krem == 0) return;
  for (int d = threadIdx.x; d < kSelBins; d += kBlock) lh[d] = 0;
  __syncthreads();
  const u64 prefix = st[T.t].prefix, mask = st[T.t].mask;
  const u64 dm = (1ull << width) - 1ull;
  u64 k[kSelRpt];
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q) {  // every key loaded before any is counted
    const int64_t i = T.r0 + q * kBlock + threadIdx.x;
    k[q] = keys[i < T.r1 ? i : T.r0];
  }
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q) {
    const int64_t i = T.r0 + q * kBlock + threadIdx.x;
    if (i < T.r1 && (k[q] & mask) == prefix) atomicAdd(&lh[(k[q] >> shift) & dm], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kSelBins; d += kBlock) {
    const unsigned c = lh[d];
    if (c) atomicAdd(&hist[(int64_t)T.t * kSelBins + d], c);
  }
}

// inclusive prefix sum of v over the 256 threads of a workgroup (wave scans, then the waves'
// totals through LDS); *total receives the sum over all threads. Every thread must call it.
__device__ __forceinline__ long long blk256_incl(long long v, long long* total) {
  __shared__ long long wt[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  long long inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wt[wave] = inc;
  __syncthreads();
  long long before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    before += w < wave ? wt[w] : 0;
    all += wt[w];
  }
  __syncthreads();
  *total = all;
  return before + inc;
}

// the digit holding the krem-th largest remaining key: per thread 8 consecutive bins, a suffix
// scan over the threads, then the thread whose bins cross krem finds the digit
__global__ void __launch_bounds__(256) k_bsel_pick(SegSel* __restrict__ st,
                                                   unsigned* __restrict__ hist, int shift,
                                                   int width) {
  constexpr int kPer = kSelBins / 256;
  const int t = blockIdx.x, d0 = threadIdx.x * kPer;
  unsigned* h = hist + (int64_t)t * kSelBins;
  const bool skip = st[t].all || st[t].krem == 0;
  unsigned c[kPer];
  long long tot = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    c[j] = skip ? 0u : h[d0 + j];
    tot += c[j];
  }
  long long all;
  const long long incl = blk256_incl(tot, &all);
  if (!skip) {
    const long long krem = st[t].krem;
    // keys in bins above this thread's; walk this thread's bins from the top
    long long above = all - incl;
    int pick = -1;
    long long above_pick = 0;
#pragma unroll
    for (int j = kPer - 1; j >= 0; --j) {
      const long long with = above + c[j];
      if (pick < 0 && with >= krem && above < krem) {
        pick = d0 + j;
        above_pick = above;
      }
      above = with;
    }
    __syncthreads();
    if (pick >= 0) {
      st[t].prefix |= ((u64)pick) << shift;
      st[t].mask |= ((1ull << width) - 1ull) << shift;
      st[t].krem = krem - above_pick;
    }
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) h[d0 + j] = 0;
}

// ---- the selection's output: the keys above each trajectory's threshold key, and the last krem
// of the keys equal to it (among equal keys the later rows are kept: the stable order of the
// reference's sort). After the first kSelFull digit passes over every row, the rows whose keys
// match the prefix fixed so far (a bucket of relative width 2^-11 around the threshold) are the
// candidates: k_bsel_cand gathers their row indices into their trajectory's stretch of a
// workspace array, and k_bsel_final (one workgroup per trajectory) runs the remaining digits over
// them, then finds the cut: the row index of the first kept tie (ties ranked in row order), so a
// row is kept iff its key is above the threshold, or equal to it at a row index >= the cut.
constexpr int kSelFull = 2;

__global__ void __launch_bounds__(kBlock) k_bsel_cand(const u64* __restrict__ keys,
                                                      const int32_t* __restrict__ seg_off,
                                                      const int64_t* __restrict__ boff, int ntraj,
                                                      const SegSel* __restrict__ st,
                                                      int32_t* __restrict__ ci,
                                                      int32_t* __restrict__ ccount) {
  const SegTile T = seg_tile(seg_off, boff, ntraj, kSelTile);
  if (T.t < 0 || st[T.t].all || st[T.t].krem == 0) return;
  const u64 prefix = st[T.t].prefix, mask = st[T.t].mask;
  const int lane = threadIdx.x & 63;
  u64 k[kSelRpt];
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q) {
    const int64_t i = T.r0 + q * kBlock + threadIdx.x;
    k[q] = keys[i < T.r1 ? i : T.r0];
  }
  int32_t* out = ci + seg_off[T.t];  // the trajectory's stretch: at most R_t candidates
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q) {
    const int64_t i = T.r0 + q * kBlock + threadIdx.x;
    const bool m = i < T.r1 && (k[q] & mask) == prefix;
    const u64 bl = __ballot(m);
    if (bl == 0) continue;
    int base = 0;
    if (lane == 0) base = atomicAdd(&ccount[T.t], (int)__popcll(bl));
    base = __shfl(base, 0, 64);
    const int pre = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bl >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((unsigned)bl, 0u));
    if (m) out[base + pre] = (int32_t)i;
  }
}

// In one 256-thread workgroup: the bin of h[0, kSelBins) holding the r-th key (r >= 1) counted
// from the top (desc) or the r-th (r >= 0, 0-based) counted from the bottom (asc); returns the
// bin and, in *before, the keys counted past before it. Every thread must call it.
__device__ __forceinline__ int blk_pick_bin(const unsigned* h, long long r, bool desc,
                                            long long* before) {
  constexpr int kPer = kSelBins / 256;
  __shared__ int s_pick;
  __shared__ long long s_before;
  const int d0 = threadIdx.x * kPer;
  unsigned c[kPer];
  long long tot = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    c[j] = h[d0 + j];
    tot += c[j];
  }
  if (threadIdx.x == 0) s_pick = -1;
  long long all;
  const long long incl = blk256_incl(tot, &all);  // (its barriers publish s_pick's reset)
  if (desc) {
    long long above = all - incl;  // keys in the bins of the threads after this one
#pragma unroll
    for (int j = kPer - 1; j >= 0; --j) {
      const long long with = above + c[j];
      if (with >= r && above < r) {
        s_pick = d0 + j;
        s_before = above;
      }
      above = with;
    }
  } else {
    long long below = incl - tot;  // keys in the bins of the threads before this one
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const long long with = below + c[j];
      if (below <= r && r < with) {
        s_pick = d0 + j;
        s_before = below;
      }
      below = with;
    }
  }
  __syncthreads();
  const int pick = s_pick;
  *before = s_before;
  __syncthreads();
  return pick;
}

// One workgroup per trajectory: digit passes kSelFull .. kSelPasses - 1 over the candidates, then
// the cut (the (ties - krem)-th smallest row index among the keys equal to the threshold, by
// digits of the row index), and the kept count added to *out_count. A trajectory that keeps
// nothing gets the threshold ~0 (no key reaches it: keys are the bits of non-negative doubles).
__global__ void __launch_bounds__(256) k_bsel_final(SegSel* __restrict__ st,
                                                    const u64* __restrict__ keys,
                                                    const int32_t* __restrict__ ci,
                                                    const int32_t* __restrict__ ccount,
                                                    const int32_t* __restrict__ seg_off,
                                                    const int64_t* __restrict__ kpt,
                                                    int32_t* __restrict__ cut,
                                                    int32_t* __restrict__ out_count) {
  __shared__ unsigned h[kSelBins];
  __shared__ int s_n;
  const int t = blockIdx.x;
  SegSel s = st[t];
  const long long Rt = seg_off[t + 1] - seg_off[t];
  if (threadIdx.x == 0) {
    const long long kept = s.all ? Rt : (kpt[t] < Rt ? kpt[t] : Rt);
    atomicAdd(out_count, (int)kept);
  }
  if (s.all) {
    if (threadIdx.x == 0) cut[t] = 0;
    return;
  }
  if (s.krem == 0) {
    if (threadIdx.x == 0) {
      st[t].prefix = ~0ull;
      st[t].mask = ~0ull;
      cut[t] = 0x7fffffff;
    }
    return;
  }
  const int n = ccount[t];
  const int32_t* c = ci + seg_off[t];
  // up to kCandLds candidates (the usual case) are read once into LDS, with their keys
  constexpr int kCandLds = 2048;
  __shared__ u64 lk[kCandLds];
  __shared__ int32_t li[kCandLds];
  const bool inl = n <= kCandLds;
  if (inl) {
    for (int q = threadIdx.x; q < n; q += 256) {
      const int32_t i = c[q];
      li[q] = i;
      lk[q] = keys[i];
    }
    __syncthreads();
  }
  auto key_at = [&](int q) -> u64 { return inl ? lk[q] : keys[c[q]]; };
  auto idx_at = [&](int q) -> uint32_t { return inl ? (uint32_t)li[q] : (uint32_t)c[q]; };
  for (int pass = kSelFull; pass < kSelPasses; ++pass) {
    const int shift = sel_shift(pass), width = sel_width(pass);
    const u64 dm = (1ull << width) - 1ull;
    for (int d = threadIdx.x; d < kSelBins; d += 256) h[d] = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += 256) {
      const u64 k = key_at(q);
      if ((k & s.mask) == s.prefix) atomicAdd(&h[(k >> shift) & dm], 1u);
    }
    __syncthreads();
    long long above;
    const int pick = blk_pick_bin(h, s.krem, true, &above);
    s.prefix |= ((u64)pick) << shift;
    s.mask |= dm << shift;
    s.krem -= above;
  }
  // ties: the keys equal to the threshold; the first kept one has tie rank th = ties - krem
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  int nt = 0;
  for (int q = threadIdx.x; q < n; q += 256) nt += key_at(q) == s.prefix ? 1 : 0;
  atomicAdd(&s_n, nt);
  __syncthreads();
  long long r = s_n - s.krem;  // 0 <= r < ties
  // the r-th smallest row index among the ties, 11 + 11 + 9 bits (row indices < 2^31)
  uint32_t ip = 0, im = 0;
  for (int pass = 0; pass < 3; ++pass) {
    const int shift = pass == 0 ? 20 : (pass == 1 ? 9 : 0);
    const int width = pass == 2 ? 9 : 11;
    const uint32_t dm = (1u << width) - 1u;
    for (int d = threadIdx.x; d < kSelBins; d += 256) h[d] = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += 256) {
      const uint32_t i = idx_at(q);
      if (key_at(q) == s.prefix && (i & im) == ip) atomicAdd(&h[(i >> shift) & dm], 1u);
    }
    __syncthreads();
    long long below;
    const int pick = blk_pick_bin(h, r, false, &below);
    ip |= (uint32_t)pick << shift;
    im |= dm << shift;
    r -= below;
  }
  if (threadIdx.x == 0) {
    st[t].prefix = s.prefix;
    st[t].mask = s.mask;
    st[t].krem = s.krem;
    cut[t] = (int32_t)ip;
  }
}

__device__ __forceinline__ bool bsel_kept(u64 k, int64_t i, const SegSel& s, int32_t cut) {
  return s.all || k > s.prefix || (k == s.prefix && i >= (int64_t)cut);
}

// per tile: the rows kept (tiles past the last trajectory keep nothing)
__global__ void __launch_bounds__(kBlock) k_bsel_count(const u64* __restrict__ keys,
                                                       const int32_t* __restrict__ seg_off,
                                                       const int64_t* __restrict__ boff, int ntraj,
                                                       const SegSel* __restrict__ st,
                                                       const int32_t* __restrict__ cut,
                                                       int32_t* __restrict__ kc) {
  __shared__ int sg[kBlock / 64];
  const SegTile T = seg_tile(seg_off, boff, ntraj, kSelTile);
  if (T.t < 0) {
    if (threadIdx.x == 0) kc[blockIdx.x] = 0;
    return;
  }
  const SegSel s = st[T.t];
  const int32_t cu = cut[T.t];
  u64 k[kSelRpt];
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q) {
    const int64_t i = T.r0 + q * kBlock + threadIdx.x;
    k[q] = keys[i < T.r1 ? i : T.r0];
  }
  int ng = 0;
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q) {
    const int64_t i = T.r0 + q * kBlock + threadIdx.x;
    ng += (i < T.r1 && bsel_kept(k[q], i, s, cu)) ? 1 : 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) ng += __shfl_xor(ng, o, 64);
  if ((threadIdx.x & 63) == 0) sg[threadIdx.x >> 6] = ng;
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) a += sg[w];
    kc[blockIdx.x] = a;
  }
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_bsel_emit(const u64* __restrict__ keys,
                                                      const u64* __restrict__ rows,
                                                      const int32_t* __restrict__ seg_off,
                                                      const int64_t* __restrict__ boff, int ntraj,
                                                      const SegSel* __restrict__ st,
                                                      const int32_t* __restrict__ cut,
                                                      const int32_t* __restrict__ KO,
                                                      u64* __restrict__ out) {
  constexpr int kWaves = kBlock / 64;
  __shared__ int wt[kSelRpt * kWaves];
  const SegTile T = seg_tile(seg_off, boff, ntraj, kSelTile);
  if (T.t < 0) return;
  const SegSel s = st[T.t];
  const int32_t cu = cut[T.t];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  u64 k[kSelRpt];
  Row<W> x[kSelRpt];
  bool keep[kSelRpt];
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q) {
    const int64_t i = T.r0 + q * kBlock + threadIdx.x;
    const int64_t ic = i < T.r1 ? i : T.r0;
    k[q] = keys[ic];
    x[q] = load_row<W>(rows, ic);
  }
  int pre[kSelRpt];
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q) {
    const int64_t i = T.r0 + q * kBlock + threadIdx.x;
    keep[q] = i < T.r1 && bsel_kept(k[q], i, s, cu);
    const u64 bl = __ballot(keep[q]);
    pre[q] = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bl >> 32),
                                            __builtin_amdgcn_mbcnt_lo((unsigned)bl, 0u));
    if (lane == 0) wt[q * kWaves + wave] = __popcll(bl);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int u = 0; u < kSelRpt * kWaves; ++u) {
      const int c = wt[u];
      wt[u] = run;
      run += c;
    }
  }
  __syncthreads();
  const int ko = KO[blockIdx.x];
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q)
    if (keep[q]) store_row<W>(out, ko + wt[q * kWaves + wave] + pre[q], x[q]);
}

// the same decisions as k_bsel_emit, written as one flag byte per row (1 kept, 0 not), with bit
// 7 set where in_be (optional) is nonzero
__global__ void __launch_bounds__(kBlock) k_bsel_flags(const u64* __restrict__ keys,
                                                       const int32_t* __restrict__ seg_off,
                                                       const int64_t* __restrict__ boff, int ntraj,
                                                       const SegSel* __restrict__ st,
                                                       const int32_t* __restrict__ cut,
                                                       const uint8_t* __restrict__ in_be,
                                                       uint8_t* __restrict__ flags) {
  const SegTile T = seg_tile(seg_off, boff, ntraj, kSelTile);
  if (T.t < 0) return;
  const SegSel s = st[T.t];
  const int32_t cu = cut[T.t];
  u64 k[kSelRpt];
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q) {
    const int64_t i = T.r0 + q * kBlock + threadIdx.x;
    k[q] = keys[i < T.r1 ? i : T.r0];
  }
#pragma unroll
  for (int q = 0; q < kSelRpt; ++q) {
    const int64_t i = T.r0 + q * kBlock + threadIdx.x;
    if (i < T.r1)
      flags[i] = (uint8_t)((bsel_kept(k[q], i, s, cu) ? 1 : 0) |
                           ((in_be && in_be[i]) ? 0x80 : 0));
  }
}

size_t bsel_ws_bytes(int64_t R, int ntraj) {
  const int64_t nt = seg_grid(R, ntraj, kSelTile);
  return a256((size_t)ntraj * sizeof(SegSel)) + a256((size_t)ntraj * kSelBins * 4) +
         a256((size_t)(ntraj + 1) * 8) + a256((size_t)R * 8) + a256((size_t)R * 4) +
         2 * a256((size_t)ntraj * 4) + a256((size_t)nt * 4) + a256((size_t)(nt + 1) * 4) +
         a256(scan_ws_bytes(nt));
}

// W == 0: the decisions as flag bytes (out_flags), no rows read or written
template <int W>
static int run_bselect(const double* psi, const uint64_t* rows, int64_t R,
                       const int32_t* seg_off, int ntraj, const int64_t* kpt, uint64_t* out_rows,
                       int32_t* out_count, void* ws, hipStream_t st, uint8_t* out_flags = nullptr,
                       const uint8_t* in_be = nullptr, const uint64_t* keys_in = nullptr) {
  const unsigned gsh = seg_grid(R, ntraj, kSelTile);
  char* p = (char*)ws;
  SegSel* s = (SegSel*)p;
  p += a256((size_t)ntraj * sizeof(SegSel));
  unsigned* hist = (unsigned*)p;
  p += a256((size_t)ntraj * kSelBins * 4);
  int64_t* boffh = (int64_t*)p;  // workgroup offsets of the kSelTile-row tiles
  p += a256((size_t)(ntraj + 1) * 8);
  u64* keys = (u64*)p;
  p += a256((size_t)R * 8);
  int32_t* ci = (int32_t*)p;  // candidates' row indices, trajectory t's at seg_off[t]
  p += a256((size_t)R * 4);
  int32_t* ccount = (int32_t*)p;
  p += a256((size_t)ntraj * 4);
  int32_t* cut = (int32_t*)p;
  p += a256((size_t)ntraj * 4);
  int32_t* kc = (int32_t*)p;
  p += a256((size_t)gsh * 4);
  int32_t* KO = (int32_t*)p;
  p += a256((size_t)(gsh + 1) * 4);
  hipLaunchKernelGGL(k_seg_blocks, dim3(1), dim3(kSegThreads), 0, st, seg_off, ntraj, kSelTile,
                     boffh);
  if (keys_in)
    keys = (u64*)keys_in;  // the keys of these amplitudes, formed by the observables pass
  else if (R > 0)
    hipLaunchKernelGGL(k_bsel_keys, dim3(grid_for(R, kBlock)), dim3(kBlock), 0, st,
                       (const double2*)psi, R, keys);
  hipLaunchKernelGGL(k_bsel_init, dim3(ntraj), dim3(256), 0, st, seg_off, kpt, ntraj, s, hist,
                     ccount, out_count);
  for (int pass = 0; pass < kSelFull; ++pass) {
    const int shift = sel_shift(pass), width = sel_width(pass);
    hipLaunchKernelGGL(k_bsel_hist, dim3(gsh), dim3(kBlock), 0, st, (const u64*)keys, seg_off,
                       (const int64_t*)boffh, ntraj, (const SegSel*)s, shift, width, hist);
    hipLaunchKernelGGL(k_bsel_pick, dim3(ntraj), dim3(256), 0, st, s, hist, shift, width);
  }
  hipLaunchKernelGGL(k_bsel_cand, dim3(gsh), dim3(kBlock), 0, st, (const u64*)keys, seg_off,
                     (const int64_t*)boffh, ntraj, (const SegSel*)s, ci, ccount);
  hipLaunchKernelGGL(k_bsel_final, dim3(ntraj), dim3(256), 0, st, s, (const u64*)keys,
                     (const int32_t*)ci, (const int32_t*)ccount, seg_off, kpt, cut, out_count);
  if constexpr (W == 0) {
    hipLaunchKernelGGL(k_bsel_flags, dim3(gsh), dim3(kBlock), 0, st, (const u64*)keys, seg_off,
                       (const int64_t*)boffh, ntraj, (const SegSel*)s, (const int32_t*)cut,
                       in_be, out_flags);
  } else {
    hipLaunchKernelGGL(k_bsel_count, dim3(gsh), dim3(kBlock), 0, st, (const u64*)keys, seg_off,
                       (const int64_t*)boffh, ntraj, (const SegSel*)s, (const int32_t*)cut, kc);
    int rc = launch_status();
    if (rc) return rc;
    rc = scan_i32(kc, gsh, KO, p, scan_ws_bytes(gsh), st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bsel_emit<W>, dim3(gsh), dim3(kBlock), 0, st, (const u64*)keys,
                       (const u64*)rows, seg_off, (const int64_t*)boffh, ntraj,
                       (const SegSel*)s, (const int32_t*)cut, (const int32_t*)KO, (u64*)out_rows);
  }
  return launch_status();
}

// ---- per-trajectory observables (Holstein, batched layout: field 0 trajectory, 1 position) -------

constexpr int kObsTileB = 256;  // the smallest tile (sizes the workspace)

template <int W>
constexpr int obs_tile() {
  // more rows per lane spread each site's wave reduction over more rows
  return W <= 3 ? 1024 : (W <= 6 ? 512 : kObsTileB);
}

// Partial sums of one tile (rows of one trajectory): the sum of weights, the weight at each
// exciton site and the weighted phonons at each site. Wave w takes sites w, w + 4, ..., so a
// site's field offset and width are uniform across the wave; lane l holds rows l, l + 64, ...
// with their weights and exciton positions in registers, and a site's sum ends in a wave
// reduction. Only sites inside the tile's range of exciton positions can hold exciton weight.
// With fin (the step's series state ts and its term buffers): the finish of the series, as
// k_btay_finish does it, with psi updated in place: mixed, psi (holding v) gets every term of
// its trajectory added in order; FP64, a trajectory whose series ended on an even term gets that
// term added. keys (optional) receives each row's select key, the bits of |psi|^2 (k_bsel_keys'
// expression), for the next select.
// The buffers of a series' terms, for its finish: FP64, t[0] is the even-term buffer; mixed,
// t[k] is term k's buffer (n of them)
struct SeriesTerms {
  const void* t[PACES_SERIES_MAX_TERMS];
  int n;
};

struct ObsFinish {
  SeriesTerms terms;
  int k0;  // mixed: the first term to add (a later k0 completes an earlier finish of psi)
  int mixed;
  const paces_batch_taylor_t* ts;
  u64* keys;
  u64* pmax;  // with vmax: each tile's largest key (workspace), reduced per trajectory into vmax
  u64* vmax;
};

static bool series_terms(const void* const* terms, int32_t nterms, SeriesTerms* st) {
  if (!terms || nterms < 1 || nterms > PACES_SERIES_MAX_TERMS) return false;
  for (int k = 0; k < PACES_SERIES_MAX_TERMS; ++k) {
    st->t[k] = k < nterms ? terms[k] : nullptr;
    if (k < nterms && !terms[k]) return false;
  }
  st->n = nterms;
  return true;
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_bobs_partial(const u64* __restrict__ rows,
                                                         double2* __restrict__ psi, ObsFinish fin,
                                                         const int32_t* __restrict__ seg_off,
                                                         const int64_t* __restrict__ boff,
                                                         int ntraj, const int32_t* __restrict__ off,
                                                         const int32_t* __restrict__ width, int N,
                                                         const int32_t* __restrict__ traj_L, int geo,
                                                         double* __restrict__ part) {
  constexpr int TB = obs_tile<W>(), RPL = TB / 64, PER = TB * W / kBlock, TPR = TB / kBlock;
  constexpr int NW = kBlock / 64;
  static_assert(TB % kBlock == 0, "whole rows per thread");
  __shared__ u64 srow[TB * W];
  __shared__ double sw[TB];
  __shared__ short spos[TB];  // positions < PACES_MAX_SITES; -1 past the tile's rows
  __shared__ int spr[2];
  __shared__ double snorm[NW];
  __shared__ u64 smax[NW];
  const SegTile T = seg_tile(seg_off, boff, ntraj, TB);
  if (T.t < 0) return;
  if (threadIdx.x == 0) {
    spr[0] = 0x7fffffff;
    spr[1] = -1;
  }
  const int S = 2 * N + 1;
  const int nr = (int)(T.r1 - T.r0);
  // the wave index is uniform: the site loop's field offsets then come from scalar loads and
  // its branches are scalar
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  {
    // the rows to LDS
    u64 tmp[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int e = (int)threadIdx.x + j * kBlock;
      tmp[j] = e < nr * W ? rows[T.r0 * W + e] : 0ull;
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int e = (int)threadIdx.x + j * kBlock;
      if (e < nr * W) srow[e] = tmp[j];
    }
  }
  {
    // the amplitudes, each row by one thread (rows t, t + kBlock, ...), with the series'
    // finish first (each row read and written once); the weights go through LDS to the lanes
    // that hold their rows in the site loop below
    const int jt = fin.ts ? fin.ts[T.t].j_ind : 0;
    const bool add = fin.ts && (fin.mixed ? jt >= fin.k0 : (jt & 1) == 0);
    double2 v[TPR];
    int64_t ii[TPR];
#pragma unroll
    for (int q = 0; q < TPR; ++q) {
      const int r = (int)threadIdx.x + kBlock * q;
      ii[q] = T.r0 + (r < nr ? r : 0);
      v[q] = psi[ii[q]];
    }
    if (add) {
      if (fin.mixed) {
        // psi holds v + t_0 + ... + t_(k0-1): t_k0, ..., t_jt added in order
        const int nk = jt + 1 < fin.terms.n ? jt + 1 : fin.terms.n;
#pragma unroll 4
        for (int k = fin.k0; k < nk; ++k) {
          const float2* x = (const float2*)fin.terms.t[k];
          float2 xv[TPR];
#pragma unroll
          for (int q = 0; q < TPR; ++q) xv[q] = x[ii[q]];
#pragma unroll
          for (int q = 0; q < TPR; ++q) {
            v[q].x = v[q].x + (double)xv[q].x;
            v[q].y = v[q].y + (double)xv[q].y;
          }
        }
      } else {
        const double2* x = (const double2*)fin.terms.t[0];
#pragma unroll
        for (int q = 0; q < TPR; ++q) {
          const double2 xv = x[ii[q]];
          v[q].x = v[q].x + xv.x;
          v[q].y = v[q].y + xv.y;
        }
      }
#pragma unroll
      for (int q = 0; q < TPR; ++q)
        if ((int)threadIdx.x + kBlock * q < nr) psi[ii[q]] = v[q];
    }
    double a = 0.0;
    u64 mx = 0;
#pragma unroll
    for (int q = 0; q < TPR; ++q) {
      const int r = (int)threadIdx.x + kBlock * q;
      const double k = v[q].x * v[q].x + v[q].y * v[q].y;
      const u64 kb = (u64)__double_as_longlong(k);
      if (fin.keys && r < nr) fin.keys[ii[q]] = kb;
      sw[r] = r < nr ? k : 0.0;
      a += r < nr ? k : 0.0;
      mx = (r < nr && kb > mx) ? kb : mx;  // keys of non-negative doubles order as the values
    }
    a = wave_sum(a);
    mx = wave_max_u64(mx);
    if (lane == 0) {
      snorm[wave] = a;
      smax[wave] = mx;
    }
  }
  __syncthreads();
  {
    // each row's exciton position (field 1, uniform offset), once per row, and the tile's range
    const int po = __builtin_amdgcn_readfirstlane(off[1]);
    const int pb = __builtin_amdgcn_readfirstlane(width[1]);
    const int pk = po >> 6, pend = (po & 63) + pb;
    const u64 pmask = (pb >= 64) ? ~0ull : ((1ull << pb) - 1ull);
    int lo = 0x7fffffff, hi = -1;
#pragma unroll
    for (int q = 0; q < TPR; ++q) {
      const int r = (int)threadIdx.x + kBlock * q;
      int p = -1;
      if (r < nr) {
        const u64 w0 = srow[r * W + pk];
        u64 x;
        if (pend <= 64) {
          x = w0 >> (64 - pend);
        } else {  // straddles words pk and pk + 1 (the layout keeps pk + 1 < W)
          const u64 w1 = srow[r * W + pk + 1];
          x = (w0 << (pend - 64)) | (w1 >> (128 - pend));
        }
        p = (int)(x & pmask);
        lo = p < lo ? p : lo;
        hi = p > hi ? p : hi;
      }
      spos[r] = (short)p;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const int l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
      lo = l2 < lo ? l2 : lo;
      hi = h2 > hi ? h2 : hi;
    }
    if (lane == 0 && hi >= 0) {
      atomicMin(&spr[0], lo);
      atomicMax(&spr[1], hi);
    }
  }
  __syncthreads();
  if (wave == 0 && lane == 0) {
    double a = 0.0;
    u64 mx = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      a += snorm[k];
      mx = smax[k] > mx ? smax[k] : mx;
    }
    part[(int64_t)blockIdx.x * S + 2 * N] = a;
    if (fin.vmax) fin.pmax[blockIdx.x] = mx;
  }
  double w[RPL];
#pragma unroll
  for (int j = 0; j < RPL; ++j) w[j] = sw[lane + 64 * j];
  const int pmin = spr[0], pmax = spr[1];
  // sites past the tile's own lattice hold no phonons and no exciton: their partials are zero
  const int L = traj_L[T.t];
  const int Nt = geo == 1 ? L : (geo == 2 ? L * L : L * L * L);
  for (int s = wave; s < N; s += NW) {
    if (s >= Nt) {
      if (lane == 0) {
        part[(int64_t)blockIdx.x * S + s] = 0.0;
        part[(int64_t)blockIdx.x * S + N + s] = 0.0;
      }
      continue;
    }
    const int fo = off[s + 2], b = width[s + 2];
    const int k = fo >> 6, o = fo & 63;
    double ap = 0.0;
    if (b > 0 && o + b > 64) {  // the field straddles two words
      const u64 mask = (b >= 64) ? ~0ull : ((1ull << b) - 1ull);
#pragma unroll
      for (int j = 0; j < RPL; ++j) {
        const int r = lane + 64 * j;
        const u64 n = ((srow[r * W + k] << (o + b - 64)) | (srow[r * W + k + 1] >> (128 - o - b))) &
                      mask;
        ap = fma(w[j], (double)n, ap);
      }
    } else if (b > 0 && b <= 32) {
      const uint32_t mask = (b >= 32) ? ~0u : ((1u << b) - 1u);
      const uint32_t* s32 = (const uint32_t*)srow;
      if (o + b <= 32 || o >= 32) {
        // the field lies in one 32-bit half of its word: a 32-bit LDS read and a bit-field
        // extract per row, one FMA (the sums are compared within 1e-10, not bit for bit)
        const int half = o >= 32 ? 0 : 1;  // little endian: the high half is at byte 4
        const int sh = (half ? 32 : 64) - o - b;
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          const int r = lane + 64 * j;
          const uint32_t n = (s32[(r * W + k) * 2 + half] >> sh) & mask;
          ap = fma(w[j], (double)n, ap);
        }
      } else {
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          const int r = lane + 64 * j;
          ap = fma(w[j], (double)(uint32_t)((srow[r * W + k] >> (64 - o - b)) & mask), ap);
        }
      }
    } else if (b > 0) {
      const u64 mask = (b >= 64) ? ~0ull : ((1ull << b) - 1ull);
#pragma unroll
      for (int j = 0; j < RPL; ++j) {
        const int r = lane + 64 * j;
        ap = fma(w[j], (double)((srow[r * W + k] >> (64 - o - b)) & mask), ap);
      }
    }
    ap = wave_sum(ap);
    double ae = 0.0;
    if (s >= pmin && s <= pmax) {
#pragma unroll
      for (int j = 0; j < RPL; ++j) ae += (spos[lane + 64 * j] == s) ? w[j] : 0.0;
      ae = wave_sum(ae);
    }
    if (lane == 0) {
      part[(int64_t)blockIdx.x * S + s] = ae;
      part[(int64_t)blockIdx.x * S + N + s] = ap;
    }
  }
}

// one wavefront per (trajectory, output): sum of the trajectory's workgroup partials
__global__ void __launch_bounds__(64) k_bobs_final(const double* __restrict__ part,
                                                   const int64_t* __restrict__ boff, int S,
                                                   double* __restrict__ out) {
  const int t = blockIdx.x / S, s = blockIdx.x % S;
  double a = 0.0;
  for (int64_t b = boff[t] + threadIdx.x; b < boff[t + 1]; b += 64) a += part[b * S + s];
  a = wave_sum(a);
  if (threadIdx.x == 0) out[(int64_t)t * S + s] = a;
}

size_t bobs_ws_bytes(int64_t R, int ntraj, int N) {
  int64_t nblk = (R + kObsTileB - 1) / kObsTileB + ntraj + 1;
  return a256((size_t)(ntraj + 1) * 8) + a256((size_t)nblk * (2 * N + 1) * 8) +
         a256((size_t)nblk * 8);
}

// one wavefront per trajectory: the largest key of its tiles (0 for an empty trajectory)
__global__ void __launch_bounds__(64) k_bobs_vmax(const u64* __restrict__ pmax,
                                                  const int64_t* __restrict__ boff,
                                                  u64* __restrict__ vmax) {
  const int t = blockIdx.x;
  u64 m = 0;
  for (int64_t b = boff[t] + threadIdx.x; b < boff[t + 1]; b += 64) m = pmax[b] > m ? pmax[b] : m;
  m = wave_max_u64(m);
  if (threadIdx.x == 0) vmax[t] = m;
}

template <int W>
static int run_bobs(const uint64_t* rows, double* psi, int64_t R, const int32_t* seg_off,
                    int ntraj, const paces_layout_t* lay, int N, const int32_t* traj_L, int geo,
                    double* out, void* ws, hipStream_t st, ObsFinish fin) {
  char* p = (char*)ws;
  int64_t* boff = (int64_t*)p;
  p += a256((size_t)(ntraj + 1) * 8);
  double* part = (double*)p;
  const int S = 2 * N + 1;
  constexpr int TB = obs_tile<W>();
  const int64_t nblk = (R + kObsTileB - 1) / kObsTileB + ntraj + 1;
  p += a256((size_t)nblk * S * 8);
  fin.pmax = (u64*)p;
  hipLaunchKernelGGL(k_seg_blocks, dim3(1), dim3(kSegThreads), 0, st, seg_off, ntraj, TB, boff);
  hipLaunchKernelGGL(k_bobs_partial<W>, dim3(seg_grid(R, ntraj, TB)), dim3(kBlock), 0, st,
                     (const u64*)rows, (double2*)psi, fin, seg_off, (const int64_t*)boff, ntraj,
                     lay->d_off, lay->d_width, N, traj_L, geo, part);
  hipLaunchKernelGGL(k_bobs_final, dim3(ntraj * S), dim3(64), 0, st, (const double*)part,
                     (const int64_t*)boff, S, out);
  if (fin.vmax)
    hipLaunchKernelGGL(k_bobs_vmax, dim3(ntraj), dim3(64), 0, st, (const u64*)fin.pmax,
                       (const int64_t*)boff, fin.vmax);
  return launch_status();
}

// ---- batched Taylor series with a per-trajectory stopping test --------------------------------------

struct CsrSetB64 {
  int nterms;
  const int32_t* indptr[PACES_MAX_TERMS];
  const int32_t* indices[PACES_MAX_TERMS];
  const void* data[PACES_MAX_TERMS];
  // coded values (paces_csr_coded_t): data unused
  int coded;
  int sqrt_len;
  const uint8_t* code[PACES_MAX_TERMS];
  const double* term_value;
  const double* sqrt_table;
  // slot matrix (paces_slots_t, coded values): indptr, indices and code unused
  const uint32_t* slots;  // slot m of row i at slots[m * slot_stride + i]
  int64_t slot_stride;
  int nslots;
  int col_bits;
  int slot_begin;  // bit m set: slot m >= 1 is the first slot of a term
  int slot_vib[PACES_MAX_TERMS];
};

// term t takes coded values of the form sqrt_table[m] * v (a vib term)
__device__ __forceinline__ bool coded_vib(const CsrSetB64& H, int t) {
  return H.slots ? H.slot_vib[t] != 0 : H.code[t] != nullptr;
}

// Coded entry values: a vib entry's raw load is its code byte (an absent entry reads a byte of
// the term's indptr; the mask keeps any byte inside the table), a hopping entry loads nothing;
// the value is the workgroup's LDS table entry of the row's trajectory.
template <typename F>
struct CodedVals {
  typedef uint32_t raw_t;
  const F* vt;  // LDS: [term][code]
  __device__ __forceinline__ uint32_t load(const CsrSetB64& H, int t, int32_t k, bool ok) const {
    const uint8_t* c = H.code[t];
    if (!c) return 0u;
    const uint8_t* pc = ok ? c + k : (const uint8_t*)H.indptr[t];
    return (uint32_t)*pc & (PACES_CODE_TABLE - 1);
  }
  __device__ __forceinline__ F value(int t, uint32_t raw) const {
    return vt[t * PACES_CODE_TABLE + raw];
  }
};

template <typename F>
struct Vec2;
template <>
struct Vec2<double> {
  typedef double2 T;
};
template <>
struct Vec2<float> {
  typedef float2 T;
};

template <typename F, typename VL>
__device__ __forceinline__ typename Vec2<F>::T hx_row_t(int64_t i, const F* __restrict__ diag,
                                                        const CsrSetB64& H,
                                                        const typename Vec2<F>::T* __restrict__ x,
                                                        typename Vec2<F>::T& xi_out,
                                                        const VL& vl) {
  typedef typename Vec2<F>::T V;
  const V xi = x[i];
  xi_out = xi;
  const F d = diag[i];
  F yr = d * xi.x, yi = d * xi.y;
  // unrolled over the compile-time maximum so every H field index is static (registers, not
  // scratch, when H is built inside the kernel); the order of the sums is the term order
#pragma unroll
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    if (t >= H.nterms) break;
    const int32_t b = H.indptr[t][i], e = H.indptr[t][i + 1];
    F tr = 0, ti = 0;
    for (int32_t k = b; k < e; ++k) {
      const F a = vl.value(t, vl.load(H, t, k, true));
      const V xj = x[H.indices[t][k]];
      tr = tr + a * xj.x;
      ti = ti + a * xj.y;
    }
    yr = yr + tr;
    yi = yi + ti;
  }
  V o;
  o.x = yr;
  o.y = yi;
  return o;
}

template <int BS>
__device__ __forceinline__ void blk_max2(u64 a, u64 b, u64* __restrict__ part, int64_t tile) {
  __shared__ u64 sa[BS / 64], sb[BS / 64];
  a = wave_max_u64(a);
  b = wave_max_u64(b);
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 ma = 0, mb = 0;
#pragma unroll
    for (int q = 0; q < BS / 64; ++q) {
      ma = sa[q] > ma ? sa[q] : ma;
      mb = sb[q] > mb ? sb[q] : mb;
    }
    part[2 * tile] = ma;
    part[2 * tile + 1] = mb;
  }
}

// Launch shape (tuned): BS threads per workgroup, RPT rows per thread (t, t + BS, ...), a
// segment-aligned tile of BS * RPT rows; begin and every term of one series use the same shape.
// WRITE_RES false (the graph's series): res is left unwritten; the first odd term reads v in its
// place (res_src of btay_term_body)
template <typename F, int BS, int RPT, bool WRITE_RES = true>
__device__ __forceinline__ void btay_begin_body(const double2* __restrict__ v,
                                                double2* __restrict__ res,
                                                typename Vec2<F>::T* __restrict__ x0,
                                                const int32_t* __restrict__ seg_off,
                                                const int64_t* __restrict__ boff, int ntraj,
                                                u64* __restrict__ part) {
  const SegTile T = seg_tile(seg_off, boff, ntraj, BS * RPT);
  if (T.t < 0) return;
  u64 m = 0;
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int64_t i = T.r0 + q * BS + threadIdx.x;
    if (i < T.r1) {
      double2 a = v[i];
      if (WRITE_RES) res[i] = a;
      typename Vec2<F>::T c;
      c.x = (F)a.x;
      c.y = (F)a.y;
      x0[i] = c;
      const u64 w = (u64)__double_as_longlong(a.x * a.x + a.y * a.y);
      m = w > m ? w : m;
    }
  }
  blk_max2<BS>(m, 0, part, T.b);
}

// per-trajectory reduction of the partials; fin == 1: initialise c1, else the stopping test
// use_prev: the term formed no partial sum (a mixed-form term): divide by ||v||_inf (rn as the
// begin set it) instead of this term's ||res||_inf
__device__ __forceinline__ void btay_reduce_body(const u64* __restrict__ part,
                                                 const int64_t* __restrict__ boff,
                                                 paces_batch_taylor_t* __restrict__ ts,
                                                 paces_batch_taylor_hdr_t* __restrict__ hdr,
                                                 int fin, int j, double tol, int use_prev) {
  const int t = blockIdx.x;
  if (!fin && ts[t].done) return;
  u64 a = 0, b = 0;
  for (int64_t q = boff[t] + threadIdx.x; q < boff[t + 1]; q += 64) {
    u64 x = part[2 * q], y = part[2 * q + 1];
    a = x > a ? x : a;
    b = y > b ? y : b;
  }
  a = wave_max_u64(a);
  b = wave_max_u64(b);
  if (threadIdx.x != 0) return;
  const double m0 = sqrt(__longlong_as_double((long long)a));
  if (fin) {
    ts[t].c1 = m0;
    ts[t].rn = m0;  // res = v
    ts[t].c1_plus_c2 = 0.0;
    ts[t].ratio = 0.0;
    ts[t].done = (boff[t + 1] == boff[t]) ? 1 : 0;  // an empty trajectory has nothing to do
    ts[t].j_ind = 0;
    return;
  }
  const double rn = use_prev ? ts[t].rn : sqrt(__longlong_as_double((long long)b));
  if (!use_prev) ts[t].rn = rn;
  const double c1pc2 = ts[t].c1 + m0;
  const double ratio = c1pc2 / rn;
  ts[t].c1_plus_c2 = c1pc2;
  ts[t].ratio = ratio;
  ts[t].j_ind = j;
  if (ratio <= tol) {
    ts[t].done = 1;
  } else {
    ts[t].c1 = m0;
    atomicAdd(&hdr->notdone, 1);
  }
}

// The term's output for one row: t = -i c (H.x)_i, the partial sum (res, plus x_i on odd terms,
// plus t; stored on odd terms) and the running maxima of |t|^2 and |res|^2. Without with_res (a
// mixed-form even term) the partial sum is neither read nor formed.
template <typename F>
__device__ __forceinline__ void btay_row_out(int64_t i, typename Vec2<F>::T h,
                                             typename Vec2<F>::T xi, double2 r, F c, bool odd,
                                             bool with_res,
                                             typename Vec2<F>::T* __restrict__ xout,
                                             double2* __restrict__ res, u64& mt, u64& mr) {
  typename Vec2<F>::T t;
  t.x = h.y * c;
  t.y = -(h.x * c);
  xout[i] = t;
  if (!with_res) {
    const double tx = t.x, ty = t.y;
    const u64 a = (u64)__double_as_longlong(tx * tx + ty * ty);
    mt = a > mt ? a : mt;
    return;
  }
  if (odd) {
    r.x = r.x + (double)xi.x;
    r.y = r.y + (double)xi.y;
  }
  r.x = r.x + (double)t.x;
  r.y = r.y + (double)t.y;
  if (odd) res[i] = r;
  const double tx = t.x, ty = t.y;
  const u64 a = (u64)__double_as_longlong(tx * tx + ty * ty);
  const u64 b = (u64)__double_as_longlong(r.x * r.x + r.y * r.y);
  mt = a > mt ? a : mt;
  mr = b > mr ? b : mr;
}

// The batched term's rows of one thread: the partial sums are loaded first (independent of H.x),
// then H.x in load phases (hx_rows_phased), then each row's output.
template <typename F, int BS, int RPT, int NT, bool WITH_RES, typename VL>
__device__ __forceinline__ void btay_rows_phased(const F* __restrict__ diag, const CsrSetB64& H,
                                                 const typename Vec2<F>::T* __restrict__ xin,
                                                 typename Vec2<F>::T* __restrict__ xout,
                                                 double2* __restrict__ res,
                                                 const double2* __restrict__ res_src, F c,
                                                 bool odd, const SegTile& T, u64& mt, u64& mr,
                                                 const VL& vl) {
  typedef typename Vec2<F>::T V;
  int64_t ii[RPT];
  bool vv[RPT];
  double2 rr[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    ii[q] = T.r0 + q * BS + threadIdx.x;
    vv[q] = ii[q] < T.r1;
    if (WITH_RES) rr[q] = res_src[vv[q] ? ii[q] : T.r0];
  }
  V h[RPT], xi[RPT];
  hx_rows_phased<F, V, RPT, NT>(ii, vv, T.r0, diag, H, xin, h, xi, vl);
#pragma unroll
  for (int q = 0; q < RPT; ++q)
    if (vv[q]) btay_row_out<F>(ii[q], h[q], xi[q], rr[q], c, odd, WITH_RES, xout, res, mt, mr);
}

// A tile's rows of one term: the phased form for one or two Hamiltonian terms, else row by row.
// WITH_RES false (a mixed-form even term): the partial sum is neither read nor formed; the
// stopping test divides by the last exact ||res||_inf (btay_reduce_body's use_prev).
template <typename F, int BS, int RPT, bool WITH_RES, typename VL>
__device__ __forceinline__ void btay_rows_all(const F* __restrict__ diag, const CsrSetB64& H,
                                              const typename Vec2<F>::T* __restrict__ xin,
                                              typename Vec2<F>::T* __restrict__ xout,
                                              double2* __restrict__ res,
                                              const double2* __restrict__ res_src, F c, bool odd,
                                              const SegTile& T, u64& mt, u64& mr, const VL& vl) {
  // the phased form needs about 40 VGPRs per FP64 row: not at 1024 threads x 4 rows (128 max)
  constexpr bool phased = !(sizeof(F) == 8 && BS * RPT > 2048);
  if (phased && H.nterms == 2) {
    btay_rows_phased<F, BS, RPT, 2, WITH_RES>(diag, H, xin, xout, res, res_src, c, odd, T, mt,
                                              mr, vl);
  } else if (phased && H.nterms == 1) {
    btay_rows_phased<F, BS, RPT, 1, WITH_RES>(diag, H, xin, xout, res, res_src, c, odd, T, mt,
                                              mr, vl);
  } else {
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int64_t i = T.r0 + q * BS + threadIdx.x;
      if (i < T.r1) {
        typename Vec2<F>::T xi;
        const typename Vec2<F>::T h = hx_row_t<F>(i, diag, H, xin, xi, vl);
        double2 r = {0.0, 0.0};
        if (WITH_RES) r = res_src[i];
        btay_row_out<F>(i, h, xi, r, c, odd, WITH_RES, xout, res, mt, mr);
      }
    }
  }
}

template <typename F, int BS, int RPT, typename VL>
__device__ __forceinline__ void btay_rows_par(const F* __restrict__ diag, const CsrSetB64& H,
                                              const typename Vec2<F>::T* __restrict__ xin,
                                              typename Vec2<F>::T* __restrict__ xout,
                                              double2* __restrict__ res,
                                              const double2* __restrict__ res_src, F c, bool odd,
                                              const SegTile& T, u64& mt, u64& mr, const VL& vl) {
  if (sizeof(F) == 8)
    btay_rows_all<F, BS, RPT, true>(diag, H, xin, xout, res, res_src, c, odd, T, mt, mr, vl);
  else
    btay_rows_all<F, BS, RPT, false>(diag, H, xin, xout, res, res_src, c, odd, T, mt, mr, vl);
}

// ---- the term over a slot matrix: a row's slots come in one load with its own x and diagonal,
// then every occupied slot's x is gathered (an empty slot gathers a row of the tile and the value
// goes unused), then the sums: the diagonal, then each term's occupied slots in slot order (its
// CSR row, columns ascending), as hx_rows_phased sums a CSR row.
template <int NS>
__device__ __forceinline__ void load_slots(const uint32_t* __restrict__ p, int64_t stride,
                                           int64_t i, uint32_t (&s)[NS]) {
  static_assert(NS == 4 || NS == 8, "slots per row");
#pragma unroll
  for (int m = 0; m < NS; ++m) s[m] = p[m * stride + i];
}

// rows c0 .. c0 + RC - 1 of the thread's RPT rows (t, t + BS, ...)
template <typename F, int BS, int RC, int NS, bool WITH_RES>
__device__ __forceinline__ void btay_rows_slots_chunk(
    int c0, const F* __restrict__ diag, const CsrSetB64& H,
    const typename Vec2<F>::T* __restrict__ xin, typename Vec2<F>::T* __restrict__ xout,
    double2* __restrict__ res, const double2* __restrict__ res_src, F c, bool odd,
    const SegTile& T, u64& mt, u64& mr, const F* vt) {
  typedef typename Vec2<F>::T V;
  const uint32_t colmask = (1u << H.col_bits) - 1u;
  int64_t ii[RC];
  bool vv[RC];
  double2 rr[RC];
  V xi[RC];
  F d[RC];
  uint32_t sw[RC][NS];
#pragma unroll
  for (int q = 0; q < RC; ++q) {
    ii[q] = T.r0 + (c0 + q) * BS + threadIdx.x;
    vv[q] = ii[q] < T.r1;
    const int64_t i = vv[q] ? ii[q] : T.r0;
    if (WITH_RES) rr[q] = res_src[i];
    xi[q] = xin[i];
    d[q] = diag[i];
    load_slots<NS>(H.slots, H.slot_stride, i, sw[q]);
  }
  V xj[RC][NS];
#pragma unroll
  for (int q = 0; q < RC; ++q)
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const bool ok = sw[q][m] != PACES_SLOT_EMPTY;
      xj[q][m] = xin[ok ? (int64_t)(sw[q][m] & colmask) : T.r0];
    }
#pragma unroll
  for (int q = 0; q < RC; ++q) {
    if (!vv[q]) continue;
    F yr = d[q] * xi[q].x, yi = d[q] * xi[q].y;
    F tr = 0, ti = 0;
    int t = 0;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      if ((H.slot_begin >> m) & 1) {  // the next term's slots begin
        yr = yr + tr;
        yi = yi + ti;
        tr = 0;
        ti = 0;
        ++t;
      }
      if (sw[q][m] != PACES_SLOT_EMPTY) {
        const F av = vt[t * PACES_CODE_TABLE + (int)(sw[q][m] >> H.col_bits)];
        tr = tr + av * xj[q][m].x;
        ti = ti + av * xj[q][m].y;
      }
    }
    yr = yr + tr;
    yi = yi + ti;
    V h;
    h.x = yr;
    h.y = yi;
    btay_row_out<F>(ii[q], h, xi[q], rr[q], c, odd, WITH_RES, xout, res, mt, mr);
  }
}

template <typename F, int BS, int RPT, int NS, bool WITH_RES>
__device__ __forceinline__ void btay_rows_slots(const F* __restrict__ diag, const CsrSetB64& H,
                                                const typename Vec2<F>::T* __restrict__ xin,
                                                typename Vec2<F>::T* __restrict__ xout,
                                                double2* __restrict__ res,
                                                const double2* __restrict__ res_src, F c,
                                                bool odd, const SegTile& T, u64& mt, u64& mr,
                                                const F* vt) {
  // rows per load phase: the gathered x of a phase within 64 registers
  constexpr int kFit = 256 / (NS * (int)sizeof(typename Vec2<F>::T));
  constexpr int RC = kFit < 1 ? 1 : (kFit < RPT ? kFit : RPT);
  static_assert(RPT % RC == 0, "row chunks");
#pragma unroll
  for (int c0 = 0; c0 < RPT; c0 += RC)
    btay_rows_slots_chunk<F, BS, RC, NS, WITH_RES>(c0, diag, H, xin, xout, res, res_src, c, odd, T,
                                                   mt, mr, vt);
}

// FP64 (complex128 terms): the partial sum res is stored on odd terms only. On an even term j
// the sum res + t_j is formed in registers for the stopping test; the next (odd) term adds its
// own input (t_j) and then its output, (res + t_j) + t_(j+1), the same roundings in the same
// order as storing every term. A trajectory whose series ends on an even term gets its last term
// added by the finish. Mixed (complex64 terms): no term reads or forms the partial sum; every
// term writes a buffer of its own, the stopping test divides by ||v||_inf, and the finish forms
// v + t_0 + ... + t_j in that order (the roundings of a running sum).
template <typename F, int BS, int RPT, int NS>
__device__ __forceinline__ void btay_term_body(
    const F* __restrict__ diag, const CsrSetB64& H, const typename Vec2<F>::T* __restrict__ xin,
    typename Vec2<F>::T* __restrict__ xout, double2* __restrict__ res,
    const double2* __restrict__ res_src, F c, int j,
    const int32_t* __restrict__ seg_off, const int64_t* __restrict__ boff, int ntraj,
    const paces_batch_taylor_t* __restrict__ ts, paces_batch_taylor_hdr_t* __restrict__ hdr,
    u64* __restrict__ part) {
  if (blockIdx.x == 0 && threadIdx.x == 0) hdr->notdone = 0;
  const SegTile T = seg_tile<true>(seg_off, boff, ntraj, BS * RPT);
  if (T.t < 0 || ts[T.t].done) return;
  const bool odd = (j & 1) != 0;
  u64 mt = 0, mr = 0;
  if (H.coded) {
    // the tile's trajectory's entry values, per term and code (the values the pairs fill
    // computes: sqrt_table[m] * v for a vib term, v for a hopping term)
    // (unrolled over the terms: every H field index is static, so H stays in registers when the
    // kernel builds it from a descriptor)
    __shared__ F vt[PACES_MAX_TERMS * PACES_CODE_TABLE];
    static_assert(BS >= PACES_CODE_TABLE, "one code per thread");
#pragma unroll
    for (int t = 0; t < PACES_MAX_TERMS; ++t) {
      if (t >= H.nterms) break;
      const double v = H.term_value[(int64_t)T.t * H.nterms + t];
      const int m = threadIdx.x;
      if (m < PACES_CODE_TABLE)
        vt[t * PACES_CODE_TABLE + m] =
            (F)(coded_vib(H, t) ? (m < H.sqrt_len ? H.sqrt_table[m] * v : 0.0) : v);
    }
    __syncthreads();
    if constexpr (NS > 0) {
      if (sizeof(F) == 8)
        btay_rows_slots<F, BS, RPT, NS, true>(diag, H, xin, xout, res, res_src, c, odd, T, mt, mr,
                                              vt);
      else
        btay_rows_slots<F, BS, RPT, NS, false>(diag, H, xin, xout, res, res_src, c, odd, T, mt,
                                               mr, vt);
    } else {
      btay_rows_par<F, BS, RPT>(diag, H, xin, xout, res, res_src, c, odd, T, mt, mr,
                                CodedVals<F>{vt});
    }
  } else if constexpr (NS == 0) {
    btay_rows_par<F, BS, RPT>(diag, H, xin, xout, res, res_src, c, odd, T, mt, mr,
                              PlainVals<F>{});
  }
  blk_max2<BS>(mt, mr, part, T.b);
}

// Launch shape (tuned): BS threads per workgroup, RPT rows per thread (t, t + BS, ...), a
// segment-aligned tile of BS * RPT rows; begin and every term of one series use the same shape.
template <typename F, int BS, int RPT>
__global__ void __launch_bounds__(BS) k_btay_begin(const double2* __restrict__ v,
                                                   double2* __restrict__ res,
                                                   typename Vec2<F>::T* __restrict__ x0,
                                                   const int32_t* __restrict__ seg_off,
                                                   const int64_t* __restrict__ boff, int ntraj,
                                                   u64* __restrict__ part) {
  btay_begin_body<F, BS, RPT, sizeof(F) == 8>(v, res, x0, seg_off, boff, ntraj, part);
}

__global__ void __launch_bounds__(64) k_btay_reduce(const u64* __restrict__ part,
                                                    const int64_t* __restrict__ boff,
                                                    paces_batch_taylor_t* __restrict__ ts,
                                                    paces_batch_taylor_hdr_t* __restrict__ hdr,
                                                    int fin, int j, double tol, int use_prev) {
  btay_reduce_body(part, boff, ts, hdr, fin, j, tol, use_prev);
}

template <typename F, int BS, int RPT, int NS>
__global__ void __launch_bounds__(BS) k_btay_term(
    const F* __restrict__ diag, CsrSetB64 H, const typename Vec2<F>::T* __restrict__ xin,
    typename Vec2<F>::T* __restrict__ xout, double2* __restrict__ res, F c, int j,
    const int32_t* __restrict__ seg_off, const int64_t* __restrict__ boff, int ntraj,
    const paces_batch_taylor_t* __restrict__ ts, paces_batch_taylor_hdr_t* __restrict__ hdr,
    u64* __restrict__ part) {
  btay_term_body<F, BS, RPT, NS>(diag, H, xin, xout, res, res, c, j, seg_off, boff, ntraj, ts,
                                 hdr, part);
}

// After the series. FP64: a trajectory whose last term j is even has res one term behind; its
// last term is still in terms.t[0] (the even-term buffer: later terms skip stopped
// trajectories). Mixed: res holds v, and gets t_0, ..., t_j added in order (terms.t[k] is term
// k's buffer).
template <typename F, int BS, int RPT>
__global__ void __launch_bounds__(BS) k_btay_finish(SeriesTerms terms, double2* __restrict__ res,
                                                    const int32_t* __restrict__ seg_off,
                                                    const int64_t* __restrict__ boff, int ntraj,
                                                    const paces_batch_taylor_t* __restrict__ ts) {
  typedef typename Vec2<F>::T V;
  const SegTile T = seg_tile(seg_off, boff, ntraj, BS * RPT);
  if (T.t < 0) return;
  const int jt = ts[T.t].j_ind;
  if (sizeof(F) == 8 && (jt & 1) != 0) return;
  const int nk = sizeof(F) == 8 ? 1 : (jt + 1 < terms.n ? jt + 1 : terms.n);
  int64_t ii[RPT];
  double2 r[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int64_t i = T.r0 + q * BS + threadIdx.x;
    ii[q] = i < T.r1 ? i : T.r0;
    r[q] = res[ii[q]];
  }
  for (int k = 0; k < nk; ++k) {
    const V* x = (const V*)terms.t[k];
    V xv[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) xv[q] = x[ii[q]];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      r[q].x = r[q].x + (double)xv[q].x;
      r[q].y = r[q].y + (double)xv[q].y;
    }
  }
#pragma unroll
  for (int q = 0; q < RPT; ++q)
    if (T.r0 + q * BS + threadIdx.x < T.r1) res[ii[q]] = r[q];
}

// ---- the series from a device descriptor (N1: captured once into a HIP graph, replayed every
// step after the caller rewrites the descriptor). Grids are sized for a row capacity; workgroups
// past the step's segments exit (seg_tile), so one graph serves every basis up to the capacity.
// The descriptor as the kernels read it: the same layout as paces_batch_series_t, with every
// pointer declared in the global address space, so loads through them are global (not flat)
// accesses, as for kernel-argument pointers.
struct SeriesDev {
  int64_t R;
  PACES_GPTR(const void) diag;
  int32_t nterms_h;
  int32_t pad0;
  PACES_GPTR(const int32_t) indptr[PACES_MAX_TERMS];
  PACES_GPTR(const int32_t) indices[PACES_MAX_TERMS];
  PACES_GPTR(const void) data[PACES_MAX_TERMS];
  PACES_GPTR(const double) v;
  PACES_GPTR(double) res;
  PACES_GPTR(void) x[3];
  PACES_GPTR(const int32_t) seg_off;
  PACES_GPTR(paces_batch_taylor_t) ts;
  PACES_GPTR(paces_batch_taylor_hdr_t) hdr;
  PACES_GPTR(void) ws;
  double tol;
  double c[PACES_SERIES_MAX_TERMS];
  int32_t coded;
  int32_t sqrt_len;
  PACES_GPTR(const uint8_t) code[PACES_MAX_TERMS];
  PACES_GPTR(const double) term_value;
  PACES_GPTR(const double) sqrt_table;
  PACES_GPTR(const uint32_t) slots;
  int32_t nslots;
  int32_t col_bits;
  int32_t slot_off[PACES_MAX_TERMS + 1];
  int32_t slot_vib[PACES_MAX_TERMS];
  int32_t slot_stride;
  PACES_GPTR(void) xt[PACES_SERIES_MAX_TERMS];
  PACES_GPTR(const u64) vmax;
};
static_assert(sizeof(SeriesDev) == sizeof(paces_batch_series_t), "descriptor layout");
static_assert(offsetof(SeriesDev, xt) == offsetof(paces_batch_series_t, xt), "descriptor layout");
static_assert(offsetof(SeriesDev, vmax) == offsetof(paces_batch_series_t, vmax),
              "descriptor layout");
static_assert(offsetof(SeriesDev, slots) == offsetof(paces_batch_series_t, slots),
              "descriptor layout");
static_assert(offsetof(SeriesDev, slot_vib) == offsetof(paces_batch_series_t, slot_vib),
              "descriptor layout");
static_assert(offsetof(SeriesDev, code) == offsetof(paces_batch_series_t, code),
              "descriptor layout");
static_assert(offsetof(SeriesDev, c) == offsetof(paces_batch_series_t, c), "descriptor layout");
static_assert(offsetof(SeriesDev, ws) == offsetof(paces_batch_series_t, ws), "descriptor layout");

template <typename T>
__device__ __forceinline__ T* gen(PACES_GPTR(T) p) {
  return (T*)p;
}

__device__ __forceinline__ const int64_t* ser_boff(const SeriesDev* d) {
  return gen((PACES_GPTR(const int64_t))d->ws);
}
__device__ __forceinline__ u64* ser_part(const SeriesDev* d, int ntraj) {
  return gen((PACES_GPTR(u64))((PACES_GPTR(char))d->ws +
                               (((size_t)(ntraj + 1) * 8 + 255) & ~(size_t)255)));
}

__global__ void __launch_bounds__(kSegThreads) k_bser_segblocks(
    const paces_batch_series_t* __restrict__ desc, int ntraj, int tile) {
  const SeriesDev* d = (const SeriesDev*)desc;
  seg_blocks_body(gen(d->seg_off), ntraj, tile, gen((PACES_GPTR(int64_t))d->ws));
}

template <typename F, int BS, int RPT>
__global__ void __launch_bounds__(BS) k_bser_begin(const paces_batch_series_t* __restrict__ desc,
                                                   int ntraj) {
  const SeriesDev* d = (const SeriesDev*)desc;
  typedef typename Vec2<F>::T V;
  btay_begin_body<F, BS, RPT, false>(gen((PACES_GPTR(const double2))d->v),
                                     gen((PACES_GPTR(double2))d->res),
                                     gen((PACES_GPTR(V))d->x[2]), gen(d->seg_off), ser_boff(d),
                                     ntraj, ser_part(d, ntraj));
}

// the begin's state from the largest key of each trajectory's v (vmax), when x[2] already holds
// v as complex64: what btay_reduce_body's fin pass sets from the begin's maxima
__global__ void __launch_bounds__(256) k_bser_init(const paces_batch_series_t* __restrict__ desc,
                                                   int ntraj) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntraj) return;
  const SeriesDev* d = (const SeriesDev*)desc;
  paces_batch_taylor_t* ts = gen(d->ts);
  const int32_t* seg = gen(d->seg_off);
  const double m0 = sqrt(__longlong_as_double((long long)gen(d->vmax)[t]));
  ts[t].c1 = m0;
  ts[t].rn = m0;
  ts[t].c1_plus_c2 = 0.0;
  ts[t].ratio = 0.0;
  ts[t].done = (seg[t + 1] == seg[t]) ? 1 : 0;
  ts[t].j_ind = 0;
}

__global__ void __launch_bounds__(64) k_bser_reduce(const paces_batch_series_t* __restrict__ desc,
                                                    int ntraj, int fin, int j, int use_prev) {
  const SeriesDev* d = (const SeriesDev*)desc;
  btay_reduce_body(ser_part(d, ntraj), ser_boff(d), gen(d->ts), gen(d->hdr), fin, j, d->tol,
                   use_prev);
}

template <typename F, int BS, int RPT, int NS>
__global__ void __launch_bounds__(BS) k_bser_term(const paces_batch_series_t* __restrict__ desc,
                                                  int ntraj, int j) {
  const SeriesDev* d = (const SeriesDev*)desc;
  CsrSetB64 H;
  H.nterms = d->nterms_h;
  H.coded = d->coded;
  H.sqrt_len = d->sqrt_len;
  H.term_value = gen(d->term_value);
  H.sqrt_table = gen(d->sqrt_table);
#pragma unroll
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    H.indptr[t] = gen(d->indptr[t]);
    H.indices[t] = gen(d->indices[t]);
    H.data[t] = gen(d->data[t]);
    H.code[t] = gen(d->code[t]);
  }
  H.slots = gen(d->slots);
  H.slot_stride = d->slot_stride;
  H.nslots = d->nslots;
  H.col_bits = d->col_bits;
  H.slot_begin = 0;
#pragma unroll
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    H.slot_vib[t] = d->slot_vib[t];
    if (t >= 1 && t < H.nterms) H.slot_begin |= 1 << d->slot_off[t];
  }
  typedef typename Vec2<F>::T V;
  // mixed: term j writes xt[j] and reads xt[j - 1] (x[2] for term 0); FP64: x[j % 2] and
  // x[(j - 1) % 2]
  const V* xin = gen((PACES_GPTR(const V))(
      j == 0 ? d->x[2] : (sizeof(F) == 4 ? d->xt[j - 1] : d->x[(j - 1) & 1])));
  V* xout = gen((PACES_GPTR(V))(sizeof(F) == 4 ? d->xt[j] : d->x[j & 1]));
  // the begin kernel left res unwritten: until term 1 has stored it, the partial sum is v
  // (term 0 reads it for the FP64 stopping test, term 1 adds to it); with a tolerance below 1
  // no series stops at term 0 ((||v|| + ||t_0||) / ||v + t_0|| is about 1), so the finish pass
  // never adds to an unwritten res
  double2* res = gen((PACES_GPTR(double2))d->res);
  const double2* res_src = j <= 1 ? gen((PACES_GPTR(const double2))d->v) : res;
  btay_term_body<F, BS, RPT, NS>(gen((PACES_GPTR(const F))d->diag), H, xin, xout, res,
                                 res_src, (F)d->c[j], j, gen(d->seg_off), ser_boff(d), ntraj,
                                 gen(d->ts), gen(d->hdr), ser_part(d, ntraj));
}

size_t btay_ws_bytes(int64_t R, int ntraj) {
  int64_t nblk = (R + 127) / 128 + ntraj + 1;  // the smallest tile of any launch shape
  return a256((size_t)(ntraj + 1) * 8) + a256((size_t)nblk * 2 * 8);
}

static void no_slots(CsrSetB64* H) {
  H->slots = nullptr;
  H->slot_stride = 0;
  H->nslots = 0;
  H->col_bits = 0;
  H->slot_begin = 0;
  for (int t = 0; t < PACES_MAX_TERMS; ++t) H->slot_vib[t] = 0;
}

// the host fields of a slot matrix, checked (its pointers are the caller's)
static bool slots_ok(const paces_slots_t* S, int64_t R) {
  if (!S || S->nterms < 1 || S->nterms > PACES_MAX_TERMS) return false;
  if (S->nslots != 4 && S->nslots != 8) return false;
  if (S->col_bits < 1 || S->col_bits > 31 || R > ((int64_t)1 << S->col_bits) - 1) return false;
  if (S->sqrt_len < 1 || S->sqrt_len > PACES_CODE_TABLE) return false;
  if ((int64_t)S->sqrt_len > ((int64_t)1 << (32 - S->col_bits))) return false;
  if (S->slot_off[0] != 0 || S->slot_off[S->nterms] > S->nslots) return false;
  for (int t = 0; t < S->nterms; ++t)
    if (S->slot_off[t + 1] < S->slot_off[t] + 2) return false;  // every term has slots
  if (!S->term_value || !S->sqrt_table || (R > 0 && !S->slots) || S->stride < R) return false;
  return true;
}

static bool csr_slots(const paces_slots_t* S, int64_t R, CsrSetB64* H) {
  if (!slots_ok(S, R)) return false;
  H->nterms = S->nterms;
  H->coded = 1;
  H->sqrt_len = S->sqrt_len;
  H->term_value = S->term_value;
  H->sqrt_table = S->sqrt_table;
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    H->indptr[t] = nullptr;
    H->indices[t] = nullptr;
    H->data[t] = nullptr;
    H->code[t] = nullptr;
  }
  H->slots = S->slots;
  H->slot_stride = S->stride;
  H->nslots = S->nslots;
  H->col_bits = S->col_bits;
  H->slot_begin = 0;
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    H->slot_vib[t] = t < S->nterms ? (S->vib[t] ? 1 : 0) : 0;
    if (t >= 1 && t < S->nterms) H->slot_begin |= 1 << S->slot_off[t];
  }
  return true;
}

static bool csr_b(const paces_csr_set_t* h64, const paces_csr32_set_t* h32, CsrSetB64* H) {
  const int n = h64 ? h64->nterms : (h32 ? h32->nterms : -1);
  if (n < 0 || n > PACES_MAX_TERMS) return false;
  H->nterms = n;
  H->coded = 0;
  H->sqrt_len = 0;
  H->term_value = nullptr;
  H->sqrt_table = nullptr;
  no_slots(H);
  for (int t = 0; t < PACES_MAX_TERMS; ++t) H->code[t] = nullptr;
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    H->indptr[t] = t < n ? (h64 ? h64->indptr[t] : h32->indptr[t]) : nullptr;
    H->indices[t] = t < n ? (h64 ? h64->indices[t] : h32->indices[t]) : nullptr;
    H->data[t] = t < n ? (h64 ? (const void*)h64->data[t] : (const void*)h32->data[t]) : nullptr;
    if (t < n && !H->indptr[t]) return false;
  }
  return true;
}

static bool csr_coded(const paces_csr_coded_t* hc, CsrSetB64* H) {
  if (!hc || hc->nterms < 0 || hc->nterms > PACES_MAX_TERMS) return false;
  if (hc->sqrt_len < 0 || hc->sqrt_len > PACES_CODE_TABLE) return false;
  if (hc->nterms > 0 && !hc->term_value) return false;
  H->nterms = hc->nterms;
  H->coded = 1;
  no_slots(H);
  H->sqrt_len = hc->sqrt_len;
  H->term_value = hc->term_value;
  H->sqrt_table = hc->sqrt_table;
  bool vib = false;
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    const bool in = t < hc->nterms;
    H->indptr[t] = in ? hc->indptr[t] : nullptr;
    H->indices[t] = in ? hc->indices[t] : nullptr;
    H->data[t] = nullptr;
    H->code[t] = in ? hc->code[t] : nullptr;
    if (in && !H->indptr[t]) return false;
    vib = vib || H->code[t];
  }
  return !vib || hc->sqrt_table;
}

}  // namespace paces

using namespace paces;

// one batched Taylor term (paces_batch_taylor_term_tuned / _coded) over the CSR set H
static int btay_term_launch(int64_t R, const void* diag, int mixed, const CsrSetB64& H,
                            const void* x_in, void* x_out, double* res, double c, int32_t j,
                            double tol, const int32_t* seg_off, int32_t ntraj,
                            paces_batch_taylor_t* ts, paces_batch_taylor_hdr_t* hdr, void* ws,
                            size_t ws_bytes, const paces_launch_cfg_t* cfg, void* stream);

extern "C" {

int paces_batch_segments(const uint64_t* rows, int64_t R, int32_t W, int32_t ntraj,
                         int32_t* seg_off, void* stream) {
  if (R < 0 || ntraj < 1 || !seg_off || (R > 0 && !rows)) return PACES_E_INVALID;
  int trajbits = 0;
  while ((1 << trajbits) < ntraj) ++trajbits;
  hipStream_t st = (hipStream_t)stream;
#define SEG_LAUNCH(WW)                                                                          \
  hipLaunchKernelGGL(k_seg_offsets<WW>, dim3(grid_for(ntraj + 1, kBlock)), dim3(kBlock), 0, st, \
                     (const u64*)rows, R, trajbits, ntraj, seg_off);                           \
  return launch_status();
  switch (W) {
    case 1: SEG_LAUNCH(1)
    case 2: SEG_LAUNCH(2)
    case 3: SEG_LAUNCH(3)
    case 4: SEG_LAUNCH(4)
    case 5: SEG_LAUNCH(5)
    case 6: SEG_LAUNCH(6)
    case 7: SEG_LAUNCH(7)
    case 8: SEG_LAUNCH(8)
    case 9: SEG_LAUNCH(9)
    case 10: SEG_LAUNCH(10)
    case 11: SEG_LAUNCH(11)
    case 12: SEG_LAUNCH(12)
    case 13: SEG_LAUNCH(13)
    case 14: SEG_LAUNCH(14)
    case 15: SEG_LAUNCH(15)
    case 16: SEG_LAUNCH(16)
    default: return PACES_E_UNSUPPORTED;
  }
#undef SEG_LAUNCH
}

size_t paces_batch_select_workspace_bytes(int64_t R, int32_t ntraj) {
  return bsel_ws_bytes(R, ntraj);
}

int paces_batch_select_topk(const double* psi, const uint64_t* rows, int64_t R, int32_t W,
                            const int32_t* seg_off, int32_t ntraj, const int64_t* k_per_traj,
                            uint64_t* out_rows, int32_t* out_count, void* ws, size_t ws_bytes,
                            void* stream) {
  return paces_batch_select_topk_keys(psi, nullptr, rows, R, W, seg_off, ntraj, k_per_traj,
                                      out_rows, out_count, ws, ws_bytes, stream);
}

int paces_batch_select_topk_keys(const double* psi, const uint64_t* keys, const uint64_t* rows,
                                 int64_t R, int32_t W, const int32_t* seg_off, int32_t ntraj,
                                 const int64_t* k_per_traj, uint64_t* out_rows,
                                 int32_t* out_count, void* ws, size_t ws_bytes, void* stream) {
  if (R < 0 || ntraj < 1 || !seg_off || !k_per_traj || !out_count || !ws) return PACES_E_INVALID;
  if (R > 0 && ((!psi && !keys) || !rows || !out_rows)) return PACES_E_INVALID;
  if (R >= 2147483647LL) return PACES_E_INVALID;
  if (ws_bytes < bsel_ws_bytes(R, ntraj)) return PACES_E_WORKSPACE;
  PACES_DISPATCH_W(W, run_bselect, psi, rows, R, seg_off, ntraj, k_per_traj, out_rows, out_count,
                   ws, (hipStream_t)stream, nullptr, nullptr, keys);
}

int paces_batch_select_flags(const double* psi, const uint64_t* keys, int64_t R,
                             const int32_t* seg_off, int32_t ntraj, const int64_t* k_per_traj,
                             const uint8_t* in_be, uint8_t* keep, int32_t* out_count, void* ws,
                             size_t ws_bytes, void* stream) {
  if (R < 0 || ntraj < 1 || !seg_off || !k_per_traj || !out_count || !ws) return PACES_E_INVALID;
  if (R > 0 && ((!psi && !keys) || !keep)) return PACES_E_INVALID;
  if (R >= 2147483647LL) return PACES_E_INVALID;
  if (ws_bytes < bsel_ws_bytes(R, ntraj)) return PACES_E_WORKSPACE;
  return run_bselect<0>(psi, nullptr, R, seg_off, ntraj, k_per_traj, nullptr, out_count, ws,
                        (hipStream_t)stream, keep, in_be, keys);
}

size_t paces_holstein_observables_batched_workspace_bytes(int64_t R, int32_t ntraj,
                                                          int32_t nsites) {
  return bobs_ws_bytes(R, ntraj, nsites);
}

int paces_holstein_observables_batched(const uint64_t* rows, const double* psi, int64_t R,
                                       const int32_t* seg_off, const paces_layout_t* lay,
                                       const paces_holstein_t* h, double* out, void* ws,
                                       size_t ws_bytes, void* stream) {
  if (!lay || !h || !h->batched || !seg_off || !out || !ws || R < 0) return PACES_E_INVALID;
  if (R > 0 && (!rows || !psi)) return PACES_E_INVALID;
  if (lay->nfields != h->nchain + 2) return PACES_E_INVALID;
  if (!h->d_traj_L || h->geo_dim < 1 || h->geo_dim > 3) return PACES_E_INVALID;
  if (ws_bytes < bobs_ws_bytes(R, h->ntraj, h->nchain)) return PACES_E_WORKSPACE;
  ObsFinish fin = {};
  PACES_DISPATCH_W(lay->W, run_bobs, rows, const_cast<double*>(psi), R, seg_off, h->ntraj, lay,
                   h->nchain, h->d_traj_L, h->geo_dim, out, ws, (hipStream_t)stream, fin);
}

int paces_holstein_observables_batched_finish(uint64_t const* rows, double* psi, int64_t R,
                                              const int32_t* seg_off, const paces_layout_t* lay,
                                              const paces_holstein_t* h,
                                              const void* const* terms, int32_t nterms,
                                              int32_t k0, int32_t mixed,
                                              const paces_batch_taylor_t* ts, uint64_t* keys,
                                              uint64_t* vmax, double* out, void* ws,
                                              size_t ws_bytes, void* stream) {
  if (!lay || !h || !h->batched || !seg_off || !out || !ws || R < 0) return PACES_E_INVALID;
  if (k0 < 0 || (k0 > 0 && !mixed)) return PACES_E_INVALID;
  if (R > 0 && (!rows || !psi)) return PACES_E_INVALID;
  if (lay->nfields != h->nchain + 2) return PACES_E_INVALID;
  if (!h->d_traj_L || h->geo_dim < 1 || h->geo_dim > 3) return PACES_E_INVALID;
  if (ws_bytes < bobs_ws_bytes(R, h->ntraj, h->nchain)) return PACES_E_WORKSPACE;
  ObsFinish fin = {};
  if (ts && R > 0 && !series_terms(terms, nterms, &fin.terms)) return PACES_E_INVALID;
  fin.k0 = k0;
  fin.mixed = mixed ? 1 : 0;
  fin.ts = ts;
  fin.keys = (u64*)keys;
  fin.vmax = (u64*)vmax;
  PACES_DISPATCH_W(lay->W, run_bobs, rows, psi, R, seg_off, h->ntraj, lay, h->nchain,
                   h->d_traj_L, h->geo_dim, out, ws, (hipStream_t)stream, fin);
}

size_t paces_batch_taylor_workspace_bytes(int64_t R, int32_t ntraj) {
  return btay_ws_bytes(R, ntraj);
}

int paces_batch_taylor_begin(int64_t R, const double* v, double* res, void* x0, int32_t mixed,
                             const int32_t* seg_off, int32_t ntraj, paces_batch_taylor_t* ts,
                             void* ws, size_t ws_bytes, void* stream) {
  return paces_batch_taylor_begin_tuned(R, v, res, x0, mixed, seg_off, ntraj, ts, ws, ws_bytes,
                                        nullptr, stream);
}

int paces_batch_taylor_begin_tuned(int64_t R, const double* v, double* res, void* x0,
                                   int32_t mixed, const int32_t* seg_off, int32_t ntraj,
                                   paces_batch_taylor_t* ts, void* ws, size_t ws_bytes,
                                   const paces_launch_cfg_t* cfg, void* stream) {
  if (R < 0 || ntraj < 1 || !seg_off || !ts || !ws || !cfg_ok(cfg)) return PACES_E_INVALID;
  if (R > 0 && (!v || (!mixed && !res) || !x0)) return PACES_E_INVALID;  // mixed: res unused
  if (ws_bytes < btay_ws_bytes(R, ntraj)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* p = (char*)ws;
  int64_t* boff = (int64_t*)p;
  p += a256((size_t)(ntraj + 1) * 8);
  u64* part = (u64*)p;
  int B, RR;
  cfg_get(cfg, &B, &RR);
  hipLaunchKernelGGL(k_seg_blocks, dim3(1), dim3(kSegThreads), 0, st, seg_off, ntraj, B * RR, boff);
  const unsigned gs = seg_grid(R, ntraj, B * RR);
#define BB_LAUNCH(BB, QQ)                                                                      \
  if (mixed)                                                                                   \
    hipLaunchKernelGGL((k_btay_begin<float, BB, QQ>), dim3(gs), dim3(BB), 0, st,               \
                       (const double2*)v, (double2*)res, (float2*)x0, seg_off,                 \
                       (const int64_t*)boff, ntraj, part);                                     \
  else                                                                                         \
    hipLaunchKernelGGL((k_btay_begin<double, BB, QQ>), dim3(gs), dim3(BB), 0, st,              \
                       (const double2*)v, (double2*)res, (double2*)x0, seg_off,                \
                       (const int64_t*)boff, ntraj, part)
  PACES_CFG_SWITCH(B, RR, BB_LAUNCH)
#undef BB_LAUNCH
  hipLaunchKernelGGL(k_btay_reduce, dim3(ntraj), dim3(64), 0, st, (const u64*)part,
                     (const int64_t*)boff, ts, (paces_batch_taylor_hdr_t*)nullptr, 1, 0, 0.0, 0);
  return launch_status();
}

size_t paces_batch_series_desc_bytes(void) { return sizeof(paces_batch_series_t); }

int paces_batch_series_launch(const paces_batch_series_t* desc, int64_t R_cap, int32_t ntraj,
                              int32_t mixed, int32_t nslots, int32_t nterms, int32_t prebegun,
                              const paces_launch_cfg_t* cfg, void* stream) {
  if (!desc || R_cap < 0 || ntraj < 1 || nterms < 2 || nterms > PACES_SERIES_MAX_TERMS ||
      !cfg_ok(cfg) || (nslots != 0 && nslots != 4 && nslots != 8))
    return PACES_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  int B, RR;
  cfg_get(cfg, &B, &RR);
  const unsigned gs = seg_grid(R_cap, ntraj, B * RR);
  const unsigned gst = xcd_grid(gs);  // the terms' grid
  hipLaunchKernelGGL(k_bser_segblocks, dim3(1), dim3(kSegThreads), 0, st, desc, ntraj, B * RR);
  if (prebegun) {
    hipLaunchKernelGGL(k_bser_init, dim3((ntraj + 255) / 256), dim3(256), 0, st, desc, ntraj);
  } else {
#define SB_LAUNCH(BB, QQ)                                                                      \
  if (mixed)                                                                                   \
    hipLaunchKernelGGL((k_bser_begin<float, BB, QQ>), dim3(gs), dim3(BB), 0, st, desc, ntraj); \
  else                                                                                         \
    hipLaunchKernelGGL((k_bser_begin<double, BB, QQ>), dim3(gs), dim3(BB), 0, st, desc, ntraj)
    PACES_CFG_SWITCH(B, RR, SB_LAUNCH)
#undef SB_LAUNCH
    hipLaunchKernelGGL(k_bser_reduce, dim3(ntraj), dim3(64), 0, st, desc, ntraj, 1, 0, 0);
  }
  for (int j = 0; j < nterms; ++j) {
#define ST_LAUNCH_NS(BB, QQ, NS)                                                               \
  if (mixed)                                                                                   \
    hipLaunchKernelGGL((k_bser_term<float, BB, QQ, NS>), dim3(gst), dim3(BB), 0, st, desc,     \
                       ntraj, j);                                                              \
  else                                                                                         \
    hipLaunchKernelGGL((k_bser_term<double, BB, QQ, NS>), dim3(gst), dim3(BB), 0, st, desc,    \
                       ntraj, j)
#define ST_LAUNCH(BB, QQ)          \
  if (nslots == 4) {               \
    ST_LAUNCH_NS(BB, QQ, 4);       \
  } else if (nslots == 8) {        \
    ST_LAUNCH_NS(BB, QQ, 8);       \
  } else {                         \
    ST_LAUNCH_NS(BB, QQ, 0);       \
  }
    PACES_CFG_SWITCH(B, RR, ST_LAUNCH)
#undef ST_LAUNCH
#undef ST_LAUNCH_NS
    hipLaunchKernelGGL(k_bser_reduce, dim3(ntraj), dim3(64), 0, st, desc, ntraj, 0, j,
                       mixed ? 1 : 0);
  }
  return launch_status();
}

int paces_batch_taylor_term(int64_t R, const void* diag, const paces_csr_set_t* H64,
                            const paces_csr32_set_t* H32, const void* x_in, void* x_out,
                            double* res, double c, int32_t j, double tol, const int32_t* seg_off,
                            int32_t ntraj, paces_batch_taylor_t* ts,
                            paces_batch_taylor_hdr_t* hdr, void* ws, size_t ws_bytes,
                            void* stream) {
  return paces_batch_taylor_term_tuned(R, diag, H64, H32, x_in, x_out, res, c, j, tol, seg_off,
                                       ntraj, ts, hdr, ws, ws_bytes, nullptr, stream);
}

int paces_batch_taylor_term_tuned(int64_t R, const void* diag, const paces_csr_set_t* H64,
                                  const paces_csr32_set_t* H32, const void* x_in, void* x_out,
                                  double* res, double c, int32_t j, double tol,
                                  const int32_t* seg_off, int32_t ntraj, paces_batch_taylor_t* ts,
                                  paces_batch_taylor_hdr_t* hdr, void* ws, size_t ws_bytes,
                                  const paces_launch_cfg_t* cfg, void* stream) {
  CsrSetB64 H;
  const int mixed = H32 != nullptr;
  if (!csr_b(H64, H32, &H)) return PACES_E_INVALID;
  return btay_term_launch(R, diag, mixed, H, x_in, x_out, res, c, j, tol, seg_off, ntraj, ts,
                          hdr, ws, ws_bytes, cfg, stream);
}

int paces_batch_taylor_term_coded(int64_t R, const void* diag, int32_t mixed,
                                  const paces_csr_coded_t* Hc, const void* x_in, void* x_out,
                                  double* res, double c, int32_t j, double tol,
                                  const int32_t* seg_off, int32_t ntraj, paces_batch_taylor_t* ts,
                                  paces_batch_taylor_hdr_t* hdr, void* ws, size_t ws_bytes,
                                  const paces_launch_cfg_t* cfg, void* stream) {
  CsrSetB64 H;
  if (!csr_coded(Hc, &H)) return PACES_E_INVALID;
  return btay_term_launch(R, diag, mixed ? 1 : 0, H, x_in, x_out, res, c, j, tol, seg_off, ntraj,
                          ts, hdr, ws, ws_bytes, cfg, stream);
}

int paces_batch_taylor_term_slots(int64_t R, const void* diag, int32_t mixed,
                                  const paces_slots_t* Hs, const void* x_in, void* x_out,
                                  double* res, double c, int32_t j, double tol,
                                  const int32_t* seg_off, int32_t ntraj, paces_batch_taylor_t* ts,
                                  paces_batch_taylor_hdr_t* hdr, void* ws, size_t ws_bytes,
                                  const paces_launch_cfg_t* cfg, void* stream) {
  CsrSetB64 H;
  if (!csr_slots(Hs, R, &H)) return PACES_E_INVALID;
  return btay_term_launch(R, diag, mixed ? 1 : 0, H, x_in, x_out, res, c, j, tol, seg_off, ntraj,
                          ts, hdr, ws, ws_bytes, cfg, stream);
}

}  // extern "C"

static int btay_term_launch(int64_t R, const void* diag, int mixed, const CsrSetB64& H,
                            const void* x_in, void* x_out, double* res, double c, int32_t j,
                            double tol, const int32_t* seg_off, int32_t ntraj,
                            paces_batch_taylor_t* ts, paces_batch_taylor_hdr_t* hdr, void* ws,
                            size_t ws_bytes, const paces_launch_cfg_t* cfg, void* stream) {
  if (R < 0 || ntraj < 1 || !seg_off || !ts || !hdr || !ws || !cfg_ok(cfg))
    return PACES_E_INVALID;
  if (R > 0 && (!diag || !x_in || !x_out || (!mixed && !res))) return PACES_E_INVALID;
  if (ws_bytes < btay_ws_bytes(R, ntraj)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* p = (char*)ws;
  const int64_t* boff = (const int64_t*)p;
  p += a256((size_t)(ntraj + 1) * 8);
  u64* part = (u64*)p;
  int B, RR;
  cfg_get(cfg, &B, &RR);
  const unsigned gs = xcd_grid(seg_grid(R, ntraj, B * RR));
  const int nslots = H.slots ? H.nslots : 0;
#define BT_LAUNCH_NS(BB, QQ, NS)                                                               \
  if (mixed)                                                                                   \
    hipLaunchKernelGGL((k_btay_term<float, BB, QQ, NS>), dim3(gs), dim3(BB), 0, st,            \
                       (const float*)diag, H, (const float2*)x_in, (float2*)x_out,             \
                       (double2*)res, (float)c, j, seg_off, boff, ntraj,                       \
                       (const paces_batch_taylor_t*)ts, hdr, part);                            \
  else                                                                                         \
    hipLaunchKernelGGL((k_btay_term<double, BB, QQ, NS>), dim3(gs), dim3(BB), 0, st,           \
                       (const double*)diag, H, (const double2*)x_in, (double2*)x_out,          \
                       (double2*)res, c, j, seg_off, boff, ntraj,                              \
                       (const paces_batch_taylor_t*)ts, hdr, part)
#define BT_LAUNCH(BB, QQ)          \
  if (nslots == 4) {               \
    BT_LAUNCH_NS(BB, QQ, 4);       \
  } else if (nslots == 8) {        \
    BT_LAUNCH_NS(BB, QQ, 8);       \
  } else {                         \
    BT_LAUNCH_NS(BB, QQ, 0);       \
  }
  PACES_CFG_SWITCH(B, RR, BT_LAUNCH)
#undef BT_LAUNCH
#undef BT_LAUNCH_NS
  hipLaunchKernelGGL(k_btay_reduce, dim3(ntraj), dim3(64), 0, st, (const u64*)part, boff, ts, hdr,
                     0, (int)j, tol, mixed ? 1 : 0);
  return launch_status();
}

extern "C" {

int paces_batch_taylor_finish(int64_t R, const void* const* terms, int32_t nterms, int32_t mixed,
                              double* res, const int32_t* seg_off, int32_t ntraj,
                              const paces_batch_taylor_t* ts, void* ws, size_t ws_bytes,
                              const paces_launch_cfg_t* cfg, void* stream) {
  if (R < 0 || ntraj < 1 || !seg_off || !ts || !ws || !cfg_ok(cfg)) return PACES_E_INVALID;
  SeriesTerms tt = {};
  if (R > 0 && (!res || !series_terms(terms, nterms, &tt))) return PACES_E_INVALID;
  if (ws_bytes < btay_ws_bytes(R, ntraj)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t* boff = (const int64_t*)ws;  // the series' tile offsets (same shape)
  int B, RR;
  cfg_get(cfg, &B, &RR);
  const unsigned gs = seg_grid(R, ntraj, B * RR);
#define BF_LAUNCH(BB, QQ)                                                                      \
  if (mixed)                                                                                   \
    hipLaunchKernelGGL((k_btay_finish<float, BB, QQ>), dim3(gs), dim3(BB), 0, st, tt,          \
                       (double2*)res, seg_off, boff, ntraj, ts);                               \
  else                                                                                         \
    hipLaunchKernelGGL((k_btay_finish<double, BB, QQ>), dim3(gs), dim3(BB), 0, st, tt,         \
                       (double2*)res, seg_off, boff, ntraj, ts)
  PACES_CFG_SWITCH(B, RR, BF_LAUNCH)
#undef BF_LAUNCH
  return launch_status();
}

}  // extern "C"
