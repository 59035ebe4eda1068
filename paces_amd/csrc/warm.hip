// Incremental adaptation of the batched step: the new basis from the previous one's slot matrix.
//
// The reference rebuilds every step's basis from scratch: the selection, its images, the images
// of those, and a sort + dedup of the whole candidate list (paces/core/hamiltonian.py:383-492,
// paces/core/time_evolution.py:735-807). Every row of the previous expanded basis (B_e) has all
// its neighbours in the previous basis R_old and lists them in the previous slot matrix (a pair
// with an endpoint in B_e is stored in both rows' slots), and the new selection and the rows its
// expansion adds almost always lie in that old B_e. So the new basis is mostly a subset of
// R_old: marked here through the old slots, with explicit images only for the few marked rows
// outside the old B_e (whose neighbours may be new rows). The result is the same set of rows,
// the same slot matrix and the same carried-over amplitudes as the full construction.
//
// Flag bytes over R_old: bit 0 selected, bit 1 reached by the first expansion, bit 2 by the
// second, bit 7 the row was in the old B_e; a row belongs to the new basis when any of bits 0-6
// is set, to the new B_e with bit 0 or 1.

#include "common.hpp"

namespace paces {

int scan_i32(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws_bytes(int64_t n);

static size_t a256w(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr int kWRpt = 8;                 // consecutive rows per thread in the tiled passes
constexpr int kWTile = kBlock * kWRpt;  // rows per workgroup

// ---- marking through the old slots ------------------------------------------------------------
// Row i is a source when its flags hold a bit of src_mask and none of skip_mask. Every neighbour
// of a source in the old B_e (bit 7) gets mark_bit; a source outside the old B_e is flagged
// explicit (its neighbours must be formed from the row). The slots are symmetric (a pair is
// stored in both rows), so each row reads its own slots and marks itself when one names a source
// in the old B_e: every row's flag byte is written by its own thread only, and the bits the
// others read (src_mask, skip_mask, bit 7) are not among those it writes.
// kWarmXcdRun > 0: workgroup b of the emit takes block (k / C) * 8C + (b % 8) * C + k % C of
// rows, k = b / 8, C = kWarmXcdRun: each XCD (workgroup b runs on XCD b % 8) takes runs of C
// consecutive blocks, so blocks whose rows' neighbours overlap share an L2 (the grid is padded
// to a multiple of 8C; 4% faster than round robin, no gain for the marks)
#ifndef PACES_WARM_XCD_RUN
#define PACES_WARM_XCD_RUN 64
#endif
constexpr int kWarmXcdRun = PACES_WARM_XCD_RUN;

__device__ __forceinline__ int64_t warm_block() {
  const int64_t b = blockIdx.x;
  if (kWarmXcdRun == 0) return b;
  const int64_t k = b >> 3;
  return (k / kWarmXcdRun) * (8 * kWarmXcdRun) + (b & 7) * kWarmXcdRun + k % kWarmXcdRun;
}

static unsigned warm_grid(int64_t g) {
  if (kWarmXcdRun == 0) return (unsigned)g;
  const int64_t q = 8 * kWarmXcdRun;
  return (unsigned)((g + q - 1) / q * q);
}

// A thread takes kMarkRpt rows (t, t + kBlock, ...): every slot of its rows is loaded, then every
// named row's flags, then the rows are decided, so the dependent loads of all rows overlap.
constexpr int kMarkRpt = 8;

template <int NS>
__global__ void __launch_bounds__(kBlock) k_slots_mark(const uint32_t* __restrict__ slots,
                                                       int64_t stride, uint32_t colmask,
                                                       int64_t R, uint8_t* flags, int src_mask,
                                                       int skip_mask, int mark_bit,
                                                       uint8_t* __restrict__ expl) {
  const int64_t i0 = (int64_t)blockIdx.x * (kBlock * kMarkRpt) + threadIdx.x;
  uint32_t nb[kMarkRpt][NS];
  int f[kMarkRpt];
#pragma unroll
  for (int q = 0; q < kMarkRpt; ++q) {
    const int64_t i = i0 + q * kBlock;
    const int64_t ic = i < R ? i : 0;
#pragma unroll
    for (int m = 0; m < NS; ++m) nb[q][m] = i < R ? slots[(int64_t)m * stride + ic] : PACES_SLOT_EMPTY;
    f[q] = flags[ic];
  }
  int g[kMarkRpt][NS];
#pragma unroll
  for (int q = 0; q < kMarkRpt; ++q)
#pragma unroll
    for (int m = 0; m < NS; ++m)
      g[q][m] = nb[q][m] != PACES_SLOT_EMPTY ? flags[nb[q][m] & colmask] : 0;
#pragma unroll
  for (int q = 0; q < kMarkRpt; ++q) {
    const int64_t i = i0 + q * kBlock;
    bool hit = false;
#pragma unroll
    for (int m = 0; m < NS; ++m)
      hit = hit || ((g[q][m] & 0x80) && (g[q][m] & src_mask) && !(g[q][m] & skip_mask));
    const bool e = i < R && (f[q] & src_mask) && !(f[q] & skip_mask) && !(f[q] & 0x80);
    if (i < R) {
      // every row's byte is written (unmarked rows unchanged), so the stores fill whole lines;
      // a concurrent reader of this byte tests only src_mask, skip_mask and 0x80
      flags[i] = (uint8_t)(f[q] | (hit ? mark_bit : 0));
      expl[i] = e ? 1 : 0;
    }
  }
}

// *count += the nonzero bytes of a[0, n): 8 bytes per load, grid-stride, one atomic per
// workgroup (a fixed grid of kCountBlocks; an atomic per wave had every wave with an explicit
// row serialise on one address)
constexpr int kCountBlocks = 2048;
__global__ void __launch_bounds__(kBlock) k_count_u8(const uint8_t* __restrict__ a, int64_t n,
                                                     int32_t* __restrict__ count) {
  __shared__ int ws[kBlock / 64];
  int c = 0;
  const int64_t n8 = n >> 3;
  const bool al = ((uintptr_t)a & 7) == 0;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n8;
       v += (int64_t)gridDim.x * kBlock) {
    if (al) {
      const u64 x = ((const u64*)a)[v];
#pragma unroll
      for (int j = 0; j < 8; ++j) c += ((x >> (8 * j)) & 0xff) ? 1 : 0;
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) c += a[v * 8 + j] ? 1 : 0;
    }
  }
  if (blockIdx.x == 0)
    for (int64_t i = (n8 << 3) + threadIdx.x; i < n; i += kBlock) c += a[i] ? 1 : 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += ws[w];
    if (t) atomicAdd(count, t);
  }
}

// ---- stable compaction of flagged indices ------------------------------------------------------
static_assert(kWRpt == 8, "a thread's flags are one 8-byte load");
__device__ __forceinline__ int tile_count(const uint8_t* __restrict__ flags, int64_t n, int mask,
                                          int64_t r0, int (&c)[kWRpt]) {
  int tot = 0;
  if (r0 + kWRpt <= n && ((uintptr_t)flags & 7) == 0) {
    // r0 is a multiple of 8: the thread's 8 flags in one aligned load
    const u64 f8 = *(const u64*)(flags + r0);
#pragma unroll
    for (int j = 0; j < kWRpt; ++j) {
      c[j] = ((int)(f8 >> (8 * j)) & mask) != 0 ? 1 : 0;
      tot += c[j];
    }
    return tot;
  }
#pragma unroll
  for (int j = 0; j < kWRpt; ++j) {
    c[j] = (r0 + j < n && (flags[r0 + j] & mask) != 0) ? 1 : 0;
    tot += c[j];
  }
  return tot;
}

// exclusive prefix of v across the workgroup (256 threads), and the workgroup's total
__device__ __forceinline__ int block_prefix(int v, int* total) {
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    before += w < wave ? wsum[w] : 0;
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

__global__ void __launch_bounds__(kBlock) k_flag_tilecount(const uint8_t* __restrict__ flags,
                                                           int64_t n, int mask,
                                                           int32_t* __restrict__ ts) {
  const int64_t r0 = (int64_t)blockIdx.x * kWTile + (int64_t)threadIdx.x * kWRpt;
  int c[kWRpt];
  const int tot = tile_count(flags, n, mask, r0, c);
  int all;
  block_prefix(tot, &all);
  if (threadIdx.x == 0) ts[blockIdx.x] = all;
}

__global__ void __launch_bounds__(kBlock) k_flag_emit(const uint8_t* __restrict__ flags, int64_t n,
                                                      int mask, const int32_t* __restrict__ TO,
                                                      int ntiles, int32_t* __restrict__ out,
                                                      int32_t* __restrict__ out_count) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *out_count = TO[ntiles];
  const int64_t r0 = (int64_t)blockIdx.x * kWTile + (int64_t)threadIdx.x * kWRpt;
  int c[kWRpt];
  const int tot = tile_count(flags, n, mask, r0, c);
  int all;
  int o = TO[blockIdx.x] + block_prefix(tot, &all);
#pragma unroll
  for (int j = 0; j < kWRpt; ++j) {
    if (c[j]) out[o] = (int32_t)(r0 + j);
    o += c[j];
  }
}

__global__ void k_set_i32(int32_t* __restrict__ p, int32_t v) { *p = v; }

template <int W>
__global__ void __launch_bounds__(kBlock) k_gather_rows(const u64* __restrict__ rows,
                                                        const int32_t* __restrict__ idx, int64_t m,
                                                        u64* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  store_row<W>(out, j, load_row<W>(rows, idx[j]));
}

// dst[dmap ? dmap[j] : j] = src[smap ? smap[j] : j] (index bookkeeping of the explicit images)
__global__ void __launch_bounds__(kBlock) k_move_i32(const int32_t* __restrict__ src,
                                                     const int32_t* __restrict__ smap,
                                                     const int32_t* __restrict__ dmap, int64_t m,
                                                     int32_t* __restrict__ dst) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  dst[dmap ? (int64_t)dmap[j] : j] = src[smap ? (int64_t)smap[j] : j];
}

// dst[dmap[j]] = v
__global__ void __launch_bounds__(kBlock) k_fill_u8_at(uint8_t* __restrict__ dst,
                                                       const int32_t* __restrict__ dmap, int64_t m,
                                                       uint8_t v) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) dst[dmap[j]] = v;
}

// flags[idx[j]] |= bit where hit[j]; miss[j] = !hit[j]
__global__ void __launch_bounds__(kBlock) k_mark_hits(const int32_t* __restrict__ idx,
                                                      const uint8_t* __restrict__ hit, int64_t m,
                                                      uint8_t* flags, int bit,
                                                      uint8_t* __restrict__ miss) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const bool h = hit[j] != 0;
  miss[j] = h ? 0 : 1;
  if (h) {
    uint8_t* p = flags + idx[j];
    *p = (uint8_t)(*p | bit);
  }
}

// ---- positions in the new basis ----------------------------------------------------------------
// The new basis is the flagged rows of R_old in order with the new rows Y (sorted, none in R_old)
// merged in: ins[j] = lower_bound(R_old, Y[j]), ascending. Row i of R_old takes position
// NP[i] = (flagged rows before i) + #{j : ins[j] <= i} (written for every i in [0, R]; it is
// row i's position when row i is flagged); Y[j] takes NP[ins[j]] - #{k : ins[k] <= ins[j]} + j.

// number of entries of the ascending ins[0, n) that are <= v
__device__ __forceinline__ int count_le(const int32_t* __restrict__ ins, int n, int64_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ins[mid] <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// number of entries of the ascending ins[lo, hi) that are <= v, plus lo
__device__ __forceinline__ int count_le_in(const int32_t* __restrict__ ins, int lo, int hi,
                                           int64_t v) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ins[mid] <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// tfirst[t] = the entries of ins below tile t's first row (t <= ntiles), one search per tile, so
// that the rows of a tile search only the tile's own insertion points
__global__ void __launch_bounds__(kBlock) k_ins_tile_first(const int32_t* __restrict__ ins, int nY,
                                                           int ntiles,
                                                           int32_t* __restrict__ tfirst) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t > ntiles) return;
  tfirst[t] = count_le(ins, nY, (int64_t)t * kWTile - 1);
}

__global__ void __launch_bounds__(kBlock) k_warm_newpos(const uint8_t* __restrict__ flags,
                                                        int64_t R, const int32_t* __restrict__ TO,
                                                        int ntiles,
                                                        const int32_t* __restrict__ ins, int nY,
                                                        const int32_t* __restrict__ tfirst,
                                                        int32_t* __restrict__ NP,
                                                        int32_t* __restrict__ out_total) {
  const int64_t r0 = (int64_t)blockIdx.x * kWTile + (int64_t)threadIdx.x * kWRpt;
  int c[kWRpt];
  const int tot = tile_count(flags, R, 0x7f, r0, c);
  int all;
  int o = TO[blockIdx.x] + block_prefix(tot, &all);
  if (r0 >= R + 1) return;
  int sh = nY ? count_le_in(ins, tfirst[blockIdx.x], tfirst[blockIdx.x + 1], r0) : 0;
  // the next insertion point past r0 - 1 in a register: loads only as insertions are passed
  int64_t nxt = sh < nY ? ins[sh] : INT64_MAX;
  int v[kWRpt];
#pragma unroll
  for (int j = 0; j < kWRpt; ++j) {
    const int64_t r = r0 + j;
    while (nxt <= r) {
      ++sh;
      nxt = sh < nY ? ins[sh] : INT64_MAX;
    }
    v[j] = o + sh;
    o += c[j];
  }
  if (r0 + kWRpt <= R + 1 && ((uintptr_t)NP & 15) == 0) {
    // r0 is a multiple of 8: two aligned 16-byte stores
    ((int4*)(NP + r0))[0] = make_int4(v[0], v[1], v[2], v[3]);
    ((int4*)(NP + r0))[1] = make_int4(v[4], v[5], v[6], v[7]);
  } else {
#pragma unroll
    for (int j = 0; j < kWRpt; ++j)
      if (r0 + j <= R) NP[r0 + j] = v[j];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *out_total = TO[ntiles] + nY;
}

__global__ void __launch_bounds__(kBlock) k_warm_ypos(const int32_t* __restrict__ NP,
                                                      const int32_t* __restrict__ ins, int nY,
                                                      int32_t* __restrict__ ypos) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nY) return;
  const int p = ins[j];
  ypos[j] = NP[p] - count_le(ins, nY, p) + j;
}

// ---- the new basis ----------------------------------------------------------------------------
// Flagged rows of R_old move to NP[i] with their amplitude (the reassignment: a row of both bases
// keeps its amplitude) and diagonal; a slot keeps its entry when the pair has an endpoint in the
// new B_e (the column renumbered), else it becomes empty. Pairs whose B_e endpoint lies outside
// the old B_e are not in the old slots: the explicit pairs fill them afterwards.
template <int W, typename D>
__global__ void __launch_bounds__(kBlock) k_warm_emit(
    const u64* __restrict__ rows, const double2* __restrict__ psi, const D* __restrict__ diag,
    int64_t R, const uint8_t* __restrict__ flags, int be_mask, const int32_t* __restrict__ NP,
    const uint32_t* __restrict__ slots, int64_t stride_old, int nslots, uint32_t colmask,
    u64* __restrict__ out_rows, double2* __restrict__ out_psi, D* __restrict__ out_diag,
    uint8_t* __restrict__ out_be, uint32_t* __restrict__ out_slots, int64_t stride_new,
    float2* __restrict__ out_x0) {
  const int64_t i = warm_block() * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int f = flags[i] & 0x7f;
  if (!f) return;
  uint32_t nb[PACES_MAX_SLOTS];
#pragma unroll
  for (int m = 0; m < PACES_MAX_SLOTS; ++m)
    nb[m] = m < nslots ? slots[(int64_t)m * stride_old + i] : PACES_SLOT_EMPTY;
  const int64_t p = NP[i];
  const Row<W> x = load_row<W>(rows, i);
  const double2 v = psi[i];
  const D d = diag[i];
  int fj[PACES_MAX_SLOTS];
  int32_t pj[PACES_MAX_SLOTS];
#pragma unroll
  for (int m = 0; m < PACES_MAX_SLOTS; ++m) {
    const bool occ = nb[m] != PACES_SLOT_EMPTY;
    const int64_t j = occ ? (int64_t)(nb[m] & colmask) : i;
    fj[m] = flags[j] & 0x7f;
    pj[m] = NP[j];
  }
  store_row<W>(out_rows, p, x);
  out_psi[p] = v;
  if (out_x0) out_x0[p] = make_float2((float)v.x, (float)v.y);  // the series' term input
  out_diag[p] = d;
  out_be[p] = (f & be_mask) ? 1 : 0;
#pragma unroll
  for (int m = 0; m < PACES_MAX_SLOTS; ++m) {
    if (m < nslots) {
      // (a kept entry's neighbour is always a row of the new basis, fj != 0: tested, so that an
      // inconsistent flag set gives an empty slot rather than a column past the new rows)
      const bool keep =
          nb[m] != PACES_SLOT_EMPTY && fj[m] != 0 && ((f | fj[m]) & be_mask) != 0;
      out_slots[(int64_t)m * stride_new + p] =
          keep ? ((uint32_t)pj[m] | (nb[m] & ~colmask)) : PACES_SLOT_EMPTY;
    }
  }
}

template <int W, typename D>
__global__ void __launch_bounds__(kBlock) k_warm_emit_new(
    const u64* __restrict__ Y, int nY, const int32_t* __restrict__ ypos,
    const D* __restrict__ ydiag, const uint8_t* __restrict__ ybe, int nslots,
    u64* __restrict__ out_rows, double2* __restrict__ out_psi, D* __restrict__ out_diag,
    uint8_t* __restrict__ out_be, uint32_t* __restrict__ out_slots, int64_t stride_new,
    float2* __restrict__ out_x0) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nY) return;
  const int64_t p = ypos[j];
  store_row<W>(out_rows, p, load_row<W>(Y, j));
  out_psi[p] = make_double2(0.0, 0.0);
  if (out_x0) out_x0[p] = make_float2(0.0f, 0.0f);
  out_diag[p] = ydiag[j];
  out_be[p] = ybe[j];
  for (int m = 0; m < nslots; ++m) out_slots[(int64_t)m * stride_new + p] = PACES_SLOT_EMPTY;
}

template <int W>
static int run_gather(const uint64_t* rows, const int32_t* idx, int64_t m, uint64_t* out,
                      hipStream_t st) {
  if (m > 0)
    hipLaunchKernelGGL(k_gather_rows<W>, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, st,
                       (const u64*)rows, idx, m, (u64*)out);
  return launch_status();
}

template <int W>
static int run_emit(const uint64_t* rows, const double* psi, const void* diag, int diag_f32,
                    int64_t R, const uint8_t* flags, int be_mask, const int32_t* NP,
                    const uint32_t* slots, int64_t stride_old, int nslots, uint32_t colmask,
                    uint64_t* out_rows, double* out_psi, void* out_diag, uint8_t* out_be,
                    uint32_t* out_slots, int64_t stride_new, void* out_x0, hipStream_t st) {
  if (R == 0) return PACES_OK;
  if (diag_f32)
    hipLaunchKernelGGL((k_warm_emit<W, float>), dim3(warm_grid(grid_for(R, kBlock))),
                       dim3(kBlock), 0, st,
                       (const u64*)rows, (const double2*)psi, (const float*)diag, R, flags,
                       be_mask, NP, slots, stride_old, nslots, colmask, (u64*)out_rows,
                       (double2*)out_psi, (float*)out_diag, out_be, out_slots, stride_new,
                       (float2*)out_x0);
  else
    hipLaunchKernelGGL((k_warm_emit<W, double>), dim3(warm_grid(grid_for(R, kBlock))),
                       dim3(kBlock), 0, st,
                       (const u64*)rows, (const double2*)psi, (const double*)diag, R, flags,
                       be_mask, NP, slots, stride_old, nslots, colmask, (u64*)out_rows,
                       (double2*)out_psi, (double*)out_diag, out_be, out_slots, stride_new,
                       (float2*)out_x0);
  return launch_status();
}

template <int W>
static int run_emit_new(const uint64_t* Y, int nY, const int32_t* ypos, const void* ydiag,
                        int diag_f32, const uint8_t* ybe, int nslots, uint64_t* out_rows,
                        double* out_psi, void* out_diag, uint8_t* out_be, uint32_t* out_slots,
                        int64_t stride_new, void* out_x0, hipStream_t st) {
  if (nY == 0) return PACES_OK;
  if (diag_f32)
    hipLaunchKernelGGL((k_warm_emit_new<W, float>), dim3(grid_for(nY, kBlock)), dim3(kBlock), 0,
                       st, (const u64*)Y, nY, ypos, (const float*)ydiag, ybe, nslots,
                       (u64*)out_rows, (double2*)out_psi, (float*)out_diag, out_be, out_slots,
                       stride_new, (float2*)out_x0);
  else
    hipLaunchKernelGGL((k_warm_emit_new<W, double>), dim3(grid_for(nY, kBlock)), dim3(kBlock), 0,
                       st, (const u64*)Y, nY, ypos, (const double*)ydiag, ybe, nslots,
                       (u64*)out_rows, (double2*)out_psi, (double*)out_diag, out_be, out_slots,
                       stride_new, (float2*)out_x0);
  return launch_status();
}

static bool colmask_of(int32_t col_bits, uint32_t* m) {
  if (col_bits < 1 || col_bits > 31) return false;
  *m = (1u << col_bits) - 1u;
  return true;
}

}  // namespace paces

using namespace paces;

extern "C" {

int paces_slots_mark(const uint32_t* slots, int64_t stride, int32_t nslots, int32_t col_bits,
                     int64_t R, uint8_t* flags, int32_t src_mask, int32_t skip_mask,
                     int32_t mark_bit, uint8_t* expl, int32_t* expl_count, void* stream) {
  uint32_t cm;
  if (R < 0 || !colmask_of(col_bits, &cm) || nslots < 1 || nslots > PACES_MAX_SLOTS ||
      stride < R || !expl_count)
    return PACES_E_INVALID;
  if (R > 0 && (!slots || !flags || !expl)) return PACES_E_INVALID;
  const int tested = src_mask | skip_mask | 0x80;
  if (mark_bit < 1 || mark_bit > 0x7f || (mark_bit & tested) || ((src_mask | skip_mask) & 0x80))
    return PACES_E_INVALID;
  if (nslots != 4 && nslots != 8) return PACES_E_INVALID;
  if (R > 0) {
    const dim3 g(grid_for(R, kBlock * kMarkRpt));
    if (nslots == 4)
      hipLaunchKernelGGL(k_slots_mark<4>, g, dim3(kBlock), 0, (hipStream_t)stream, slots, stride,
                         cm, R, flags, (int)src_mask, (int)skip_mask, (int)mark_bit, expl);
    else
      hipLaunchKernelGGL(k_slots_mark<8>, g, dim3(kBlock), 0, (hipStream_t)stream, slots, stride,
                         cm, R, flags, (int)src_mask, (int)skip_mask, (int)mark_bit, expl);
    hipLaunchKernelGGL(k_count_u8, dim3(kCountBlocks), dim3(kBlock), 0, (hipStream_t)stream,
                       (const uint8_t*)expl, R, expl_count);
  }
  return launch_status();
}

size_t paces_flag_indices_workspace_bytes(int64_t n) {
  const int64_t nt = (n + kWTile - 1) / kWTile;
  return a256w((size_t)(nt + 1) * 4) * 2 + a256w(scan_ws_bytes(nt + 1));
}

int paces_flag_indices(const uint8_t* flags, int64_t n, int32_t mask, int32_t* out_idx,
                       int32_t* out_count, void* ws, size_t ws_bytes, void* stream) {
  if (n < 0 || n >= 2147483647LL || !out_count || !ws || (n > 0 && (!flags || !out_idx)))
    return PACES_E_INVALID;
  if (ws_bytes < paces_flag_indices_workspace_bytes(n)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) {
    hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, st, out_count, 0);
    return launch_status();
  }
  const int64_t nt = (n + kWTile - 1) / kWTile;
  char* p = (char*)ws;
  int32_t* ts = (int32_t*)p;
  p += a256w((size_t)(nt + 1) * 4);
  int32_t* TO = (int32_t*)p;
  p += a256w((size_t)(nt + 1) * 4);
  hipLaunchKernelGGL(k_flag_tilecount, dim3(nt), dim3(kBlock), 0, st, flags, n, (int)mask, ts);
  int rc = launch_status();
  if (rc) return rc;
  rc = scan_i32(ts, nt, TO, p, scan_ws_bytes(nt + 1), st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_flag_emit, dim3(nt), dim3(kBlock), 0, st, flags, n, (int)mask,
                     (const int32_t*)TO, (int)nt, out_idx, out_count);
  return launch_status();
}

int paces_gather_rows(const uint64_t* rows, int32_t W, const int32_t* idx, int64_t m,
                      uint64_t* out, void* stream) {
  if (m < 0 || (m > 0 && (!rows || !idx || !out))) return PACES_E_INVALID;
  PACES_DISPATCH_W(W, run_gather, rows, idx, m, out, (hipStream_t)stream);
}

int paces_mark_hits(const int32_t* idx, const uint8_t* hit, int64_t m, uint8_t* flags,
                    int32_t bit, uint8_t* miss, void* stream) {
  if (m < 0 || bit < 1 || bit > 255 || (m > 0 && (!idx || !hit || !flags || !miss)))
    return PACES_E_INVALID;
  if (m > 0)
    hipLaunchKernelGGL(k_mark_hits, dim3(grid_for(m, kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, idx, hit, m, flags, (int)bit, miss);
  return launch_status();
}

int paces_move_i32(const int32_t* src, const int32_t* smap, const int32_t* dmap, int64_t m,
                   int32_t* dst, void* stream) {
  if (m < 0 || (m > 0 && (!src || !dst))) return PACES_E_INVALID;
  if (m > 0)
    hipLaunchKernelGGL(k_move_i32, dim3(grid_for(m, kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, src, smap, dmap, m, dst);
  return launch_status();
}

int paces_fill_u8_at(uint8_t* dst, const int32_t* dmap, int64_t m, int32_t v, void* stream) {
  if (m < 0 || v < 0 || v > 255 || (m > 0 && (!dst || !dmap))) return PACES_E_INVALID;
  if (m > 0)
    hipLaunchKernelGGL(k_fill_u8_at, dim3(grid_for(m, kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, dst, dmap, m, (uint8_t)v);
  return launch_status();
}

size_t paces_warm_positions_workspace_bytes(int64_t R) {
  const int64_t nt = (R + 1 + kWTile - 1) / kWTile;
  return a256w((size_t)(nt + 1) * 4) * 3 + a256w(scan_ws_bytes(nt + 1));
}

int paces_warm_positions(const uint8_t* flags, int64_t R, const int32_t* ins, int64_t nY,
                         int32_t* NP, int32_t* ypos, int32_t* out_total, void* ws,
                         size_t ws_bytes, void* stream) {
  if (R < 0 || nY < 0 || R + nY >= 2147483647LL || !NP || !out_total || !ws ||
      (R > 0 && !flags) || (nY > 0 && (!ins || !ypos)))
    return PACES_E_INVALID;
  if (ws_bytes < paces_warm_positions_workspace_bytes(R)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nt = (R + 1 + kWTile - 1) / kWTile;  // NP has R + 1 entries
  char* p = (char*)ws;
  int32_t* ts = (int32_t*)p;
  p += a256w((size_t)(nt + 1) * 4);
  int32_t* TO = (int32_t*)p;
  p += a256w((size_t)(nt + 1) * 4);
  int32_t* tfirst = (int32_t*)p;
  p += a256w((size_t)(nt + 1) * 4);
  hipLaunchKernelGGL(k_flag_tilecount, dim3(nt), dim3(kBlock), 0, st, flags, R, 0x7f, ts);
  int rc = launch_status();
  if (rc) return rc;
  if (nY > 0)
    hipLaunchKernelGGL(k_ins_tile_first, dim3(grid_for(nt + 1, kBlock)), dim3(kBlock), 0, st, ins,
                       (int)nY, (int)nt, tfirst);
  rc = scan_i32(ts, nt, TO, p, scan_ws_bytes(nt + 1), st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_warm_newpos, dim3(nt), dim3(kBlock), 0, st, flags, R, (const int32_t*)TO,
                     (int)nt, ins, (int)nY, (const int32_t*)tfirst, NP, out_total);
  if (nY > 0)
    hipLaunchKernelGGL(k_warm_ypos, dim3(grid_for(nY, kBlock)), dim3(kBlock), 0, st,
                       (const int32_t*)NP, ins, (int)nY, ypos);
  return launch_status();
}

int paces_warm_emit(const uint64_t* rows, int32_t W, const double* psi, const void* diag,
                    int32_t diag_f32, int64_t R, const uint8_t* flags, int32_t be_mask,
                    const int32_t* NP, const uint32_t* slots, int64_t stride_old, int32_t nslots,
                    int32_t col_bits, uint64_t* out_rows, double* out_psi, void* out_diag,
                    uint8_t* out_be, uint32_t* out_slots, int64_t stride_new, void* out_x0,
                    void* stream) {
  uint32_t cm;
  if (R < 0 || !colmask_of(col_bits, &cm) || nslots < 1 || nslots > PACES_MAX_SLOTS ||
      stride_old < R || stride_new < 0)
    return PACES_E_INVALID;
  if (R > 0 && (!rows || !psi || !diag || !flags || !NP || !slots || !out_rows || !out_psi ||
                !out_diag || !out_be || !out_slots))
    return PACES_E_INVALID;
  PACES_DISPATCH_W(W, run_emit, rows, psi, diag, diag_f32, R, flags, (int)be_mask, NP, slots,
                   stride_old, (int)nslots, cm, out_rows, out_psi, out_diag, out_be, out_slots,
                   stride_new, out_x0, (hipStream_t)stream);
}

int paces_warm_emit_new(const uint64_t* Y, int32_t W, int64_t nY, const int32_t* ypos,
                        const void* ydiag, int32_t diag_f32, const uint8_t* ybe, int32_t nslots,
                        uint64_t* out_rows, double* out_psi, void* out_diag, uint8_t* out_be,
                        uint32_t* out_slots, int64_t stride_new, void* out_x0, void* stream) {
  if (nY < 0 || nY >= 2147483647LL || nslots < 1 || nslots > PACES_MAX_SLOTS) return PACES_E_INVALID;
  if (nY > 0 && (!Y || !ypos || !ydiag || !ybe || !out_rows || !out_psi || !out_diag ||
                 !out_be || !out_slots))
    return PACES_E_INVALID;
  PACES_DISPATCH_W(W, run_emit_new, Y, (int)nY, ypos, ydiag, diag_f32, ybe, (int)nslots,
                   out_rows, out_psi, out_diag, out_be, out_slots, stride_new, out_x0,
                   (hipStream_t)stream);
}

}  // extern "C"
