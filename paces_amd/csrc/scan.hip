// Device-wide exclusive prefix sums: the compaction primitive behind every boolean-mask index
// of the reference step (a[mask] in paces/models/holstein.py:187, 255 and the position
// partitions of paces/aux/cupy_search.py:16-51).
//
// Reduce-then-scan: tiles of kTile elements are reduced, the tile sums are scanned recursively,
// and each tile is rescanned with its offset. Inputs are read twice, outputs written once.
#include "common.hpp"

namespace paces {

constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;  // 2048 elements per workgroup
static_assert(kTile == kScanTile, "scan tile");

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    int o = __shfl_up(v, s, 64);
    if (lane >= s) v += o;
  }
  return v;
}

// exclusive scan of one value per thread across the workgroup; returns the block total in *tot
__device__ __forceinline__ int block_excl_scan(int v, int* tot) {
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int off = 0, total = 0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k) {
    int s = wsum[k];
    off += (k < wid) ? s : 0;
    total += s;
  }
  __syncthreads();
  *tot = total;
  return off + inc - v;
}

// a thread's kItems consecutive inputs: one 8-byte load (u8) or two 16-byte loads (int) when the
// run is whole and aligned, else element by element
__device__ __forceinline__ void load_items(const uint8_t* __restrict__ in, int64_t base, int64_t n,
                                           int (&v)[kItems]) {
  const uint8_t* p = in + base;
  if (base + kItems <= n && (((uintptr_t)p) & 7) == 0) {
    const uint64_t w = *(const uint64_t*)p;
#pragma unroll
    for (int k = 0; k < kItems; ++k) v[k] = (int)((w >> (8 * k)) & 255u);
  } else {
#pragma unroll
    for (int k = 0; k < kItems; ++k) v[k] = (base + k < n) ? (int)p[k] : 0;
  }
}

__device__ __forceinline__ void load_items(const int* __restrict__ in, int64_t base, int64_t n,
                                           int (&v)[kItems]) {
  const int* p = in + base;
  if (base + kItems <= n && (((uintptr_t)p) & 15) == 0) {
    const int4 a = *(const int4*)p, b = *(const int4*)(p + 4);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
    v[4] = b.x;
    v[5] = b.y;
    v[6] = b.z;
    v[7] = b.w;
  } else {
#pragma unroll
    for (int k = 0; k < kItems; ++k) v[k] = (base + k < n) ? p[k] : 0;
  }
}

// a thread's kItems consecutive codes reduced to their low bits (two 16-byte loads when aligned)
__device__ __forceinline__ void load_items(const uint32_t* __restrict__ in, int64_t base, int64_t n,
                                           int (&v)[kItems]) {
  const uint32_t* p = in + base;
  if (base + kItems <= n && (((uintptr_t)p) & 15) == 0) {
    const uint4 a = *(const uint4*)p, b = *(const uint4*)(p + 4);
    v[0] = (int)(a.x & 1u);
    v[1] = (int)(a.y & 1u);
    v[2] = (int)(a.z & 1u);
    v[3] = (int)(a.w & 1u);
    v[4] = (int)(b.x & 1u);
    v[5] = (int)(b.y & 1u);
    v[6] = (int)(b.z & 1u);
    v[7] = (int)(b.w & 1u);
  } else {
#pragma unroll
    for (int k = 0; k < kItems; ++k) v[k] = (base + k < n) ? (int)(p[k] & 1u) : 0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_tile_reduce(const T* __restrict__ in, int64_t n,
                                                        int* __restrict__ sums) {
  int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  int v[kItems];
  load_items(in, base, n, v);
  int s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) s += v[k];
  int tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// rescan every tile; tile_off == nullptr means a single tile with offset 0.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_tile_scan(const T* __restrict__ in, int64_t n,
                                                      const int* __restrict__ tile_off,
                                                      int* __restrict__ out) {
  int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  int v[kItems];
  load_items(in, base, n, v);
  int s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) s += v[k];
  int tot;
  int run = block_excl_scan(s, &tot) + (tile_off ? tile_off[blockIdx.x] : 0);
  int o[kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    o[k] = run;
    run += v[k];
  }
  int* q = out + base;
  if (base + kItems <= n && (((uintptr_t)q) & 15) == 0) {
    *(int4*)q = make_int4(o[0], o[1], o[2], o[3]);
    *(int4*)(q + 4) = make_int4(o[4], o[5], o[6], o[7]);
  } else {
#pragma unroll
    for (int k = 0; k < kItems; ++k)
      if (base + k < n) q[k] = o[k];
  }
  // the thread owning element n-1 (or thread 0 of an empty range) writes the total
  if (n == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
  } else if (base <= n - 1 && n - 1 < base + kItems) {
    out[n] = run;
  }
}

static int64_t n_tiles(int64_t n) { return n <= 0 ? 1 : (n + kTile - 1) / kTile; }

static size_t scan_ws_ints(int64_t n) {
  // tile sums and their exclusive scan (nt + 1) at every recursion level
  size_t total = 0;
  int64_t m = n;
  while (m > kTile) {
    int64_t nt = n_tiles(m);
    total += (size_t)nt + (size_t)(nt + 1);
    m = nt;
  }
  return total + 64;
}

template <typename T>
static int scan_rec(const T* in, int64_t n, int* out, int* ws, hipStream_t st) {
  if (n <= kTile) {
    hipLaunchKernelGGL(k_tile_scan<T>, dim3(1), dim3(kBlock), 0, st, in, n, (const int*)nullptr,
                       out);
    return launch_status();
  }
  int64_t nt = n_tiles(n);
  int* sums = ws;
  int* offs = ws + nt;
  hipLaunchKernelGGL(k_tile_reduce<T>, dim3((unsigned)nt), dim3(kBlock), 0, st, in, n, sums);
  int rc = launch_status();
  if (rc) return rc;
  rc = scan_rec<int>(sums, nt, offs, ws + nt + (nt + 1), st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_tile_scan<T>, dim3((unsigned)nt), dim3(kBlock), 0, st, in, n,
                     (const int*)offs, out);
  return launch_status();
}

int scan_u8(const uint8_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes,
            hipStream_t st) {
  if (n < 0 || !out || (n > 0 && !in)) return PACES_E_INVALID;
  if (n >= 2147483647LL) return PACES_E_INVALID;
  if (ws_bytes < scan_ws_ints(n) * sizeof(int)) return PACES_E_WORKSPACE;
  return scan_rec<uint8_t>(in, n, out, (int*)ws, st);
}

int scan_i32(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes,
             hipStream_t st) {
  if (n < 0 || !out || (n > 0 && !in)) return PACES_E_INVALID;
  if (n >= 2147483647LL) return PACES_E_INVALID;
  if (ws_bytes < scan_ws_ints(n) * sizeof(int)) return PACES_E_WORKSPACE;
  return scan_rec<int>(in, n, out, (int*)ws, st);
}

size_t scan_ws_bytes(int64_t n) { return scan_ws_ints(n) * sizeof(int); }

// The scan's first half for a caller that rescans the tiles itself: offs[t] = number of flags
// set before tile t (kScanTile elements each), for t in [0, tiles]; offs[tiles] is the total.
size_t scan_tile_offsets_ws_bytes(int64_t n) {
  const int64_t nt = n_tiles(n);
  return ((size_t)nt + scan_ws_ints(nt)) * sizeof(int);
}

// (the low bit of each 32-bit code is its flag)
int scan_tile_offsets_bit0(const uint32_t* in, int64_t n, int32_t* offs, void* ws, size_t ws_bytes,
                           hipStream_t st) {
  if (n < 0 || !offs || (n > 0 && !in)) return PACES_E_INVALID;
  if (n >= 2147483647LL) return PACES_E_INVALID;
  if (ws_bytes < scan_tile_offsets_ws_bytes(n)) return PACES_E_WORKSPACE;
  const int64_t nt = n_tiles(n);
  int* sums = (int*)ws;
  hipLaunchKernelGGL(k_tile_reduce<uint32_t>, dim3((unsigned)nt), dim3(kBlock), 0, st, in, n,
                     sums);
  int rc = launch_status();
  if (rc) return rc;
  return scan_rec<int>(sums, nt, offs, sums + nt, st);
}

}  // namespace paces

extern "C" {

size_t paces_scan_workspace_bytes(int64_t n) { return paces::scan_ws_bytes(n); }

int paces_exclusive_scan_u8(const uint8_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes,
                            void* stream) {
  return paces::scan_u8(in, n, out, ws, ws_bytes, (hipStream_t)stream);
}

int paces_exclusive_scan_i32(const int32_t* in, int64_t n, int32_t* out, void* ws,
                             size_t ws_bytes, void* stream) {
  return paces::scan_i32(in, n, out, ws, ws_bytes, (hipStream_t)stream);
}

}  // extern "C"
