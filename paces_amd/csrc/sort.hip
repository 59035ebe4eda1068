// A7 general sort + dedup of packed rows (cupy_unique, paces/aux/helpers.py:63-75, as used by
// enlarge_basis_set, paces/core/hamiltonian.py:550-577): LSD radix sort on 4-bit digits, from
// the least significant bit of the last word to the most significant bit of word 0, with a
// stable scatter per pass, so the result is lexicographic row order. Digit positions at which
// every row agrees are skipped (paces_rows_varying_bits). The step itself merges sorted runs
// (paces_merge_unique); this is the route for rows in no particular order.
#include "common.hpp"

namespace paces {

int scan_i32(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
int scan_u8(const uint8_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws_bytes(int64_t n);

static size_t a256s(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr int kSortKpt = 8;                      // keys per thread
constexpr int kSortTile = kBlock * kSortKpt;     // keys per workgroup

__global__ void __launch_bounds__(kBlock) k_zero_words(u64* __restrict__ p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// mask[w] |= row[i][w] ^ row[0][w]: the bits that are not the same in every row
template <int W>
__global__ void __launch_bounds__(kBlock) k_varying(const u64* __restrict__ rows, int64_t n,
                                                    u64* __restrict__ mask) {
  __shared__ u64 sm[W];
  if ((int)threadIdx.x < W) sm[threadIdx.x] = 0;
  __syncthreads();
  u64 acc[W];
#pragma unroll
  for (int k = 0; k < W; ++k) acc[k] = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
#pragma unroll
    for (int k = 0; k < W; ++k) acc[k] |= rows[i * W + k] ^ rows[k];
  }
#pragma unroll
  for (int k = 0; k < W; ++k) {
    u64 v = acc[k];
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v |= __shfl_xor(v, s, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicOr(&sm[k], v);
  }
  __syncthreads();
  if ((int)threadIdx.x < W && sm[threadIdx.x]) atomicOr(&mask[threadIdx.x], sm[threadIdx.x]);
}

// digit of row i at (word w, shift); rows past n get 16 (no digit)
__device__ __forceinline__ int digit_of(const u64* __restrict__ rows, int64_t i, int64_t n, int W,
                                        int w, int shift) {
  return i < n ? (int)((rows[i * W + w] >> shift) & 15ull) : 16;
}

// per-workgroup digit counts, digit-major: hist[d * nblk + b]
__global__ void __launch_bounds__(kBlock) k_sort_hist(const u64* __restrict__ rows, int64_t n, int W,
                                                      int w, int shift, int64_t nblk,
                                                      int32_t* __restrict__ hist) {
  __shared__ int32_t cnt[16];
  if (threadIdx.x < 16) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)threadIdx.x * kSortKpt;
  u64 c = 0;  // 16 nibble counters (at most 8 keys per thread)
#pragma unroll
  for (int q = 0; q < kSortKpt; ++q) {
    const int d = digit_of(rows, base + q, n, W, w, shift);
    if (d < 16) c += 1ull << (4 * d);
  }
#pragma unroll
  for (int d = 0; d < 16; ++d) {
    const int v = (int)((c >> (4 * d)) & 15ull);
    if (v) atomicAdd(&cnt[d], v);
  }
  __syncthreads();
  if (threadIdx.x < 16) hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter: key k of thread t (keys in thread order within the tile) goes to
// G[d * nblk + b] + (keys of digit d in threads before t) + (earlier keys of digit d in t).
template <int W>
__global__ void __launch_bounds__(kBlock) k_sort_scatter(const u64* __restrict__ in,
                                                         const int32_t* __restrict__ pin,
                                                         int64_t n, int w, int shift,
                                                         const int32_t* __restrict__ G,
                                                         int64_t nblk, u64* __restrict__ out,
                                                         int32_t* __restrict__ pout) {
  __shared__ int32_t s[16][kBlock];
  __shared__ int32_t seg[16][16];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)t * kSortKpt;
  int dig[kSortKpt];
  u64 c = 0;
#pragma unroll
  for (int q = 0; q < kSortKpt; ++q) {
    dig[q] = digit_of(in, base + q, n, W, w, shift);
    if (dig[q] < 16) c += 1ull << (4 * dig[q]);
  }
#pragma unroll
  for (int d = 0; d < 16; ++d) s[d][t] = (int32_t)((c >> (4 * d)) & 15ull);
  __syncthreads();
  // exclusive scan over threads of each digit's counts: 16 segments of 16 threads per digit
  {
    const int d = t & 15, g = t >> 4;
    int32_t sum = 0;
    for (int k = 0; k < 16; ++k) sum += s[d][16 * g + k];
    seg[d][g] = sum;
  }
  __syncthreads();
  if (t < 16) {
    int32_t run = 0;
    for (int g = 0; g < 16; ++g) {
      const int32_t v = seg[t][g];
      seg[t][g] = run;
      run += v;
    }
  }
  __syncthreads();
  {
    const int d = t & 15, g = t >> 4;
    int32_t run = seg[d][g];
    for (int k = 0; k < 16; ++k) {
      const int32_t v = s[d][16 * g + k];
      s[d][16 * g + k] = run;
      run += v;
    }
  }
  __syncthreads();
  u64 seen = 0;
#pragma unroll
  for (int q = 0; q < kSortKpt; ++q) {
    const int d = dig[q];
    if (d >= 16) continue;
    const int32_t local = (int32_t)((seen >> (4 * d)) & 15ull);
    seen += 1ull << (4 * d);
    const int64_t pos = (int64_t)G[(int64_t)d * nblk + blockIdx.x] + s[d][t] + local;
    const int64_t i = base + q;
#pragma unroll
    for (int k = 0; k < W; ++k) out[pos * W + k] = in[i * W + k];
    pout[pos] = pin ? pin[i] : (int32_t)i;
  }
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_adjacent_new(const u64* __restrict__ rows, int64_t n,
                                                         uint8_t* __restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool nw = i == 0;
  if (!nw) {
#pragma unroll
    for (int k = 0; k < W; ++k) nw |= rows[i * W + k] != rows[(i - 1) * W + k];
  }
  flag[i] = nw ? 1 : 0;
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_compact_sorted(const u64* __restrict__ rows,
                                                           const int32_t* __restrict__ perm,
                                                           int64_t n,
                                                           const uint8_t* __restrict__ flag,
                                                           const int32_t* __restrict__ G,
                                                           u64* __restrict__ out,
                                                           int32_t* __restrict__ pout,
                                                           int32_t* __restrict__ count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *count = G[n];
  if (i >= n || !flag[i]) return;
  const int32_t j = G[i];
#pragma unroll
  for (int k = 0; k < W; ++k) out[(int64_t)j * W + k] = rows[i * W + k];
  if (pout) pout[j] = perm[i];
}

__global__ void k_write_count(int32_t* __restrict__ count, int32_t v) { *count = v; }

__global__ void __launch_bounds__(kBlock) k_iota_sort(int32_t* __restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int32_t)i;
}

struct SortWs {
  u64* rows;    // n * W
  int32_t* p0;  // n
  int32_t* p1;  // n
  int32_t* hist;
  int32_t* G;
  uint8_t* flag;
  int32_t* F;
  void* scan;
  size_t scan_bytes;
};

static size_t sort_ws(int64_t n, int W, SortWs* s, char* base) {
  const int64_t nblk = (n + kSortTile - 1) / kSortTile;
  const int64_t m = 16 * (nblk > 0 ? nblk : 1);
  const int64_t nn = n > 0 ? n : 1;
  size_t off = 0;
  auto take = [&](size_t b) {
    char* r = base ? base + off : nullptr;
    off += a256s(b);
    return r;
  };
  char* rows = take((size_t)nn * W * 8);
  char* p0 = take((size_t)nn * 4);
  char* p1 = take((size_t)nn * 4);
  char* hist = take((size_t)m * 4);
  char* G = take((size_t)(m + 1) * 4);
  char* flag = take((size_t)nn);
  char* F = take((size_t)(nn + 1) * 4);
  const size_t sb = scan_ws_bytes(m > nn ? m : nn);
  char* sc = take(sb);
  if (s) {
    s->rows = (u64*)rows;
    s->p0 = (int32_t*)p0;
    s->p1 = (int32_t*)p1;
    s->hist = (int32_t*)hist;
    s->G = (int32_t*)G;
    s->flag = (uint8_t*)flag;
    s->F = (int32_t*)F;
    s->scan = sc;
    s->scan_bytes = sb;
  }
  return off;
}

template <int W>
static int run_varying(const uint64_t* rows, int64_t n, uint64_t* mask, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(kBlock), 0, st, (u64*)mask, W);
  if (n > 1) {
    int64_t g = (n + kBlock - 1) / kBlock;
    if (g > 1024) g = 1024;
    hipLaunchKernelGGL(k_varying<W>, dim3((unsigned)g), dim3(kBlock), 0, st, (const u64*)rows, n,
                       (u64*)mask);
  }
  return launch_status();
}

template <int W>
static int run_sort(const uint64_t* rows, int64_t n, const uint64_t* varying, int unique,
                    uint64_t* out_rows, int32_t* out_perm, int32_t* out_count, void* ws,
                    hipStream_t st) {
  SortWs s;
  sort_ws(n, W, &s, (char*)ws);
  const int64_t nblk = (n + kSortTile - 1) / kSortTile;
  // the passes that can change the order
  int wl[16 * PACES_MAX_W], sl[16 * PACES_MAX_W], np = 0;
  for (int w = W - 1; w >= 0; --w)
    for (int d = 0; d < 16; ++d)
      if (!varying || ((varying[w] >> (4 * d)) & 15ull)) {
        wl[np] = w;
        sl[np] = 4 * d;
        ++np;
      }
  // ping-pong so the last pass writes the final buffer: out_rows when no dedup follows,
  // else the workspace copy that the compaction reads
  u64* fin = unique ? s.rows : (u64*)out_rows;
  u64* oth = unique ? (u64*)out_rows : s.rows;
  int32_t* pfin = s.p0;
  int32_t* poth = s.p1;
  const u64* src = (const u64*)rows;
  const int32_t* psrc = nullptr;
  int rc;
  if (np == 0) {
    if (hipMemcpyAsync(fin, rows, (size_t)n * W * 8, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return PACES_E_HIP;
    hipLaunchKernelGGL(k_iota_sort, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, pfin, n);
  }
  for (int p = 0; p < np; ++p) {
    // pass p writes fin when (np - 1 - p) is even
    const bool to_fin = ((np - 1 - p) & 1) == 0;
    u64* dst = to_fin ? fin : oth;
    int32_t* pdst = to_fin ? pfin : poth;
    hipLaunchKernelGGL(k_sort_hist, dim3((unsigned)nblk), dim3(kBlock), 0, st, src, n, W, wl[p],
                       sl[p], nblk, s.hist);
    rc = launch_status();
    if (rc) return rc;
    rc = scan_i32(s.hist, 16 * nblk, s.G, s.scan, s.scan_bytes, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sort_scatter<W>, dim3((unsigned)nblk), dim3(kBlock), 0, st, src, psrc, n,
                       wl[p], sl[p], (const int32_t*)s.G, nblk, dst, pdst);
    rc = launch_status();
    if (rc) return rc;
    src = dst;
    psrc = pdst;
  }
  if (!unique) {
    if (out_perm &&
        hipMemcpyAsync(out_perm, pfin, (size_t)n * 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return PACES_E_HIP;
    if (out_count) hipLaunchKernelGGL(k_write_count, dim3(1), dim3(1), 0, st, out_count, (int32_t)n);
    return launch_status();
  }
  hipLaunchKernelGGL(k_adjacent_new<W>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st,
                     (const u64*)fin, n, s.flag);
  rc = launch_status();
  if (rc) return rc;
  rc = scan_u8(s.flag, n, s.F, s.scan, s.scan_bytes, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_compact_sorted<W>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st,
                     (const u64*)fin, (const int32_t*)pfin, n, (const uint8_t*)s.flag,
                     (const int32_t*)s.F, (u64*)out_rows, out_perm, out_count);
  return launch_status();
}

}  // namespace paces

using namespace paces;

extern "C" {

int paces_rows_varying_bits(const uint64_t* rows, int64_t n, int32_t W, uint64_t* mask,
                            void* stream) {
  if (n < 0 || !mask || (n > 0 && !rows)) return PACES_E_INVALID;
  PACES_DISPATCH_W(W, run_varying, rows, n, mask, (hipStream_t)stream);
}

size_t paces_sort_rows_workspace_bytes(int64_t n, int32_t W) {
  if (n < 0 || W < 1) return 0;
  return sort_ws(n, W, nullptr, nullptr);
}

int paces_sort_rows(const uint64_t* rows, int64_t n, int32_t W, const uint64_t* varying,
                    int32_t unique, uint64_t* out_rows, int32_t* out_perm, int32_t* out_count,
                    void* ws, size_t ws_bytes, void* stream) {
  if (n < 0 || W < 1 || !ws) return PACES_E_INVALID;
  if (n > 0 && (!rows || !out_rows)) return PACES_E_INVALID;
  if (unique && !out_count) return PACES_E_INVALID;
  if (n >= 2147483647LL) return PACES_E_INVALID;
  if (ws_bytes < sort_ws(n, W, nullptr, nullptr)) return PACES_E_WORKSPACE;
  if (n == 0) {
    if (out_count) hipLaunchKernelGGL(k_write_count, dim3(1), dim3(1), 0, (hipStream_t)stream,
                                      out_count, 0);
    return launch_status();
  }
  PACES_DISPATCH_W(W, run_sort, rows, n, varying, unique ? 1 : 0, out_rows, out_perm, out_count,
                   ws, (hipStream_t)stream);
}

}  // extern "C"
