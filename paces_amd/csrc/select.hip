// F1: top-k select of the basis (_determine_select_whoami, paces/core/time_evolution.py:833-910,
// shuffle_seed=None) as a radix select plus a stable compaction.
//
// The key of row i is the bit pattern of |psi_i|^2 = re*re + im*im (a non-negative double, so
// its bits order like the value). Eight 8-bit digit passes find the k-th largest key T and the
// number `need` of rows equal to T that are kept; the reference's stable ascending argsort keeps
// the last `need` of those ties, i.e. the ones with the highest indices. Kept rows are written in
// index order, which for a sorted basis is lexicographic order (the lexsort at 902-903).
#include "common.hpp"

namespace paces {

int scan_u8(const uint8_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws_bytes(int64_t n);

struct SelState {
  u64 prefix;
  u64 mask;
  long long krem;  // rows still to take among those matching prefix under mask
  long long pad;
};

__global__ void __launch_bounds__(kBlock) k_sel_keys(const double2* __restrict__ psi, int64_t R,
                                                     u64* __restrict__ keys) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  double2 v = psi[i];
  double k = v.x * v.x + v.y * v.y;  // compiled without FMA contraction
  keys[i] = (u64)__double_as_longlong(k);
}

__global__ void k_sel_init(SelState* s, unsigned* hist, long long k) {
  if (threadIdx.x == 0) {
    s->prefix = 0;
    s->mask = 0;
    s->krem = k;
    s->pad = 0;
  }
  hist[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(kBlock) k_sel_hist(const u64* __restrict__ keys, int64_t R,
                                                     const SelState* __restrict__ s, int shift,
                                                     unsigned* __restrict__ hist) {
  __shared__ unsigned lh[256];
  lh[threadIdx.x] = 0;
  __syncthreads();
  const u64 prefix = s->prefix, mask = s->mask;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < R;
       i += (int64_t)gridDim.x * blockDim.x) {
    u64 k = keys[i];
    if ((k & mask) == prefix) atomicAdd(&lh[(k >> shift) & 255u], 1u);
  }
  __syncthreads();
  unsigned c = lh[threadIdx.x];
  if (c) atomicAdd(&hist[threadIdx.x], c);
}

// one workgroup of 256 threads: pick the digit holding the krem-th largest matching key
__global__ void __launch_bounds__(256) k_sel_pick(SelState* s, unsigned* hist, int shift) {
  __shared__ long long suf[257];
  const int t = threadIdx.x;
  long long c = hist[t];
  // suffix sums S(t) = sum_{u >= t} hist[u], computed with a simple LDS loop (256 entries)
  suf[t] = c;
  if (t == 0) suf[256] = 0;
  __syncthreads();
  if (t == 0) {
    for (int u = 255; u >= 0; --u) suf[u] += suf[u + 1];
  }
  __syncthreads();
  const long long krem = s->krem;
  const bool mine = (suf[t] >= krem) && (suf[t + 1] < krem);
  __syncthreads();
  if (mine) {
    s->prefix |= ((u64)t) << shift;
    s->mask |= ((u64)255) << shift;
    s->krem = krem - suf[t + 1];
  }
  hist[t] = 0;
}

__global__ void __launch_bounds__(kBlock) k_sel_ties(const u64* __restrict__ keys, int64_t R,
                                                     const SelState* __restrict__ s,
                                                     uint8_t* __restrict__ tie) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  tie[i] = (keys[i] == s->prefix) ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_sel_keep(const u64* __restrict__ keys, int64_t R,
                                                     const SelState* __restrict__ s,
                                                     const int32_t* __restrict__ TG,
                                                     uint8_t* __restrict__ keep) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const u64 T = s->prefix;
  const long long need = s->krem;
  const long long nt = TG[R];
  u64 k = keys[i];
  bool kp = (k > T) || (k == T && (long long)TG[i] >= nt - need);
  keep[i] = kp ? 1 : 0;
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_sel_compact(const u64* __restrict__ rows, int64_t R,
                                                        const uint8_t* __restrict__ keep,
                                                        const int32_t* __restrict__ G,
                                                        int64_t k, u64* __restrict__ out,
                                                        int32_t* __restrict__ out_index) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R || !keep[i]) return;
  int64_t p = G[i];
  if (p >= k) return;  // exactly k rows are kept; never write past the caller's buffer
  store_row<W>(out, p, load_row<W>(rows, i));
  if (out_index) out_index[p] = (int32_t)i;
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_copy_all(const u64* __restrict__ rows, int64_t R,
                                                     u64* __restrict__ out,
                                                     int32_t* __restrict__ out_index) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  store_row<W>(out, i, load_row<W>(rows, i));
  if (out_index) out_index[i] = (int32_t)i;
}

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t select_ws_bytes(int64_t R) {
  return a256(sizeof(SelState)) + a256(256 * 4) + a256((size_t)R * 8) + 2 * a256((size_t)R) +
         2 * a256((size_t)(R + 1) * 4) + a256(scan_ws_bytes(R));
}

template <int W>
static int run_select(const double* psi, const uint64_t* rows, int64_t R, int64_t k,
                      uint64_t* out_rows, int32_t* out_index, void* ws, size_t ws_bytes,
                      hipStream_t st) {
  if (R == 0 || k == 0) return PACES_OK;
  if (k >= R) {
    hipLaunchKernelGGL(k_copy_all<W>, dim3(grid_for(R, kBlock)), dim3(kBlock), 0, st,
                       (const u64*)rows, R, (u64*)out_rows, out_index);
    return launch_status();
  }
  if (ws_bytes < select_ws_bytes(R)) return PACES_E_WORKSPACE;
  char* p = (char*)ws;
  SelState* s = (SelState*)p;
  p += a256(sizeof(SelState));
  unsigned* hist = (unsigned*)p;
  p += a256(256 * 4);
  u64* keys = (u64*)p;
  p += a256((size_t)R * 8);
  uint8_t* tie = (uint8_t*)p;
  p += a256((size_t)R);
  uint8_t* keep = (uint8_t*)p;
  p += a256((size_t)R);
  int32_t* TG = (int32_t*)p;
  p += a256((size_t)(R + 1) * 4);
  int32_t* KG = (int32_t*)p;
  p += a256((size_t)(R + 1) * 4);
  void* sws = p;
  size_t sws_bytes = scan_ws_bytes(R);

  const unsigned gR = grid_for(R, kBlock);
  const unsigned gH = gR < 2048u ? gR : 2048u;
  hipLaunchKernelGGL(k_sel_keys, dim3(gR), dim3(kBlock), 0, st, (const double2*)psi, R, keys);
  hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(256), 0, st, s, hist, (long long)k);
  for (int pass = 0; pass < 8; ++pass) {
    int shift = 56 - 8 * pass;
    hipLaunchKernelGGL(k_sel_hist, dim3(gH), dim3(kBlock), 0, st, (const u64*)keys, R,
                       (const SelState*)s, shift, hist);
    hipLaunchKernelGGL(k_sel_pick, dim3(1), dim3(256), 0, st, s, hist, shift);
  }
  hipLaunchKernelGGL(k_sel_ties, dim3(gR), dim3(kBlock), 0, st, (const u64*)keys, R,
                     (const SelState*)s, tie);
  int rc = launch_status();
  if (rc) return rc;
  rc = scan_u8(tie, R, TG, sws, sws_bytes, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sel_keep, dim3(gR), dim3(kBlock), 0, st, (const u64*)keys, R,
                     (const SelState*)s, (const int32_t*)TG, keep);
  rc = launch_status();
  if (rc) return rc;
  rc = scan_u8(keep, R, KG, sws, sws_bytes, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sel_compact<W>, dim3(gR), dim3(kBlock), 0, st, (const u64*)rows, R,
                     (const uint8_t*)keep, (const int32_t*)KG, k, (u64*)out_rows, out_index);
  return launch_status();
}

}  // namespace paces

using namespace paces;

extern "C" {

size_t paces_select_workspace_bytes(int64_t R) { return select_ws_bytes(R); }

int paces_select_topk(const double* psi, const uint64_t* rows, int64_t R, int32_t W, int64_t k,
                      uint64_t* out_rows, int32_t* out_index, void* ws, size_t ws_bytes,
                      void* stream) {
  if (R < 0 || k < 0 || W < 1 || R >= 2147483647LL) return PACES_E_INVALID;
  if (R > 0 && k > 0 && (!psi || !rows || !out_rows)) return PACES_E_INVALID;
  PACES_DISPATCH_W(W, run_select, psi, rows, R, k, out_rows, out_index, ws, ws_bytes,
                   (hipStream_t)stream);
}

}  // extern "C"
