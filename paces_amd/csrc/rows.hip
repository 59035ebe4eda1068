// Packed-row primitives: pack / field read / field add (A2-A5), lexicographic lower_bound (B1),
// union of sorted unique runs (A7) and amplitude reassignment (G1).
#include <mutex>

#include "common.hpp"

namespace paces {

int scan_u8(const uint8_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
int scan_i32(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws_bytes(int64_t n);
size_t scan_tile_offsets_ws_bytes(int64_t n);
int scan_tile_offsets_bit0(const uint32_t* in, int64_t n, int32_t* offs, void* ws, size_t ws_bytes,
                           hipStream_t st);

// ------------------------------------------------------------------------------------------------
// A2-A5 (paces/core/hamiltonian.py:218-360, paces/models/_kernels/_holstein_kernels.py:96-179)
// ------------------------------------------------------------------------------------------------

template <int W>
__global__ void __launch_bounds__(kBlock) k_pack(const int32_t* __restrict__ raw, int64_t n,
                                                 int nfields, const int32_t* __restrict__ off,
                                                 const int32_t* __restrict__ width,
                                                 u64* __restrict__ rows) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Row<W> r;
#pragma unroll
  for (int k = 0; k < W; ++k) r.w[k] = 0;
  for (int f = 0; f < nfields; ++f) {
    int b = width[f];
    if (b) add_field<W>(r, off[f], b, (int64_t)(uint32_t)raw[i * nfields + f]);
  }
  store_row<W>(rows, i, r);
}

// field >= 0: that field; field == -1: the phonon field at the exciton (field pos+1)
template <int W>
__global__ void __launch_bounds__(kBlock) k_get_field(const u64* __restrict__ rows, int64_t n,
                                                      int field, const int32_t* __restrict__ off,
                                                      const int32_t* __restrict__ width,
                                                      int32_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Row<W> r = load_row<W>(rows, i);
  int f = field;
  if (f < 0) f = (int)get_field<W>(r, off[0], width[0]) + 1;
  out[i] = (int32_t)get_field<W>(r, off[f], width[f]);
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_add_field(u64* __restrict__ rows, int64_t n, int field,
                                                      int delta, const int32_t* __restrict__ off,
                                                      const int32_t* __restrict__ width) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Row<W> r = load_row<W>(rows, i);
  int f = field;
  if (f < 0) f = (int)get_field<W>(r, off[0], width[0]) + 1;
  add_field<W>(r, off[f], width[f], delta);
  store_row<W>(rows, i, r);
}

template <int W>
static int run_pack(const int32_t* raw, int64_t n, const paces_layout_t* L, uint64_t* rows,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_pack<W>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, raw, n,
                     L->nfields, L->d_off, L->d_width, (u64*)rows);
  return launch_status();
}

template <int W>
static int run_get(const uint64_t* rows, int64_t n, const paces_layout_t* L, int field,
                   int32_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_get_field<W>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st,
                     (const u64*)rows, n, field, L->d_off, L->d_width, out);
  return launch_status();
}

template <int W>
static int run_add(uint64_t* rows, int64_t n, const paces_layout_t* L, int field, int delta,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_add_field<W>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, (u64*)rows,
                     n, field, delta, L->d_off, L->d_width);
  return launch_status();
}

static bool layout_ok(const paces_layout_t* L) {
  return L && L->W >= 1 && L->W <= PACES_MAX_ROW_WORDS && L->nfields >= 1 && L->d_off &&
         L->d_width;
}

// ------------------------------------------------------------------------------------------------
// B1 lower_bound (paces/aux/cupy_search.py:149-228)
// ------------------------------------------------------------------------------------------------

template <int W>
__global__ void __launch_bounds__(kBlock) k_lower_bound(const u64* __restrict__ book, int64_t M,
                                                        const u64* __restrict__ q, int64_t Q,
                                                        int32_t* __restrict__ out_idx,
                                                        uint8_t* __restrict__ out_hit,
                                                        int32_t* __restrict__ miss) {
  __shared__ LookupSmem<W> sm;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool v = i < Q;
  Row<W> r;
#pragma unroll
  for (int k = 0; k < W; ++k) r.w[k] = 0;
  if (v) r = load_row<W>(q, i);
  int64_t lb = block_lower_bound<W>(r, v, book, M, sm);
  if (!v) return;
  bool hit = (lb < M) && eq_row<W>(r, book + lb * W);
  if (out_idx) out_idx[i] = (int32_t)lb;
  if (out_hit) out_hit[i] = hit ? 1 : 0;
  if (miss && !hit) atomicAdd(miss, 1);
}

// Partition pass for sorted queries: the lower bounds of the first and last query of every tile
// of kBlock queries, one thread each, so all the searches run concurrently.
template <int W>
__global__ void __launch_bounds__(kBlock) k_tile_bounds(const u64* __restrict__ q, int64_t Q,
                                                        const u64* __restrict__ book, int64_t M,
                                                        int64_t ntiles,
                                                        int64_t* __restrict__ lohi) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * ntiles) return;
  int64_t t = e >> 1;
  int64_t qi = (e & 1) ? ((t + 1) * kBlock - 1 < Q - 1 ? (t + 1) * kBlock - 1 : Q - 1)
                       : t * kBlock;
  Row<W> r = load_row<W>(q, qi);
  lohi[e] = lower_bound_row<W>(book, 0, M, r);
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_lower_bound_sorted(
    const u64* __restrict__ book, int64_t M, const u64* __restrict__ q, int64_t Q,
    const int64_t* __restrict__ lohi, int32_t* __restrict__ out_idx, uint8_t* __restrict__ out_hit,
    int32_t* __restrict__ miss) {
  __shared__ LookupSmem<W> sm;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool v = i < Q;
  Row<W> r;
#pragma unroll
  for (int k = 0; k < W; ++k) r.w[k] = 0;
  if (v) r = load_row<W>(q, i);
  int64_t lb = block_lower_bound_range<W>(r, v, book, lohi[2 * blockIdx.x],
                                          lohi[2 * blockIdx.x + 1], sm);
  if (!v) return;
  bool hit = (lb < M) && eq_row<W>(r, book + lb * W);
  if (out_idx) out_idx[i] = (int32_t)lb;
  if (out_hit) out_hit[i] = hit ? 1 : 0;
  if (miss && !hit) atomicAdd(miss, 1);
}

size_t lb_ws_bytes(int64_t Q) { return (size_t)(2 * ((Q + kBlock - 1) / kBlock) + 2) * 8; }

template <int W>
static int run_lb(const uint64_t* book, int64_t M, const uint64_t* q, int64_t Q, int q_sorted,
                  int32_t* idx, uint8_t* hit, int32_t* miss, void* ws, size_t ws_bytes,
                  hipStream_t st) {
  if (Q == 0) return PACES_OK;
  if (q_sorted && ws && ws_bytes >= lb_ws_bytes(Q)) {
    int64_t nt = (Q + kBlock - 1) / kBlock;
    int64_t* lohi = (int64_t*)ws;
    hipLaunchKernelGGL(k_tile_bounds<W>, dim3(grid_for(2 * nt, kBlock)), dim3(kBlock), 0, st,
                       (const u64*)q, Q, (const u64*)book, M, nt, lohi);
    hipLaunchKernelGGL(k_lower_bound_sorted<W>, dim3((unsigned)nt), dim3(kBlock), 0, st,
                       (const u64*)book, M, (const u64*)q, Q, (const int64_t*)lohi, idx, hit,
                       miss);
    return launch_status();
  }
  hipLaunchKernelGGL(k_lower_bound<W>, dim3(grid_for(Q, kBlock)), dim3(kBlock), 0, st,
                     (const u64*)book, M, (const u64*)q, Q, idx, hit, miss);
  return launch_status();
}

// ------------------------------------------------------------------------------------------------
// A7 union of K sorted unique runs (paces/aux/helpers.py:63-75 on the concatenation of
// paces/core/hamiltonian.py:434-452). Every element gets its position in the merge of all runs
// (ties by run order); an element equal to one of an earlier run is dropped, and the output index
// of a position is the number of kept elements before it. Dropped elements report the index of
// their kept copy. Passes: partition bounds, merged positions (k_mu_rank), tile sums of the kept
// flags, and the output in merged order (k_mu_emit).
// ------------------------------------------------------------------------------------------------

struct RunsArg {
  const u64* ptr[PACES_MAX_RUNS];
  int64_t off[PACES_MAX_RUNS + 1];
  int64_t boff[PACES_MAX_RUNS + 1];  // first workgroup of each run in k_mu_rank's grid
  int K;
  uint8_t code[PACES_MAX_RUNS];  // out_code[row] = code[run of the row's kept copy]
};

__device__ __forceinline__ int run_of(const RunsArg& R, int64_t g) {
  int r = 0;
  for (int s = 1; s < R.K; ++s) r += (g >= R.off[s]) ? 1 : 0;
  return r;
}

// the address of candidate g's row and its run, by a uniform loop of selects over the runs (no
// per-lane indexing of the argument struct, which would be a dependent load)
template <int W>
__device__ __forceinline__ const u64* cand_row(const RunsArg& R, int64_t g, int& run) {
  const u64* p = R.ptr[0] + g * W;
  int r = 0;
  for (int s = 1; s < R.K; ++s) {
    const bool in = g >= R.off[s];
    p = in ? R.ptr[s] + (g - R.off[s]) * W : p;
    r = in ? s : r;
  }
  run = r;
  return p;
}

__device__ __forceinline__ int run_of_block(const RunsArg& R, int64_t b) {
  int r = 0;
  for (int s = 1; s < R.K; ++s) r += (b >= R.boff[s]) ? 1 : 0;
  return r;
}

// Partition pass of the union: for every tile of every run and every other run s, the lower
// bounds in s of the tile's first and last rows (one thread each).
template <int W>
__global__ void __launch_bounds__(kBlock) k_mu_bounds(RunsArg R, int64_t* __restrict__ lohi) {
  const int64_t nblk = R.boff[R.K];
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nblk * R.K * 2) return;
  const int64_t b = e / (2 * R.K);
  const int s = (int)((e / 2) % R.K);
  const int r = run_of_block(R, b);
  if (s == r) {
    lohi[e] = 0;
    return;
  }
  const int64_t len_r = R.off[r + 1] - R.off[r];
  const int64_t t0 = (b - R.boff[r]) * kBlock;
  const int64_t own = (e & 1) ? ((t0 + kBlock - 1 < len_r - 1) ? t0 + kBlock - 1 : len_r - 1) : t0;
  Row<W> x = load_row<W>(R.ptr[r], own);
  lohi[e] = lower_bound_row<W>(R.ptr[s], 0, R.off[s + 1] - R.off[s], x);
}

// Merged-order schedule of k_mu_rank's tiles. A tile's rows land in the merged order near
// est = its first row's index + sum over the other runs of that row's lower bound there. Tiles
// sorted by est (buckets of 256 positions; order inside a bucket does not matter) and dealt to
// the XCDs in contiguous slices put the tiles that stage the same ranges of the other runs, and
// write the same stretch of the merged order, on one L2 at the same time.
constexpr int kXcds = 8;
constexpr int kSchedShift = 8;

__global__ void __launch_bounds__(kBlock) k_mu_sched_est(RunsArg R, const int64_t* __restrict__ lohi,
                                                         int32_t* __restrict__ est,
                                                         int32_t* __restrict__ cnt) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= R.boff[R.K]) return;
  const int r = run_of_block(R, b);
  int64_t e = (b - R.boff[r]) * kBlock;
  for (int s = 0; s < R.K; ++s)
    if (s != r) e += lohi[b * R.K * 2 + 2 * s];
  est[b] = (int32_t)e;
  atomicAdd(&cnt[e >> kSchedShift], 1);
}

__global__ void __launch_bounds__(kBlock) k_mu_sched_place(int64_t nblk,
                                                           const int32_t* __restrict__ est,
                                                           const int32_t* __restrict__ start,
                                                           int32_t* __restrict__ cnt,
                                                           int32_t* __restrict__ sched) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int bk = est[b] >> kSchedShift;
  sched[start[bk] + atomicSub(&cnt[bk], 1) - 1] = (int32_t)b;
}

__global__ void __launch_bounds__(kBlock) k_zero_i32r(int32_t* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// One workgroup per tile of one run: the tile's rows are sorted, so the rows of run s they can
// land among form the narrow range [lo, hi] the partition pass found (hi = lower bound of the
// tile's last row); that range, with the row at hi, is staged in LDS.
// Each row x of run r gets its position in the merge of all runs with ties broken by run order:
//   m(x) = own index + sum_{s < r} upper_bound_s(x) + sum_{s > r} lower_bound_s(x),
// a permutation of 0 .. C-1 since every run is sorted unique. x is kept (its first occurrence)
// when no earlier run holds it. Position m(x) of the merged order receives one 32-bit code,
// (candidate index << 1) | kept, which the output pass then walks in order.
// The other runs' ranges are packed into groups that fit the LDS buffer together; each group is
// copied global -> LDS by direct loads (16 bytes per lane, no registers), then searched, within
// one barrier pair. A range is copied as the 16-byte blocks that hold it: a block never crosses
// a page, so the few bytes read beyond the range cannot fault.
constexpr int kMuWords = 2048;  // 16 KiB: with the run table, LDS leaves room for 9 workgroups per CU

struct MuTable {
  PACES_GPTR(const unsigned char) blk[PACES_MAX_RUNS];  // first 16-byte block of each range
  long long lohi[2 * PACES_MAX_RUNS];                   // this tile's partition row
  int q[PACES_MAX_RUNS];       // slot -> run (every run but the tile's own, in order)
  int lo[PACES_MAX_RUNS];      // first staged row (index in its run)
  int n[PACES_MAX_RUNS];       // rows staged
  int nblk[PACES_MAX_RUNS];    // 16-byte blocks copied
  int base[PACES_MAX_RUNS];    // LDS word offset of the first block (even)
  int row0[PACES_MAX_RUNS];    // LDS word offset of row lo
  int gs[PACES_MAX_RUNS + 1];  // group g = slots [gs[g], gs[g + 1])
  int ng;
};

template <int W>
__global__ void __launch_bounds__(kBlock) k_mu_rank(RunsArg R, const int64_t* __restrict__ lohi,
                                                    const int32_t* __restrict__ sched, int per,
                                                    uint32_t* __restrict__ code) {
  // one LDS buffer for a group of ranges (a range larger than the buffer is searched in global
  // memory instead: LDS per workgroup stays at the buffer, for occupancy)
  __shared__ __align__(16) u64 buf[kMuWords];
  __shared__ MuTable T;
  const PACES_LPTR(const unsigned char) lbuf = (PACES_LPTR(const unsigned char))buf;
  // workgroups go to the XCDs round robin: XCD x runs the x-th slice of the merged-order schedule
  const int j = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  if (j >= (int)R.boff[R.K]) return;
  const int64_t tile = sched[j];
  const int r = run_of_block(R, tile);
  const int K = R.K;
  const int64_t* bl = lohi + tile * K * 2;
  // the tile's own row is loaded first, so its latency overlaps the table's
  const int64_t len_r = R.off[r + 1] - R.off[r];
  const int own = (int)((tile - R.boff[r]) * kBlock) + (int)threadIdx.x;
  const bool v = own < len_r;
  Row<W> x;
#pragma unroll
  for (int k = 0; k < W; ++k) x.w[k] = 0;
  if (v) x = load_row<W>(R.ptr[r], own);
  if ((int)threadIdx.x < 2 * K) T.lohi[threadIdx.x] = bl[threadIdx.x];
  __syncthreads();
  if (threadIdx.x == 0) {
    int i = 0, g = 0, words = 0;
    T.gs[0] = 0;
    for (int s = 0; s < K; ++s) {
      if (s == r) continue;
      const int lo = (int)T.lohi[2 * s], hi = (int)T.lohi[2 * s + 1];
      const int len = (int)(R.off[s + 1] - R.off[s]);
      const int n = (hi < len ? hi + 1 : len) - lo;
      const uintptr_t a0 = (uintptr_t)(R.ptr[s] + (int64_t)lo * W);
      const uintptr_t b0 = a0 & ~(uintptr_t)15, b1 = (a0 + (uintptr_t)n * W * 8 + 15) & ~(uintptr_t)15;
      const int nb = n > 0 ? (int)((b1 - b0) >> 4) : 0;
      const int w = 2 * nb;
      if (i > T.gs[g] && words + w > kMuWords) {
        T.gs[++g] = i;
        words = 0;
      }
      T.q[i] = s;
      T.lo[i] = lo;
      T.n[i] = n;
      T.nblk[i] = nb;
      T.blk[i] = (PACES_GPTR(const unsigned char))b0;
      T.base[i] = words;
      T.row0[i] = words + (int)((a0 - b0) >> 3);
      words += w;
      ++i;
      if (w > kMuWords) {  // a range too large for the buffer stands alone
        T.gs[++g] = i;
        words = 0;
      }
    }
    if (i > T.gs[g]) T.gs[++g] = i;
    T.ng = g;
  }
  __syncthreads();
  bool keep = true;
  int m = own;
  const int ng = T.ng;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int g = 0; g < ng; ++g) {
    const int a = T.gs[g], b = T.gs[g + 1];
    if (2 * T.nblk[a] > kMuWords) {  // a per-thread search over the partition range
      const int s = T.q[a];
      const int64_t len = R.off[s + 1] - R.off[s];
      const int64_t lb = v ? lower_bound_row<W>(R.ptr[s], T.lohi[2 * s], T.lohi[2 * s + 1], x)
                           : 0;
      if (v && s < r && lb < len && eq_row<W>(x, R.ptr[s] + lb * W)) {
        keep = false;
        m += (int)lb + 1;  // the equal row of the earlier run comes first
      } else {
        m += (int)lb;
      }
      continue;
    }
    // each wave copies 64 blocks (1 KiB) per load: lane l's block lands at the wave's base + 16 l
    for (int k = a; k < b; ++k) {
      const int nb = T.nblk[k];
      PACES_GPTR(const unsigned char) src = T.blk[k];
      const int base = T.base[k];
      for (int c0 = wave * 64; c0 < nb; c0 += kBlock) {
        if (c0 + lane < nb)
          __builtin_amdgcn_global_load_lds((PACES_GPTR(const void))(src + (size_t)(c0 + lane) * 16),
                                           (PACES_LPTR(void))(buf + base + 2 * c0), 16, 0, 0);
      }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (v) {
      for (int k = a; k < b; ++k) {
        const PACES_LPTR(const unsigned char) rows = lbuf + 8 * T.row0[k];
        const int n = T.n[k];
        if (T.q[k] < r) {  // earlier run: upper bound, and an equal row there drops x
          const int pos = bound_lds<W, true>(rows, n, x);
          if (pos > 0) {
            u64 w[W];
            load_lds_row<W>(rows, (unsigned)(pos - 1), w);
            bool eq = true;
#pragma unroll
            for (int j = 0; j < W; ++j) eq = eq & (w[j] == x.w[j]);
            keep = keep & !eq;
          }
          m += T.lo[k] + pos;
        } else {
          m += T.lo[k] + bound_lds<W, false>(rows, n, x);
        }
      }
    }
    __syncthreads();
  }
  if (v) code[m] = ((uint32_t)(R.off[r] + own) << 1) | (keep ? 1u : 0u);
}

// The union's output in merged order. Workgroup t takes merged positions [kScanTile t,
// kScanTile (t + 1)); thread x takes positions kScanTile t + kBlock k + x for k < kSteps, so a
// wave covers 64 consecutive positions at each step: the flag and index loads and the row stores
// are contiguous. A position's exclusive prefix of kept flags is the tile's offset, plus the wave
// totals before the wave's step, plus the kept lanes below it (mbcnt of the ballot). That prefix
// is a kept row's output index; a repeat's kept copy sits just before it in merged order, so a
// repeat's index is the prefix - 1.
constexpr uint32_t kMuNoCode = 0xFFFFFFFEu;  // flag 0, index 2^31 - 1: above every candidate

// Optional epilogue of the emit (DG): the Holstein diagonal of every output row, written with
// the row: eps + hw * (sum of the phonon fields, a weighted popcount) + de * field pf, in FP64
// or rounded to FP32 -- the expression and order of paces_holstein_diag.
struct RowDiag {
  const u64* sum_mask;
  const int32_t* off;
  const int32_t* width;
  int sum_bits, pf;
  double eps, hw, de;
  double* d64;
  float* d32;
};

template <int W>
__device__ __forceinline__ double row_diag(const Row<W>& r, const RowDiag& D) {
  long long s = 0;
#pragma unroll
  for (int w = 0; w < W; ++w)
    for (int k = 0; k < D.sum_bits; ++k)
      s += (long long)__popcll(r.w[w] & D.sum_mask[w * D.sum_bits + k]) << k;
  double d = D.eps + D.hw * (double)s;
  if (D.de != 0.0) d = d + D.de * (double)get_field<W>(r, D.off[D.pf], D.width[D.pf]);
  return d;
}

template <int W, bool DG>
__global__ void __launch_bounds__(kBlock) k_mu_emit(RunsArg R, const uint32_t* __restrict__ code,
                                                    const int32_t* __restrict__ offs,
                                                    u64* __restrict__ out,
                                                    uint8_t* __restrict__ from0,
                                                    int32_t* __restrict__ cand_pos,
                                                    uint8_t* __restrict__ out_code,
                                                    int32_t* __restrict__ out_count, RowDiag D) {
  constexpr int kSteps = kScanTile / kBlock;
  constexpr int kWaves = kBlock / 64;
  __shared__ int wpre[kSteps * kWaves];
  const int64_t C = R.off[R.K];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t t0 = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
  int f[kSteps], g[kSteps], pre[kSteps];
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    const int64_t m = t0 + k * kBlock;
    const uint32_t c = m < C ? code[m] : kMuNoCode;
    f[k] = (int)(c & 1u);
    g[k] = c == kMuNoCode ? -1 : (int)(c >> 1);
  }
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    const u64 b = __ballot(f[k] != 0);
    pre[k] = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32),
                                            __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
    if (lane == 0) wpre[k * kWaves + wave] = __popcll(b);
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // exclusive prefix of the wave totals, in position order
    int run = 0;
    for (int i = 0; i < kSteps * kWaves; ++i) {
      const int c = wpre[i];
      wpre[i] = run;
      run += c;
    }
  }
  __syncthreads();
  const int base = offs[blockIdx.x];
  // every step's row is loaded before any is stored (a position without a kept row reads
  // candidate 0's row, unused), so the gathers overlap
  Row<W> x[kSteps];
  int rr[kSteps];
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    const int gc = (f[k] && g[k] >= 0) ? g[k] : 0;
    x[k] = load_row<W>(cand_row<W>(R, gc, rr[k]), 0);
  }
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    if (g[k] < 0) continue;  // no candidate landed here (only for runs that repeat a row)
    const int G = base + wpre[k * kWaves + wave] + pre[k];
    const int pos = f[k] ? G : G - 1;
    if (cand_pos) cand_pos[g[k]] = pos;
    if (f[k]) {
      store_row<W>(out, pos, x[k]);
      if (from0) from0[pos] = (rr[k] == 0) ? 1 : 0;
      if (out_code) out_code[pos] = R.code[rr[k]];
      if (DG) {
        const double d = row_diag<W>(x[k], D);
        if (D.d64)
          D.d64[pos] = d;
        else
          D.d32[pos] = (float)d;  // the FP64 value rounded to nearest
      }
    }
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *out_count = offs[gridDim.x];
}

__global__ void k_write_zero(int32_t* p) { *p = 0; }

// merged-order codes start empty, so a position no row lands on (possible only for runs that
// repeat a row) adds nothing and is skipped
__global__ void __launch_bounds__(kBlock) k_mu_fill(uint32_t* __restrict__ code, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) code[i] = kMuNoCode;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static int64_t mu_scan_tiles(int64_t C) { return C <= 0 ? 1 : (C + kScanTile - 1) / kScanTile; }

static int64_t mu_sched_buckets(int64_t C) { return (C >> kSchedShift) + 1; }

size_t mu_ws_bytes(int64_t C, int K) {
  // the partition table holds 2K int64 per tile; runs add at most K partial tiles
  int64_t ntiles = (C + kBlock - 1) / kBlock + K;
  const int64_t nbk = mu_sched_buckets(C);
  const size_t sched = 2 * align256((size_t)ntiles * 4) + align256((size_t)nbk * 4) +
                       align256((size_t)(nbk + 1) * 4) + align256(scan_ws_bytes(nbk));
  const size_t scan = scan_tile_offsets_ws_bytes(C);
  return align256((size_t)C * 4) + align256((size_t)(mu_scan_tiles(C) + 1) * 4) +
         align256((size_t)ntiles * K * 2 * 8) + align256(sched > scan ? sched : scan);
}

template <int W>
static int run_mu(const RunsArg& R, uint64_t* out, int32_t* out_count, uint8_t* from0,
                  int32_t* cand_pos, uint8_t* out_code, const RowDiag& D, void* ws,
                  size_t ws_bytes, hipStream_t st) {
  int64_t C = R.off[R.K];
  if (ws_bytes < mu_ws_bytes(C, R.K)) return PACES_E_WORKSPACE;
  if (C == 0) {
    hipLaunchKernelGGL(k_write_zero, dim3(1), dim3(1), 0, st, out_count);
    return launch_status();
  }
  char* p = (char*)ws;
  uint32_t* code = (uint32_t*)p;
  p += align256((size_t)C * 4);
  const int64_t nst = mu_scan_tiles(C);
  int32_t* offs = (int32_t*)p;
  p += align256((size_t)(nst + 1) * 4);
  int64_t* lohi = (int64_t*)p;
  p += align256((size_t)((C + kBlock - 1) / kBlock + R.K) * R.K * 2 * 8);
  const int64_t nblk = R.boff[R.K];
  hipLaunchKernelGGL(k_mu_fill, dim3(grid_for(C, kBlock)), dim3(kBlock), 0, st, code, C);
  hipLaunchKernelGGL(k_mu_bounds<W>, dim3(grid_for(nblk * R.K * 2, kBlock)), dim3(kBlock), 0, st,
                     R, lohi);
  {  // the schedule shares the area the tile-offset scan uses afterwards
    char* q = p;
    int32_t* est = (int32_t*)q;
    q += align256((size_t)nblk * 4);
    int32_t* sched = (int32_t*)q;
    q += align256((size_t)nblk * 4);
    const int64_t nbk = mu_sched_buckets(C);
    int32_t* cnt = (int32_t*)q;
    q += align256((size_t)nbk * 4);
    int32_t* start = (int32_t*)q;
    q += align256((size_t)(nbk + 1) * 4);
    hipLaunchKernelGGL(k_zero_i32r, dim3(grid_for(nbk, kBlock)), dim3(kBlock), 0, st, cnt, nbk);
    hipLaunchKernelGGL(k_mu_sched_est, dim3(grid_for(nblk, kBlock)), dim3(kBlock), 0, st, R,
                       (const int64_t*)lohi, est, cnt);
    int rc = launch_status();
    if (rc) return rc;
    rc = scan_i32(cnt, nbk, start, q, scan_ws_bytes(nbk), st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mu_sched_place, dim3(grid_for(nblk, kBlock)), dim3(kBlock), 0, st, nblk,
                       (const int32_t*)est, (const int32_t*)start, cnt, sched);
    const int per = (int)((nblk + kXcds - 1) / kXcds);
    hipLaunchKernelGGL(k_mu_rank<W>, dim3((unsigned)(per * kXcds)), dim3(kBlock), 0, st, R,
                       (const int64_t*)lohi, (const int32_t*)sched, per, code);
  }
  int rc = launch_status();
  if (rc) return rc;
  rc = scan_tile_offsets_bit0(code, C, offs, p, scan_tile_offsets_ws_bytes(C), st);
  if (rc) return rc;
  if (D.d64 || D.d32)
    hipLaunchKernelGGL((k_mu_emit<W, true>), dim3((unsigned)nst), dim3(kBlock), 0, st, R,
                       (const uint32_t*)code, (const int32_t*)offs, (u64*)out, from0, cand_pos,
                       out_code, out_count, D);
  else
    hipLaunchKernelGGL((k_mu_emit<W, false>), dim3((unsigned)nst), dim3(kBlock), 0, st, R,
                       (const uint32_t*)code, (const int32_t*)offs, (u64*)out, from0, cand_pos,
                       out_code, out_count, D);
  return launch_status();
}

// ------------------------------------------------------------------------------------------------
// G1 reassignment (paces/core/time_evolution.py:781-797)
// ------------------------------------------------------------------------------------------------

template <int W>
__global__ void __launch_bounds__(kBlock) k_reassign(const u64* __restrict__ old_rows, int64_t M,
                                                     const double2* __restrict__ old_psi,
                                                     const u64* __restrict__ new_rows, int64_t R,
                                                     const int64_t* __restrict__ lohi,
                                                     double2* __restrict__ new_psi) {
  __shared__ LookupSmem<W> sm;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool v = i < R;
  Row<W> x;
#pragma unroll
  for (int k = 0; k < W; ++k) x.w[k] = 0;
  if (v) x = load_row<W>(new_rows, i);
  int64_t lb = block_lower_bound_range<W>(x, v, old_rows, lohi[2 * blockIdx.x],
                                          lohi[2 * blockIdx.x + 1], sm);
  if (!v) return;
  double2 val = make_double2(0.0, 0.0);
  if (lb < M && eq_row<W>(x, old_rows + lb * W)) val = old_psi[lb];
  new_psi[i] = val;
}

template <int W>
static int run_reassign(const uint64_t* old_rows, int64_t M, const double* old_psi,
                        const uint64_t* new_rows, int64_t R, double* new_psi, void* ws,
                        hipStream_t st) {
  if (R == 0) return PACES_OK;
  int64_t nt = (R + kBlock - 1) / kBlock;
  int64_t* lohi = (int64_t*)ws;
  hipLaunchKernelGGL(k_tile_bounds<W>, dim3(grid_for(2 * nt, kBlock)), dim3(kBlock), 0, st,
                     (const u64*)new_rows, R, (const u64*)old_rows, M, nt, lohi);
  hipLaunchKernelGGL(k_reassign<W>, dim3((unsigned)nt), dim3(kBlock), 0, st,
                     (const u64*)old_rows, M, (const double2*)old_psi, (const u64*)new_rows, R,
                     (const int64_t*)lohi, (double2*)new_psi);
  return launch_status();
}

// ------------------------------------------------------------------------------------------------
// N3 pruning (no reference counterpart): which union rows to keep after scoring, and the
// compaction of the kept rows with the candidate positions remapped into the kept numbering.
// ------------------------------------------------------------------------------------------------

// finaliser of MurmurHash3 (Appleby), 64-bit variant
__device__ __forceinline__ uint64_t fmix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

__global__ void __launch_bounds__(kBlock) k_prune_plan(const float* __restrict__ prob, int64_t n,
                                                       float thr, int sample, uint64_t smask,
                                                       uint64_t seed, uint8_t* __restrict__ keep) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t k = prob[i] >= thr ? 1 : 0;
  if (!k && sample && (fmix64((uint64_t)i ^ seed) & smask) == 0) k = 2;
  keep[i] = k;
}

// the added rows' decisions scattered to their union positions (the rest keep in_basis's 1)
__global__ void __launch_bounds__(kBlock) k_prune_rows(const float* __restrict__ prob,
                                                       const int32_t* __restrict__ row_index,
                                                       int64_t m, float thr, int sample,
                                                       uint64_t smask, uint64_t seed,
                                                       uint8_t* __restrict__ keep) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int64_t i = row_index[j];
  uint8_t k = prob[j] >= thr ? 1 : 0;
  if (!k && sample && (fmix64((uint64_t)i ^ seed) & smask) == 0) k = 2;
  keep[i] = k;
}

__global__ void __launch_bounds__(kBlock) k_force_keep(const int32_t* __restrict__ idx, int64_t m,
                                                       uint8_t* __restrict__ keep) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) keep[idx[j]] = 1;
}

__global__ void __launch_bounds__(kBlock) k_keep_flags(const uint8_t* __restrict__ keep, int64_t n,
                                                       uint8_t* __restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = keep[i] != 0 ? 1 : 0;
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_compact(const uint8_t* __restrict__ keep, int64_t n,
                                                    const int32_t* __restrict__ G,
                                                    const u64* __restrict__ rows,
                                                    const double2* __restrict__ psi,
                                                    const uint8_t* __restrict__ inb,
                                                    u64* __restrict__ orows,
                                                    double2* __restrict__ opsi,
                                                    uint8_t* __restrict__ oinb,
                                                    uint8_t* __restrict__ osample,
                                                    int32_t* __restrict__ ocount) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *ocount = G[n];
  if (i >= n || !keep[i]) return;
  const int32_t j = G[i];
  store_row<W>(orows, j, load_row<W>(rows, i));
  if (opsi) opsi[j] = psi[i];
  if (oinb) oinb[j] = inb[i];
  if (osample) osample[j] = keep[i] == 2 ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_remap_pos(const int32_t* __restrict__ cand_pos,
                                                      int64_t C, const uint8_t* __restrict__ keep,
                                                      const int32_t* __restrict__ G,
                                                      int32_t* __restrict__ out) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= C) return;
  const int32_t p = cand_pos[g];
  out[g] = keep[p] ? G[p] : -1;
}

__global__ void __launch_bounds__(kBlock) k_mark_indices(const int32_t* __restrict__ idx,
                                                         int64_t m, const uint8_t* __restrict__ src,
                                                         uint8_t* __restrict__ out) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int32_t i = idx[j];
  out[i] = src ? src[i] : 1;
}

__global__ void __launch_bounds__(kBlock) k_scatter_u8(const int32_t* __restrict__ idx, int64_t m,
                                                       const uint8_t* __restrict__ src,
                                                       uint8_t* __restrict__ out) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) out[idx[j]] = src[j];
}

size_t compact_ws_bytes(int64_t n) {
  return align256((size_t)(n > 0 ? n : 1)) + align256((size_t)(n + 1) * 4) +
         align256(scan_ws_bytes(n));
}

template <int W>
static int run_compact(const uint8_t* keep, int64_t n, const uint64_t* rows, const double* psi,
                       const uint8_t* inb, const int32_t* cand_pos, int64_t C, uint64_t* orows,
                       double* opsi, uint8_t* oinb, uint8_t* osample, int32_t* ocand,
                       int32_t* ocount, void* ws, hipStream_t st) {
  char* p = (char*)ws;
  uint8_t* flag = (uint8_t*)p;
  p += align256((size_t)(n > 0 ? n : 1));
  int32_t* G = (int32_t*)p;
  p += align256((size_t)(n + 1) * 4);
  hipLaunchKernelGGL(k_keep_flags, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, keep, n, flag);
  int rc = launch_status();
  if (rc) return rc;
  rc = scan_u8(flag, n, G, p, scan_ws_bytes(n), st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_compact<W>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, keep, n,
                     (const int32_t*)G, (const u64*)rows, (const double2*)psi, inb, (u64*)orows,
                     (double2*)opsi, oinb, osample, ocount);
  if (C > 0)
    hipLaunchKernelGGL(k_remap_pos, dim3(grid_for(C, kBlock)), dim3(kBlock), 0, st, cand_pos, C,
                       keep, (const int32_t*)G, ocand);
  return launch_status();
}

// ---- paces_reassign without allocation: the tile bounds in a static device array. Host code may
// not use the array's name as an address (it names a host-side stand-in): each device's address
// comes from hipGetSymbolAddress, fetched once per device with that device current.
static __device__ __align__(256) unsigned char scratch[PACES_REASSIGN_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static void* scratch_ptr[kMaxDevices];
static std::mutex scratch_mu;

static int scratch_for_current_device(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return PACES_E_HIP;
  if (dev < 0 || dev >= kMaxDevices) return PACES_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PACES_E_HIP;
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return PACES_OK;
}

}  // namespace paces

using namespace paces;

extern "C" {

int paces_pack_rows(const int32_t* raw, int64_t n, const paces_layout_t* lay, uint64_t* rows,
                    void* stream) {
  if (!layout_ok(lay) || n < 0 || (n > 0 && (!raw || !rows))) return PACES_E_INVALID;
  if (n == 0) return PACES_OK;
  PACES_DISPATCH_W(lay->W, run_pack, raw, n, lay, rows, (hipStream_t)stream);
}

int paces_get_field(const uint64_t* rows, int64_t n, const paces_layout_t* lay, int32_t field,
                    int32_t* out, void* stream) {
  if (!layout_ok(lay) || n < 0 || field < -1 || field >= lay->nfields) return PACES_E_INVALID;
  if (n == 0) return PACES_OK;
  if (!rows || !out) return PACES_E_INVALID;
  PACES_DISPATCH_W(lay->W, run_get, rows, n, lay, field, out, (hipStream_t)stream);
}

int paces_add_field(uint64_t* rows, int64_t n, const paces_layout_t* lay, int32_t field,
                    int32_t delta, void* stream) {
  if (!layout_ok(lay) || n < 0 || field < -1 || field >= lay->nfields) return PACES_E_INVALID;
  if (n == 0 || delta == 0) return PACES_OK;
  if (!rows) return PACES_E_INVALID;
  PACES_DISPATCH_W(lay->W, run_add, rows, n, lay, field, delta, (hipStream_t)stream);
}

size_t paces_lower_bound_workspace_bytes(int64_t Q) { return lb_ws_bytes(Q); }

int paces_lower_bound(const uint64_t* book, int64_t M, const uint64_t* q, int64_t Q, int32_t W,
                      int32_t q_sorted, int32_t* out_idx, uint8_t* out_hit, int32_t* miss_count,
                      void* ws, size_t ws_bytes, void* stream) {
  if (M < 0 || Q < 0 || M >= 2147483647LL || W < 1) return PACES_E_INVALID;
  if (Q > 0 && (!q || (M > 0 && !book))) return PACES_E_INVALID;
  PACES_DISPATCH_W(W, run_lb, book, M, q, Q, q_sorted, out_idx, out_hit, miss_count, ws, ws_bytes,
                   (hipStream_t)stream);
}

size_t paces_merge_unique_workspace_bytes(int64_t total, int32_t K, int32_t W) {
  if (total < 0 || K < 1 || W < 1) return 0;
  return mu_ws_bytes(total, K);
}

static int merge_unique_impl(const paces_runs_t* runs, int32_t W, uint64_t* out_rows,
                             int32_t* out_count, uint8_t* out_from_run0, int32_t* cand_pos,
                             const uint8_t* run_codes, uint8_t* out_code, const RowDiag& D,
                             void* ws, size_t ws_bytes, void* stream) {
  if (!runs || runs->K < 1 || runs->K > PACES_MAX_RUNS || !out_count) return PACES_E_INVALID;
  if (out_code && !run_codes) return PACES_E_INVALID;
  RunsArg R;
  R.K = runs->K;
  for (int s = 0; s < PACES_MAX_RUNS; ++s) R.code[s] = (run_codes && s < R.K) ? run_codes[s] : 0;
  R.off[0] = 0;
  for (int s = 0; s < R.K; ++s) {
    if (runs->len[s] < 0 || (runs->len[s] > 0 && !runs->ptr[s])) return PACES_E_INVALID;
    R.ptr[s] = (const u64*)runs->ptr[s];
    R.off[s + 1] = R.off[s] + runs->len[s];
  }
  for (int s = R.K; s < PACES_MAX_RUNS; ++s) {
    R.ptr[s] = nullptr;
    R.off[s + 1] = R.off[R.K];
  }
  R.boff[0] = 0;
  for (int s = 0; s < PACES_MAX_RUNS; ++s)
    R.boff[s + 1] = R.boff[s] + (R.off[s + 1] - R.off[s] + kBlock - 1) / kBlock;
  if (R.off[R.K] >= 2147483647LL) return PACES_E_INVALID;
  if (R.off[R.K] > 0 && !out_rows) return PACES_E_INVALID;
  if (W < 1 || W > PACES_MAX_ROW_WORDS) return PACES_E_UNSUPPORTED;
  PACES_DISPATCH_W(W, run_mu, R, out_rows, out_count, out_from_run0, cand_pos, out_code, D, ws,
                   ws_bytes, (hipStream_t)stream);
}

int paces_merge_unique(const paces_runs_t* runs, int32_t W, uint64_t* out_rows,
                       int32_t* out_count, uint8_t* out_from_run0, int32_t* cand_pos, void* ws,
                       size_t ws_bytes, void* stream) {
  RowDiag D = {};
  return merge_unique_impl(runs, W, out_rows, out_count, out_from_run0, cand_pos, nullptr, nullptr,
                           D, ws, ws_bytes, stream);
}

int paces_merge_unique_coded(const paces_runs_t* runs, int32_t W, uint64_t* out_rows,
                             int32_t* out_count, uint8_t* out_from_run0, int32_t* cand_pos,
                             const uint8_t* run_codes, uint8_t* out_code, void* ws,
                             size_t ws_bytes, void* stream) {
  if (!out_code || !run_codes) return PACES_E_INVALID;
  RowDiag D = {};
  return merge_unique_impl(runs, W, out_rows, out_count, out_from_run0, cand_pos, run_codes,
                           out_code, D, ws, ws_bytes, stream);
}

int paces_merge_unique_diag(const paces_runs_t* runs, int32_t W, uint64_t* out_rows,
                            int32_t* out_count, uint8_t* out_from_run0, int32_t* cand_pos,
                            const paces_layout_t* lay, const paces_holstein_t* h, double* diag,
                            float* diag32, void* ws, size_t ws_bytes, void* stream) {
  if (!lay || !h || lay->W != W || (!diag) == (!diag32)) return PACES_E_INVALID;
  if (!lay->d_sum_mask || lay->sum_bits < 1 || lay->sum_bits > 32) return PACES_E_INVALID;
  if (h->delta_eps != 0.0 && (!lay->d_off || !lay->d_width || lay->nfields < 2))
    return PACES_E_INVALID;
  RowDiag D;
  D.sum_mask = (const u64*)lay->d_sum_mask;
  D.off = lay->d_off;
  D.width = lay->d_width;
  D.sum_bits = lay->sum_bits;
  D.pf = h->batched ? 1 : 0;
  D.eps = h->eps_sys;
  D.hw = h->hbar_omega;
  D.de = h->delta_eps;
  D.d64 = diag;
  D.d32 = diag32;
  return merge_unique_impl(runs, W, out_rows, out_count, out_from_run0, cand_pos, nullptr, nullptr,
                           D, ws, ws_bytes, stream);
}

size_t paces_reassign_workspace_bytes(int64_t R) { return lb_ws_bytes(R); }

int paces_reassign(const uint64_t* old_rows, int64_t M, const double* old_psi,
                   const uint64_t* new_rows, int64_t R, int32_t W, double* new_psi, void* ws,
                   size_t ws_bytes, void* stream) {
  if (M < 0 || R < 0 || W < 1) return PACES_E_INVALID;
  if (R > 0 && (!new_rows || !new_psi || !ws || (M > 0 && (!old_rows || !old_psi))))
    return PACES_E_INVALID;
  if (ws_bytes < lb_ws_bytes(R)) return PACES_E_WORKSPACE;
  PACES_DISPATCH_W(W, run_reassign, old_rows, M, old_psi, new_rows, R, new_psi, ws,
                   (hipStream_t)stream);
}

int paces_reassign_noalloc(const uint64_t* old_rows, int64_t M, const double* old_psi,
                           const uint64_t* new_rows, int64_t R, int32_t W, double* new_psi,
                           void* stream) {
  if (M < 0 || R < 0 || W < 1) return PACES_E_INVALID;
  if (R > PACES_REASSIGN_NOALLOC_MAX_ROWS) return PACES_E_SCRATCH;
  if (R > 0 && (!new_rows || !new_psi || (M > 0 && (!old_rows || !old_psi))))
    return PACES_E_INVALID;
  if (R == 0) return PACES_OK;
  if (lb_ws_bytes(R) > PACES_REASSIGN_NOALLOC_SCRATCH_BYTES) return PACES_E_SCRATCH;
  void* ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc != PACES_OK) return rc;
  // the tile bounds are written in full before they are read: nothing to clear
  PACES_DISPATCH_W(W, run_reassign, old_rows, M, old_psi, new_rows, R, new_psi, ws,
                   (hipStream_t)stream);
}

int paces_prune_plan(const float* prob, int64_t n, float threshold, const int32_t* forced_idx,
                     int64_t nforced, int32_t sample_bits, uint64_t seed, uint8_t* keep,
                     void* stream) {
  if (n < 0 || nforced < 0 || sample_bits > 62) return PACES_E_INVALID;
  if (n > 0 && (!prob || !keep)) return PACES_E_INVALID;
  if (nforced > 0 && !forced_idx) return PACES_E_INVALID;
  if (n == 0) return PACES_OK;
  hipStream_t st = (hipStream_t)stream;
  const uint64_t smask = sample_bits > 0 ? ((1ULL << sample_bits) - 1ULL) : 0ULL;
  hipLaunchKernelGGL(k_prune_plan, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, prob, n,
                     threshold, sample_bits >= 0 ? 1 : 0, smask, seed, keep);
  if (nforced > 0)
    hipLaunchKernelGGL(k_force_keep, dim3(grid_for(nforced, kBlock)), dim3(kBlock), 0, st,
                       forced_idx, nforced, keep);
  return launch_status();
}

int paces_prune_plan_rows(const float* prob, const int32_t* row_index, int64_t m,
                          const uint8_t* in_basis, int64_t n, float threshold,
                          int32_t sample_bits, uint64_t seed, uint8_t* keep, void* stream) {
  if (n < 0 || m < 0 || m > n || sample_bits > 62) return PACES_E_INVALID;
  if (n > 0 && (!in_basis || !keep)) return PACES_E_INVALID;
  if (m > 0 && (!prob || !row_index)) return PACES_E_INVALID;
  if (n == 0) return PACES_OK;
  hipStream_t st = (hipStream_t)stream;
  if (hipMemcpyAsync(keep, in_basis, (size_t)n, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return PACES_E_HIP;
  if (m == 0) return PACES_OK;
  const uint64_t smask = sample_bits > 0 ? ((1ULL << sample_bits) - 1ULL) : 0ULL;
  hipLaunchKernelGGL(k_prune_rows, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, st, prob,
                     row_index, m, threshold, sample_bits >= 0 ? 1 : 0, smask, seed, keep);
  return launch_status();
}

int paces_scatter_u8(const int32_t* idx, int64_t m, const uint8_t* src, uint8_t* out,
                     void* stream) {
  if (m < 0 || (m > 0 && (!idx || !src || !out))) return PACES_E_INVALID;
  if (m == 0) return PACES_OK;
  hipLaunchKernelGGL(k_scatter_u8, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, (hipStream_t)stream,
                     idx, m, src, out);
  return launch_status();
}

int paces_mark_indices(const int32_t* idx, int64_t m, const uint8_t* src, uint8_t* out,
                       void* stream) {
  if (m < 0 || (m > 0 && (!idx || !out))) return PACES_E_INVALID;
  if (m == 0) return PACES_OK;
  hipLaunchKernelGGL(k_mark_indices, dim3(grid_for(m, kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, idx, m, src, out);
  return launch_status();
}

size_t paces_compact_workspace_bytes(int64_t n) { return compact_ws_bytes(n); }

int paces_compact_kept(const uint8_t* keep, int64_t n, const uint64_t* rows, int32_t W,
                       const double* psi, const uint8_t* in_basis, const int32_t* cand_pos,
                       int64_t C, uint64_t* out_rows, double* out_psi, uint8_t* out_in_basis,
                       uint8_t* out_sample, int32_t* out_cand_pos, int32_t* out_count, void* ws,
                       size_t ws_bytes, void* stream) {
  if (n < 0 || C < 0 || !out_count || !ws) return PACES_E_INVALID;
  if (n > 0 && (!keep || !rows || !out_rows)) return PACES_E_INVALID;
  if ((psi && !out_psi) || (in_basis && !out_in_basis)) return PACES_E_INVALID;
  if (C > 0 && (!cand_pos || !out_cand_pos)) return PACES_E_INVALID;
  if (n >= 2147483647LL) return PACES_E_INVALID;
  if (ws_bytes < compact_ws_bytes(n)) return PACES_E_WORKSPACE;
  PACES_DISPATCH_W(W, run_compact, keep, n, rows, psi, in_basis, cand_pos, C, out_rows, out_psi,
                   out_in_basis, out_sample, out_cand_pos, out_count, ws, (hipStream_t)stream);
}

}  // extern "C"

