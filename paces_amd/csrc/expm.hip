// Pair-vector kernels of the block 1-norm estimator behind the scaling-and-squaring propagator
// (paces/aux/_scipy/_onenormest.py:11-103, 202-345: Higham and Tisseur 2000, Algorithm 2.4,
// with t = 2 columns). The Hamiltonian is real, so the estimator's two real columns travel as
// one complex128 vector (column 0 in the real parts, column 1 in the imaginary parts) and one
// complex H.x applies H to both. Reductions use a fixed number of blocks and a fixed order, so
// every result depends only on the inputs.
#include "common.hpp"

namespace paces {

constexpr int kPairBlock = 256;
constexpr int kPairBlocks = 1024;  // partial-sum blocks of a reduction (fixed)

static int pair_blocks(int64_t n) {
  const int64_t b = (n + kPairBlock - 1) / kPairBlock;
  return (int)(b < kPairBlocks ? (b > 0 ? b : 1) : kPairBlocks);
}

// per block: sums over rows i = b*256 + t, stepping by the grid, of f(a_i, b_i) into nq lanes
template <int NQ, typename F>
__device__ __forceinline__ void pair_partial(int64_t n, double* __restrict__ part, F f) {
  __shared__ double sm[NQ][kPairBlock / 64];
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kPairBlock;
  for (int64_t i = (int64_t)blockIdx.x * kPairBlock + threadIdx.x; i < n; i += stride) {
    double v[NQ];
    f(i, v);
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = acc[q] + v[q];
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const double w = wave_sum(acc[q]);
    if ((threadIdx.x & 63) == 0) sm[q][threadIdx.x >> 6] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kPairBlock / 64; ++w) s = s + sm[q][w];
      part[(int64_t)q * kPairBlocks + blockIdx.x] = s;
    }
  }
}

// out[q] = sum over the nb partials of lane q, in block order
__global__ void __launch_bounds__(64) k_pair_final(const double* __restrict__ part, int nb, int nq,
                                                   double* __restrict__ out) {
  const int q = blockIdx.x;
  if (q >= nq || threadIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s = s + part[(int64_t)q * kPairBlocks + b];
  out[q] = s;
}

__global__ void __launch_bounds__(kPairBlock) k_pair_abs_sums(const double2* __restrict__ y,
                                                              int64_t n, double* __restrict__ part) {
  pair_partial<2>(n, part, [&](int64_t i, double* v) {
    const double2 a = y[i];
    v[0] = fabs(a.x);
    v[1] = fabs(a.y);
  });
}

__global__ void __launch_bounds__(kPairBlock) k_pair_dots(const double2* __restrict__ a,
                                                          const double2* __restrict__ b, int64_t n,
                                                          double* __restrict__ part) {
  pair_partial<4>(n, part, [&](int64_t i, double* v) {
    const double2 u = a[i], w = b[i];
    v[0] = u.x * w.x;
    v[1] = u.x * w.y;
    v[2] = u.y * w.x;
    v[3] = u.y * w.y;
  });
}

// sign with sign(0) = 1, per column
__global__ void __launch_bounds__(kPairBlock) k_pair_signs(const double2* __restrict__ y, int64_t n,
                                                           double2* __restrict__ s) {
  const int64_t i = (int64_t)blockIdx.x * kPairBlock + threadIdx.x;
  if (i >= n) return;
  const double2 a = y[i];
  s[i] = make_double2(a.x < 0.0 ? -1.0 : 1.0, a.y < 0.0 ? -1.0 : 1.0);
}

__global__ void __launch_bounds__(kPairBlock) k_pair_rowmax(const double2* __restrict__ z, int64_t n,
                                                            double* __restrict__ h) {
  const int64_t i = (int64_t)blockIdx.x * kPairBlock + threadIdx.x;
  if (i >= n) return;
  const double2 a = z[i];
  h[i] = fmax(fabs(a.x), fabs(a.y));
}

// (value, index) ordered by value, then index: a before b when a is larger
__device__ __forceinline__ bool vi_before(double av, int64_t ai, double bv, int64_t bi) {
  return av > bv || (av == bv && ai > bi);
}

// the first (value desc, index desc) entry of h strictly after (bv, bi) in that order (all
// entries when bi < 0): per block, then k_argmax_final
__global__ void __launch_bounds__(kPairBlock) k_argmax_after(const double* __restrict__ h, int64_t n,
                                                             double bv, int64_t bi,
                                                             double* __restrict__ pv,
                                                             int64_t* __restrict__ pi) {
  __shared__ double sv[kPairBlock];
  __shared__ int64_t si[kPairBlock];
  double mv = 0.0;
  int64_t mi = -1;
  const int64_t stride = (int64_t)gridDim.x * kPairBlock;
  for (int64_t i = (int64_t)blockIdx.x * kPairBlock + threadIdx.x; i < n; i += stride) {
    const double v = h[i];
    const bool after = bi < 0 || vi_before(bv, bi, v, i);
    if (after && (mi < 0 || vi_before(v, i, mv, mi))) {
      mv = v;
      mi = i;
    }
  }
  sv[threadIdx.x] = mv;
  si[threadIdx.x] = mi;
  __syncthreads();
  for (int w = kPairBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double ov = sv[threadIdx.x + w];
      const int64_t oi = si[threadIdx.x + w];
      if (oi >= 0 && (si[threadIdx.x] < 0 || vi_before(ov, oi, sv[threadIdx.x], si[threadIdx.x]))) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

__global__ void __launch_bounds__(64) k_argmax_final(const double* __restrict__ pv,
                                                     const int64_t* __restrict__ pi, int nb,
                                                     int64_t* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double mv = 0.0;
  int64_t mi = -1;
  for (int b = 0; b < nb; ++b) {
    if (pi[b] >= 0 && (mi < 0 || vi_before(pv[b], pi[b], mv, mi))) {
      mv = pv[b];
      mi = pi[b];
    }
  }
  *out = mi;
}

// column col (0: real parts, 1: imaginary parts) of x: +-scale from bits of a hash of
// (seed, i) (splitmix64 finaliser)
__global__ void __launch_bounds__(kPairBlock) k_pair_random_column(double2* __restrict__ x,
                                                                   int64_t n, int col, u64 seed,
                                                                   double scale) {
  const int64_t i = (int64_t)blockIdx.x * kPairBlock + threadIdx.x;
  if (i >= n) return;
  u64 z = seed + 0x9E3779B97F4A7C15ull * (u64)(i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  const double v = (z >> 63) ? -scale : scale;
  double2 a = x[i];
  if (col == 0)
    a.x = v;
  else
    a.y = v;
  x[i] = a;
}

// x = (re, im) everywhere, then entry i0 of column 0 and entry i1 of column 1 set to one when
// those indices are >= 0 (the estimator's unit vectors)
__global__ void __launch_bounds__(kPairBlock) k_pair_fill(double2* __restrict__ x, int64_t n,
                                                          double re, double im, int64_t i0,
                                                          int64_t i1) {
  const int64_t i = (int64_t)blockIdx.x * kPairBlock + threadIdx.x;
  if (i >= n) return;
  x[i] = make_double2(i == i0 ? 1.0 : re, i == i1 ? 1.0 : im);
}

// sum of a float64 vector (partials, then k_pair_final)
__global__ void __launch_bounds__(kPairBlock) k_sum_f64(const double* __restrict__ x, int64_t n,
                                                        double* __restrict__ part) {
  pair_partial<1>(n, part, [&](int64_t i, double* v) { v[0] = x[i]; });
}

// out_i = x_i - m
__global__ void __launch_bounds__(kPairBlock) k_shift_f64(const double* __restrict__ x, int64_t n,
                                                          double m, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kPairBlock + threadIdx.x;
  if (i < n) out[i] = x[i] - m;
}

// per block: the largest row sum |d_i - m| + sum over terms and entries |a|
__global__ void __launch_bounds__(kPairBlock) k_shifted_row_abs_max(const double* __restrict__ d,
                                                                    double m, paces_csr_set_t H,
                                                                    int64_t n,
                                                                    double* __restrict__ part) {
  __shared__ double sm[kPairBlock / 64];
  double best = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kPairBlock;
  for (int64_t i = (int64_t)blockIdx.x * kPairBlock + threadIdx.x; i < n; i += stride) {
    double r = fabs(d[i] - m);
    for (int t = 0; t < H.nterms; ++t)
      for (int32_t k = H.indptr[t][i]; k < H.indptr[t][i + 1]; ++k) r = r + fabs(H.data[t][k]);
    best = fmax(best, r);
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) best = fmax(best, __shfl_xor(best, s, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
#pragma unroll
    for (int w = 0; w < kPairBlock / 64; ++w) b = fmax(b, sm[w]);
    part[blockIdx.x] = b;
  }
}

__global__ void __launch_bounds__(64) k_max_final(const double* __restrict__ part, int nb,
                                                  double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double b = 0.0;
  for (int k = 0; k < nb; ++k) b = fmax(b, part[k]);
  *out = b;
}

}  // namespace paces

using namespace paces;

extern "C" {

size_t paces_pair_workspace_bytes(void) {
  return (size_t)kPairBlocks * (4 * sizeof(double)) + (size_t)kPairBlocks * sizeof(int64_t);
}

int paces_pair_abs_sums(const double* y, int64_t n, double* out, void* ws, size_t ws_bytes,
                        void* stream) {
  if (n < 1 || !y || !out || !ws) return PACES_E_INVALID;
  if (ws_bytes < paces_pair_workspace_bytes()) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int nb = pair_blocks(n);
  hipLaunchKernelGGL(k_pair_abs_sums, dim3(nb), dim3(kPairBlock), 0, st, (const double2*)y, n,
                     (double*)ws);
  hipLaunchKernelGGL(k_pair_final, dim3(2), dim3(64), 0, st, (const double*)ws, nb, 2, out);
  return launch_status();
}

int paces_pair_dots(const double* a, const double* b, int64_t n, double* out, void* ws,
                    size_t ws_bytes, void* stream) {
  if (n < 1 || !a || !b || !out || !ws) return PACES_E_INVALID;
  if (ws_bytes < paces_pair_workspace_bytes()) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int nb = pair_blocks(n);
  hipLaunchKernelGGL(k_pair_dots, dim3(nb), dim3(kPairBlock), 0, st, (const double2*)a,
                     (const double2*)b, n, (double*)ws);
  hipLaunchKernelGGL(k_pair_final, dim3(4), dim3(64), 0, st, (const double*)ws, nb, 4, out);
  return launch_status();
}

int paces_pair_signs(const double* y, int64_t n, double* s, void* stream) {
  if (n < 1 || !y || !s) return PACES_E_INVALID;
  hipLaunchKernelGGL(k_pair_signs, dim3(grid_for(n, kPairBlock)), dim3(kPairBlock), 0,
                     (hipStream_t)stream, (const double2*)y, n, (double2*)s);
  return launch_status();
}

int paces_pair_rowmax(const double* z, int64_t n, double* h, void* stream) {
  if (n < 1 || !z || !h) return PACES_E_INVALID;
  hipLaunchKernelGGL(k_pair_rowmax, dim3(grid_for(n, kPairBlock)), dim3(kPairBlock), 0,
                     (hipStream_t)stream, (const double2*)z, n, h);
  return launch_status();
}

int paces_argmax_after(const double* h, int64_t n, double after_value, int64_t after_index,
                       int64_t* out, void* ws, size_t ws_bytes, void* stream) {
  if (n < 1 || !h || !out || !ws || after_index >= n) return PACES_E_INVALID;
  if (ws_bytes < paces_pair_workspace_bytes()) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int nb = pair_blocks(n);
  double* pv = (double*)ws;
  int64_t* pi = (int64_t*)((char*)ws + (size_t)kPairBlocks * 4 * sizeof(double));
  hipLaunchKernelGGL(k_argmax_after, dim3(nb), dim3(kPairBlock), 0, st, h, n, after_value,
                     after_index, pv, pi);
  hipLaunchKernelGGL(k_argmax_final, dim3(1), dim3(64), 0, st, (const double*)pv,
                     (const int64_t*)pi, nb, out);
  return launch_status();
}

int paces_pair_random_column(double* x, int64_t n, int32_t col, uint64_t seed, double scale,
                             void* stream) {
  if (n < 1 || !x || (col != 0 && col != 1)) return PACES_E_INVALID;
  hipLaunchKernelGGL(k_pair_random_column, dim3(grid_for(n, kPairBlock)), dim3(kPairBlock), 0,
                     (hipStream_t)stream, (double2*)x, n, (int)col, (u64)seed, scale);
  return launch_status();
}

int paces_sum_f64(const double* x, int64_t n, double* out, void* ws, size_t ws_bytes,
                  void* stream) {
  if (n < 1 || !x || !out || !ws) return PACES_E_INVALID;
  if (ws_bytes < paces_pair_workspace_bytes()) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int nb = pair_blocks(n);
  hipLaunchKernelGGL(k_sum_f64, dim3(nb), dim3(kPairBlock), 0, st, x, n, (double*)ws);
  hipLaunchKernelGGL(k_pair_final, dim3(1), dim3(64), 0, st, (const double*)ws, nb, 1, out);
  return launch_status();
}

int paces_shift_f64(const double* x, int64_t n, double m, double* out, void* stream) {
  if (n < 1 || !x || !out) return PACES_E_INVALID;
  hipLaunchKernelGGL(k_shift_f64, dim3(grid_for(n, kPairBlock)), dim3(kPairBlock), 0,
                     (hipStream_t)stream, x, n, m, out);
  return launch_status();
}

int paces_csr_shifted_onenorm(int64_t n, const double* diag, double m, const paces_csr_set_t* h,
                              double* out, void* ws, size_t ws_bytes, void* stream) {
  if (n < 1 || !diag || !h || !out || !ws || h->nterms < 0 || h->nterms > PACES_MAX_TERMS)
    return PACES_E_INVALID;
  for (int t = 0; t < h->nterms; ++t)
    if (!h->indptr[t] || !h->indices[t] || !h->data[t]) return PACES_E_INVALID;
  if (ws_bytes < paces_pair_workspace_bytes()) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int nb = pair_blocks(n);
  hipLaunchKernelGGL(k_shifted_row_abs_max, dim3(nb), dim3(kPairBlock), 0, st, diag, m, *h, n,
                     (double*)ws);
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(64), 0, st, (const double*)ws, nb, out);
  return launch_status();
}

int paces_pair_fill(double* x, int64_t n, double re, double im, int64_t i0, int64_t i1,
                    void* stream) {
  if (n < 1 || !x || i0 >= n || i1 >= n) return PACES_E_INVALID;
  hipLaunchKernelGGL(k_pair_fill, dim3(grid_for(n, kPairBlock)), dim3(kPairBlock), 0,
                     (hipStream_t)stream, (double2*)x, n, re, im, i0, i1);
  return launch_status();
}

}  // extern "C"
