// E1/E2/E3: H.x, the Taylor series for exp(-i dt H) v, and the deterministic reductions behind
// norms, vdot and the Lanczos recurrence (paces/core/time_evolution.py:382-400, 482-549).
//
// FP64 parity order (compiled with -ffp-contract=off): per row
//   y = d*x;  for each term: t = 0; t = t + a_k * x_{j_k} (ascending columns); y = y + t
// which is the reference's res = diag*x; res += M_term.dot(x) in use_terms order, with each
// row of a term summed left to right.
#include "common.hpp"

namespace paces {

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

struct CsrSet {
  int nterms;
  const int32_t* indptr[PACES_MAX_TERMS];
  const int32_t* indices[PACES_MAX_TERMS];
  const double* data[PACES_MAX_TERMS];
};

__device__ __forceinline__ double2 hx_row(int64_t i, const double* __restrict__ diag,
                                          const CsrSet& H, const double2* __restrict__ x) {
  const double2 xi = x[i];
  const double d = diag[i];
  double yr = d * xi.x, yi = d * xi.y;
  for (int t = 0; t < H.nterms; ++t) {
    const int32_t b = H.indptr[t][i], e = H.indptr[t][i + 1];
    double tr = 0.0, ti = 0.0;
    for (int32_t k = b; k < e; ++k) {
      const double a = H.data[t][k];
      const double2 xj = x[H.indices[t][k]];
      tr = tr + a * xj.x;
      ti = ti + a * xj.y;
    }
    yr = yr + tr;
    yi = yi + ti;
  }
  return make_double2(yr, yi);
}

__global__ void __launch_bounds__(kBlock) k_spmv(int64_t R, const double* __restrict__ diag,
                                                 CsrSet H, const double2* __restrict__ x,
                                                 double2* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (H.nterms == 1 || H.nterms == 2) {  // one row per thread, its loads in phases
    const int64_t ii[1] = {i};
    const bool vv[1] = {i < R};
    double2 h[1], xi[1];
    const int64_t spare = (int64_t)blockIdx.x * blockDim.x;  // the block's first row (< R)
    if (H.nterms == 2)
      hx_rows_phased<double, double2, 1, 2>(ii, vv, spare, diag, H, x, h, xi);
    else
      hx_rows_phased<double, double2, 1, 1>(ii, vv, spare, diag, H, x, h, xi);
    if (vv[0]) y[i] = h[0];
    return;
  }
  if (i >= R) return;
  y[i] = hx_row(i, diag, H, x);
}

// ---- Taylor ---------------------------------------------------------------------------------------

__global__ void k_tay_zero(paces_taylor_state_t* s) {
  s->max_term2 = 0;
  s->max_res2 = 0;
  s->c1 = 0.0;
  s->c1_plus_c2 = 0.0;
  s->ratio = 0.0;
  s->done = 0;
  s->j_ind = 0;
}

// Block-level max of two non-negative doubles given as bit patterns; thread 0 writes the pair
// to part[2*blockIdx.x]. Non-negative doubles order like their bit patterns.
template <int BS>
__device__ __forceinline__ void block_max2_store_t(u64 a, u64 b, u64* __restrict__ part) {
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
    part[2 * blockIdx.x] = ma;
    part[2 * blockIdx.x + 1] = mb;
  }
}

__device__ __forceinline__ void block_max2_store(u64 a, u64 b, u64* __restrict__ part) {
  __shared__ u64 sa[kBlock / 64], sb[kBlock / 64];
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
    for (int q = 0; q < kBlock / 64; ++q) {
      ma = sa[q] > ma ? sa[q] : ma;
      mb = sb[q] > mb ? sb[q] : mb;
    }
    part[2 * blockIdx.x] = ma;
    part[2 * blockIdx.x + 1] = mb;
  }
}

// one workgroup: max over nblk partial pairs -> (m0, m1)
__device__ __forceinline__ void reduce_partials(const u64* __restrict__ part, int nblk, u64* m0,
                                                u64* m1) {
  __shared__ u64 sa[1024 / 64], sb[1024 / 64];
  u64 a = 0, b = 0;
  // 8 independent loads in flight per thread
  int k = threadIdx.x;
  for (; k + 7 * (int)blockDim.x < nblk; k += 8 * blockDim.x) {
    u64 x[8], y[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      x[q] = part[2 * (k + q * blockDim.x)];
      y[q] = part[2 * (k + q * blockDim.x) + 1];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      a = x[q] > a ? x[q] : a;
      b = y[q] > b ? y[q] : b;
    }
  }
  for (; k < nblk; k += blockDim.x) {
    u64 x = part[2 * k], y = part[2 * k + 1];
    a = x > a ? x : a;
    b = y > b ? y : b;
  }
  a = wave_max_u64(a);
  b = wave_max_u64(b);
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 ma = 0, mb = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      ma = sa[q] > ma ? sa[q] : ma;
      mb = sb[q] > mb ? sb[q] : mb;
    }
    *m0 = ma;
    *m1 = mb;
  }
}

__global__ void __launch_bounds__(kBlock) k_tay_begin(int64_t R, const double2* __restrict__ v,
                                                      double2* __restrict__ res,
                                                      u64* __restrict__ part) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  u64 m = 0;
  if (i < R) {
    double2 a = v[i];
    res[i] = a;
    double q = a.x * a.x + a.y * a.y;
    m = (u64)__double_as_longlong(q);
  }
  block_max2_store(m, 0, part);
}

__global__ void __launch_bounds__(1024) k_tay_begin_fin(paces_taylor_state_t* s,
                                                        const u64* __restrict__ part, int nblk) {
  u64 m0 = 0, m1 = 0;
  reduce_partials(part, nblk, &m0, &m1);
  if (threadIdx.x == 0) {
    s->c1 = sqrt(__longlong_as_double((long long)m0));
    s->rn = s->c1;  // res = v (the mixed form's first even term divides by it)
  }
}

// One Taylor term j: x_out = (-i c) H x_in, res += x_out, per-workgroup maxima of |x_out|^2 and
// |res|^2 for the stopping test. (a + ib) * (-0 - ic) = (b*c, -(a*c)) exactly. Launch shape: BS
// threads per workgroup, each taking RPT rows (t, t + BS, ...) of the workgroup's BS * RPT rows;
// per-row arithmetic does not depend on the shape and the block maxima are order-independent, so
// every shape gives identical results.
template <int BS, int RPT>
__global__ void __launch_bounds__(BS) k_tay_term_t(int64_t R, const double* __restrict__ diag,
                                                   CsrSet H, const double2* __restrict__ xin,
                                                   double2* __restrict__ xout,
                                                   double2* __restrict__ res, double c,
                                                   const paces_taylor_state_t* __restrict__ s,
                                                   u64* __restrict__ part) {
  if (s->done) return;
  const int64_t base = (int64_t)blockIdx.x * (BS * RPT) + threadIdx.x;
  u64 mt = 0, mr = 0;
  // rows in load phases (hx_rows_phased) for one or two terms, where the registers allow
  constexpr bool phased = BS * RPT <= 2048;
  if (phased && (H.nterms == 1 || H.nterms == 2)) {
    int64_t ii[RPT];
    bool vv[RPT];
    double2 rr[RPT], h[RPT], xi[RPT];
    const int64_t spare = (int64_t)blockIdx.x * (BS * RPT);  // the block's first row (< R)
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      ii[q] = base + (int64_t)q * BS;
      vv[q] = ii[q] < R;
      rr[q] = res[vv[q] ? ii[q] : spare];
    }
    if (H.nterms == 2)
      hx_rows_phased<double, double2, RPT, 2>(ii, vv, spare, diag, H, xin, h, xi);
    else
      hx_rows_phased<double, double2, RPT, 1>(ii, vv, spare, diag, H, xin, h, xi);
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      if (!vv[q]) continue;
      const double2 t = make_double2(h[q].y * c, -(h[q].x * c));
      xout[ii[q]] = t;
      double2 r = rr[q];
      r.x = r.x + t.x;
      r.y = r.y + t.y;
      res[ii[q]] = r;
      const u64 a = (u64)__double_as_longlong(t.x * t.x + t.y * t.y);
      const u64 b = (u64)__double_as_longlong(r.x * r.x + r.y * r.y);
      mt = a > mt ? a : mt;
      mr = b > mr ? b : mr;
    }
    block_max2_store_t<BS>(mt, mr, part);
    return;
  }
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int64_t i = base + (int64_t)q * BS;
    if (i < R) {
      const double2 h = hx_row(i, diag, H, xin);
      const double2 t = make_double2(h.y * c, -(h.x * c));
      xout[i] = t;
      double2 r = res[i];
      r.x = r.x + t.x;
      r.y = r.y + t.y;
      res[i] = r;
      const u64 a = (u64)__double_as_longlong(t.x * t.x + t.y * t.y);
      const u64 b = (u64)__double_as_longlong(r.x * r.x + r.y * r.y);
      mt = a > mt ? a : mt;
      mr = b > mr ? b : mr;
    }
  }
  block_max2_store_t<BS>(mt, mr, part);
}

// use_prev: the term formed no partial sum (a mixed-form even term): divide by the last exact
// ||res||_inf (s->rn) instead of this term's, which is then kept in s->rn
__global__ void __launch_bounds__(1024) k_tay_check(paces_taylor_state_t* s,
                                                    const u64* __restrict__ part, int nblk, int j,
                                                    double tol, int use_prev) {
  if (s->done) return;
  u64 m0 = 0, m1 = 0;
  reduce_partials(part, nblk, &m0, &m1);
  if (threadIdx.x != 0) return;
  const double c2 = sqrt(__longlong_as_double((long long)m0));
  const double rn = use_prev ? s->rn : sqrt(__longlong_as_double((long long)m1));
  if (!use_prev) s->rn = rn;
  const double c1pc2 = s->c1 + c2;
  const double ratio = c1pc2 / rn;
  s->c1_plus_c2 = c1pc2;
  s->ratio = ratio;
  s->j_ind = j;
  if (ratio <= tol)
    s->done = 1;
  else
    s->c1 = c2;
}

// ---- mixed precision: complex64 terms, FP32 matrix and accumulation, complex128 sum ----------

struct CsrSet32 {
  int nterms;
  const int32_t* indptr[PACES_MAX_TERMS];
  const int32_t* indices[PACES_MAX_TERMS];
  const float* data[PACES_MAX_TERMS];
};

__device__ __forceinline__ float2 hx_row32(int64_t i, const float* __restrict__ diag,
                                           const CsrSet32& H, const float2* __restrict__ x) {
  const float2 xi = x[i];
  const float d = diag[i];
  float yr = d * xi.x, yi = d * xi.y;
  for (int t = 0; t < H.nterms; ++t) {
    const int32_t b = H.indptr[t][i], e = H.indptr[t][i + 1];
    float tr = 0.0f, ti = 0.0f;
    for (int32_t k = b; k < e; ++k) {
      const float a = H.data[t][k];
      const float2 xj = x[H.indices[t][k]];
      tr = tr + a * xj.x;
      ti = ti + a * xj.y;
    }
    yr = yr + tr;
    yi = yi + ti;
  }
  return make_float2(yr, yi);
}

// res = v (complex128), x0 = v rounded to complex64, partial maxima of |v|^2
__global__ void __launch_bounds__(kBlock) k_tay_begin_mixed(int64_t R,
                                                            const double2* __restrict__ v,
                                                            double2* __restrict__ res,
                                                            float2* __restrict__ x0,
                                                            u64* __restrict__ part) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  u64 m = 0;
  if (i < R) {
    double2 a = v[i];
    res[i] = a;
    x0[i] = make_float2((float)a.x, (float)a.y);
    m = (u64)__double_as_longlong(a.x * a.x + a.y * a.y);
  }
  block_max2_store(m, 0, part);
}

// One mixed term j. The complex128 partial sum is formed and stored on odd terms only: an odd
// term adds its input (t_(j-1), the even term's output) and then its own output,
// (res + t_(j-1)) + t_j, so a series that stops on an even term needs
// paces_taylor_finish_mixed to add that term. An even term neither reads nor writes res; its
// stopping test divides by the last exact ||res||_inf (k_tay_check's use_prev).
template <int BS, int RPT, bool ODD>
__device__ __forceinline__ void tay_term_mixed_body(int64_t R, const float* __restrict__ diag,
                                                    const CsrSet32& H,
                                                    const float2* __restrict__ xin,
                                                    float2* __restrict__ xout,
                                                    double2* __restrict__ res, float c,
                                                    u64* __restrict__ part) {
  const int64_t base = (int64_t)blockIdx.x * (BS * RPT) + threadIdx.x;
  u64 mt = 0, mr = 0;
  int64_t ii[RPT];
  bool vv[RPT];
  double2 rr[RPT];
  float2 h[RPT], xi[RPT];
  const int64_t spare = (int64_t)blockIdx.x * (BS * RPT);  // the block's first row (< R)
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    ii[q] = base + (int64_t)q * BS;
    vv[q] = ii[q] < R;
    if (ODD) rr[q] = res[vv[q] ? ii[q] : spare];
  }
  if (H.nterms == 2) {  // rows in load phases (hx_rows_phased)
    hx_rows_phased<float, float2, RPT, 2>(ii, vv, spare, diag, H, xin, h, xi);
  } else if (H.nterms == 1) {
    hx_rows_phased<float, float2, RPT, 1>(ii, vv, spare, diag, H, xin, h, xi);
  } else {
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int64_t i = vv[q] ? ii[q] : spare;
      h[q] = hx_row32(i, diag, H, xin);
      xi[q] = xin[i];
    }
  }
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    if (!vv[q]) continue;
    const float2 t = make_float2(h[q].y * c, -(h[q].x * c));
    xout[ii[q]] = t;
    const double tx = t.x, ty = t.y;
    const u64 a = (u64)__double_as_longlong(tx * tx + ty * ty);
    mt = a > mt ? a : mt;
    if (ODD) {
      double2 r = rr[q];
      r.x = r.x + (double)xi[q].x;
      r.y = r.y + (double)xi[q].y;
      r.x = r.x + (double)t.x;
      r.y = r.y + (double)t.y;
      res[ii[q]] = r;
      const u64 b = (u64)__double_as_longlong(r.x * r.x + r.y * r.y);
      mr = b > mr ? b : mr;
    }
  }
  block_max2_store_t<BS>(mt, mr, part);
}

template <int BS, int RPT>
__global__ void __launch_bounds__(BS) k_tay_term_mixed_t(
    int64_t R, const float* __restrict__ diag, CsrSet32 H, const float2* __restrict__ xin,
    float2* __restrict__ xout, double2* __restrict__ res, float c, int odd,
    const paces_taylor_state_t* __restrict__ s, u64* __restrict__ part) {
  if (s->done) return;
  if (odd)
    tay_term_mixed_body<BS, RPT, true>(R, diag, H, xin, xout, res, c, part);
  else
    tay_term_mixed_body<BS, RPT, false>(R, diag, H, xin, xout, res, c, part);
}

// after the series: a series that stopped on an even term j has res one term behind; that term
// is still in x_even (even terms write there, and later terms skip a stopped series)
__global__ void __launch_bounds__(kBlock) k_tay_finish_mixed(int64_t R,
                                                             const float2* __restrict__ x_even,
                                                             double2* __restrict__ res,
                                                             const paces_taylor_state_t* s) {
  if ((s->j_ind & 1) != 0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const float2 x = x_even[i];
  double2 r = res[i];
  r.x = r.x + (double)x.x;
  r.y = r.y + (double)x.y;
  res[i] = r;
}

__global__ void __launch_bounds__(kBlock) k_f64_to_f32(int64_t n, const double* __restrict__ in,
                                                       float* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

static bool csr_set32(const paces_csr32_set_t* h, CsrSet32* H) {
  if (!h || h->nterms < 0 || h->nterms > PACES_MAX_TERMS) return false;
  H->nterms = h->nterms;
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    H->indptr[t] = t < h->nterms ? h->indptr[t] : nullptr;
    H->indices[t] = t < h->nterms ? h->indices[t] : nullptr;
    H->data[t] = t < h->nterms ? h->data[t] : nullptr;
    if (t < h->nterms && !H->indptr[t]) return false;
  }
  return true;
}

// ---- deterministic reductions ------------------------------------------------------------------

constexpr int kRedTile = kBlock * 4;

// mode 0: sum |x|^2 -> part[b]; mode 1: sum conj(x) y -> part[2b], part[2b+1]
__global__ void __launch_bounds__(kBlock) k_red_partial(const double2* __restrict__ x,
                                                        const double2* __restrict__ y, int64_t R,
                                                        int mode, double* __restrict__ part,
                                                        const uint8_t* __restrict__ mask) {
  __shared__ double red[2][kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kRedTile;
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
    if (i < R) {
      double2 u = x[i];
      if (mode == 0) {
        if (!mask || mask[i]) a += u.x * u.x + u.y * u.y;
      } else {
        double2 v = y[i];
        a += u.x * v.x + u.y * v.y;
        b += u.x * v.y - u.y * v.x;
      }
    }
  }
  a = wave_sum(a);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) {
      s0 += red[0][q];
      s1 += red[1][q];
    }
    if (mode == 0) {
      part[blockIdx.x] = s0;
    } else {
      part[2 * blockIdx.x] = s0;
      part[2 * blockIdx.x + 1] = s1;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_red_final(const double* __restrict__ part, int nblk,
                                                      int width, double* __restrict__ out) {
  __shared__ double red[kBlock / 64];
  const int c = blockIdx.x;
  double a = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) a += part[(int64_t)b * width + c];
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) t += red[q];
    out[c] = t;
  }
}

static int64_t red_blocks(int64_t R) {
  int64_t n = (R + kRedTile - 1) / kRedTile;
  return n < 1 ? 1 : n;
}

__global__ void __launch_bounds__(kBlock) k_axpby(int64_t R, double ar, double ai,
                                                  const double2* __restrict__ x, double br,
                                                  double bi, double2* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  double2 u = x[i];
  double2 o;
  o.x = ar * u.x - ai * u.y;
  o.y = ar * u.y + ai * u.x;
  if (br != 0.0 || bi != 0.0) {  // b == 0: y is write-only (may be uninitialised)
    double2 v = y[i];
    o.x = o.x + (br * v.x - bi * v.y);
    o.y = o.y + (br * v.y + bi * v.x);
  }
  y[i] = o;
}

static bool csr_set(const paces_csr_set_t* h, CsrSet* H) {
  if (!h || h->nterms < 0 || h->nterms > PACES_MAX_TERMS) return false;
  H->nterms = h->nterms;
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    H->indptr[t] = t < h->nterms ? h->indptr[t] : nullptr;
    H->indices[t] = t < h->nterms ? h->indices[t] : nullptr;
    H->data[t] = t < h->nterms ? h->data[t] : nullptr;
    if (t < h->nterms && !H->indptr[t]) return false;
  }
  return true;
}

}  // namespace paces

using namespace paces;

extern "C" {

int paces_spmv(int64_t R, const double* diag, const paces_csr_set_t* h, const double* x,
               double* y, void* stream) {
  CsrSet H;
  if (R < 0 || !csr_set(h, &H)) return PACES_E_INVALID;
  if (R == 0) return PACES_OK;
  if (!diag || !x || !y) return PACES_E_INVALID;
  hipLaunchKernelGGL(k_spmv, dim3(grid_for(R, kBlock)), dim3(kBlock), 0, (hipStream_t)stream, R,
                     diag, H, (const double2*)x, (double2*)y);
  return launch_status();
}

size_t paces_taylor_workspace_bytes(int64_t R) {
  int64_t nb = (R + 127) / 128;  // partial pairs for the smallest tile of any launch shape
  return a256((size_t)(nb < 1 ? 1 : nb) * 2 * sizeof(u64));
}

int paces_taylor_begin(int64_t R, const double* v, double* res, paces_taylor_state_t* state,
                       void* ws, size_t ws_bytes, void* stream) {
  if (R < 0 || !state || !ws || (R > 0 && (!v || !res))) return PACES_E_INVALID;
  if (ws_bytes < paces_taylor_workspace_bytes(R)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_tay_zero, dim3(1), dim3(1), 0, st, state);
  int nblk = 0;
  if (R > 0) {
    nblk = (int)grid_for(R, kBlock);
    hipLaunchKernelGGL(k_tay_begin, dim3(nblk), dim3(kBlock), 0, st, R, (const double2*)v,
                       (double2*)res, (u64*)ws);
  }
  hipLaunchKernelGGL(k_tay_begin_fin, dim3(1), dim3(1024), 0, st, state, (const u64*)ws, nblk);
  return launch_status();
}

int paces_taylor_term(int64_t R, const double* diag, const paces_csr_set_t* h, const double* x_in,
                      double* x_out, double* res, double c, int32_t j, double tol,
                      paces_taylor_state_t* state, void* ws, size_t ws_bytes, void* stream) {
  return paces_taylor_term_tuned(R, diag, h, x_in, x_out, res, c, j, tol, state, ws, ws_bytes,
                                 nullptr, stream);
}

int paces_taylor_term_tuned(int64_t R, const double* diag, const paces_csr_set_t* h,
                            const double* x_in, double* x_out, double* res, double c, int32_t j,
                            double tol, paces_taylor_state_t* state, void* ws, size_t ws_bytes,
                            const paces_launch_cfg_t* cfg, void* stream) {
  CsrSet H;
  if (R < 0 || !state || !ws || !csr_set(h, &H) || !cfg_ok(cfg)) return PACES_E_INVALID;
  if (R > 0 && (!diag || !x_in || !x_out || !res)) return PACES_E_INVALID;
  if (ws_bytes < paces_taylor_workspace_bytes(R)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  int nblk = 0;
  if (R > 0) {
    int B, RR;
    cfg_get(cfg, &B, &RR);
    nblk = (int)grid_for(R, B * RR);
#define TAY_LAUNCH(BB, QQ)                                                                  \
  hipLaunchKernelGGL((k_tay_term_t<BB, QQ>), dim3(nblk), dim3(BB), 0, st, R, diag, H,       \
                     (const double2*)x_in, (double2*)x_out, (double2*)res, c,              \
                     (const paces_taylor_state_t*)state, (u64*)ws)
    PACES_CFG_SWITCH(B, RR, TAY_LAUNCH)
#undef TAY_LAUNCH
  }
  hipLaunchKernelGGL(k_tay_check, dim3(1), dim3(1024), 0, st, state, (const u64*)ws, nblk,
                     (int)j, tol, 0);
  return launch_status();
}

int paces_convert_f64_to_f32(int64_t n, const double* in, float* out, void* stream) {
  if (n < 0 || (n > 0 && (!in || !out))) return PACES_E_INVALID;
  if (n == 0) return PACES_OK;
  hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, (hipStream_t)stream,
                     n, in, out);
  return launch_status();
}

int paces_taylor_begin_mixed(int64_t R, const double* v, double* res, float* x0,
                             paces_taylor_state_t* state, void* ws, size_t ws_bytes,
                             void* stream) {
  if (R < 0 || !state || !ws || (R > 0 && (!v || !res || !x0))) return PACES_E_INVALID;
  if (ws_bytes < paces_taylor_workspace_bytes(R)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_tay_zero, dim3(1), dim3(1), 0, st, state);
  int nblk = 0;
  if (R > 0) {
    nblk = (int)grid_for(R, kBlock);
    hipLaunchKernelGGL(k_tay_begin_mixed, dim3(nblk), dim3(kBlock), 0, st, R, (const double2*)v,
                       (double2*)res, (float2*)x0, (u64*)ws);
  }
  hipLaunchKernelGGL(k_tay_begin_fin, dim3(1), dim3(1024), 0, st, state, (const u64*)ws, nblk);
  return launch_status();
}

int paces_taylor_term_mixed(int64_t R, const float* diag, const paces_csr32_set_t* h,
                            const float* x_in, float* x_out, double* res, float c, int32_t j,
                            double tol, paces_taylor_state_t* state, void* ws, size_t ws_bytes,
                            void* stream) {
  return paces_taylor_term_mixed_tuned(R, diag, h, x_in, x_out, res, c, j, tol, state, ws,
                                       ws_bytes, nullptr, stream);
}

int paces_taylor_term_mixed_tuned(int64_t R, const float* diag, const paces_csr32_set_t* h,
                                  const float* x_in, float* x_out, double* res, float c,
                                  int32_t j, double tol, paces_taylor_state_t* state, void* ws,
                                  size_t ws_bytes, const paces_launch_cfg_t* cfg, void* stream) {
  CsrSet32 H;
  if (R < 0 || !state || !ws || !csr_set32(h, &H) || !cfg_ok(cfg)) return PACES_E_INVALID;
  if (R > 0 && (!diag || !x_in || !x_out || !res)) return PACES_E_INVALID;
  if (ws_bytes < paces_taylor_workspace_bytes(R)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  int nblk = 0;
  if (R > 0) {
    int B, RR;
    cfg_get(cfg, &B, &RR);
    nblk = (int)grid_for(R, B * RR);
#define TAYM_LAUNCH(BB, QQ)                                                                 \
  hipLaunchKernelGGL((k_tay_term_mixed_t<BB, QQ>), dim3(nblk), dim3(BB), 0, st, R, diag, H, \
                     (const float2*)x_in, (float2*)x_out, (double2*)res, c, (int)(j & 1),  \
                     (const paces_taylor_state_t*)state, (u64*)ws)
    PACES_CFG_SWITCH(B, RR, TAYM_LAUNCH)
#undef TAYM_LAUNCH
  }
  hipLaunchKernelGGL(k_tay_check, dim3(1), dim3(1024), 0, st, state, (const u64*)ws, nblk,
                     (int)j, tol, (j & 1) ? 0 : 1);
  return launch_status();
}

int paces_taylor_finish_mixed(int64_t R, const float* x_even, double* res,
                              const paces_taylor_state_t* state, void* stream) {
  if (R < 0 || !state || (R > 0 && (!x_even || !res))) return PACES_E_INVALID;
  if (R == 0) return PACES_OK;
  hipLaunchKernelGGL(k_tay_finish_mixed, dim3(grid_for(R, kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, R, (const float2*)x_even, (double2*)res, state);
  return launch_status();
}

size_t paces_reduce_workspace_bytes(int64_t R) { return a256((size_t)red_blocks(R) * 2 * 8); }

int paces_norm2(const double* x, int64_t R, double* out, void* ws, size_t ws_bytes, void* stream) {
  if (R < 0 || !out || !ws || (R > 0 && !x)) return PACES_E_INVALID;
  if (ws_bytes < paces_reduce_workspace_bytes(R)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  int64_t nb = red_blocks(R);
  hipLaunchKernelGGL(k_red_partial, dim3((unsigned)nb), dim3(kBlock), 0, st, (const double2*)x,
                     (const double2*)nullptr, R, 0, (double*)ws, (const uint8_t*)nullptr);
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(kBlock), 0, st, (const double*)ws, (int)nb, 1,
                     out);
  return launch_status();
}

int paces_norm2_masked(const double* x, const uint8_t* mask, int64_t R, double* out, void* ws,
                       size_t ws_bytes, void* stream) {
  if (R < 0 || !out || !ws || (R > 0 && (!x || !mask))) return PACES_E_INVALID;
  if (ws_bytes < paces_reduce_workspace_bytes(R)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  int64_t nb = red_blocks(R);
  hipLaunchKernelGGL(k_red_partial, dim3((unsigned)nb), dim3(kBlock), 0, st, (const double2*)x,
                     (const double2*)nullptr, R, 0, (double*)ws, mask);
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(kBlock), 0, st, (const double*)ws, (int)nb, 1,
                     out);
  return launch_status();
}

int paces_vdot(const double* x, const double* y, int64_t R, double* out, void* ws,
               size_t ws_bytes, void* stream) {
  if (R < 0 || !out || !ws || (R > 0 && (!x || !y))) return PACES_E_INVALID;
  if (ws_bytes < paces_reduce_workspace_bytes(R)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  int64_t nb = red_blocks(R);
  hipLaunchKernelGGL(k_red_partial, dim3((unsigned)nb), dim3(kBlock), 0, st, (const double2*)x,
                     (const double2*)y, R, 1, (double*)ws, (const uint8_t*)nullptr);
  hipLaunchKernelGGL(k_red_final, dim3(2), dim3(kBlock), 0, st, (const double*)ws, (int)nb, 2,
                     out);
  return launch_status();
}

int paces_axpby(int64_t R, double ar, double ai, const double* x, double br, double bi, double* y,
                void* stream) {
  if (R < 0 || (R > 0 && (!x || !y))) return PACES_E_INVALID;
  if (R == 0) return PACES_OK;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(R, kBlock)), dim3(kBlock), 0, (hipStream_t)stream, R,
                     ar, ai, (const double2*)x, br, bi, (double2*)y);
  return launch_status();
}

const char* paces_strerror(int code) {
  switch (code) {
    case PACES_OK: return "ok";
    case PACES_E_INVALID: return "invalid argument";
    case PACES_E_UNSUPPORTED: return "unsupported size (row width, term or run count)";
    case PACES_E_HIP: return "HIP launch or runtime error";
    case PACES_E_WORKSPACE: return "workspace too small";
    case PACES_E_SCRATCH: return "size beyond the fixed scratch of a no-allocation entry point";
    default: return "unknown paces status";
  }
}

int paces_version(void) { return 10000; }

}  // extern "C"
