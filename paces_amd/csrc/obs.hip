// The Holstein observables beyond the site densities (paces/models/holstein.py:325-377): sums of
// conj(x) * y over the rows of each exciton position (bond hopping and vibronic coupling
// energies), and the sum of the amplitudes of the phonon-vacuum rows (<mu(t) mu(0)>).
//
// Rows of a plain basis are sorted with the exciton position as the leading field, so the rows
// of position p are one segment [seg_off[p], seg_off[p + 1]) (paces_batch_segments with the
// position count in place of the trajectory count gives the offsets), and the vacuum row of
// position p, all phonon fields zero, is the first row of its segment when present.
#include "common.hpp"

namespace paces {

constexpr int kSegBlock = 256;

// out[2s], out[2s+1] = sum over segment s of conj(x_i) * y_i. One workgroup per segment; a thread
// sums rows s0 + t, s0 + t + 256, ... in order, then a fixed tree: the result does not depend
// on anything but the segment's values.
__global__ void __launch_bounds__(kSegBlock) k_seg_vdot(const double2* __restrict__ x,
                                                        const double2* __restrict__ y,
                                                        const int32_t* __restrict__ seg_off,
                                                        double* __restrict__ out) {
  __shared__ double sr[kSegBlock / 64], si[kSegBlock / 64];
  const int s = blockIdx.x;
  const int64_t b = seg_off[s], e = seg_off[s + 1];
  double ar = 0.0, ai = 0.0;
  for (int64_t i = b + threadIdx.x; i < e; i += kSegBlock) {
    const double2 u = x[i], v = y[i];
    // conj(u) * v = (u.x v.x + u.y v.y) + i (u.x v.y - u.y v.x)
    ar = ar + (u.x * v.x + u.y * v.y);
    ai = ai + (u.x * v.y - u.y * v.x);
  }
  ar = wave_sum(ar);
  ai = wave_sum(ai);
  if ((threadIdx.x & 63) == 0) {
    sr[threadIdx.x >> 6] = ar;
    si[threadIdx.x >> 6] = ai;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0.0, ti = 0.0;
#pragma unroll
    for (int w = 0; w < kSegBlock / 64; ++w) {
      tr = tr + sr[w];
      ti = ti + si[w];
    }
    out[2 * s] = tr;
    out[2 * s + 1] = ti;
  }
}

// out[0..1] = sum over positions p (in order) of x at the first row of segment p when every bit
// of that row after the first lead_bits is zero (the phonon vacuum at position p).
__global__ void k_vacuum_sum(const u64* __restrict__ rows, int W, int lead_bits,
                             const double2* __restrict__ x, const int32_t* __restrict__ seg_off,
                             int nseg, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double sr = 0.0, si = 0.0;
  for (int p = 0; p < nseg; ++p) {
    const int64_t i = seg_off[p];
    if (i >= seg_off[p + 1]) continue;
    const u64* r = rows + i * W;
    // word 0 without its lead bits, then the remaining words
    u64 rest = lead_bits >= 64 ? 0ull : (r[0] << lead_bits);
    for (int k = 1; k < W; ++k) rest |= r[k];
    if (rest == 0ull) {
      sr = sr + x[i].x;
      si = si + x[i].y;
    }
  }
  out[0] = sr;
  out[1] = si;
}

}  // namespace paces

using namespace paces;

extern "C" {

int paces_segment_vdot(const double* x, const double* y, const int32_t* seg_off, int32_t nseg,
                       double* out, void* stream) {
  if (nseg < 1 || !seg_off || !out || !x || !y) return PACES_E_INVALID;
  hipLaunchKernelGGL(k_seg_vdot, dim3((unsigned)nseg), dim3(kSegBlock), 0, (hipStream_t)stream,
                     (const double2*)x, (const double2*)y, seg_off, out);
  return launch_status();
}

int paces_vacuum_sum(const uint64_t* rows, int32_t W, int32_t lead_bits, const double* x,
                     const int32_t* seg_off, int32_t nseg, double* out, void* stream) {
  if (W < 1 || W > PACES_MAX_ROW_WORDS || lead_bits < 0 || lead_bits > 64 || nseg < 1 ||
      !seg_off || !out || !rows || !x)
    return PACES_E_INVALID;
  hipLaunchKernelGGL(k_vacuum_sum, dim3(1), dim3(64), 0, (hipStream_t)stream, (const u64*)rows,
                     (int)W, (int)lead_bits, (const double2*)x, seg_off, (int)nseg, out);
  return launch_status();
}

}  // extern "C"
