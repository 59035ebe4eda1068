// Candidate scorer: a 64 -> 256 -> 256 -> 256 -> 1 network (GELU hidden layers, sigmoid output)
// with BF16 matrix operands and FP32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16). No counterpart in the reference: it scores the candidate rows of
// an expansion so that rows unlikely to survive the next selection are not materialised.
//
// Parameters live in one FP32 vector (the master copy Adam updates):
//   Wt1 [256 x 64] | b1 [256] | Wt2 [256 x 256] | b2 | Wt3 [256 x 256] | b3 | w4 [256] | b4 [1]
// Wt is output-major: row n holds the weights into output n, so a lane's 8 consecutive k of one
// output are one 16-byte load. The three matrices also have BF16 copies (the matrix operands).
//
// Kernels:
//   k_gemm<AK, BK>  generic 64x64 tile product through LDS with three epilogues (forward layer,
//                   input gradient times GELU', weight-gradient partials for split-K)
//   k_head_train    output layer, loss, and the gradient into the last hidden layer
//   k_colsum / k_sum_parts   deterministic bias-gradient and split-K reductions
//   k_adam          Adam update of the FP32 master and refresh of the BF16 copies
//   k_mlp_infer<RB> all layers fused for a tile of 64 RB rows: activations kept in LDS, weight
//                   chunks staged once per workgroup in LDS, transposed products

#include "common.hpp"

namespace paces {
namespace mlp {

constexpr int IN = PACES_MLP_IN, HID = PACES_MLP_HIDDEN;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + HID * IN, OFF_W2 = OFF_B1 + HID,
              OFF_B2 = OFF_W2 + HID * HID, OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID * HID,
              OFF_W4 = OFF_B3 + HID, OFF_B4 = OFF_W4 + HID, NPARAM = OFF_B4 + 1;
constexpr int BOFF_W1 = 0, BOFF_W2 = HID * IN, BOFF_W3 = BOFF_W2 + HID * HID,
              NBF = BOFF_W3 + HID * HID;
static_assert(NPARAM == PACES_MLP_NPARAM, "parameter count");
static_assert(NBF == PACES_MLP_NBF16, "bf16 weight count");

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (the conversion gfx950 does in hardware)
__device__ __forceinline__ uint16_t f2bf(float f) {
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// two floats -> packed bf16 bits (lo = a), one v_cvt_pk_bf16_f32
typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
__device__ __forceinline__ uint32_t bf16_pair(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// GELU(a) = a Phi(a) with Phi the standard normal CDF, in a form without transcendental functions
// or branches (the epilogue of every layer evaluates it on each activation, and v_exp / v_rcp run
// at a quarter of the VALU rate): GELU(a) = a / 2 + h(a) with h even, h = u P(u) (u = a^2) for
// |a| <= 3.5 and h = |a| / 2 beyond. P has degree 5, fitted to minimise the largest error of GELU
// under the constraint 3.5^2 P(3.5^2) = 3.5 / 2, so the pieces meet at |a| = 3.5 (a above, 0
// below). With c = a clamped to [-3.5, 3.5] and u = c^2:
//   GELU(a) = fma(1/2, |a| - |c|, fma(1/2, a, u P(u)))
// in 11 operations (|a| - |c| is 0 inside and |a| - 3.5 beyond; no max, whose IEEE form would
// add NaN-quieting operations). Largest absolute error against the erf form: 8.2e-4 (well below a bf16 unit of
// the outputs it feeds); GELU(a) - GELU(-a) = a holds exactly in real arithmetic (P is even).
constexpr float kGeluCut = 3.5f;
constexpr float kG0 = 0.39518553018569946f, kG1 = -0.061023153364658356f,
                kG2 = 0.0072402348741889f, kG3 = -0.0005149735952727497f,
                kG4 = 1.9077817341894843e-05f, kG5 = -2.691017755296343e-07f;

__device__ __forceinline__ float gelu_f(float a) {
  const float c = __builtin_amdgcn_fmed3f(a, -kGeluCut, kGeluCut);
  const float u = c * c;
  float p = fmaf(kG5, u, kG4);
  p = fmaf(p, u, kG3);
  p = fmaf(p, u, kG2);
  p = fmaf(p, u, kG1);
  p = fmaf(p, u, kG0);
  return fmaf(0.5f, fabsf(a) - fabsf(c), fmaf(0.5f, a, p * u));
}

// GELU(a) = 0.5 a + sum c_k a^(2k+2), so GELU'(a) = 0.5 + a sum (2k + 2) c_k a^(2k)
__device__ __forceinline__ float gelu_grad_f(float a) {
  const float u = a * a;
  const float q = 2.0f * kG0 + u * (4.0f * kG1 + u * (6.0f * kG2 + u * (8.0f * kG3 +
                  u * (10.0f * kG4 + u * (12.0f * kG5)))));
  const float d = 0.5f + a * q;
  return a > kGeluCut ? 1.0f : (a < -kGeluCut ? 0.0f : d);
}

__device__ __forceinline__ bf16x8 lds_frag(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *(const u32x4*)p);
}
__device__ __forceinline__ bf16x8 glb_frag(const uint16_t* __restrict__ p) {
  return __builtin_bit_cast(bf16x8, *(const u32x4*)p);
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
  return v;
}

// ---- generic tile product --------------------------------------------------------------------
// C[m][n] = sum_k A(m,k) B(k,n); A(m,k) = A[row(m)*sam + k*sak], B(k,n) = B[row(k)*sbk + n*sbn]
// with optional row gathers aidx (A rows) and bidx (B's k rows). AK: A is k-contiguous (sak = 1),
// else m-contiguous (sam = 1); BK: B is k-contiguous (sbk = 1), else n-contiguous (sbn = 1).
// Requires M, N multiples of 8 where the matching operand is loaded along them, K multiples of
// 8 for k-contiguous operands, and 16-byte aligned rows.
enum { kModeFwd = 0, kModeDx = 1, kModeDw = 2 };

struct GemmArg {
  const uint16_t* A;
  int64_t sam, sak;
  const int32_t* aidx;
  const uint16_t* B;
  int64_t sbk, sbn;
  const int32_t* bidx;
  int64_t M, N, K, kchunk;
  int mode;
  const float* bias;    // fwd
  float* pre;           // fwd: FP32 pre-activation [M x N]
  uint16_t* out;        // fwd: bf16 GELU(pre); dx: bf16 acc * GELU'(pre_in)
  const float* pre_in;  // dx
  float* part;          // dw: [gridDim.z][M x N]
};

constexpr int kLdsK = 40;  // 32 k values + padding (80-byte rows keep 16-byte alignment)

template <bool AK, bool BK>
__global__ void __launch_bounds__(256) k_gemm(GemmArg g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[64 * kLdsK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64 * kLdsK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * 64, n0 = (int64_t)blockIdx.y * 64;
  const int64_t kb = (int64_t)blockIdx.z * g.kchunk;
  const int64_t ke = (kb + g.kchunk < g.K) ? kb + g.kchunk : g.K;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int64_t k0 = kb; k0 < ke; k0 += 32) {
    // ---- stage A (64 x 32) as As[m][k]
    if (AK) {
      const int r = t >> 2, c8 = (t & 3) * 8;
      const int64_t m = m0 + r, k = k0 + c8;
      u32x4 v = {0u, 0u, 0u, 0u};
      if (m < g.M && k < ke) {
        const int64_t row = g.aidx ? (int64_t)g.aidx[m] : m;
        const uint16_t* p = g.A + row * g.sam + k;
        if (k + 8 <= ke) {
          v = *(const u32x4*)p;
        } else {
          uint16_t tmp[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) tmp[q] = (k + q < ke) ? p[q] : (uint16_t)0;
          v = __builtin_bit_cast(u32x4, tmp);
        }
      }
      *(u32x4*)&As[r * kLdsK + c8] = v;
    } else {
      const int c = t >> 3, r8 = (t & 7) * 8;
      const int64_t k = k0 + c, m = m0 + r8;
      uint16_t tmp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (k < ke && m < g.M) {
        const uint16_t* p = g.A + k * g.sak + m;
        if (m + 8 <= g.M) {
          u32x4 v = *(const u32x4*)p;
          const uint16_t* s = (const uint16_t*)&v;
#pragma unroll
          for (int q = 0; q < 8; ++q) tmp[q] = s[q];
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q) tmp[q] = (m + q < g.M) ? p[q] : (uint16_t)0;
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) As[(r8 + q) * kLdsK + c] = tmp[q];
    }
    // ---- stage B (32 x 64) as Bs[n][k]
    if (BK) {
      const int r = t >> 2, c8 = (t & 3) * 8;
      const int64_t n = n0 + r, k = k0 + c8;
      u32x4 v = {0u, 0u, 0u, 0u};
      if (n < g.N && k < ke) {
        const uint16_t* p = g.B + n * g.sbn + k;
        if (k + 8 <= ke) {
          v = *(const u32x4*)p;
        } else {
          uint16_t tmp[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) tmp[q] = (k + q < ke) ? p[q] : (uint16_t)0;
          v = __builtin_bit_cast(u32x4, tmp);
        }
      }
      *(u32x4*)&Bs[r * kLdsK + c8] = v;
    } else {
      const int c = t >> 3, r8 = (t & 7) * 8;
      const int64_t k = k0 + c, n = n0 + r8;
      uint16_t tmp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (k < ke && n < g.N) {
        const int64_t row = g.bidx ? (int64_t)g.bidx[k] : k;
        const uint16_t* p = g.B + row * g.sbk + n;
        if (n + 8 <= g.N) {
          u32x4 v = *(const u32x4*)p;
          const uint16_t* s = (const uint16_t*)&v;
#pragma unroll
          for (int q = 0; q < 8; ++q) tmp[q] = s[q];
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q) tmp[q] = (n + q < g.N) ? p[q] : (uint16_t)0;
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) Bs[(r8 + q) * kLdsK + c] = tmp[q];
    }
    __syncthreads();
    const int lr = lane & 15, lk = 8 * (lane >> 4);
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = lds_frag(&As[(wm + 16 * i + lr) * kLdsK + lk]);
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = lds_frag(&Bs[(wn + 16 * j + lr) * kLdsK + lk]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    __syncthreads();
  }

  // ---- epilogue: lane holds D[4*(lane/16) + r][lane%16] of each 16x16 block
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int64_t n = n0 + wn + 16 * j + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        const float v = acc[i][j][r];
        const int64_t o = m * g.N + n;
        if (g.mode == kModeFwd) {
          const float a = v + g.bias[n];
          g.pre[o] = a;
          g.out[o] = f2bf(gelu_f(a));
        } else if (g.mode == kModeDx) {
          g.out[o] = f2bf(v * gelu_grad_f(g.pre_in[o]));
        } else {
          g.part[(int64_t)blockIdx.z * g.M * g.N + o] = v;
        }
      }
}

// ---- output layer + loss + gradient into the last hidden layer --------------------------------
// One wave per row (64 lanes x 4 columns), 16 rows per wave, 64 rows per block. Per block:
// partial dW4 [256] and (db4, loss) partials, summed later in a fixed order.
__global__ void __launch_bounds__(256) k_head_train(const uint16_t* __restrict__ h3,
                                                    const float* __restrict__ a3,
                                                    const float* __restrict__ params,
                                                    const uint8_t* __restrict__ labels,
                                                    const int32_t* __restrict__ idx, int64_t B,
                                                    float pos_weight, uint16_t* __restrict__ da3,
                                                    float* __restrict__ part_w4,
                                                    float* __restrict__ part_s) {
  __shared__ float sw4[4][HID];
  __shared__ float ss[4][2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const float* w4 = params + OFF_W4;
  const float b4 = params[OFF_B4];
  float wl[4], acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < 4; ++c) wl[c] = w4[4 * lane + c];
  float sdz = 0.f, sloss = 0.f;
  const float invB = 1.0f / (float)B;
  for (int q = 0; q < 16; ++q) {
    const int64_t b = (int64_t)blockIdx.x * 64 + w * 16 + q;
    if (b >= B) break;
    float hv[4], zp = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      hv[c] = bf2f(h3[b * HID + 4 * lane + c]);
      zp += hv[c] * wl[c];
    }
    const float z = wave_sum_f(zp) + b4;
    const float y = (float)labels[idx ? idx[b] : b];
    const float wt = y > 0.f ? pos_weight : 1.0f;
    const float p = 1.0f / (1.0f + expf(-z));
    const float sp = fmaxf(z, 0.f) + log1pf(expf(-fabsf(z)));
    const float dz = wt * (p - y) * invB;
    sdz += dz;
    sloss += wt * (sp - y * z) * invB;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t o = b * HID + 4 * lane + c;
      da3[o] = f2bf(dz * wl[c] * gelu_grad_f(a3[o]));
      acc[c] += dz * hv[c];
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) sw4[w][4 * lane + c] = acc[c];
  if (lane == 0) {
    ss[w][0] = sdz;
    ss[w][1] = sloss;
  }
  __syncthreads();
  const int n = threadIdx.x;
  part_w4[(int64_t)blockIdx.x * HID + n] = sw4[0][n] + sw4[1][n] + sw4[2][n] + sw4[3][n];
  if (threadIdx.x < 2)
    part_s[(int64_t)blockIdx.x * 2 + threadIdx.x] =
        ss[0][threadIdx.x] + ss[1][threadIdx.x] + ss[2][threadIdx.x] + ss[3][threadIdx.x];
}

// column sums of a bf16 [rows x 256] matrix over chunks of 256 rows: part[chunk][256]
__global__ void __launch_bounds__(256) k_colsum(const uint16_t* __restrict__ a, int64_t rows,
                                                float* __restrict__ part) {
  const int n = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 256;
  const int64_t r1 = (r0 + 256 < rows) ? r0 + 256 : rows;
  float s = 0.f;
  for (int64_t r = r0; r < r1; ++r) s += bf2f(a[r * HID + n]);
  part[(int64_t)blockIdx.x * HID + n] = s;
}

// out[j * ostride] = sum over s of part[s * n + j], in order of s
__global__ void __launch_bounds__(256) k_sum_parts(const float* __restrict__ part, int nparts,
                                                   int64_t n, float* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float s = 0.f;
  for (int q = 0; q < nparts; ++q) s += part[(int64_t)q * n + j];
  out[j] = s;
}

__device__ __forceinline__ int bf_index(int j) {
  if (j >= OFF_W1 && j < OFF_B1) return BOFF_W1 + (j - OFF_W1);
  if (j >= OFF_W2 && j < OFF_B2) return BOFF_W2 + (j - OFF_W2);
  if (j >= OFF_W3 && j < OFF_B3) return BOFF_W3 + (j - OFF_W3);
  return -1;
}

__global__ void __launch_bounds__(256) k_to_bf16(const float* __restrict__ params,
                                                 uint16_t* __restrict__ wbf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NPARAM) return;
  const int q = bf_index(j);
  if (q >= 0) wbf[q] = f2bf(params[j]);
}

// Adam (Kingma & Ba 2015, Algorithm 1); c1 = 1/(1-beta1^t), c2 = 1/(1-beta2^t) from the host
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ params,
                                              const float* __restrict__ grads,
                                              float* __restrict__ m, float* __restrict__ v,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2, uint16_t* __restrict__ wbf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NPARAM) return;
  const float g = grads[j];
  const float mj = b1 * m[j] + (1.0f - b1) * g;
  const float vj = b2 * v[j] + (1.0f - b2) * g * g;
  m[j] = mj;
  v[j] = vj;
  const float p = params[j] - lr * (mj * c1) / (sqrtf(vj * c2) + eps);
  params[j] = p;
  const int q = bf_index(j);
  if (q >= 0) wbf[q] = f2bf(p);
}

// ---- fused inference: all layers for a tile of 64 RB rows, 2 RB waves ---------------------------
// Wave w takes rows [64 (w / 2), +64) and outputs [128 (w % 2), +128) of each layer. Each 32-k
// chunk of a layer's weights (256 outputs x 32 k, 16 KB) is loaded once per workgroup with
// coalesced 16-byte loads into LDS and read there by every wave; the next chunk is loaded into
// registers while the current one feeds the matrix cores. The activations of the tile stay in LDS
// (row-major, the features in the first IN columns until the first layer replaces them):
// 64 RB x 512 bytes, plus the 16 KB chunk.
//
// Each product is formed transposed, D[n][m] = sum_k W[n][k] H[m][k] (the weight fragment as the
// matrix core's A operand), so a lane's four accumulator values are four consecutive outputs of
// one row: the epilogue writes them as one 8-byte store. The accumulators start at the bias. The
// head sums, per row, over j (16-output blocks) then r within a lane, then across the four lane
// groups, then the two output halves: the same order for every RB, so the tiles give identical
// outputs.
//
// LDS layouts without padding, swizzled at 16-byte granularity so that the eight lanes a
// ds_read_b128 serves together (eight consecutive rows, one 16-byte column) fall in distinct banks:
//   activations: row m of 256 values, 16-byte column q stored at column q ^ (m % 8)
//   weight chunk: row n of 32 k values, 16-byte column q stored at column q ^ ((n / 2) % 4)
// Unpadded, a 128-row tile and its weight chunk take 80 KB, so two such workgroups share a CU.
__device__ __forceinline__ int hs_at(int m, int k) {
  return m * HID + ((((k >> 3) ^ (m & 7))) << 3) + (k & 7);
}
__device__ __forceinline__ int wb_at(int n, int k) {
  return n * 32 + ((((k >> 3) ^ ((n >> 1) & 3))) << 3) + (k & 7);
}

template <int K, int NT>
__device__ __forceinline__ void stage_chunk_load(const uint16_t* __restrict__ Wt, int c,
                                                 u32x4 (&v)[1024 / NT]) {
#pragma unroll
  for (int q = 0; q < 1024 / NT; ++q) {
    const int p = (int)threadIdx.x + NT * q, n = p >> 2, kk = (p & 3) * 8;
    v[q] = *(const u32x4*)(Wt + (int64_t)n * K + 32 * c + kk);
  }
}

template <int NT>
__device__ __forceinline__ void stage_chunk_store(uint16_t* Wb, const u32x4 (&v)[1024 / NT]) {
#pragma unroll
  for (int q = 0; q < 1024 / NT; ++q) {
    const int p = (int)threadIdx.x + NT * q, n = p >> 2, kk = (p & 3) * 8;
    *(u32x4*)&Wb[wb_at(n, kk)] = v[q];
  }
}

// acc[i][j] += W[128 ch + 16 j + .][:] . H[64 rb + 16 i + .][:] over K, weights through LDS
template <int K, int NT>
__device__ __forceinline__ void infer_layer(const uint16_t* Hs, uint16_t* Wb,
                                            const uint16_t* __restrict__ Wt,
                                            f32x4 (&acc)[4][8]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rb = w >> 1, ch = w & 1;
  const int lr = lane & 15, lk = 8 * (lane >> 4);
  u32x4 nxt[1024 / NT];
  stage_chunk_load<K, NT>(Wt, 0, nxt);
  stage_chunk_store<NT>(Wb, nxt);
  __syncthreads();
  // hs_at(64 rb + 16 i + lr, 32 c + lk) = row base + 8 ((g ^ lr) % 4) + 32 (c ^ (lr % 8 / 4))
  // + 4096 i (g = lane / 16): two per-lane bases (even and odd c) plus immediate offsets
  const int m = 64 * rb + lr, g = lane >> 4, s2 = (lr >> 2) & 1;
  const uint16_t* a_even = Hs + m * HID + 8 * ((g ^ lr) & 3) + 32 * s2;
  const uint16_t* a_odd = Hs + m * HID + 8 * ((g ^ lr) & 3) - 32 * s2;
  const uint16_t* bw = Wb + wb_at(128 * ch + lr, lk);  // + 512 j: rows 16 j further
#pragma unroll
  for (int c = 0; c < K / 32; ++c) {
    if (c + 1 < K / 32) stage_chunk_load<K, NT>(Wt, c + 1, nxt);
    bf16x8 a[4], b[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = lds_frag((c & 1 ? a_odd : a_even) + 16 * i * HID + 32 * c);
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = lds_frag(bw + 16 * j * 32);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i][j] = mfma(b[j], a[i], acc[i][j]);
    __syncthreads();  // every wave has read this chunk (and, after the last, the activations)
    if (c + 1 < K / 32) {
      stage_chunk_store<NT>(Wb, nxt);
      __syncthreads();
    }
  }
}

// accumulators := the bias of the outputs they hold (lane: outputs 16 j + 4 (lane / 16) + 0..3)
__device__ __forceinline__ void infer_bias(const float* __restrict__ bias, f32x4 (&acc)[4][8]) {
  const int lane = threadIdx.x & 63, ch = (threadIdx.x >> 6) & 1;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const f32x4 b = *(const f32x4*)(bias + 128 * ch + 16 * j + 4 * (lane >> 4));
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i][j] = b;
  }
}

// H[m][n..n+3] = bf16(GELU(acc)), one 8-byte store per accumulator
__device__ __forceinline__ void infer_store(const f32x4 (&acc)[4][8], uint16_t* Hs) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rb = w >> 1, ch = w & 1;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = 64 * rb + 16 * i + (lane & 15);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int n = 128 * ch + 16 * j + 4 * (lane >> 4);
      uint2 v;
      v.x = bf16_pair(gelu_f(acc[i][j][0]), gelu_f(acc[i][j][1]));
      v.y = bf16_pair(gelu_f(acc[i][j][2]), gelu_f(acc[i][j][3]));
      *(uint2*)&Hs[hs_at(m, n)] = v;
    }
  }
}

template <int RB>
__global__ void __launch_bounds__(128 * RB, 2) k_mlp_infer(const uint16_t* __restrict__ X, int64_t n,
                                                        const float* __restrict__ params,
                                                        const uint16_t* __restrict__ wbf,
                                                        float* __restrict__ prob,
                                                        float* __restrict__ logit) {
  constexpr int ROWS = 64 * RB, NT = 128 * RB;
  __shared__ __attribute__((aligned(16))) uint16_t Hs[ROWS * HID];
  __shared__ __attribute__((aligned(16))) uint16_t Wb[256 * 32];
  float* zp = (float*)Wb;  // [2][ROWS] after the last layer (at most 2 KB of the chunk's 16 KB)
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int rb = w >> 1, ch = w & 1;
  const int64_t m0 = (int64_t)blockIdx.x * ROWS;
  for (int e = t; e < ROWS * IN / 8; e += NT) {
    const int r = e / (IN / 8), c8 = (e % (IN / 8)) * 8;
    u32x4 v = {0u, 0u, 0u, 0u};
    if (m0 + r < n) v = *(const u32x4*)(X + (m0 + r) * IN + c8);
    *(u32x4*)&Hs[hs_at(r, c8)] = v;
  }
  f32x4 acc[4][8];
  infer_bias(params + OFF_B1, acc);
  infer_layer<IN, NT>(Hs, Wb, wbf + BOFF_W1, acc);  // its first barrier also orders the features
  infer_store(acc, Hs);
  __syncthreads();
  infer_bias(params + OFF_B2, acc);
  infer_layer<HID, NT>(Hs, Wb, wbf + BOFF_W2, acc);
  infer_store(acc, Hs);
  __syncthreads();
  infer_bias(params + OFF_B3, acc);
  infer_layer<HID, NT>(Hs, Wb, wbf + BOFF_W3, acc);
  // head: z = bf16(GELU(a3)) . w4 + b4
  const float* w4 = params + OFF_W4;
  float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const f32x4 wn = *(const f32x4*)(w4 + 128 * ch + 16 * j + 4 * (lane >> 4));
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[i] += bf2f(f2bf(gelu_f(acc[i][j][r]))) * wn[r];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float v = part[i];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (lane < 16) zp[ch * ROWS + 64 * rb + 16 * i + lane] = v;
  }
  __syncthreads();
  for (int r = t; r < ROWS; r += NT) {
    if (m0 + r >= n) continue;
    const float z = zp[r] + zp[ROWS + r] + params[OFF_B4];
    if (logit) logit[m0 + r] = z;
    if (prob) prob[m0 + r] = 1.0f / (1.0f + __expf(-z));
  }
}

// ---- host side ---------------------------------------------------------------------------------
static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr int64_t kSplitK = 512;  // batch rows per weight-gradient partial
constexpr int kInferTileDefault = 128;  // rows per workgroup of the fused inference (two per CU)

struct TrainWs {
  float *a1, *a2, *a3;
  uint16_t *h1, *h2, *h3, *d0, *d1;
  float *part, *colpart, *hpart_w4, *hpart_s, *scal;
};

static size_t train_ws(int64_t B, TrainWs* p, char* base) {
  const int64_t nsplit = (B + kSplitK - 1) / kSplitK;
  const int64_t ncol = (B + 255) / 256, nhead = (B + 63) / 64;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* r = base ? base + off : nullptr;
    off += a256(bytes);
    return r;
  };
  const size_t f = (size_t)B * HID * 4, h = (size_t)B * HID * 2;
  char* a1 = take(f);
  char* a2 = take(f);
  char* a3 = take(f);
  char* h1 = take(h);
  char* h2 = take(h);
  char* h3 = take(h);
  char* d0 = take(h);
  char* d1 = take(h);
  char* part = take((size_t)nsplit * HID * HID * 4);
  char* colpart = take((size_t)ncol * HID * 4);
  char* hw4 = take((size_t)nhead * HID * 4);
  char* hs = take((size_t)nhead * 2 * 4);
  char* scal = take(2 * 4);
  if (p) {
    p->a1 = (float*)a1;
    p->a2 = (float*)a2;
    p->a3 = (float*)a3;
    p->h1 = (uint16_t*)h1;
    p->h2 = (uint16_t*)h2;
    p->h3 = (uint16_t*)h3;
    p->d0 = (uint16_t*)d0;
    p->d1 = (uint16_t*)d1;
    p->part = (float*)part;
    p->colpart = (float*)colpart;
    p->hpart_w4 = (float*)hw4;
    p->hpart_s = (float*)hs;
    p->scal = (float*)scal;
  }
  return off;
}

static GemmArg gemm_base() {
  GemmArg g;
  g.A = nullptr;
  g.sam = g.sak = 0;
  g.aidx = nullptr;
  g.B = nullptr;
  g.sbk = g.sbn = 0;
  g.bidx = nullptr;
  g.M = g.N = g.K = g.kchunk = 0;
  g.mode = kModeFwd;
  g.bias = nullptr;
  g.pre = nullptr;
  g.out = nullptr;
  g.pre_in = nullptr;
  g.part = nullptr;
  return g;
}

// forward layer: out = GELU(in . Wt^T + b), in [B x K] rows (optionally gathered)
static void fwd_layer(const uint16_t* in, int64_t K, const int32_t* idx, const uint16_t* Wt,
                      const float* bias, int64_t B, float* pre, uint16_t* out, hipStream_t st) {
  GemmArg g = gemm_base();
  g.A = in;
  g.sam = K;
  g.sak = 1;
  g.aidx = idx;
  g.B = Wt;
  g.sbk = 1;
  g.sbn = K;
  g.M = B;
  g.N = HID;
  g.K = K;
  g.kchunk = K;
  g.mode = kModeFwd;
  g.bias = bias;
  g.pre = pre;
  g.out = out;
  hipLaunchKernelGGL((k_gemm<true, true>), dim3(grid_for(B, 64), HID / 64, 1), dim3(256), 0, st,
                     g);
}

// input gradient: out = (dA . Wt) * GELU'(pre_in), dA [B x 256], Wt [256 x Kin]
static void dx_layer(const uint16_t* dA, const uint16_t* Wt, int64_t Kin, const float* pre_in,
                     int64_t B, uint16_t* out, hipStream_t st) {
  GemmArg g = gemm_base();
  g.A = dA;
  g.sam = HID;
  g.sak = 1;
  g.B = Wt;
  g.sbk = Kin;
  g.sbn = 1;
  g.M = B;
  g.N = Kin;
  g.K = HID;
  g.kchunk = HID;
  g.mode = kModeDx;
  g.pre_in = pre_in;
  g.out = out;
  hipLaunchKernelGGL((k_gemm<true, false>), dim3(grid_for(B, 64), (unsigned)(Kin / 64), 1),
                     dim3(256), 0, st, g);
}

// weight gradient dWt[n][k] = sum_b dA[b][n] Hin[row(b)][k] into grads (split-K, fixed order)
static void dw_layer(const uint16_t* dA, const uint16_t* Hin, int64_t Kin, const int32_t* hidx,
                     int64_t B, float* part, float* grad, hipStream_t st) {
  const int64_t nsplit = (B + kSplitK - 1) / kSplitK;
  GemmArg g = gemm_base();
  g.A = dA;
  g.sam = 1;
  g.sak = HID;
  g.B = Hin;
  g.sbk = Kin;
  g.sbn = 1;
  g.bidx = hidx;
  g.M = HID;
  g.N = Kin;
  g.K = B;
  g.kchunk = kSplitK;
  g.mode = kModeDw;
  g.part = part;
  hipLaunchKernelGGL((k_gemm<false, false>), dim3(HID / 64, (unsigned)(Kin / 64), (unsigned)nsplit),
                     dim3(256), 0, st, g);
  hipLaunchKernelGGL(k_sum_parts, dim3(grid_for(HID * Kin, 256)), dim3(256), 0, st,
                     (const float*)part, (int)nsplit, (int64_t)HID * Kin, grad);
}

static void bias_grad(const uint16_t* dA, int64_t B, float* colpart, float* grad,
                      hipStream_t st) {
  const int64_t nchunk = (B + 255) / 256;
  hipLaunchKernelGGL(k_colsum, dim3((unsigned)nchunk), dim3(256), 0, st, dA, B, colpart);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (const float*)colpart, (int)nchunk,
                     (int64_t)HID, grad);
}

}  // namespace mlp
}  // namespace paces

using namespace paces;
using namespace paces::mlp;

extern "C" {

int paces_mlp_weights_bf16(const float* params, uint16_t* wbf16, void* stream) {
  if (!params || !wbf16) return PACES_E_INVALID;
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(NPARAM, 256)), dim3(256), 0, (hipStream_t)stream,
                     params, wbf16);
  return launch_status();
}

size_t paces_mlp_train_workspace_bytes(int64_t B) {
  if (B < 1) return 0;
  return train_ws(B, nullptr, nullptr);
}

int paces_mlp_grad(const uint16_t* X, const int32_t* idx, const uint8_t* labels, int64_t B,
                   float pos_weight, const float* params, const uint16_t* wbf16, float* grads,
                   float* loss, void* ws, size_t ws_bytes, void* stream) {
  if (!X || !labels || B < 1 || !params || !wbf16 || !grads || !ws) return PACES_E_INVALID;
  if (B > (1LL << 24)) return PACES_E_UNSUPPORTED;
  if (ws_bytes < train_ws(B, nullptr, nullptr)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  TrainWs p;
  train_ws(B, &p, (char*)ws);
  // forward
  fwd_layer(X, IN, idx, wbf16 + BOFF_W1, params + OFF_B1, B, p.a1, p.h1, st);
  fwd_layer(p.h1, HID, nullptr, wbf16 + BOFF_W2, params + OFF_B2, B, p.a2, p.h2, st);
  fwd_layer(p.h2, HID, nullptr, wbf16 + BOFF_W3, params + OFF_B3, B, p.a3, p.h3, st);
  // head: da3 -> d0
  const int64_t nhead = (B + 63) / 64;
  hipLaunchKernelGGL(k_head_train, dim3((unsigned)nhead), dim3(256), 0, st, (const uint16_t*)p.h3,
                     (const float*)p.a3, params, labels, idx, B, pos_weight, p.d0, p.hpart_w4,
                     p.hpart_s);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (const float*)p.hpart_w4, (int)nhead,
                     (int64_t)HID, grads + OFF_W4);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (const float*)p.hpart_s, (int)nhead,
                     (int64_t)2, p.scal);
  // layer 3
  dw_layer(p.d0, p.h2, HID, nullptr, B, p.part, grads + OFF_W3, st);
  bias_grad(p.d0, B, p.colpart, grads + OFF_B3, st);
  dx_layer(p.d0, wbf16 + BOFF_W3, HID, p.a2, B, p.d1, st);  // da2 -> d1
  // layer 2
  dw_layer(p.d1, p.h1, HID, nullptr, B, p.part, grads + OFF_W2, st);
  bias_grad(p.d1, B, p.colpart, grads + OFF_B2, st);
  dx_layer(p.d1, wbf16 + BOFF_W2, HID, p.a1, B, p.d0, st);  // da1 -> d0
  // layer 1 (the inputs are gathered rows of X)
  dw_layer(p.d0, X, IN, idx, B, p.part, grads + OFF_W1, st);
  bias_grad(p.d0, B, p.colpart, grads + OFF_B1, st);
  // db4 and the loss
  if (hipMemcpyAsync(grads + OFF_B4, p.scal, 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return PACES_E_HIP;
  if (loss && hipMemcpyAsync(loss, p.scal + 1, 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return PACES_E_HIP;
  return launch_status();
}

int paces_mlp_adam(float* params, const float* grads, float* m, float* v, int32_t step, float lr,
                   float beta1, float beta2, float eps, uint16_t* wbf16, void* stream) {
  if (!params || !grads || !m || !v || !wbf16 || step < 1) return PACES_E_INVALID;
  const float c1 = (float)(1.0 / (1.0 - pow((double)beta1, (double)step)));
  const float c2 = (float)(1.0 / (1.0 - pow((double)beta2, (double)step)));
  hipLaunchKernelGGL(k_adam, dim3(grid_for(NPARAM, 256)), dim3(256), 0, (hipStream_t)stream,
                     params, grads, m, v, lr, beta1, beta2, eps, c1, c2, wbf16);
  return launch_status();
}

int paces_mlp_infer(const uint16_t* X, int64_t n, const float* params, const uint16_t* wbf16,
                    float* prob, float* logit, void* stream) {
  return paces_mlp_infer_tiled(X, n, params, wbf16, prob, logit, 0, stream);
}

int paces_mlp_infer_tiled(const uint16_t* X, int64_t n, const float* params,
                          const uint16_t* wbf16, float* prob, float* logit, int32_t tile_rows,
                          void* stream) {
  if (n < 0 || !params || !wbf16 || (!prob && !logit)) return PACES_E_INVALID;
  if (tile_rows == 0) tile_rows = kInferTileDefault;
  if (tile_rows != 64 && tile_rows != 128 && tile_rows != 256) return PACES_E_INVALID;
  if (n == 0) return PACES_OK;
  if (!X) return PACES_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  if (tile_rows == 256)
    hipLaunchKernelGGL(k_mlp_infer<4>, dim3(grid_for(n, 256)), dim3(512), 0, st, X, n, params,
                       wbf16, prob, logit);
  else if (tile_rows == 128)
    hipLaunchKernelGGL(k_mlp_infer<2>, dim3(grid_for(n, 128)), dim3(256), 0, st, X, n, params,
                       wbf16, prob, logit);
  else
    hipLaunchKernelGGL(k_mlp_infer<1>, dim3(grid_for(n, 64)), dim3(128), 0, st, X, n, params,
                       wbf16, prob, logit);
  return launch_status();
}


}  // extern "C"
