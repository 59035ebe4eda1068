// Holstein-model kernels: candidate images (C1/C2/A8), per-term CSR (D1/D2), diagonal (A6) and
// observables (H1). References: paces/models/holstein.py:151-322,
// paces/models/_kernels/_holstein_kernels.py:96-255, paces/core/hamiltonian.py:363-547.
#include "common.hpp"

namespace paces {

int scan_u8(const uint8_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
int scan_i32(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws_bytes(int64_t n);

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

// ------------------------------------------------------------------------------------------------
// Operation tables built on the host from paces_holstein_t
// ------------------------------------------------------------------------------------------------

constexpr int kMaxOps = 32;

// An image op maps a row to one neighbour. Hop ops shift the exciton position inside one lattice
// dimension: with mod = pos % span, target = pos + dir*oe when the move stays inside the span
// (wrap == 0), or pos + dir*oe - dir*span when it crosses the periodic boundary (wrap == 1).
// Vib ops add dir to the phonon count at the exciton site.
struct OpTable {
  int n;
  int has_vib;  // some op is a vib op (the phonon fields exist)
  // 32-bit fields: a uniform dynamic index then reads them with scalar loads
  int32_t kind[kMaxOps];
  int32_t dir[kMaxOps];
  int32_t wrap[kMaxOps];
  int32_t term[kMaxOps];
  int32_t dim[kMaxOps];
  int32_t order[kMaxOps];
  int32_t oe[kMaxOps];
  int32_t span[kMaxOps];
  double value[kMaxOps];
};

// Model parameters beyond the op table: where the exciton position lives and, for batched
// trajectories, the per-trajectory lattice length and term values.
struct ModelArg {
  int pf;  // field index of the exciton position (0, or 1 when field 0 is the trajectory id)
  int batched;
  int nterms;
  const int32_t* traj_L;
  const double* traj_val;
};

// width of the trajectory-id field: ceil(log2(ntraj))
static int tbits(int ntraj) {
  int b = 0;
  while ((1 << b) < ntraj) ++b;
  return b;
}

// trajectory id of a batched row from its first word: the top trajbits bits (a one-trajectory
// batch has a zero-width field, where a shift by 64 would be undefined)
__device__ __forceinline__ int traj_of(u64 w0, int trajbits) {
  return trajbits ? (int)(w0 >> (64 - trajbits)) : 0;
}

static ModelArg marg(const paces_holstein_t* h) {
  ModelArg M;
  M.batched = h->batched ? 1 : 0;
  M.pf = M.batched;
  M.nterms = h->nterms;
  M.traj_L = h->d_traj_L;
  M.traj_val = h->d_traj_value;
  return M;
}

// Image runs (enlarge_basis_set / generate_mel_*(raw_map_to=True)): OBC keeps the in-span moves
// (holstein.py:185-199); periodic adds the wrapped moves as separate runs so every run stays
// sorted.
static int build_image_ops(const paces_holstein_t* h, OpTable* T) {
  T->n = 0;
  T->has_vib = 0;
  for (int t = 0; t < h->nterms; ++t) {
    const paces_term_t& tm = h->terms[t];
    if (tm.kind == PACES_TERM_HOP) {
      long long Ld = 1;
      for (int d = 0; d < h->geo_dim; ++d) {
        long long oe = (long long)tm.order * Ld, span = Ld * h->lattice_l;
        for (int wrap = 0; wrap <= (h->periodic ? 1 : 0); ++wrap) {
          for (int dir = +1; dir >= -1; dir -= 2) {
            if (T->n >= kMaxOps) return PACES_E_UNSUPPORTED;
            int k = T->n++;
            T->kind[k] = PACES_TERM_HOP;
            T->dir[k] = (int8_t)dir;
            T->wrap[k] = (int8_t)wrap;
            T->term[k] = (int8_t)t;
            T->dim[k] = (int8_t)d;
            T->order[k] = (int8_t)tm.order;
            T->oe[k] = (int32_t)oe;
            T->span[k] = (int32_t)span;
            T->value[k] = tm.value;
          }
        }
        Ld *= h->lattice_l;
      }
    } else if (tm.kind == PACES_TERM_VIB) {
      for (int dir = +1; dir >= -1; dir -= 2) {
        if (T->n >= kMaxOps) return PACES_E_UNSUPPORTED;
        int k = T->n++;
        T->kind[k] = PACES_TERM_VIB;
        T->has_vib = 1;
        T->dir[k] = (int8_t)dir;
        T->wrap[k] = 0;
        T->term[k] = (int8_t)t;
        T->dim[k] = 0;
        T->order[k] = 0;
        T->oe[k] = 0;
        T->span[k] = 0;
        T->value[k] = tm.value;
      }
    } else {
      return PACES_E_INVALID;
    }
  }
  return PACES_OK;
}

// Matrix-element neighbours (one per direction; the periodic wrap is folded in), grouped by term.
static int build_nb_ops(const paces_holstein_t* h, OpTable* T) {
  T->n = 0;
  T->has_vib = 0;
  for (int t = 0; t < h->nterms; ++t) {
    const paces_term_t& tm = h->terms[t];
    if (tm.kind == PACES_TERM_HOP) {
      long long Ld = 1;
      for (int d = 0; d < h->geo_dim; ++d) {
        for (int dir = +1; dir >= -1; dir -= 2) {
          if (T->n >= kMaxOps) return PACES_E_UNSUPPORTED;
          int k = T->n++;
          T->kind[k] = PACES_TERM_HOP;
          T->dir[k] = (int8_t)dir;
          T->wrap[k] = (int8_t)(h->periodic ? 2 : 0);  // 2: either side, wrap when needed
          T->term[k] = (int8_t)t;
          T->dim[k] = (int8_t)d;
          T->order[k] = (int8_t)tm.order;
          T->oe[k] = (int32_t)(tm.order * Ld);
          T->span[k] = (int32_t)(Ld * h->lattice_l);
          T->value[k] = tm.value;
        }
        Ld *= h->lattice_l;
      }
    } else {
      for (int dir = +1; dir >= -1; dir -= 2) {
        if (T->n >= kMaxOps) return PACES_E_UNSUPPORTED;
        int k = T->n++;
        T->kind[k] = PACES_TERM_VIB;
        T->has_vib = 1;
        T->dir[k] = (int8_t)dir;
        T->wrap[k] = 0;
        T->term[k] = (int8_t)t;
        T->dim[k] = 0;
        T->order[k] = 0;
        T->oe[k] = 0;
        T->span[k] = 0;
        T->value[k] = tm.value;
      }
    }
  }
  return PACES_OK;
}

struct LayoutArg {
  const int32_t* off;
  const int32_t* width;
  const int32_t* maxdim;
  const u64* sum_mask;
  int sum_bits;
};

// What every op of a row reads: the exciton position, the row's lattice length (batched) and the
// phonon field at the exciton. A kernel that applies several ops to one row forms it once.
struct RowCtx {
  int pos;
  int Lt;
  int voff, vb, vmax, n;
};

// Offset and width of the exciton-position and trajectory-id fields, read once per kernel as
// wave-uniform values. (The compiler cannot keep table reads in scalar registers across the
// kernel's own stores, which might alias the table, so kernels that store read them up front.)
struct RowFields {
  int po, pb, to, tb;
};

__device__ __forceinline__ RowFields row_fields(const LayoutArg& L, const ModelArg& M) {
  RowFields F;
  F.po = __builtin_amdgcn_readfirstlane(L.off[M.pf]);
  F.pb = __builtin_amdgcn_readfirstlane(L.width[M.pf]);
  F.to = M.batched ? __builtin_amdgcn_readfirstlane(L.off[0]) : 0;
  F.tb = M.batched ? __builtin_amdgcn_readfirstlane(L.width[0]) : 0;
  return F;
}

template <int W>
__device__ __forceinline__ RowCtx row_ctx(const OpTable& T, const LayoutArg& L, const ModelArg& M,
                                          const RowFields& F, const Row<W>& r) {
  RowCtx c;
  c.pos = (int)get_field<W>(r, F.po, F.pb);
  c.Lt = M.batched ? M.traj_L[get_field<W>(r, F.to, F.tb)] : 0;
  c.voff = c.vb = c.vmax = c.n = 0;
  if (T.has_vib) {  // the phonon fields exist only for models with a vib term
    const int f = M.pf + 1 + c.pos;
    c.voff = L.off[f];
    c.vb = L.width[f];
    c.vmax = L.maxdim[f];
    c.n = (int)get_field<W>(r, c.voff, c.vb);
  }
  return c;
}

// row_ctx for NQ rows at once, in phases (field extraction, then every table gather, then the
// phonon counts), so that the gathers of all NQ rows are in flight together
template <int W, int NQ>
__device__ __forceinline__ void row_ctx_n(const OpTable& T, const LayoutArg& L, const ModelArg& M,
                                          const RowFields& F, const Row<W> (&r)[NQ],
                                          RowCtx (&c)[NQ]) {
  int tr[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    c[q].pos = (int)get_field<W>(r[q], F.po, F.pb);
    tr[q] = (int)get_field<W>(r[q], F.to, F.tb);
    c[q].Lt = 0;
    c[q].voff = c[q].vb = c[q].vmax = c[q].n = 0;
  }
  if (M.batched) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) c[q].Lt = M.traj_L[tr[q]];
  }
  if (T.has_vib) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int f = M.pf + 1 + c[q].pos;
      c[q].voff = L.off[f];
      c[q].vb = L.width[f];
      c[q].vmax = L.maxdim[f];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) c[q].n = (int)get_field<W>(r[q], c[q].voff, c[q].vb);
  }
}

// The fields of op k, read once per op as wave-uniform values (scalar registers), so a loop over
// rows does not reload them.
struct OpView {
  int kind, dir, wrap, dim, order, oe, span;
  int po, pb;  // offset and width of the exciton-position field
};

__device__ __forceinline__ OpView op_view(const OpTable& T, int k, const RowFields& F) {
  OpView o;
  o.po = F.po;
  o.pb = F.pb;
  o.kind = __builtin_amdgcn_readfirstlane((int)T.kind[k]);
  o.dir = __builtin_amdgcn_readfirstlane((int)T.dir[k]);
  o.wrap = __builtin_amdgcn_readfirstlane((int)T.wrap[k]);
  o.dim = __builtin_amdgcn_readfirstlane((int)T.dim[k]);
  o.order = __builtin_amdgcn_readfirstlane((int)T.order[k]);
  o.oe = __builtin_amdgcn_readfirstlane(T.oe[k]);
  o.span = __builtin_amdgcn_readfirstlane(T.span[k]);
  return o;
}

// whether op o has an image of the row whose context is c, and (for a hop) the new position
__device__ __forceinline__ bool op_target(const OpView& o, const ModelArg& M, const RowCtx& c,
                                          int& npos) {
  npos = c.pos;
  if (o.kind == PACES_TERM_HOP) {
    int oe, span;
    if (M.batched) {
      int Ld = 1;
      for (int d = 0; d < o.dim; ++d) Ld *= c.Lt;
      oe = o.order * Ld;
      span = Ld * c.Lt;
      if (oe >= span) return false;  // hop order longer than this trajectory's lattice
    } else {
      oe = o.oe;
      span = o.span;
    }
    int mod = c.pos;
    if (mod >= span) mod = c.pos % span;  // the outermost dimension needs no division
    const bool inside = (o.dir > 0) ? (mod < span - oe) : (mod >= oe);
    const int step = o.dir * oe;
    const int wrapped = step - o.dir * span;
    if (o.wrap == 0) {
      npos = c.pos + step;
      return inside;
    }
    if (o.wrap == 1) {
      npos = c.pos + wrapped;
      return !inside;
    }
    npos = c.pos + (inside ? step : wrapped);
    return true;
  }
  return o.dir > 0 ? c.n + 1 < c.vmax : c.n > 0;
}

__device__ __forceinline__ bool op_exists(const OpView& o, const ModelArg& M, const RowCtx& c) {
  int npos;
  return op_target(o, M, c, npos);
}

// Apply op o to row r, whose context is c. Returns false when the image does not exist (boundary
// or ceiling). `nup` receives the upper phonon count of a vib pair (for the matrix element).
template <int W>
__device__ __forceinline__ bool apply_op_ctx(const OpView& o, const LayoutArg& L,
                                             const ModelArg& M, const RowCtx& c, Row<W>& r,
                                             int& nup) {
  int npos;
  if (!op_target(o, M, c, npos)) return false;
  if (o.kind == PACES_TERM_HOP) {
    add_field<W>(r, o.po, o.pb, (int64_t)(npos - c.pos));
    nup = 0;
  } else if (o.dir > 0) {
    add_field<W>(r, c.voff, c.vb, 1);
    nup = c.n + 1;
  } else {
    add_field<W>(r, c.voff, c.vb, -1);
    nup = c.n;
  }
  return true;
}

template <int W>
__device__ __forceinline__ bool apply_op(const OpTable& T, int k, const LayoutArg& L,
                                         const ModelArg& M, Row<W>& r, int& nup) {
  const RowFields F = row_fields(L, M);
  return apply_op_ctx<W>(op_view(T, k, F), L, M, row_ctx<W>(T, L, M, F, r), r, nup);
}

// matrix-element value of op k for a row of trajectory traj
__device__ __forceinline__ double op_value(const OpTable& T, int k, const ModelArg& M, int traj,
                                           const double* __restrict__ sqrt_tab, int nup) {
  const double v = M.batched ? M.traj_val[traj * M.nterms + T.term[k]] : T.value[k];
  return (T.kind[k] == PACES_TERM_HOP) ? v : sqrt_tab[nup] * v;
}

// ------------------------------------------------------------------------------------------------
// Image runs: tile counts -> scan -> tile emit
// ------------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_zero_i32(int32_t* __restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Image runs in wave tiles of kImgTile basis rows (lane x of the wave takes rows
// t0 + q 64 + x, q < kImgSteps, so the tile's row order is step-major). The count pass counts each
// tile's valid images per op; a scan of those counts in op-major order gives each (op, tile) its
// first output index, and the run offsets. The emit pass applies the ops again and writes each
// tile's images of op k in row order from there, with in-wave prefixes from ballots: no barrier
// in either pass. The per-row reads every op shares (exciton position, lattice length, phonon
// field) are formed once per row. One row per lane measured fastest (more waves in flight to
// hide the latencies; eight or two rows per lane were slower).
constexpr int kImgSteps = 1;
constexpr int kImgTile = 64 * kImgSteps;
constexpr int kImgWaves = kBlock / 64;
// rows of up to kImgHoldWords words are loaded for all steps up front; wider ones step by step
constexpr int kImgHoldWords = 4;

static int64_t img_tiles(int64_t n) { return (n + kImgTile - 1) / kImgTile; }

static size_t images_ws(int64_t n, int nops) {
  const int64_t m = (int64_t)nops * (img_tiles(n) > 0 ? img_tiles(n) : 1);
  return a256((size_t)m * 4) + a256((size_t)(m + 1) * 4) + a256(scan_ws_bytes(m));
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_img_tcount(const u64* __restrict__ basis, int64_t n,
                                                       const uint8_t* __restrict__ skip,
                                                       int64_t ntl, OpTable T, LayoutArg L,
                                                       ModelArg M, int32_t* __restrict__ tcnt) {
  const int64_t wt = (int64_t)blockIdx.x * kImgWaves + (threadIdx.x >> 6);
  if (wt >= ntl) return;  // whole waves
  const int lane = threadIdx.x & 63;
  const RowFields F = row_fields(L, M);
  RowCtx cx[kImgSteps];
  bool vv[kImgSteps];
  int so[kImgSteps];  // the one op a row skips (skip code 2 + k), else -1
  if (W <= kImgHoldWords) {  // every row's load in flight before any is used
    Row<W> rr[kImgSteps];
#pragma unroll
    for (int q = 0; q < kImgSteps; ++q) {
      const int64_t i = wt * kImgTile + q * 64 + lane;
      const int sc = (skip && i < n) ? skip[i] : 0;
      vv[q] = i < n && sc != 1;
      so[q] = sc >= 2 ? sc - 2 : -1;
      rr[q] = load_row<W>(basis, i < n ? i : 0);
    }
    row_ctx_n<W, kImgSteps>(T, L, M, F, rr, cx);
  } else {
#pragma unroll
    for (int q = 0; q < kImgSteps; ++q) {
      const int64_t i = wt * kImgTile + q * 64 + lane;
      const int sc = (skip && i < n) ? skip[i] : 0;
      vv[q] = i < n && sc != 1;
      so[q] = sc >= 2 ? sc - 2 : -1;
      cx[q] = row_ctx<W>(T, L, M, F, load_row<W>(basis, i < n ? i : 0));
    }
  }
  for (int k = 0; k < T.n; ++k) {
    const OpView o = op_view(T, k, F);
    int c = 0;
#pragma unroll
    for (int q = 0; q < kImgSteps; ++q)
      c += __popcll(__ballot(op_exists(o, M, cx[q]) & vv[q] & (so[q] != k)));
    if (lane == 0) tcnt[(int64_t)k * ntl + wt] = c;
  }
}

__global__ void k_img_run_offsets(const int32_t* __restrict__ TS, int64_t ntl, int nops,
                                  int32_t* __restrict__ run_offsets) {
  const int k = threadIdx.x;
  if (k <= nops) run_offsets[k] = TS[(int64_t)k * ntl];
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_img_temit(const u64* __restrict__ basis, int64_t n,
                                                      const uint8_t* __restrict__ skip,
                                                      int64_t ntl, OpTable T, LayoutArg L,
                                                      ModelArg M, const int32_t* __restrict__ TS,
                                                      u64* __restrict__ out,
                                                      int32_t* __restrict__ out_src,
                                                      uint8_t* __restrict__ out_nup) {
  constexpr bool kHold = W <= kImgHoldWords;
  constexpr int kH = kHold ? kImgSteps : 1;
  // per wave: the next output index of each op, and the staging of one step's images (so the
  // global stores are lane-contiguous words)
  __shared__ int nxt[kImgWaves][kMaxOps];
  __shared__ u64 stage[kImgWaves][64 * W];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t wt = (int64_t)blockIdx.x * kImgWaves + wave;
  if (wt >= ntl) return;  // whole waves
  const RowFields F = row_fields(L, M);
  if (lane < T.n) nxt[wave][lane] = TS[(int64_t)lane * ntl + wt];
  Row<W> rh[kH];
  RowCtx ch[kH];
  if (kHold) {  // every row and its context before the first store
#pragma unroll
    for (int q = 0; q < kImgSteps; ++q) {
      const int64_t i = wt * kImgTile + q * 64 + lane;
      rh[kHold ? q : 0] = load_row<W>(basis, i < n ? i : 0);
    }
    row_ctx_n<W, kH>(T, L, M, F, rh, ch);
  }
  u64* const sw = stage[wave];
#pragma unroll
  for (int q = 0; q < kImgSteps; ++q) {
    const int64_t i = wt * kImgTile + q * 64 + lane;
    const int sc = (skip && i < n) ? skip[i] : 0;
    const bool v = i < n && sc != 1;
    const int so = sc >= 2 ? sc - 2 : -1;  // the one op this row skips
    const Row<W> r0 = kHold ? rh[kHold ? q : 0] : load_row<W>(basis, i < n ? i : 0);
    const RowCtx cx = kHold ? ch[kHold ? q : 0] : row_ctx<W>(T, L, M, F, r0);
    for (int k = 0; k < T.n; ++k) {
      const OpView o = op_view(T, k, F);
      const bool ok = op_exists(o, M, cx) & v & (so != k);
      const u64 bl = __ballot(ok);
      if (bl == 0) continue;  // wave-uniform
      const int cnt = __popcll(bl);
      const int pre = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bl >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((unsigned)bl, 0u));
      const int o0 = nxt[wave][k];
      if (ok) {
        Row<W> r = r0;
        int nup = 0;
        apply_op_ctx<W>(o, L, M, cx, r, nup);
#pragma unroll
        for (int w = 0; w < W; ++w) sw[pre * W + w] = r.w[w];
        if (out_src) out_src[o0 + pre] = (int32_t)i;
        if (out_nup) out_nup[o0 + pre] = (uint8_t)nup;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      u64* const dst = out + (int64_t)o0 * W;
      for (int j = lane; j < cnt * W; j += 64) dst[j] = sw[j];
      if (lane == 0) nxt[wave][k] = o0 + cnt;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
}

template <int W>
static int run_img_count(const uint64_t* basis, int64_t n, const uint8_t* skip, const LayoutArg& L,
                         const ModelArg& M, const OpTable& T, int32_t* run_offsets, void* ws,
                         size_t ws_bytes, hipStream_t st) {
  const int64_t ntl = img_tiles(n) > 0 ? img_tiles(n) : 1;
  const int64_t m = (int64_t)T.n * ntl;
  char* p = (char*)ws;
  int32_t* tcnt = (int32_t*)p;
  p += a256((size_t)m * 4);
  int32_t* TS = (int32_t*)p;
  p += a256((size_t)(m + 1) * 4);
  if (n > 0) {
    hipLaunchKernelGGL(k_img_tcount<W>, dim3(grid_for(ntl, kImgWaves)), dim3(kBlock), 0, st,
                       (const u64*)basis, n, skip, ntl, T, L, M, tcnt);
  } else {
    hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, st, tcnt, m);
  }
  int rc = launch_status();
  if (rc) return rc;
  rc = scan_i32(tcnt, m, TS, p, scan_ws_bytes(m), st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_img_run_offsets, dim3(1), dim3(64), 0, st, (const int32_t*)TS, ntl, T.n,
                     run_offsets);
  return launch_status();
}

template <int W>
static int run_img_emit(const uint64_t* basis, int64_t n, const uint8_t* skip, const LayoutArg& L,
                        const ModelArg& M, const OpTable& T, uint64_t* out, int32_t* out_src,
                        uint8_t* out_nup, void* ws, hipStream_t st) {
  if (n == 0) return PACES_OK;
  const int64_t ntl = img_tiles(n);
  const int64_t m = (int64_t)T.n * ntl;
  char* p = (char*)ws;
  p += a256((size_t)m * 4);
  const int32_t* TS = (const int32_t*)p;
  hipLaunchKernelGGL(k_img_temit<W>, dim3(grid_for(ntl, kImgWaves)), dim3(kBlock), 0, st,
                     (const u64*)basis, n, skip, ntl, T, L, M, TS, (u64*)out, out_src, out_nup);
  return launch_status();
}

// ------------------------------------------------------------------------------------------------
// Per-term CSR, built row by row: row i of term t holds every neighbour j of row i under t with
// j in the new basis and i or j in the pre-expansion basis. This is the entry set the reference's
// MelTriples + o_star masks + H.c. produce (hamiltonian.py:363-374, 495-547; holstein.py:201-206,
// 266-277), each unordered pair once per direction.
// ------------------------------------------------------------------------------------------------

template <int W>
__global__ void __launch_bounds__(kBlock) k_csr_count(const u64* __restrict__ rows, int64_t R,
                                                      const uint8_t* __restrict__ inB, OpTable T,
                                                      LayoutArg L, ModelArg M, int nterms,
                                                      int32_t* __restrict__ nb,
                                                      uint8_t* __restrict__ nbup,
                                                      int32_t* __restrict__ cnt) {
  __shared__ LookupSmem<W> sm;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool v = i < R;
  Row<W> r0;
#pragma unroll
  for (int k = 0; k < W; ++k) r0.w[k] = 0;
  if (v) r0 = load_row<W>(rows, i);
  const bool bi = v && inB[i] != 0;
  int c[PACES_MAX_TERMS] = {0, 0, 0, 0};
  for (int k = 0; k < T.n; ++k) {
    Row<W> r = r0;
    int nup = 0;
    const bool has = v && apply_op<W>(T, k, L, M, r, nup);
    // the images of consecutive rows under one op are clustered, so the lookup stays local
    const int64_t lb = block_lower_bound<W>(r, has, rows, R, sm);
    int32_t j = -1;
    if (has && lb < R && eq_row<W>(r, rows + lb * W) && (bi || inB[lb])) j = (int32_t)lb;
    if (v) {
      nb[(int64_t)k * R + i] = j;
      nbup[(int64_t)k * R + i] = (uint8_t)nup;
    }
    if (j >= 0) {
#pragma unroll
      for (int t = 0; t < PACES_MAX_TERMS; ++t) c[t] += (T.term[k] == t) ? 1 : 0;
    }
  }
  if (v)
    for (int t = 0; t < nterms; ++t) cnt[(int64_t)t * R + i] = c[t];
}


__global__ void k_csr_nnz(const int32_t* __restrict__ G, int64_t R, int nterms,
                          int32_t* __restrict__ nnz) {
  int t = threadIdx.x;
  if (t < nterms) nnz[t] = G[(int64_t)(t + 1) * R] - G[(int64_t)t * R];
}

struct CsrOut {
  int32_t* indptr[PACES_MAX_TERMS];
  int32_t* indices[PACES_MAX_TERMS];
  double* data[PACES_MAX_TERMS];
  float* data32[PACES_MAX_TERMS];  // when set, FP32 values (rounded from the FP64 value) instead
  uint8_t* code[PACES_MAX_TERMS];  // coded: the vib pair's upper phonon count (hopping: none)
  int coded;                       // values as codes, not numbers (data, data32 unused)
};

// entry value v of term t at position i, in the output's precision
__device__ __forceinline__ void put_val(const CsrOut& o, int t, int64_t i, double v) {
  if (o.data32[t])
    o.data32[t][i] = (float)v;
  else
    o.data[t][i] = v;
}

// an entry's value (v, or the code nup of a vib pair for a coded output) at position i of term t
__device__ __forceinline__ void put_entry(const CsrOut& o, int t, int64_t i, double v, int nup) {
  if (o.coded) {
    if (o.code[t]) o.code[t][i] = (uint8_t)nup;
  } else {
    put_val(o, t, i, v);
  }
}

// Emits each row's entries of every term in ascending column order by repeated minimum
// selection over the (at most 8) neighbour slots of the term; the slots were filled by
// k_csr_count with the neighbour index (or -1) and the upper phonon count of a vib pair.
__global__ void __launch_bounds__(kBlock) k_csr_fill(int64_t R, OpTable T, ModelArg M, int nterms,
                                                     const u64* __restrict__ rows, int W,
                                                     int trajbits,
                                                     const double* __restrict__ sqrt_tab,
                                                     const int32_t* __restrict__ nb,
                                                     const uint8_t* __restrict__ nbup,
                                                     const int32_t* __restrict__ G, CsrOut out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int traj = M.batched ? traj_of(rows[i * W], trajbits) : 0;
  for (int t = 0; t < nterms; ++t) {
    const int32_t base0 = G[(int64_t)t * R];
    const int32_t base = G[(int64_t)t * R + i] - base0;
    out.indptr[t][i] = base;
    if (i == R - 1) out.indptr[t][R] = G[(int64_t)(t + 1) * R] - base0;
    int32_t last = -1;
    for (int e = 0; e < 8; ++e) {
      int32_t best = 0x7fffffff;
      double bv = 0.0;
      for (int k = 0; k < T.n; ++k) {
        if (T.term[k] != t) continue;
        const int32_t j = nb[(int64_t)k * R + i];
        if (j > last && j < best) {
          best = j;
          bv = op_value(T, k, M, traj, sqrt_tab, nbup[(int64_t)k * R + i]);
        }
      }
      if (best == 0x7fffffff) break;
      out.indices[t][base + e] = best;
      put_val(out, t, base + e, bv);
      last = best;
    }
  }
}

__global__ void k_csr_empty(CsrOut out, int nterms) {
  int t = threadIdx.x;
  if (t < nterms) out.indptr[t][0] = 0;
}

static size_t csr_ws(int64_t R, int nops, int nterms) {
  return a256((size_t)nops * R * 4) + a256((size_t)nops * R) + a256((size_t)nterms * R * 4) +
         a256((size_t)(nterms * R + 1) * 4) + a256(scan_ws_bytes((int64_t)nterms * R));
}

template <int W>
static int run_csr_count(const uint64_t* rows, int64_t R, const uint8_t* inB, const LayoutArg& L,
                         const ModelArg& M, const OpTable& T, int nterms, int32_t* nnz, void* ws,
                         hipStream_t st) {
  char* p = (char*)ws;
  int32_t* nb = (int32_t*)p;
  p += a256((size_t)T.n * R * 4);
  uint8_t* nbup = (uint8_t*)p;
  p += a256((size_t)T.n * R);
  int32_t* cnt = (int32_t*)p;
  p += a256((size_t)nterms * R * 4);
  int32_t* G = (int32_t*)p;
  p += a256((size_t)(nterms * R + 1) * 4);
  if (R > 0) {
    hipLaunchKernelGGL(k_csr_count<W>, dim3(grid_for(R, kBlock)), dim3(kBlock), 0, st,
                       (const u64*)rows, R, inB, T, L, M, nterms, nb, nbup, cnt);
    int rc = launch_status();
    if (rc) return rc;
  }
  int rc = scan_i32(cnt, (int64_t)nterms * R, G, p, scan_ws_bytes((int64_t)nterms * R), st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_csr_nnz, dim3(1), dim3(64), 0, st, (const int32_t*)G, R, nterms, nnz);
  return launch_status();
}

static int run_csr_fill(int64_t R, const OpTable& T, const ModelArg& M, int nterms,
                        const uint64_t* rows, int W, int trajbits, const double* sqrt_tab,
                        const CsrOut& out, void* ws, hipStream_t st) {
  if (R == 0) {
    hipLaunchKernelGGL(k_csr_empty, dim3(1), dim3(64), 0, st, out, nterms);
    return launch_status();
  }
  char* p = (char*)ws;
  const int32_t* nb = (const int32_t*)p;
  p += a256((size_t)T.n * R * 4);
  const uint8_t* nbup = (const uint8_t*)p;
  p += a256((size_t)T.n * R);
  p += a256((size_t)nterms * R * 4);
  const int32_t* G = (const int32_t*)p;
  hipLaunchKernelGGL(k_csr_fill, dim3(grid_for(R, kBlock)), dim3(kBlock), 0, st, R, T, M, nterms,
                     (const u64*)rows, W, trajbits, sqrt_tab, nb, nbup, G, out);
  return launch_status();
}

// ------------------------------------------------------------------------------------------------
// Per-term CSR from the union's candidate positions (no lookups). Every image candidate e of a
// basis row b under op k is the pair {b, e}; it is an entry (b, e) and its H.c. (e, b) of term
// term[k], except for a minus op whose image is itself a basis row: that pair is emitted by the
// plus op of the image (the o_star rule, hamiltonian.py:363-374; holstein.py:201-206, 266-268).
// Each row's entries come out in ascending column order, the canonical coo_matrix(...).tocsr()
// layout (time_evolution.py:369-371): through fixed slots where the column order of a term's ops
// is strict for every row (k_pairs_mark / k_pairs_place), else counted and filled through atomic
// cursors and then sorted per row (k_pairs_count / k_pairs_fill / k_pairs_sort).
// ------------------------------------------------------------------------------------------------

struct ImgRuns {
  int n;
  int64_t off[kMaxOps + 1];  // image index range of each run (host copy of the run offsets)
};

__device__ __forceinline__ int img_run_of(const ImgRuns& IR, int64_t e) {
  int k = 0;
  for (int q = 1; q < IR.n; ++q) k += (e >= IR.off[q]) ? 1 : 0;
  return k;
}

// A split image list (paces_prior_images_t): images e < A.n are the previous basis' images (the
// last expansion's), then come the images of the rows that expansion added. A prior image's
// position in the expanded basis is A.cpos[A.nprev + ea] and its source's is A.cpos[A.src[ea]]
// (ea its index in the prior arrays); both map into the new basis through cand_pos (the union
// with the expanded basis as run 0). Every prior image is a row of that basis, so the image of
// an op of negative direction gives no entry (the pair comes from the inverse op's image of it,
// as for any image inside the basis): A's index space covers the positive-direction runs only
// (A.IR, with the other runs empty), and A.base[k] is where op k's run starts in the prior
// arrays. A.n == 0: a plain list.
struct PriorImgs {
  int64_t n, nprev;
  const int32_t* cpos;
  const int32_t* src;
  const uint8_t* nup;
  ImgRuns IR;
  int64_t base[kMaxOps];
};

// image e's op k, the new-basis positions p of the image and q of its source, the source's
// expanded-basis index s (its trajectory id) and the image's upper phonon count nu (unused
// outputs cost no load)
__device__ __forceinline__ void image_ends(int64_t e, const PriorImgs& A, const ImgRuns& IR,
                                           int64_t nbasis, const int32_t* __restrict__ cand_pos,
                                           const int32_t* __restrict__ src,
                                           const uint8_t* __restrict__ nup, int& k, int32_t& p,
                                           int32_t& q, int32_t& s, int& nu) {
  if (e < A.n) {
    k = img_run_of(A.IR, e);
    const int64_t ea = e - A.IR.off[k] + A.base[k];
    const int32_t ca = A.cpos[A.nprev + ea];
    s = A.cpos[A.src[ea]];
    nu = A.nup[ea];
    // an image pruned from the expanded basis has position -1 (and so has no pair)
    p = ca >= 0 ? cand_pos[ca] : -1;
    q = s >= 0 ? cand_pos[s] : -1;
  } else {
    const int64_t e2 = e - A.n;
    k = img_run_of(IR, e2);
    p = cand_pos[nbasis + e2];
    s = src[e2];
    nu = nup ? nup[e2] : 0;
    q = cand_pos[s];
  }
}

__global__ void __launch_bounds__(kBlock) k_pairs_count(int64_t R, int64_t nbasis, ImgRuns IR,
                                                        PriorImgs A, OpTable T,
                                                        const int32_t* __restrict__ cand_pos,
                                                        const uint8_t* __restrict__ inB,
                                                        const int32_t* __restrict__ src,
                                                        int32_t* __restrict__ cnt) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= A.n + IR.off[IR.n]) return;
  int k, nu;
  int32_t p, q, s;
  image_ends(e, A, IR, nbasis, cand_pos, src, nullptr, k, p, q, s, nu);
  if (p < 0 || q < 0) return;  // an endpoint pruned from the basis
  if (T.dir[k] < 0 && inB[p]) return;
  const int64_t t = T.term[k];
  atomicAdd(&cnt[t * R + q], 1);
  atomicAdd(&cnt[t * R + p], 1);
}

// indptr of every term from the scan, and (when cursor is given) the fill cursors
// (grid: x over the rows, y over the terms)
__global__ void __launch_bounds__(kBlock) k_pairs_indptr(int64_t R, int nterms,
                                                         const int32_t* __restrict__ G,
                                                         CsrOut out, int32_t* __restrict__ cursor) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (i >= R) return;
  const int64_t e = (int64_t)t * R + i;
  const int32_t base = G[e] - G[(int64_t)t * R];
  out.indptr[t][i] = base;
  if (cursor) cursor[e] = base;
  if (i == R - 1) out.indptr[t][R] = G[(int64_t)(t + 1) * R] - G[(int64_t)t * R];
}

__global__ void __launch_bounds__(kBlock) k_pairs_fill(int64_t R, int64_t nbasis, ImgRuns IR,
                                                       PriorImgs A, OpTable T, ModelArg M,
                                                       const u64* __restrict__ basis, int W,
                                                       int trajbits,
                                                       const int32_t* __restrict__ cand_pos,
                                                       const uint8_t* __restrict__ inB,
                                                       const int32_t* __restrict__ src,
                                                       const uint8_t* __restrict__ nup,
                                                       const double* __restrict__ sqrt_tab,
                                                       int32_t* __restrict__ cursor, CsrOut out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= A.n + IR.off[IR.n]) return;
  int k, nu;
  int32_t p, q, s;
  image_ends(e, A, IR, nbasis, cand_pos, src, nup, k, p, q, s, nu);
  if (p < 0 || q < 0) return;
  if (T.dir[k] < 0 && inB[p]) return;
  const int t = T.term[k];
  double v = 0.0;
  if (!out.coded) {
    const int traj = M.batched ? traj_of(basis[(int64_t)s * W], trajbits) : 0;
    v = op_value(T, k, M, traj, sqrt_tab, nu);
  }
  const int32_t a = atomicAdd(&cursor[(int64_t)t * R + q], 1);
  out.indices[t][a] = p;
  put_entry(out, t, a, v, nu);
  const int32_t b = atomicAdd(&cursor[(int64_t)t * R + p], 1);
  out.indices[t][b] = q;
  put_entry(out, t, b, v, nu);
}

// selection sort of each row's (few) entries by column, in place
__global__ void __launch_bounds__(kBlock) k_pairs_sort(int64_t R, int nterms, CsrOut out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  for (int t = 0; t < nterms; ++t) {
    const int32_t b = out.indptr[t][i], e = out.indptr[t][i + 1];
    int32_t* ix = out.indices[t];
    double* dv = out.data[t];
    float* fv = out.data32[t];
    uint8_t* cv = out.code[t];
    for (int32_t a = b; a + 1 < e; ++a) {
      int32_t m = a;
      for (int32_t c = a + 1; c < e; ++c)
        if (ix[c] < ix[m]) m = c;
      if (m != a) {
        const int32_t tj = ix[a];
        ix[a] = ix[m];
        ix[m] = tj;
        if (out.coded) {
          if (cv) {
            const uint8_t tv = cv[a];
            cv[a] = cv[m];
            cv[m] = tv;
          }
        } else if (fv) {
          const float tv = fv[a];
          fv[a] = fv[m];
          fv[m] = tv;
        } else {
          const double tv = dv[a];
          dv[a] = dv[m];
          dv[m] = tv;
        }
      }
    }
  }
}

// Fixed slots instead of atomics and the sort, where the column order is known in advance. An
// entry (row, neighbour) belongs to the op that maps row to neighbour: the image pair {b, e} of
// op k is slot k of row q = pos(b) and slot k^1 (the inverse op: ops come in (+, -) pairs) of row
// p = pos(e). Each (row, slot) is written by one pair at most, so a byte per slot marks the
// occupied ones without atomics. A hop moves only the exciton position and a vib op only the
// phonon count at the exciton site, so the column order of a row's neighbours is the order of
// the ops' key changes: rank[k] is op k's place in that order within its term. An entry's
// position is its row's start plus its occupied slots of lower rank; rows come out sorted.
struct SlotArg {
  int8_t k0[PACES_MAX_TERMS];     // first op of each term
  int8_t S[PACES_MAX_TERMS];      // ops (slots) of each term
  int8_t rank[kMaxOps];           // column rank of op k's neighbour among its term's slots
  int32_t foff[PACES_MAX_TERMS];  // slots of the terms before t (flag offset = R * foff[t])
};

// the slot table, or false when some term's key changes do not give a strict column order for
// every row (periodic batched lattices, whose spans differ by trajectory; ties on tiny rings)
static bool build_slots(const OpTable& T, const paces_holstein_t* h, SlotArg* SL) {
  if (h->batched && h->periodic) return false;
  int k = 0, f = 0;
  for (int t = 0; t < h->nterms; ++t) {
    SL->k0[t] = (int8_t)k;
    SL->foff[t] = f;
    int S = 0;
    while (k + S < T.n && T.term[k + S] == t) ++S;
    if (S % 2 != 0 || S > 8) return false;
    for (int j = 0; j < S; j += 2)
      if (T.dir[k + j] != 1 || T.dir[k + j + 1] != -1) return false;
    long long d[8];
    for (int j = 0; j < S; ++j) {
      const int q = k + j;
      d[j] = T.kind[q] == PACES_TERM_HOP
                 ? (long long)T.dir[q] * T.oe[q] - (T.wrap[q] ? (long long)T.dir[q] * T.span[q] : 0)
                 : (long long)T.dir[q];
    }
    for (int j = 0; j < S; ++j) {
      int r = 0;
      for (int i = 0; i < S; ++i) {
        if (i != j && d[i] == d[j]) return false;
        r += d[i] < d[j] ? 1 : 0;
      }
      SL->rank[k + j] = (int8_t)r;
    }
    SL->S[t] = (int8_t)S;
    k += S;
    f += S;
  }
  return k == T.n;
}

__global__ void __launch_bounds__(kBlock) k_pairs_mark(int64_t R, int64_t nbasis, ImgRuns IR,
                                                       PriorImgs A, OpTable T, SlotArg SL,
                                                       const int32_t* __restrict__ cand_pos,
                                                       const uint8_t* __restrict__ inB,
                                                       const int32_t* __restrict__ src,
                                                       uint8_t* __restrict__ flags) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= A.n + IR.off[IR.n]) return;
  int k, nu;
  int32_t p, q, s;
  image_ends(e, A, IR, nbasis, cand_pos, src, nullptr, k, p, q, s, nu);
  const bool pin = inB[p < 0 ? 0 : p] != 0;
  if (p < 0 || q < 0) return;
  if (T.dir[k] < 0 && pin) return;
  const int t = T.term[k], S = SL.S[t];
  uint8_t* f = flags + R * SL.foff[t];
  f[(int64_t)q * S + (k - SL.k0[t])] = 1;
  f[(int64_t)p * S + ((k ^ 1) - SL.k0[t])] = 1;
}

// Row counts of the slot path in tiles of kPairTile rows: the count pass sums each tile's slot
// flags per term (a scan of those sums gives every tile's first entry and each term's nnz), the
// indptr pass then writes each row's start from its tile's offset plus the in-tile prefix. A
// thread takes kPairRpt consecutive rows. Grids: x over the tiles, y over the terms.
constexpr int kPairRpt = 8;  // (the indptr stores below are written for 8)
constexpr int kPairTile = kBlock * kPairRpt;
static_assert(kPairRpt == 8, "indptr stores");

__device__ __forceinline__ int slot_row_count(const uint8_t* f, int S) {
  if (S == 2) {  // slot counts are even, so a term's flags start 2-byte aligned
    const uint16_t v = *(const uint16_t*)f;
    return (v & 0xff) + (v >> 8);
  }
  int c = 0;
  for (int j = 0; j < S; ++j) c += f[j];
  return c;
}

__global__ void __launch_bounds__(kBlock) k_pairs_tilesum(int64_t R, SlotArg SL,
                                                          const uint8_t* __restrict__ flags,
                                                          int32_t* __restrict__ ts) {
  __shared__ int ws[kBlock / 64];
  const int t = blockIdx.y;
  const int S = SL.S[t];
  const uint8_t* f = flags + R * SL.foff[t];
  const int64_t r0 = (int64_t)blockIdx.x * kPairTile + (int64_t)threadIdx.x * kPairRpt;
  int c = 0;
#pragma unroll
  for (int j = 0; j < kPairRpt; ++j)
    if (r0 + j < R) c += slot_row_count(f + (r0 + j) * S, S);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) a += ws[w];
    ts[(int64_t)t * gridDim.x + blockIdx.x] = a;
  }
}

__global__ void k_pairs_tile_nnz(const int32_t* __restrict__ TS, int ntl, int nterms,
                                 int32_t* __restrict__ nnz) {
  const int t = threadIdx.x;
  if (t < nterms) nnz[t] = TS[(int64_t)(t + 1) * ntl] - TS[(int64_t)t * ntl];
}

__global__ void __launch_bounds__(kBlock) k_pairs_indptr_tiles(int64_t R, SlotArg SL,
                                                               const uint8_t* __restrict__ flags,
                                                               const int32_t* __restrict__ TS,
                                                               CsrOut out) {
  __shared__ int wsum[kBlock / 64];
  const int t = blockIdx.y;
  const int ntl = gridDim.x;
  const int S = SL.S[t];
  const uint8_t* f = flags + R * SL.foff[t];
  const int64_t r0 = (int64_t)blockIdx.x * kPairTile + (int64_t)threadIdx.x * kPairRpt;
  int c[kPairRpt];
  int tot = 0;
#pragma unroll
  for (int j = 0; j < kPairRpt; ++j) {
    c[j] = (r0 + j < R) ? slot_row_count(f + (r0 + j) * S, S) : 0;
    tot += c[j];
  }
  // exclusive prefix of the thread totals across the workgroup
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = tot;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int before = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) before += w < wave ? wsum[w] : 0;
  int run = TS[(int64_t)t * ntl + blockIdx.x] - TS[(int64_t)t * ntl] + before + inc - tot;
  int32_t* ip = out.indptr[t];
  int o[kPairRpt];
#pragma unroll
  for (int j = 0; j < kPairRpt; ++j) {
    o[j] = run;
    run += c[j];
  }
  if (r0 + kPairRpt <= R && (((uintptr_t)(ip + r0)) & 15) == 0) {  // two 16-byte stores
    *(int4*)(ip + r0) = make_int4(o[0], o[1], o[2], o[3]);
    *(int4*)(ip + r0 + 4) = make_int4(o[4], o[5], o[6], o[7]);
  } else {
#pragma unroll
    for (int j = 0; j < kPairRpt; ++j)
      if (r0 + j < R) ip[r0 + j] = o[j];
  }
  if (r0 <= R - 1 && R - 1 < r0 + kPairRpt)
    ip[R] = TS[(int64_t)(t + 1) * ntl] - TS[(int64_t)t * ntl];
}

__device__ __forceinline__ int32_t slot_place(const SlotArg& SL, int t, const uint8_t* f, int S,
                                              int rk, int32_t start) {
  int32_t a = start;
  for (int j = 0; j < S; ++j) a += (SL.rank[SL.k0[t] + j] < rk) ? f[j] : 0;
  return a;
}

__global__ void __launch_bounds__(kBlock) k_pairs_place(int64_t R, int64_t nbasis, ImgRuns IR,
                                                        PriorImgs A, OpTable T, SlotArg SL,
                                                        ModelArg M,
                                                        const u64* __restrict__ basis, int W,
                                                        int trajbits,
                                                        const int32_t* __restrict__ cand_pos,
                                                        const uint8_t* __restrict__ inB,
                                                        const int32_t* __restrict__ src,
                                                        const uint8_t* __restrict__ nup,
                                                        const double* __restrict__ sqrt_tab,
                                                        const uint8_t* __restrict__ flags,
                                                        CsrOut out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= A.n + IR.off[IR.n]) return;
  // every load is issued before the validity tests (indices clamped in bounds), so the
  // dependent levels overlap: (p, src, nup) -> (q, traj word, inB[p], p's slots and start) ->
  // (q's slots and start); a prior image has one more level in front (its expanded-basis rows)
  int k, nu;
  int32_t p, q, s;
  image_ends(e, A, IR, nbasis, cand_pos, src, nup, k, p, q, s, nu);
  const int t = T.term[k], S = SL.S[t];
  const int32_t pc = p < 0 ? 0 : p, qc = q < 0 ? 0 : q;
  const u64 w0 = (M.batched && !out.coded) ? basis[(int64_t)(s < 0 ? 0 : s) * W] : 0;
  const bool pin = inB[pc] != 0;
  const uint8_t* f = flags + R * SL.foff[t];
  uint8_t fp[8], fq[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    fp[j] = j < S ? f[(int64_t)pc * S + j] : 0;
    fq[j] = j < S ? f[(int64_t)qc * S + j] : 0;
  }
  const int32_t sp = out.indptr[t][pc], sq = out.indptr[t][qc];
  if (p < 0 || q < 0) return;
  if (T.dir[k] < 0 && pin) return;
  const int traj = M.batched ? traj_of(w0, trajbits) : 0;
  const double v = out.coded ? 0.0 : op_value(T, k, M, traj, sqrt_tab, nu);
  int32_t a = sq, b = sp;
  const int rk = SL.rank[k], rinv = SL.rank[k ^ 1];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (j < S) {
      const int rj = SL.rank[SL.k0[t] + j];
      a += rj < rk ? fq[j] : 0;
      b += rj < rinv ? fp[j] : 0;
    }
  }
  out.indices[t][a] = p;
  put_entry(out, t, a, v, nu);
  out.indices[t][b] = q;
  put_entry(out, t, b, v, nu);
}

// Slot matrices (paces_slots_t): the slot of (row, op) is fixed, so a pair writes its two
// entries straight into their slots; an image pair {b, e} of op k fills slot rank[k] of row
// q = pos(b) and slot rank[k ^ 1] of row p = pos(e) within their term's slots. Each slot has one
// writer at most (as the flags of k_pairs_mark), and k_slots_clear has set every slot empty first.
// Slots are stored slot-major: an image run is monotone in both q and p, so its writes sweep two
// slot columns in order and each cache line of a column is completed within a short window (a
// row-major layout had every line written piecemeal by four runs far apart in time).
// (n slots from a 16-byte aligned base: a 16-byte store per thread, the tail one slot at a time)
__global__ void __launch_bounds__(kBlock) k_slots_clear(uint32_t* __restrict__ p, int64_t n) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i + 4 <= n) {
    *(uint4*)(p + i) = make_uint4(PACES_SLOT_EMPTY, PACES_SLOT_EMPTY, PACES_SLOT_EMPTY,
                                  PACES_SLOT_EMPTY);
  } else {
    for (int64_t j = i; j < n; ++j) p[j] = PACES_SLOT_EMPTY;
  }
}

__global__ void __launch_bounds__(kBlock) k_pairs_slot_place(int64_t R, int64_t nbasis, ImgRuns IR,
                                                             PriorImgs A, OpTable T, SlotArg SL,
                                                             int col_bits,
                                                             const int32_t* __restrict__ cand_pos,
                                                             const uint8_t* __restrict__ inB,
                                                             const int32_t* __restrict__ src,
                                                             const uint8_t* __restrict__ nup,
                                                             uint32_t* __restrict__ slots) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= A.n + IR.off[IR.n]) return;
  int k, nu;
  int32_t p, q, s;
  image_ends(e, A, IR, nbasis, cand_pos, src, nup, k, p, q, s, nu);
  const bool pin = inB[p < 0 ? 0 : p] != 0;
  if (p < 0 || q < 0) return;
  if (T.dir[k] < 0 && pin) return;
  const int t = T.term[k];
  const uint32_t code = T.kind[k] == PACES_TERM_VIB ? (uint32_t)nu << col_bits : 0u;
  slots[(int64_t)(SL.foff[t] + SL.rank[k]) * R + q] = (uint32_t)p | code;
  slots[(int64_t)(SL.foff[t] + SL.rank[k ^ 1]) * R + p] = (uint32_t)q | code;
}

// Explicit pairs of a slot matrix: images e of source rows (img_src[e] indexes the source list),
// run k of the image list from op k, with src_pos / img_pos the positions of the sources and the
// images in the slot matrix's rows: slot rank[k] of the source's row and slot rank[k ^ 1] of the
// image's row, as k_pairs_slot_place writes them.
__global__ void __launch_bounds__(kBlock) k_slots_place_explicit(
    int64_t R, ImgRuns IR, OpTable T, SlotArg SL, int col_bits,
    const int32_t* __restrict__ src_pos, const int32_t* __restrict__ img_pos,
    const int32_t* __restrict__ img_map, const int32_t* __restrict__ src,
    const uint8_t* __restrict__ nup, uint32_t* __restrict__ slots) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= IR.off[IR.n]) return;
  const int k = img_run_of(IR, e);
  const int32_t q = src_pos[src[e]], p = img_pos[img_map ? img_map[e] : e];
  if (p < 0 || q < 0 || p >= R || q >= R) return;  // a row not among the R rows has no slots
  const int t = T.term[k];
  const uint32_t code = T.kind[k] == PACES_TERM_VIB ? (uint32_t)nup[e] << col_bits : 0u;
  slots[(int64_t)(SL.foff[t] + SL.rank[k]) * R + q] = (uint32_t)p | code;
  slots[(int64_t)(SL.foff[t] + SL.rank[k ^ 1]) * R + p] = (uint32_t)q | code;
}

// the largest argument list of the pairs kernels (k_pairs_place) stays within 4 KiB
static_assert(sizeof(ImgRuns) + sizeof(PriorImgs) + sizeof(OpTable) + sizeof(SlotArg) +
                  sizeof(ModelArg) + sizeof(CsrOut) + 16 * 8 <= 4096,
              "pairs kernel arguments");

static size_t pairs_ws(int64_t R, int nterms, int nops) {
  int64_t m = (int64_t)nterms * R;
  return a256((size_t)m * 4) + a256((size_t)(m + 1) * 4) + a256(scan_ws_bytes(m)) +
         a256((size_t)R * nops);
}

// the slot flags sit after the count, prefix and scan areas of the workspace
static uint8_t* pairs_flags(void* ws, int64_t R, int nterms) {
  int64_t m = (int64_t)nterms * R;
  return (uint8_t*)ws + a256((size_t)m * 4) + a256((size_t)(m + 1) * 4) + a256(scan_ws_bytes(m));
}

// ------------------------------------------------------------------------------------------------
// A6 diagonal: eps + hbar_omega * sum_i n_i (+ delta_eps * pos), holstein.py:280-287. The integer
// sum is exact in FP64, so the result is the reference's rounding of the same expression.
// ------------------------------------------------------------------------------------------------

template <int W>
__device__ __forceinline__ double row_diag(const Row<W>& r, const LayoutArg& L, int pf,
                                           double eps, double hw, double de) {
  // sum over all phonon fields as a weighted popcount of the bits of each significance
  long long s = 0;
#pragma unroll
  for (int w = 0; w < W; ++w)
    for (int k = 0; k < L.sum_bits; ++k)
      s += (long long)__popcll(r.w[w] & L.sum_mask[w * L.sum_bits + k]) << k;
  double d = eps + hw * (double)s;
  if (de != 0.0) d = d + de * (double)get_field<W>(r, L.off[pf], L.width[pf]);
  return d;
}

template <int W, typename D>
__global__ void __launch_bounds__(kBlock) k_diag(const u64* __restrict__ rows, int64_t R,
                                                 LayoutArg L, int pf, double eps, double hw,
                                                 double de, D* __restrict__ diag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  diag[i] = (D)row_diag<W>(load_row<W>(rows, i), L, pf, eps, hw, de);  // FP32: rounded to nearest
}

// ------------------------------------------------------------------------------------------------
// H1 observables: a workgroup stages a tile of rows and |psi|^2 in LDS; each thread then owns one
// or more sites and sums over the tile's rows (fixed order), so every partial is deterministic.
// out[0..N) = n_exc, out[N..2N) = n_pho, out[2N] = sum |psi|^2.
// ------------------------------------------------------------------------------------------------

constexpr int kObsLdsBytes = 32768;

template <int W>
struct ObsTile {
  static constexpr int rows = (kObsLdsBytes / (8 * W + 12)) >= 1024 ? 1024
                              : (kObsLdsBytes / (8 * W + 12)) >= 512 ? 512 : 256;
};

// Partial sums of one tile of TR rows: the sum of weights, the weight at each exciton site and
// the weighted phonons at each site. Wave w takes sites w, w + 4, ..., so a site's field offset and
// width are uniform across the wave; lane l holds rows l, l + 64, ... with their weights and
// exciton positions in registers, and a site's sum ends in a wave reduction. Only sites inside
// the tile's range of exciton positions can hold exciton weight.
template <int W>
__global__ void __launch_bounds__(kBlock) k_obs_partial(const u64* __restrict__ rows,
                                                        const double2* __restrict__ psi, int64_t R,
                                                        LayoutArg L, int N,
                                                        double* __restrict__ part) {
  constexpr int TR = ObsTile<W>::rows, RPL = TR / 64, PER = TR * W / kBlock;
  __shared__ u64 srow[TR * W];
  __shared__ int spr[2];
  const int S = 2 * N + 1;
  const int64_t base = (int64_t)blockIdx.x * TR;
  const int nr = (int)((R - base) < TR ? (R - base) : TR);
  if (threadIdx.x == 0) {
    spr[0] = 0x7fffffff;
    spr[1] = -1;
  }
  {
    u64 tmp[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int e = (int)threadIdx.x + j * kBlock;
      tmp[j] = e < nr * W ? rows[base * W + e] : 0ull;
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int e = (int)threadIdx.x + j * kBlock;
      if (e < nr * W) srow[e] = tmp[j];
    }
  }
  __syncthreads();
  // a uniform wave index: the site loop's field offsets are scalar and its branches scalar
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double w[RPL];
  int pos[RPL];
  int lo = 0x7fffffff, hi = -1;
#pragma unroll
  for (int j = 0; j < RPL; ++j) {
    const int r = lane + 64 * j;
    w[j] = 0.0;
    pos[j] = -1;
    if (r < nr) {
      const double2 v = psi[base + r];
      w[j] = v.x * v.x + v.y * v.y;
      Row<W> rr;
#pragma unroll
      for (int k = 0; k < W; ++k) rr.w[k] = srow[r * W + k];
      pos[j] = (int)get_field<W>(rr, 0, L.width[0]);
      lo = pos[j] < lo ? pos[j] : lo;
      hi = pos[j] > hi ? pos[j] : hi;
    }
  }
  if (hi >= 0) {
    atomicMin(&spr[0], lo);
    atomicMax(&spr[1], hi);
  }
  if (wave == 0) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < RPL; ++j) a += w[j];
    a = wave_sum(a);
    if (lane == 0) part[(int64_t)blockIdx.x * S + 2 * N] = a;
  }
  __syncthreads();
  const int pmin = spr[0], pmax = spr[1];
  for (int s = wave; s < N; s += kBlock / 64) {
    const int fo = L.off[s + 1], b = L.width[s + 1];
    const int k = fo >> 6, o = fo & 63;
    double ap = 0.0;
    if (b > 0 && o + b > 64) {  // the field straddles two words
      const u64 mask = (b >= 64) ? ~0ull : ((1ull << b) - 1ull);
#pragma unroll
      for (int j = 0; j < RPL; ++j) {
        const int r = lane + 64 * j;  // rows past the tile read stale LDS, weighted by 0
        const u64 n = ((srow[r * W + k] << (o + b - 64)) | (srow[r * W + k + 1] >> (128 - o - b))) &
                      mask;
        ap += w[j] * (double)n;
      }
    } else if (b > 0 && b <= 32) {
      const u64 mask = (1ull << b) - 1ull;
#pragma unroll
      for (int j = 0; j < RPL; ++j) {
        const int r = lane + 64 * j;
        ap += w[j] * (double)(uint32_t)((srow[r * W + k] >> (64 - o - b)) & mask);
      }
    } else if (b > 0) {
      const u64 mask = (b >= 64) ? ~0ull : ((1ull << b) - 1ull);
#pragma unroll
      for (int j = 0; j < RPL; ++j) {
        const int r = lane + 64 * j;
        ap += w[j] * (double)((srow[r * W + k] >> (64 - o - b)) & mask);
      }
    }
    ap = wave_sum(ap);
    double ae = 0.0;
    if (s >= pmin && s <= pmax) {
#pragma unroll
      for (int j = 0; j < RPL; ++j) ae += (pos[j] == s) ? w[j] : 0.0;
      ae = wave_sum(ae);
    }
    if (lane == 0) {
      part[(int64_t)blockIdx.x * S + s] = ae;
      part[(int64_t)blockIdx.x * S + N + s] = ap;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_obs_final(const double* __restrict__ part, int nblk,
                                                      int S, double* __restrict__ out) {
  __shared__ double red[kBlock / 64];
  const int s = blockIdx.x;
  double a = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) a += part[(int64_t)b * S + s];
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) t += red[q];
    out[s] = t;
  }
}

template <int W>
static int64_t obs_blocks(int64_t R) {
  return (R + ObsTile<W>::rows - 1) / ObsTile<W>::rows;
}

static size_t obs_ws(int64_t R, int N) {
  // sized for the smallest tile (256 rows), valid for every W
  int64_t nblk = (R + 255) / 256;
  if (nblk < 1) nblk = 1;
  return a256((size_t)nblk * (2 * N + 1) * 8);
}

template <int W>
static int run_obs(const uint64_t* rows, const double* psi, int64_t R, const LayoutArg& L, int N,
                   double* out, void* ws, hipStream_t st) {
  int64_t nblk = obs_blocks<W>(R);
  const int S = 2 * N + 1;
  double* part = (double*)ws;
  if (nblk > 0) {
    hipLaunchKernelGGL(k_obs_partial<W>, dim3((unsigned)nblk), dim3(kBlock), 0, st,
                       (const u64*)rows, (const double2*)psi, R, L, N, part);
    int rc = launch_status();
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_obs_final, dim3(S), dim3(kBlock), 0, st, (const double*)part, (int)nblk, S,
                     out);
  return launch_status();
}

template <int W>
static int run_diag(const uint64_t* rows, int64_t R, const LayoutArg& L, int pf, double eps,
                    double hw, double de, double* diag, float* diag32, hipStream_t st) {
  if (R == 0) return PACES_OK;
  if (diag32)
    hipLaunchKernelGGL((k_diag<W, float>), dim3(grid_for(R, kBlock)), dim3(kBlock), 0, st,
                       (const u64*)rows, R, L, pf, eps, hw, de, diag32);
  else
    hipLaunchKernelGGL((k_diag<W, double>), dim3(grid_for(R, kBlock)), dim3(kBlock), 0, st,
                       (const u64*)rows, R, L, pf, eps, hw, de, diag);
  return launch_status();
}

static int check_h(const paces_layout_t* lay, const paces_holstein_t* h) {
  if (!lay || !h || !lay->d_off || !lay->d_width || !lay->d_maxdim) return PACES_E_INVALID;
  if (lay->W < 1 || lay->W > PACES_MAX_ROW_WORDS) return PACES_E_UNSUPPORTED;
  const int extra = h->batched ? 2 : 1;
  if (h->nchain < 1 || h->nchain > PACES_MAX_SITES || lay->nfields != h->nchain + extra)
    return PACES_E_INVALID;
  if (h->nterms < 0 || h->nterms > PACES_MAX_TERMS) return PACES_E_UNSUPPORTED;
  if (h->geo_dim < 1) return PACES_E_INVALID;
  if (h->batched) {
    if (h->ntraj < 1 || !h->d_traj_L || !h->d_traj_value) return PACES_E_INVALID;
    return PACES_OK;
  }
  if (h->lattice_l < 1) return PACES_E_INVALID;
  if (h->periodic && h->lattice_l < 3) return PACES_E_INVALID;
  for (int t = 0; t < h->nterms; ++t) {
    if (h->terms[t].kind == PACES_TERM_HOP &&
        (h->terms[t].order < 1 || h->terms[t].order > h->lattice_l - 1))
      return PACES_E_INVALID;
  }
  return PACES_OK;
}

static LayoutArg larg(const paces_layout_t* lay) {
  LayoutArg L;
  L.off = lay->d_off;
  L.width = lay->d_width;
  L.maxdim = lay->d_maxdim;
  L.sum_mask = (const u64*)lay->d_sum_mask;
  L.sum_bits = lay->sum_bits;
  return L;
}

}  // namespace paces

using namespace paces;

// ------------------------------------------------------------------------------------------------
// N3 candidate features (no reference counterpart): one row of PACES_MLP_IN bf16 values per union
// row, the input of the candidate scorer.
//   [0, 33)  phonon count at sites pos-16 .. pos+16 over (dim - 1); -1 outside an open lattice
//   33       log10 of the largest |psi| among the row's H-neighbours through the image pairs, / 10
//   34       (diag - eps_sys) / (hbar_omega * nchain)
//   35       log10 |psi| of the row itself, / 10 (amplitudes below 1e-20 count as 1e-20)
//   36, 37   the vibronic and hopping term values; 38 pos / nchain; the rest 0
// ------------------------------------------------------------------------------------------------

constexpr int kFeatWin = 16;

__global__ void __launch_bounds__(kBlock) k_zero_u64(u64* __restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// par[i] = max |psi_j|^2 over the pairs {i, j} of the images (bit patterns of non-negative
// doubles order like the values). With image_side_only, only the image end of each pair (p) is
// updated: the rows of the expanded basis (q) get no value (the added-rows form never reads them).
__global__ void __launch_bounds__(kBlock) k_parent_amp(int64_t nbasis, int64_t nimg,
                                                       const int32_t* __restrict__ cand_pos,
                                                       const int32_t* __restrict__ src,
                                                       const double2* __restrict__ psi,
                                                       int image_side_only,
                                                       u64* __restrict__ par) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nimg) return;
  const int32_t p = cand_pos[nbasis + e];
  const int32_t q = cand_pos[src[e]];
  if (p < 0 || q < 0) return;
  const double2 a = psi[q];
  atomicMax(&par[p], (u64)__double_as_longlong(a.x * a.x + a.y * a.y));
  if (!image_side_only) {
    const double2 b = psi[p];
    atomicMax(&par[q], (u64)__double_as_longlong(b.x * b.x + b.y * b.y));
  }
}

__device__ __forceinline__ uint16_t feat_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float feat_log_amp(double w2) {
  double l = (w2 > 1e-40) ? 0.5 * log10(w2) : -20.0;
  return (float)(l / 10.0);
}

// ADDED: only the rows with in_basis[i] == 0, written at i - G[i] (G the exclusive prefix count of
// in_basis), with row_index[i - G[i]] = i, and the diagonal computed from the row (hw, de, pf as
// for k_diag) instead of read
template <int W, bool ADDED>
__global__ void __launch_bounds__(kBlock) k_features(const u64* __restrict__ rows, int64_t R,
                                                     const double2* __restrict__ psi,
                                                     const double* __restrict__ diag,
                                                     const u64* __restrict__ par, LayoutArg L,
                                                     int N, int periodic, double eps,
                                                     double inv_scale, float gfeat, float jfeat,
                                                     const uint8_t* __restrict__ in_basis,
                                                     const int32_t* __restrict__ G, int pf,
                                                     double hw, double de,
                                                     int32_t* __restrict__ row_index,
                                                     uint16_t* __restrict__ X) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  if (ADDED && in_basis[i]) return;
  const Row<W> r = load_row<W>(rows, i);
  const int64_t o = ADDED ? i - G[i] : i;
  if (ADDED) row_index[o] = (int32_t)i;
  const double dg = ADDED ? row_diag<W>(r, L, pf, eps, hw, de) : diag[i];
  const int pos = (int)get_field<W>(r, L.off[0], L.width[0]);
  const double2 a = psi[i];
  uint32_t packed[PACES_MLP_IN / 2];
#pragma unroll
  for (int c2 = 0; c2 < PACES_MLP_IN / 2; ++c2) packed[c2] = 0;
  for (int c = 0; c < PACES_MLP_IN; ++c) {
    float v = 0.f;
    if (c < 2 * kFeatWin + 1) {
      int s = pos + c - kFeatWin;
      bool in = true;
      if (periodic) {
        s %= N;
        if (s < 0) s += N;
      } else {
        in = s >= 0 && s < N;
      }
      if (in) {
        const int md = L.maxdim[s + 1];
        const int n = (int)get_field<W>(r, L.off[s + 1], L.width[s + 1]);
        v = md > 1 ? (float)n / (float)(md - 1) : 0.f;
      } else {
        v = -1.f;
      }
    } else if (c == 33) {
      v = feat_log_amp(__longlong_as_double((long long)par[i]));
    } else if (c == 34) {
      v = (float)((dg - eps) * inv_scale);
    } else if (c == 35) {
      v = feat_log_amp(a.x * a.x + a.y * a.y);
    } else if (c == 36) {
      v = gfeat;
    } else if (c == 37) {
      v = jfeat;
    } else if (c == 38) {
      v = (float)pos / (float)N;
    }
    packed[c >> 1] |= (uint32_t)feat_bf16(v) << (16 * (c & 1));
  }
  uint4* dst = (uint4*)(X + o * PACES_MLP_IN);
#pragma unroll
  for (int q = 0; q < PACES_MLP_IN / 8; ++q)
    dst[q] = make_uint4(packed[4 * q], packed[4 * q + 1], packed[4 * q + 2], packed[4 * q + 3]);
}

template <int W>
static int run_features(const uint64_t* rows, int64_t R, const double* psi, const double* diag,
                        const u64* par, LayoutArg L, int N, int periodic, double eps,
                        double inv_scale, float gfeat, float jfeat, const uint8_t* in_basis,
                        const int32_t* G, int pf, double hw, double de, int32_t* row_index,
                        uint16_t* X, hipStream_t st) {
  if (in_basis)
    hipLaunchKernelGGL((k_features<W, true>), dim3(grid_for(R, kBlock)), dim3(kBlock), 0, st,
                       (const u64*)rows, R, (const double2*)psi, diag, par, L, N, periodic, eps,
                       inv_scale, gfeat, jfeat, in_basis, G, pf, hw, de, row_index, X);
  else
    hipLaunchKernelGGL((k_features<W, false>), dim3(grid_for(R, kBlock)), dim3(kBlock), 0, st,
                       (const u64*)rows, R, (const double2*)psi, diag, par, L, N, periodic, eps,
                       inv_scale, gfeat, jfeat, in_basis, G, pf, hw, de, row_index, X);
  return launch_status();
}

extern "C" {

int paces_holstein_num_runs(const paces_holstein_t* h) {
  OpTable T;
  if (!h) return PACES_E_INVALID;
  int rc = build_image_ops(h, &T);
  return rc ? rc : T.n;
}

// Ops come in inverse pairs (2m of positive direction, 2m + 1 its inverse). A row added to a
// basis as op k's image of a row s has op k^1's image s, already in the basis; when k is of
// positive direction that image is a pure duplicate: its pair with s is the one op k's image
// marked from s's side (a negative-direction image whose other end is a basis row gives no CSR
// entry). Codes for the union of [basis, op 0 images, op 1 images, ...]: 1 for the basis run
// (skip the row), 2 + (k ^ 1) for a positive-direction op k's run (skip that op only), else 0.
int paces_holstein_inverse_skip_codes(const paces_holstein_t* h, uint8_t* codes, int32_t cap) {
  OpTable T;
  if (!h || !codes) return PACES_E_INVALID;
  int rc = build_image_ops(h, &T);
  if (rc) return rc;
  if (cap < T.n + 1) return PACES_E_INVALID;
  codes[0] = 1;
  for (int k = 0; k < T.n; ++k) codes[1 + k] = T.dir[k] > 0 ? (uint8_t)(2 + (k ^ 1)) : 0;
  return T.n + 1;
}

size_t paces_holstein_images_workspace_bytes(int64_t n, const paces_holstein_t* h) {
  OpTable T;
  if (!h || build_image_ops(h, &T)) return 0;
  return images_ws(n, T.n);
}

int paces_holstein_images_count(const uint64_t* basis, int64_t n, const paces_layout_t* lay,
                                const paces_holstein_t* h, int32_t* run_offsets, void* ws,
                                size_t ws_bytes, void* stream) {
  return paces_holstein_images_count_skip(basis, n, nullptr, lay, h, run_offsets, ws, ws_bytes,
                                          stream);
}

int paces_holstein_images_count_skip(const uint64_t* basis, int64_t n, const uint8_t* skip,
                                     const paces_layout_t* lay, const paces_holstein_t* h,
                                     int32_t* run_offsets, void* ws, size_t ws_bytes,
                                     void* stream) {
  int rc = check_h(lay, h);
  if (rc) return rc;
  OpTable T;
  rc = build_image_ops(h, &T);
  if (rc) return rc;
  if (n < 0 || (n > 0 && !basis) || !run_offsets || !ws) return PACES_E_INVALID;
  if ((int64_t)T.n * n >= 2147483647LL) return PACES_E_INVALID;
  if (ws_bytes < images_ws(n, T.n)) return PACES_E_WORKSPACE;
  PACES_DISPATCH_W(lay->W, run_img_count, basis, n, skip, larg(lay), marg(h), T, run_offsets, ws,
                   ws_bytes, (hipStream_t)stream);
}

int paces_holstein_images_emit(const uint64_t* basis, int64_t n, const paces_layout_t* lay,
                               const paces_holstein_t* h, uint64_t* out_images, int32_t* out_src,
                               uint8_t* out_nup, void* ws, size_t ws_bytes, void* stream) {
  return paces_holstein_images_emit_skip(basis, n, nullptr, lay, h, out_images, out_src, out_nup,
                                         ws, ws_bytes, stream);
}

int paces_holstein_images_emit_skip(const uint64_t* basis, int64_t n, const uint8_t* skip,
                                    const paces_layout_t* lay, const paces_holstein_t* h,
                                    uint64_t* out_images, int32_t* out_src, uint8_t* out_nup,
                                    void* ws, size_t ws_bytes, void* stream) {
  int rc = check_h(lay, h);
  if (rc) return rc;
  OpTable T;
  rc = build_image_ops(h, &T);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!basis || !out_images)) || !ws) return PACES_E_INVALID;
  if (ws_bytes < images_ws(n, T.n)) return PACES_E_WORKSPACE;
  PACES_DISPATCH_W(lay->W, run_img_emit, basis, n, skip, larg(lay), marg(h), T, out_images,
                   out_src, out_nup, ws, (hipStream_t)stream);
}

static int img_runs_arg(const OpTable& T, const int64_t* run_offsets, ImgRuns* IR) {
  IR->n = T.n;
  for (int k = 0; k <= T.n; ++k) IR->off[k] = run_offsets[k];
  for (int k = T.n + 1; k <= kMaxOps; ++k) IR->off[k] = run_offsets[T.n];
  for (int k = 0; k < T.n; ++k)
    if (IR->off[k + 1] < IR->off[k]) return PACES_E_INVALID;
  return PACES_OK;
}

// the prior part of the image list (none without prior): run offsets on the host, checked as
// the second part's are
static int prior_arg(const OpTable& T, const paces_prior_images_t* pr, PriorImgs* A) {
  A->n = 0;
  A->nprev = 0;
  A->cpos = nullptr;
  A->src = nullptr;
  A->nup = nullptr;
  A->IR.n = T.n;
  for (int k = 0; k <= kMaxOps; ++k) A->IR.off[k] = 0;
  for (int k = 0; k < kMaxOps; ++k) A->base[k] = 0;
  if (!pr) return PACES_OK;
  if (!pr->run_offsets || pr->nprev < 0) return PACES_E_INVALID;
  ImgRuns full;
  int rc = img_runs_arg(T, pr->run_offsets, &full);
  if (rc) return rc;
  if (full.off[0] != 0) return PACES_E_INVALID;
  // the positive-direction runs only (the other ops' prior images give no entry)
  int64_t acc = 0;
  for (int k = 0; k < T.n; ++k) {
    A->base[k] = full.off[k];
    A->IR.off[k] = acc;
    if (T.dir[k] > 0) acc += full.off[k + 1] - full.off[k];
  }
  for (int k = T.n; k <= kMaxOps; ++k) A->IR.off[k] = acc;
  for (int k = T.n; k < kMaxOps; ++k) A->base[k] = full.off[T.n];
  A->n = acc;
  A->nprev = pr->nprev;
  A->cpos = pr->cand_pos;
  A->src = pr->img_src;
  A->nup = pr->img_nup;
  if (A->n > 0 && (!A->cpos || !A->src || !A->nup)) return PACES_E_INVALID;
  if (A->nprev + full.off[T.n] >= 2147483647LL) return PACES_E_INVALID;
  return PACES_OK;
}

size_t paces_holstein_pairs_workspace_bytes(int64_t R, const paces_holstein_t* h) {
  OpTable T;
  if (!h || build_image_ops(h, &T)) return 0;
  return pairs_ws(R, h->nterms, T.n);
}

static int pairs_count_impl(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                            const int32_t* cand_pos, const uint8_t* in_basis,
                            const int32_t* img_src, const paces_prior_images_t* prior,
                            const paces_holstein_t* h, int32_t* term_nnz, void* ws,
                            size_t ws_bytes, void* stream) {
  OpTable T;
  if (!h || !run_offsets || !term_nnz || !ws || R < 0 || nbasis < 0) return PACES_E_INVALID;
  if (h->nterms < 1 || h->nterms > PACES_MAX_TERMS) return PACES_E_UNSUPPORTED;
  int rc = build_image_ops(h, &T);
  if (rc) return rc;
  ImgRuns IR;
  rc = img_runs_arg(T, run_offsets, &IR);
  if (rc) return rc;
  PriorImgs A;
  rc = prior_arg(T, prior, &A);
  if (rc) return rc;
  const int64_t nimg = A.n + IR.off[IR.n];
  if (nimg > 0 && (!cand_pos || !in_basis || (IR.off[IR.n] > 0 && !img_src)))
    return PACES_E_INVALID;
  if ((int64_t)h->nterms * R >= 2147483647LL) return PACES_E_INVALID;
  if (ws_bytes < pairs_ws(R, h->nterms, T.n)) return PACES_E_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t m = (int64_t)h->nterms * R;
  char* p = (char*)ws;
  int32_t* cnt = (int32_t*)p;
  p += a256((size_t)m * 4);
  int32_t* G = (int32_t*)p;
  p += a256((size_t)(m + 1) * 4);
  SlotArg SL;
  if (build_slots(T, h, &SL)) {
    uint8_t* flags = pairs_flags(ws, R, h->nterms);
    const int64_t nw = ((int64_t)R * T.n + 7) / 8;
    if (nw > 0)
      hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(nw, kBlock)), dim3(kBlock), 0, st, (u64*)flags,
                         nw);
    if (nimg > 0)
      hipLaunchKernelGGL(k_pairs_mark, dim3(grid_for(nimg, kBlock)), dim3(kBlock), 0, st, R,
                         nbasis, IR, A, T, SL, cand_pos, in_basis, img_src, flags);
    // tile sums (in the count area) and their scan (in the prefix area) stand in for the row
    // counts and their scan; the fill call reads the same places
    const int ntl = (int)((R + kPairTile - 1) / kPairTile);
    if (R > 0) {
      hipLaunchKernelGGL(k_pairs_tilesum, dim3(ntl, h->nterms), dim3(kBlock), 0, st, R, SL,
                         (const uint8_t*)flags, cnt);
      rc = launch_status();
      if (rc) return rc;
      rc = scan_i32(cnt, (int64_t)ntl * h->nterms, G, p,
                    scan_ws_bytes((int64_t)ntl * h->nterms), st);
      if (rc) return rc;
      hipLaunchKernelGGL(k_pairs_tile_nnz, dim3(1), dim3(64), 0, st, (const int32_t*)G, ntl,
                         h->nterms, term_nnz);
    } else {
      hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(kBlock), 0, st, term_nnz, (int64_t)h->nterms);
    }
    return launch_status();
  } else {
    if (m > 0) {
      hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, st, cnt, m);
    }
    if (nimg > 0) {
      hipLaunchKernelGGL(k_pairs_count, dim3(grid_for(nimg, kBlock)), dim3(kBlock), 0, st, R,
                         nbasis, IR, A, T, cand_pos, in_basis, img_src, cnt);
    }
  }
  rc = launch_status();
  if (rc) return rc;
  rc = scan_i32(cnt, m, G, p, scan_ws_bytes(m), st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_csr_nnz, dim3(1), dim3(64), 0, st, (const int32_t*)G, R, h->nterms,
                     term_nnz);
  return launch_status();
}

int paces_holstein_pairs_count(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                               const int32_t* cand_pos, const uint8_t* in_basis,
                               const int32_t* img_src, const paces_holstein_t* h,
                               int32_t* term_nnz, void* ws, size_t ws_bytes, void* stream) {
  return pairs_count_impl(R, nbasis, run_offsets, cand_pos, in_basis, img_src, nullptr, h,
                          term_nnz, ws, ws_bytes, stream);
}

int paces_holstein_pairs_count_split(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                                     const int32_t* cand_pos, const uint8_t* in_basis,
                                     const int32_t* img_src, const paces_prior_images_t* prior,
                                     const paces_holstein_t* h, int32_t* term_nnz, void* ws,
                                     size_t ws_bytes, void* stream) {
  if (!prior) return PACES_E_INVALID;
  return pairs_count_impl(R, nbasis, run_offsets, cand_pos, in_basis, img_src, prior, h,
                          term_nnz, ws, ws_bytes, stream);
}

static int pairs_fill_impl(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                           const int32_t* cand_pos, const uint8_t* in_basis,
                           const int32_t* img_src, const uint8_t* img_nup,
                           const paces_prior_images_t* prior, const uint64_t* basis,
                           const paces_holstein_t* h, const paces_layout_t* lay,
                           int32_t* const* indptr, int32_t* const* indices, double* const* data,
                           float* const* data32, uint8_t* const* code, void* ws, size_t ws_bytes,
                           void* stream) {
  OpTable T;
  if (!h || !run_offsets || !ws || !indptr || !indices || (!data && !data32 && !code) || R < 0)
    return PACES_E_INVALID;
  if (h->batched && !code && (!basis || !lay || !h->d_traj_value)) return PACES_E_INVALID;
  if (code && (!h->batched || h->sqrt_table_len > PACES_CODE_TABLE)) return PACES_E_UNSUPPORTED;
  const int W = lay ? lay->W : 1;
  const int trajbits = h->batched ? tbits(h->ntraj) : 0;
  if (h->nterms < 1 || h->nterms > PACES_MAX_TERMS) return PACES_E_UNSUPPORTED;
  if (!h->d_sqrt_table) return PACES_E_INVALID;
  int rc = build_image_ops(h, &T);
  if (rc) return rc;
  ImgRuns IR;
  rc = img_runs_arg(T, run_offsets, &IR);
  if (rc) return rc;
  PriorImgs A;
  rc = prior_arg(T, prior, &A);
  if (rc) return rc;
  const int64_t nimg = A.n + IR.off[IR.n];
  if (nimg > 0 && (!cand_pos || !in_basis || (IR.off[IR.n] > 0 && (!img_src || !img_nup))))
    return PACES_E_INVALID;
  if (ws_bytes < pairs_ws(R, h->nterms, T.n)) return PACES_E_WORKSPACE;
  CsrOut out;
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    out.indptr[t] = t < h->nterms ? indptr[t] : nullptr;
    out.indices[t] = t < h->nterms ? indices[t] : nullptr;
    out.data[t] = (data && t < h->nterms) ? data[t] : nullptr;
    out.data32[t] = (data32 && t < h->nterms) ? data32[t] : nullptr;
    out.code[t] = (code && t < h->nterms && h->terms[t].kind == PACES_TERM_VIB) ? code[t] : nullptr;
    if (t < h->nterms && !out.indptr[t]) return PACES_E_INVALID;
    if (code && t < h->nterms && h->terms[t].kind == PACES_TERM_VIB && !out.code[t])
      return PACES_E_INVALID;
  }
  out.coded = code ? 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  if (R == 0) {
    hipLaunchKernelGGL(k_csr_empty, dim3(1), dim3(64), 0, st, out, h->nterms);
    return launch_status();
  }
  const int64_t m = (int64_t)h->nterms * R;
  char* p = (char*)ws;
  int32_t* cursor = (int32_t*)p;  // the count array is no longer needed
  p += a256((size_t)m * 4);
  const int32_t* G = (const int32_t*)p;
  SlotArg SL;
  if (build_slots(T, h, &SL)) {  // the same choice as the count call
    const int ntl = (int)((R + kPairTile - 1) / kPairTile);
    hipLaunchKernelGGL(k_pairs_indptr_tiles, dim3(ntl, h->nterms), dim3(kBlock), 0, st, R, SL,
                       (const uint8_t*)pairs_flags(ws, R, h->nterms), G, out);
    if (nimg > 0)
      hipLaunchKernelGGL(k_pairs_place, dim3(grid_for(nimg, kBlock)), dim3(kBlock), 0, st, R,
                         nbasis, IR, A, T, SL, marg(h), (const u64*)basis, W, trajbits, cand_pos,
                         in_basis, img_src, img_nup, h->d_sqrt_table,
                         (const uint8_t*)pairs_flags(ws, R, h->nterms), out);
    return launch_status();
  }
  hipLaunchKernelGGL(k_pairs_indptr, dim3(grid_for(R, kBlock), h->nterms), dim3(kBlock), 0, st, R,
                     h->nterms, G, out, cursor);
  if (nimg > 0)
    hipLaunchKernelGGL(k_pairs_fill, dim3(grid_for(nimg, kBlock)), dim3(kBlock), 0, st, R, nbasis,
                       IR, A, T, marg(h), (const u64*)basis, W, trajbits, cand_pos, in_basis,
                       img_src, img_nup, h->d_sqrt_table, cursor, out);
  hipLaunchKernelGGL(k_pairs_sort, dim3(grid_for(R, kBlock)), dim3(kBlock), 0, st, R, h->nterms,
                     out);
  return launch_status();
}

int paces_holstein_pairs_fill(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                              const int32_t* cand_pos, const uint8_t* in_basis,
                              const int32_t* img_src, const uint8_t* img_nup,
                              const uint64_t* basis, const paces_layout_t* lay,
                              const paces_holstein_t* h, int32_t* const* indptr,
                              int32_t* const* indices, double* const* data, void* ws,
                              size_t ws_bytes, void* stream) {
  if (!data) return PACES_E_INVALID;
  return pairs_fill_impl(R, nbasis, run_offsets, cand_pos, in_basis, img_src, img_nup, nullptr,
                         basis, h, lay, indptr, indices, data, nullptr, nullptr, ws, ws_bytes,
                         stream);
}

int paces_holstein_pairs_fill32(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                                const int32_t* cand_pos, const uint8_t* in_basis,
                                const int32_t* img_src, const uint8_t* img_nup,
                                const uint64_t* basis, const paces_layout_t* lay,
                                const paces_holstein_t* h, int32_t* const* indptr,
                                int32_t* const* indices, float* const* data, void* ws,
                                size_t ws_bytes, void* stream) {
  if (!data) return PACES_E_INVALID;
  return pairs_fill_impl(R, nbasis, run_offsets, cand_pos, in_basis, img_src, img_nup, nullptr,
                         basis, h, lay, indptr, indices, nullptr, data, nullptr, ws, ws_bytes,
                         stream);
}

int paces_holstein_pairs_fill_coded(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                                    const int32_t* cand_pos, const uint8_t* in_basis,
                                    const int32_t* img_src, const uint8_t* img_nup,
                                    const paces_holstein_t* h, int32_t* const* indptr,
                                    int32_t* const* indices, uint8_t* const* code, void* ws,
                                    size_t ws_bytes, void* stream) {
  if (!code) return PACES_E_INVALID;
  return pairs_fill_impl(R, nbasis, run_offsets, cand_pos, in_basis, img_src, img_nup, nullptr,
                         nullptr, h, nullptr, indptr, indices, nullptr, nullptr, code, ws,
                         ws_bytes, stream);
}

int paces_holstein_pairs_fill_split(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                                    const int32_t* cand_pos, const uint8_t* in_basis,
                                    const int32_t* img_src, const uint8_t* img_nup,
                                    const paces_prior_images_t* prior, const uint64_t* basis,
                                    const paces_layout_t* lay, const paces_holstein_t* h,
                                    int32_t* const* indptr, int32_t* const* indices,
                                    double* const* data, float* const* data32,
                                    uint8_t* const* code, void* ws, size_t ws_bytes,
                                    void* stream) {
  if (!prior || (!data) + (!data32) + (!code) != 2) return PACES_E_INVALID;
  return pairs_fill_impl(R, nbasis, run_offsets, cand_pos, in_basis, img_src, img_nup, prior,
                         code ? nullptr : basis, h, code ? nullptr : lay, indptr, indices, data,
                         data32, code, ws, ws_bytes, stream);
}

// the slot table of h (SL) and the host fields of S, or a nonzero status
static int slot_layout(const paces_holstein_t* h, int64_t R, OpTable* T, SlotArg* SL,
                       paces_slots_t* S) {
  if (!h || !S || R < 0) return PACES_E_INVALID;
  if (!h->batched || h->periodic) return PACES_E_UNSUPPORTED;
  if (h->nterms < 1 || h->nterms > PACES_MAX_TERMS) return PACES_E_UNSUPPORTED;
  if (h->sqrt_table_len < 1 || h->sqrt_table_len > PACES_CODE_TABLE) return PACES_E_UNSUPPORTED;
  int rc = build_image_ops(h, T);
  if (rc) return rc;
  if (T->n < 2 || T->n > PACES_MAX_SLOTS || (T->n & 1)) return PACES_E_UNSUPPORTED;
  if (!build_slots(*T, h, SL)) return PACES_E_UNSUPPORTED;
  for (int t = 0; t < h->nterms; ++t)
    if (SL->S[t] < 2) return PACES_E_UNSUPPORTED;  // every term starts a run of slots
  // the codes (upper phonon counts 0 .. sqrt_len - 1) take the top bits; a column is at most
  // R - 1 <= 2^col_bits - 2, so no entry is all ones (PACES_SLOT_EMPTY)
  int cb = 1;
  while ((1 << cb) < h->sqrt_table_len) ++cb;
  const int col_bits = 32 - cb;
  if (R > ((int64_t)1 << col_bits) - 1 || R > 2147483647LL) return PACES_E_UNSUPPORTED;
  S->nterms = h->nterms;
  S->nslots = T->n <= 4 ? 4 : 8;
  S->col_bits = col_bits;
  S->sqrt_len = h->sqrt_table_len;
  for (int t = 0; t <= PACES_MAX_TERMS; ++t)
    S->slot_off[t] = t < h->nterms ? SL->foff[t] : T->n;
  for (int t = 0; t < PACES_MAX_TERMS; ++t)
    S->vib[t] = (t < h->nterms && h->terms[t].kind == PACES_TERM_VIB) ? 1 : 0;
  S->stride = (int32_t)R;
  return PACES_OK;
}

int paces_holstein_slot_layout(const paces_holstein_t* h, int64_t R, paces_slots_t* S) {
  OpTable T;
  SlotArg SL;
  return slot_layout(h, R, &T, &SL, S);
}

int paces_holstein_pairs_slots(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                               const int32_t* cand_pos, const uint8_t* in_basis,
                               const int32_t* img_src, const uint8_t* img_nup,
                               const paces_prior_images_t* prior, const paces_holstein_t* h,
                               uint32_t* slots, void* stream) {
  OpTable T;
  SlotArg SL;
  paces_slots_t S;
  if (!run_offsets || nbasis < 0) return PACES_E_INVALID;
  int rc = slot_layout(h, R, &T, &SL, &S);
  if (rc) return rc;
  if (R > 0 && (!slots || ((uintptr_t)slots & 15))) return PACES_E_INVALID;
  ImgRuns IR;
  rc = img_runs_arg(T, run_offsets, &IR);
  if (rc) return rc;
  PriorImgs A;
  rc = prior_arg(T, prior, &A);
  if (rc) return rc;
  const int64_t nimg = A.n + IR.off[IR.n];
  if (nimg > 0 && (!cand_pos || !in_basis || (IR.off[IR.n] > 0 && (!img_src || !img_nup))))
    return PACES_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = R * S.nslots;
  if (n > 0)
    hipLaunchKernelGGL(k_slots_clear, dim3(grid_for((n + 3) / 4, kBlock)), dim3(kBlock), 0, st,
                       slots, n);
  if (nimg > 0 && R > 0)
    hipLaunchKernelGGL(k_pairs_slot_place, dim3(grid_for(nimg, kBlock)), dim3(kBlock), 0, st, R,
                       nbasis, IR, A, T, SL, S.col_bits, cand_pos, in_basis, img_src,
                       img_nup, slots);
  return launch_status();
}

int paces_holstein_slots_place_explicit(int64_t R, const int64_t* run_offsets,
                                        const int32_t* src_pos, const int32_t* img_pos,
                                        const int32_t* img_map, const int32_t* img_src,
                                        const uint8_t* img_nup, const paces_holstein_t* h,
                                        uint32_t* slots, void* stream) {
  OpTable T;
  SlotArg SL;
  paces_slots_t S;
  if (!run_offsets) return PACES_E_INVALID;
  int rc = slot_layout(h, R, &T, &SL, &S);
  if (rc) return rc;
  ImgRuns IR;
  rc = img_runs_arg(T, run_offsets, &IR);
  if (rc) return rc;
  const int64_t nimg = IR.off[IR.n];
  if (IR.off[0] != 0) return PACES_E_INVALID;
  if (nimg > 0 && (!src_pos || !img_pos || !img_src || !img_nup || !slots || R <= 0))
    return PACES_E_INVALID;
  if (nimg > 0)
    hipLaunchKernelGGL(k_slots_place_explicit, dim3(grid_for(nimg, kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, R, IR, T, SL, S.col_bits, src_pos, img_pos, img_map,
                       img_src, img_nup, slots);
  return launch_status();
}

size_t paces_holstein_csr_workspace_bytes(int64_t R, const paces_holstein_t* h) {
  OpTable T;
  if (!h || build_nb_ops(h, &T)) return 0;
  return csr_ws(R, T.n, h->nterms);
}

int paces_holstein_csr_count(const uint64_t* rows, int64_t R, const uint8_t* in_basis,
                             const paces_layout_t* lay, const paces_holstein_t* h,
                             int32_t* term_nnz, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_h(lay, h);
  if (rc) return rc;
  OpTable T;
  rc = build_nb_ops(h, &T);
  if (rc) return rc;
  for (int t = 0; t < h->nterms; ++t) {
    int m = 0;
    for (int k = 0; k < T.n; ++k) m += (T.term[k] == t);
    if (m > 8) return PACES_E_UNSUPPORTED;
  }
  if (R < 0 || (R > 0 && (!rows || !in_basis)) || !term_nnz || !ws) return PACES_E_INVALID;
  if ((int64_t)h->nterms * R >= 2147483647LL || (int64_t)T.n * R >= 2147483647LL)
    return PACES_E_INVALID;
  if (ws_bytes < csr_ws(R, T.n, h->nterms)) return PACES_E_WORKSPACE;
  PACES_DISPATCH_W(lay->W, run_csr_count, rows, R, in_basis, larg(lay), marg(h), T, h->nterms,
                   term_nnz, ws, (hipStream_t)stream);
}

int paces_holstein_csr_fill(const uint64_t* rows, int64_t R, const paces_layout_t* lay,
                            const paces_holstein_t* h, int32_t* const* indptr,
                            int32_t* const* indices, double* const* data, void* ws,
                            size_t ws_bytes, void* stream) {
  int rc = check_h(lay, h);
  if (rc) return rc;
  OpTable T;
  rc = build_nb_ops(h, &T);
  if (rc) return rc;
  if (R < 0 || (R > 0 && !rows) || !indptr || !indices || !data || !ws) return PACES_E_INVALID;
  if (ws_bytes < csr_ws(R, T.n, h->nterms)) return PACES_E_WORKSPACE;
  if (!h->d_sqrt_table) return PACES_E_INVALID;
  CsrOut out;
  for (int t = 0; t < PACES_MAX_TERMS; ++t) {
    out.indptr[t] = t < h->nterms ? indptr[t] : nullptr;
    out.indices[t] = t < h->nterms ? indices[t] : nullptr;
    out.data[t] = t < h->nterms ? data[t] : nullptr;
    out.data32[t] = nullptr;
    out.code[t] = nullptr;
    if (t < h->nterms && !out.indptr[t]) return PACES_E_INVALID;
  }
  out.coded = 0;
  return run_csr_fill(R, T, marg(h), h->nterms, rows, lay->W, h->batched ? tbits(h->ntraj) : 0,
                      h->d_sqrt_table, out, ws, (hipStream_t)stream);
}

int paces_holstein_diag(const uint64_t* rows, int64_t R, const paces_layout_t* lay,
                        const paces_holstein_t* h, double* diag, void* stream) {
  int rc = check_h(lay, h);
  if (rc) return rc;
  if (R < 0 || (R > 0 && (!rows || !diag))) return PACES_E_INVALID;
  if (!lay->d_sum_mask || lay->sum_bits < 1 || lay->sum_bits > 32) return PACES_E_INVALID;
  PACES_DISPATCH_W(lay->W, run_diag, rows, R, larg(lay), h->batched ? 1 : 0, h->eps_sys,
                   h->hbar_omega, h->delta_eps, diag, nullptr, (hipStream_t)stream);
}

int paces_holstein_diag32(const uint64_t* rows, int64_t R, const paces_layout_t* lay,
                          const paces_holstein_t* h, float* diag, void* stream) {
  int rc = check_h(lay, h);
  if (rc) return rc;
  if (R < 0 || (R > 0 && (!rows || !diag))) return PACES_E_INVALID;
  if (!lay->d_sum_mask || lay->sum_bits < 1 || lay->sum_bits > 32) return PACES_E_INVALID;
  PACES_DISPATCH_W(lay->W, run_diag, rows, R, larg(lay), h->batched ? 1 : 0, h->eps_sys,
                   h->hbar_omega, h->delta_eps, nullptr, diag, (hipStream_t)stream);
}

size_t paces_holstein_observables_workspace_bytes(int64_t R, const paces_holstein_t* h) {
  if (!h) return 0;
  return obs_ws(R, h->nchain);
}

int paces_holstein_observables(const uint64_t* rows, const double* psi, int64_t R,
                               const paces_layout_t* lay, const paces_holstein_t* h, double* out,
                               void* ws, size_t ws_bytes, void* stream) {
  int rc = check_h(lay, h);
  if (rc) return rc;
  if (R < 0 || (R > 0 && (!rows || !psi)) || !out || !ws) return PACES_E_INVALID;
  if (ws_bytes < obs_ws(R, h->nchain)) return PACES_E_WORKSPACE;
  PACES_DISPATCH_W(lay->W, run_obs, rows, psi, R, larg(lay), h->nchain, out, ws,
                   (hipStream_t)stream);
}

size_t paces_holstein_features_workspace_bytes(int64_t R) { return a256((size_t)(R > 0 ? R : 1) * 8); }

static int features_common(const uint64_t* rows, int64_t R, const double* psi, const double* diag,
                           int64_t nbasis, const int64_t* run_offsets, const int32_t* cand_pos,
                           const int32_t* img_src, const uint8_t* in_basis,
                           const paces_layout_t* lay, const paces_holstein_t* h, uint16_t* X,
                           int32_t* row_index, void* ws, hipStream_t st) {
  OpTable T;
  int rc = build_image_ops(h, &T);
  if (rc) return rc;
  const int64_t nimg = run_offsets[T.n] - run_offsets[0];
  if (nimg < 0) return PACES_E_INVALID;
  if (nimg > 0 && (!cand_pos || !img_src)) return PACES_E_INVALID;
  char* p = (char*)ws;
  u64* par = (u64*)p;
  p += a256((size_t)R * 8);
  int32_t* G = nullptr;
  if (in_basis) {
    G = (int32_t*)p;
    p += a256((size_t)(R + 1) * 4);
    rc = scan_u8(in_basis, R, G, p, scan_ws_bytes(R), st);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(R, kBlock)), dim3(kBlock), 0, st, par, R);
  if (nimg > 0)
    hipLaunchKernelGGL(k_parent_amp, dim3(grid_for(nimg, kBlock)), dim3(kBlock), 0, st, nbasis,
                       nimg, cand_pos, img_src, (const double2*)psi, in_basis ? 1 : 0, par);
  rc = launch_status();
  if (rc) return rc;
  float gfeat = 0.f, jfeat = 0.f;
  for (int t = h->nterms - 1; t >= 0; --t) {
    if (h->terms[t].kind == PACES_TERM_VIB) gfeat = (float)h->terms[t].value;
    if (h->terms[t].kind == PACES_TERM_HOP && h->terms[t].order == 1)
      jfeat = (float)h->terms[t].value;
  }
  const double inv_scale = 1.0 / (h->hbar_omega * (double)h->nchain);
  PACES_DISPATCH_W(lay->W, run_features, rows, R, psi, diag, (const u64*)par, larg(lay),
                   h->nchain, h->periodic ? 1 : 0, h->eps_sys, inv_scale, gfeat, jfeat, in_basis,
                   (const int32_t*)G, 0, h->hbar_omega, h->delta_eps, row_index, X, st);
}

int paces_holstein_features(const uint64_t* rows, int64_t R, const double* psi,
                            const double* diag, int64_t nbasis, const int64_t* run_offsets,
                            const int32_t* cand_pos, const int32_t* img_src,
                            const paces_layout_t* lay, const paces_holstein_t* h, uint16_t* X,
                            void* ws, size_t ws_bytes, void* stream) {
  int rc = check_h(lay, h);
  if (rc) return rc;
  if (h->batched) return PACES_E_UNSUPPORTED;
  if (R < 0 || nbasis < 0 || !run_offsets || !ws) return PACES_E_INVALID;
  if (R == 0) return PACES_OK;
  if (!rows || !psi || !diag || !X) return PACES_E_INVALID;
  if (ws_bytes < paces_holstein_features_workspace_bytes(R)) return PACES_E_WORKSPACE;
  return features_common(rows, R, psi, diag, nbasis, run_offsets, cand_pos, img_src, nullptr, lay,
                         h, X, nullptr, ws, (hipStream_t)stream);
}

size_t paces_holstein_features_added_workspace_bytes(int64_t R) {
  const int64_t n = R > 0 ? R : 1;
  return a256((size_t)n * 8) + a256((size_t)(n + 1) * 4) + a256(scan_ws_bytes(n));
}

int paces_holstein_features_added(const uint64_t* rows, int64_t R, const double* psi,
                                  int64_t nbasis, const int64_t* run_offsets,
                                  const int32_t* cand_pos, const int32_t* img_src,
                                  const uint8_t* in_basis, const paces_layout_t* lay,
                                  const paces_holstein_t* h, uint16_t* X, int32_t* row_index,
                                  void* ws, size_t ws_bytes, void* stream) {
  int rc = check_h(lay, h);
  if (rc) return rc;
  if (h->batched) return PACES_E_UNSUPPORTED;
  if (R < 0 || nbasis < 0 || !run_offsets || !ws) return PACES_E_INVALID;
  if (R >= 2147483647LL) return PACES_E_INVALID;
  if (R == 0) return PACES_OK;
  if (!rows || !psi || !in_basis || !X || !row_index) return PACES_E_INVALID;
  if (!lay->d_sum_mask || lay->sum_bits < 1 || lay->sum_bits > 32) return PACES_E_INVALID;
  if (ws_bytes < paces_holstein_features_added_workspace_bytes(R)) return PACES_E_WORKSPACE;
  return features_common(rows, R, psi, nullptr, nbasis, run_offsets, cand_pos, img_src, in_basis,
                         lay, h, X, row_index, ws, (hipStream_t)stream);
}

}  // extern "C"

