/* paces_hip.h -- C ABI of libpaces_hip.so, the MI355X (gfx950) engine behind paces_amd.
 *
 * Every entry point:
 *   - returns an int status (PACES_OK == 0, negative PACES_E_* on error; paces_strerror()),
 *     and never throws across the ABI;
 *   - takes device pointers owned by the caller, element counts as int64 and a hipStream_t
 *     passed as void* (0 = the null stream);
 *   - enqueues stream-ordered work and does not synchronise the host;
 *   - allocates no device memory: temporaries live in a caller workspace whose size the
 *     matching *_workspace_bytes() function returns.
 * Rows are packed Fock states: W uint64 words per row, row-major, field 0 (exciton position) in
 * the most significant bits of word 0. Build word k equals the reference's
 * (w32[2k] << 32) | w32[2k+1], so lexicographic row order is the reference's.
 * Complex128 vectors are interleaved (re, im) doubles.
 */
#ifndef PACES_HIP_H
#define PACES_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PACES_OK 0
#define PACES_E_INVALID -1     /* bad argument (size, null pointer, out-of-range value) */
#define PACES_E_UNSUPPORTED -2 /* row width or term count beyond the compiled limits */
#define PACES_E_HIP -3         /* a HIP launch or runtime call failed */
#define PACES_E_WORKSPACE -4   /* workspace smaller than *_workspace_bytes() */
#define PACES_E_SCRATCH -5     /* size beyond the fixed scratch of a no-allocation entry point */

#define PACES_MAX_ROW_WORDS 16
#define PACES_MAX_RUNS 32
#define PACES_MAX_TERMS 4
#define PACES_MAX_SITES 512

#define PACES_TERM_HOP 0
#define PACES_TERM_VIB 1

const char* paces_strerror(int code);
int paces_version(void);

/* Packed-row layout: per-field bit offset from the MSB of word 0, width and dimension
 * (replaces HamiltonianFramework's bitwidths/totwordwidth, paces/core/hamiltonian.py:134-160). */
typedef struct {
  int32_t W;               /* uint64 words per row, 1..PACES_MAX_ROW_WORDS */
  int32_t nfields;         /* 1 + number of sites */
  const int32_t* d_off;    /* device [nfields] */
  const int32_t* d_width;  /* device [nfields] */
  const int32_t* d_maxdim; /* device [nfields] */
  /* device [W * sum_bits]: mask of the bits of word w that carry significance 2^k inside a
   * field 1..nfields-1 (all fields but the first); sum_i n_i = sum_{w,k} popcount(row[w] &
   * mask[w*sum_bits+k]) << k. May be null for calls that do not sum fields. */
  const uint64_t* d_sum_mask;
  int32_t sum_bits;
} paces_layout_t;

/* One off-diagonal Holstein term, in use_terms order (paces/models/holstein.py:151-277). */
typedef struct {
  int32_t kind;  /* PACES_TERM_HOP or PACES_TERM_VIB */
  int32_t order; /* hopping base order (1: hopping, 2: hopping2) */
  double value;  /* J (hopping) or g (vib_coupling); real (paces/core/time_evolution.py:190-196) */
} paces_term_t;

/* Holstein lattice and terms (paces/models/holstein.py:24-66, 280-287). */
typedef struct {
  int32_t nchain, lattice_l, geo_dim, periodic;
  int32_t nterms;
  paces_term_t terms[PACES_MAX_TERMS];
  double eps_sys, hbar_omega, delta_eps;
  const double* d_sqrt_table; /* device: sqrt(m) for m = 0..sqrt_table_len-1, correctly rounded */
  int32_t sqrt_table_len;
  /* Batched trajectories (no reference counterpart): field 0 is the trajectory id, field 1 the
   * exciton position and fields 2.. the phonons of nchain (= largest) sites. Each trajectory has
   * its own lattice length and term values; geo_dim, periodic, the diagonal and the dims are
   * shared. */
  int32_t batched;
  int32_t ntraj;
  const int32_t* d_traj_L;    /* device [ntraj] */
  const double* d_traj_value; /* device [ntraj * nterms], term values in use_terms order */
} paces_holstein_t;

/* ---- scans (prefix sums behind every mask compaction) ----------------------------------------- */
size_t paces_scan_workspace_bytes(int64_t n);
/* out[0..n] = exclusive prefix sums of in[0..n-1]; out[n] = total. */
int paces_exclusive_scan_u8(const uint8_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes,
                            void* stream);
int paces_exclusive_scan_i32(const int32_t* in, int64_t n, int32_t* out, void* ws,
                             size_t ws_bytes, void* stream);

/* ---- A2-A5: pack / field read / field add --------------------------------------------------------
 * paces_pack_rows replaces compress_ind_arr (paces/core/hamiltonian.py:218-248);
 * paces_get_field replaces get_n_at_site (hamiltonian.py:291-318) and, with field == -1,
 * get_phonon_at_exc (paces/models/_kernels/_holstein_kernels.py:153-179);
 * paces_add_field replaces add_n_at_site (hamiltonian.py:321-360) and, with field == -1,
 * add/rem_phonon_at_exc (_holstein_kernels.py:96-149). */
int paces_pack_rows(const int32_t* raw, int64_t n, const paces_layout_t* lay, uint64_t* rows,
                    void* stream);
int paces_get_field(const uint64_t* rows, int64_t n, const paces_layout_t* lay, int32_t field,
                    int32_t* out, void* stream);
int paces_add_field(uint64_t* rows, int64_t n, const paces_layout_t* lay, int32_t field,
                    int32_t delta, void* stream);

/* ---- B1: lexicographic lower_bound (searchsorted_full_binary, paces/aux/cupy_search.py:78-253)
 * out_idx[i] in [0, M]; out_hit (optional) = row equality at out_idx; miss_count (optional,
 * device int32) is incremented once per query not found -- the host-visible replacement of the
 * device assert of full_binary_noesc (cupy_search.py:205-228). q_sorted != 0 promises ascending
 * queries (every call on the step): a partition pass then bounds each workgroup's range, which
 * is staged in LDS; it needs a workspace of paces_lower_bound_workspace_bytes(Q) (ws may be null
 * for unsorted queries). */
size_t paces_lower_bound_workspace_bytes(int64_t Q);
int paces_lower_bound(const uint64_t* book, int64_t M, const uint64_t* q, int64_t Q, int32_t W,
                      int32_t q_sorted, int32_t* out_idx, uint8_t* out_hit, int32_t* miss_count,
                      void* ws, size_t ws_bytes, void* stream);

/* ---- A7: union of K sorted unique runs (cupy_unique, paces/aux/helpers.py:63-75, applied to
 * the concatenation built in hamiltonian.py:434-452 and 563-577). out_rows needs capacity for
 * sum(len) rows; *out_count (device) receives the number of unique rows; out_from_run0[p] = 1
 * when output row p came from run 0; cand_pos (optional, sum(len) entries) receives the output
 * index of every input row, duplicates included. The workspace size depends on the total, the
 * run count and the row width. */
typedef struct {
  const uint64_t* ptr[PACES_MAX_RUNS];
  int64_t len[PACES_MAX_RUNS];
  int32_t K;
} paces_runs_t;
size_t paces_merge_unique_workspace_bytes(int64_t total, int32_t K, int32_t W);
int paces_merge_unique(const paces_runs_t* runs, int32_t W, uint64_t* out_rows,
                       int32_t* out_count, uint8_t* out_from_run0, int32_t* cand_pos, void* ws,
                       size_t ws_bytes, void* stream);
/* The same union; every output row also gets out_code[row] = run_codes[the run of its kept copy]
 * (run_codes: K host bytes). With the codes of paces_holstein_inverse_skip_codes this is the skip
 * input of the next expansion (paces_holstein_images_*_skip). No counterpart in the reference. */
int paces_merge_unique_coded(const paces_runs_t* runs, int32_t W, uint64_t* out_rows,
                             int32_t* out_count, uint8_t* out_from_run0, int32_t* cand_pos,
                             const uint8_t* run_codes, uint8_t* out_code, void* ws,
                             size_t ws_bytes, void* stream);

/* ---- A7 general sort + dedup (cupy_unique, paces/aux/helpers.py:63-75, as called by
 * enlarge_basis_set, paces/core/hamiltonian.py:550-577), for rows in no particular order:
 * LSD radix sort on 4-bit digits with a stable scatter per pass. paces_rows_varying_bits
 * writes (device, W words) the bits that differ between rows; handing a HOST copy of it to
 * paces_sort_rows skips every digit at which all rows agree (null: all 16 W digits). out_rows
 * (capacity n, not aliasing rows) receives the sorted rows, out_perm (optional) the source index
 * of each; with unique != 0 only the first of equal rows is kept and *out_count (device)
 * receives the number of rows written. */
int paces_rows_varying_bits(const uint64_t* rows, int64_t n, int32_t W, uint64_t* mask,
                            void* stream);
size_t paces_sort_rows_workspace_bytes(int64_t n, int32_t W);
int paces_sort_rows(const uint64_t* rows, int64_t n, int32_t W, const uint64_t* varying,
                    int32_t unique, uint64_t* out_rows, int32_t* out_perm, int32_t* out_count,
                    void* ws, size_t ws_bytes, void* stream);

/* ---- F1: top-k select (_determine_select_whoami, paces/core/time_evolution.py:833-910,
 * shuffle_seed=None). Keeps the k rows of largest |psi|^2 = re*re + im*im; ties at the cut keep
 * the highest indices (stable ascending argsort, last k). Writes min(k, R) rows in index order,
 * which for a sorted basis is lexicographic order; out_index (optional) gets their indices. */
size_t paces_select_workspace_bytes(int64_t R);
int paces_select_topk(const double* psi, const uint64_t* rows, int64_t R, int32_t W, int64_t k,
                      uint64_t* out_rows, int32_t* out_index, void* ws, size_t ws_bytes,
                      void* stream);

/* ---- C1/C2/A8: Holstein candidate images (enlarge_basis_set with raw_map_to=True,
 * paces/core/hamiltonian.py:550-577; paces/models/holstein.py:151-277). Each image run of a
 * sorted unique basis is itself sorted and unique; the runs are written back to back.
 * Two phases: _count fills the workspace and run_offsets (device, nruns+1, [nruns] = total);
 * the caller reads the total, sizes out_images, then calls _emit with the same workspace. */
int paces_holstein_num_runs(const paces_holstein_t* h);
size_t paces_holstein_images_workspace_bytes(int64_t n, const paces_holstein_t* h);
int paces_holstein_images_count(const uint64_t* basis, int64_t n, const paces_layout_t* lay,
                                const paces_holstein_t* h, int32_t* run_offsets, void* ws,
                                size_t ws_bytes, void* stream);
/* out_src (optional) gets the basis index each image came from, out_nup (optional) the upper
 * phonon count of a vib pair (0 for hops). */
int paces_holstein_images_emit(const uint64_t* basis, int64_t n, const paces_layout_t* lay,
                               const paces_holstein_t* h, uint64_t* out_images, int32_t* out_src,
                               uint8_t* out_nup, void* ws, size_t ws_bytes, void* stream);
/* The same with a skip code per row (the images of the rows an expansion added: the other rows'
 * images are already in the expanded basis; no reference counterpart): skip[i] == 0 gives every
 * image of row i, 1 none, 2 + k all but op k's (a duplicate of a basis row; see
 * paces_holstein_inverse_skip_codes). out_src is the row's index in basis. Same workspace. */
int paces_holstein_images_count_skip(const uint64_t* basis, int64_t n, const uint8_t* skip,
                                     const paces_layout_t* lay, const paces_holstein_t* h,
                                     int32_t* run_offsets, void* ws, size_t ws_bytes,
                                     void* stream);
int paces_holstein_images_emit_skip(const uint64_t* basis, int64_t n, const uint8_t* skip,
                                    const paces_layout_t* lay, const paces_holstein_t* h,
                                    uint64_t* out_images, int32_t* out_src, uint8_t* out_nup,
                                    void* ws, size_t ws_bytes, void* stream);
/* codes[0..K) for the union of [basis, op 0's images, op 1's images, ...] (K = number of ops + 1,
 * returned): 1 for the basis run, 2 + (k ^ 1) for the run of a positive-direction op k, else 0.
 * Ops come in inverse pairs (k, k ^ 1): a row added as op k's image of a basis row s has s as its
 * op k ^ 1 image; for positive k that image adds neither a row nor a CSR entry. */
int paces_holstein_inverse_skip_codes(const paces_holstein_t* h, uint8_t* codes, int32_t cap);

/* ---- D1/D2 from the union (no lookups): given the union of a basis (run 0, nbasis rows) and its
 * image runs (paces_merge_unique with cand_pos), the per-term CSR on the R union rows. Each image
 * of basis row b is the pair {b, image}: entries (b, image) and (image, b), except a minus-op
 * image that is itself a basis row (the o_star rule of hamiltonian.py:363-374, holstein.py:201-206,
 * 266-268; those pairs come from the image's plus op). run_offsets is a HOST array of nruns+1
 * image offsets (from paces_holstein_images_count); in_basis is the union's out_from_run0.
 * Same two-phase protocol as paces_holstein_csr_*; columns ascend within a row. */
size_t paces_holstein_pairs_workspace_bytes(int64_t R, const paces_holstein_t* h);
int paces_holstein_pairs_count(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                               const int32_t* cand_pos, const uint8_t* in_basis,
                               const int32_t* img_src, const paces_holstein_t* h,
                               int32_t* term_nnz, void* ws, size_t ws_bytes, void* stream);
/* basis/lay are read only for batched trajectories (the trajectory of each source row). */
int paces_holstein_pairs_fill(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                              const int32_t* cand_pos, const uint8_t* in_basis,
                              const int32_t* img_src, const uint8_t* img_nup,
                              const uint64_t* basis, const paces_layout_t* lay,
                              const paces_holstein_t* h, int32_t* const* indptr,
                              int32_t* const* indices, double* const* data, void* ws,
                              size_t ws_bytes, void* stream);
/* The same entries with coded values (paces_csr_coded_t): code[t] receives the upper phonon count
 * of each entry of a vib term t (hopping terms write none; their code[t] may be null). Batched
 * trajectories only, at most PACES_CODE_TABLE sqrt_table entries. */
int paces_holstein_pairs_fill_coded(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                                    const int32_t* cand_pos, const uint8_t* in_basis,
                                    const int32_t* img_src, const uint8_t* img_nup,
                                    const paces_holstein_t* h, int32_t* const* indptr,
                                    int32_t* const* indices, uint8_t* const* code, void* ws,
                                    size_t ws_bytes, void* stream);
/* The same entries with FP32 values (each the FP64 value rounded to nearest), for the
 * mixed-precision series, which reads FP32 values only (no reference counterpart: the reference's
 * matrices are FP64). */
int paces_holstein_pairs_fill32(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                                const int32_t* cand_pos, const uint8_t* in_basis,
                                const int32_t* img_src, const uint8_t* img_nup,
                                const uint64_t* basis, const paces_layout_t* lay,
                                const paces_holstein_t* h, int32_t* const* indptr,
                                int32_t* const* indices, float* const* data, void* ws,
                                size_t ws_bytes, void* stream);
/* The image list in two parts (no reference counterpart: the reference forms the images of the
 * whole expanded basis again). When the nbasis rows are themselves the union of a previous basis
 * (nprev rows, run 0) with that basis's images, the images of those nprev rows are already
 * known. The second union then takes only the images of the rows the expansion added
 * (paces_holstein_images_*_skip with the expansion's out_from_run0 as skip); run_offsets,
 * cand_pos, img_src and img_nup describe that union as above. prior describes the first:
 * run_offsets (host) and img_src/img_nup of the previous basis's images, and cand_pos, that
 * union's candidate positions (in the nbasis rows). The pairs then take every image of the
 * nbasis rows; the output is the same matrix as from one full image list. For the fill, exactly
 * one of data (FP64), data32 (FP32) or code (coded; basis and lay are then not read) is given. */
typedef struct {
  int64_t nprev;
  const int64_t* run_offsets;
  const int32_t* cand_pos;
  const int32_t* img_src;
  const uint8_t* img_nup;
} paces_prior_images_t;
int paces_holstein_pairs_count_split(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                                     const int32_t* cand_pos, const uint8_t* in_basis,
                                     const int32_t* img_src, const paces_prior_images_t* prior,
                                     const paces_holstein_t* h, int32_t* term_nnz, void* ws,
                                     size_t ws_bytes, void* stream);
int paces_holstein_pairs_fill_split(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                                    const int32_t* cand_pos, const uint8_t* in_basis,
                                    const int32_t* img_src, const uint8_t* img_nup,
                                    const paces_prior_images_t* prior, const uint64_t* basis,
                                    const paces_layout_t* lay, const paces_holstein_t* h,
                                    int32_t* const* indptr, int32_t* const* indices,
                                    double* const* data, float* const* data32,
                                    uint8_t* const* code, void* ws, size_t ws_bytes,
                                    void* stream);

/* ---- Slot matrices for batched trajectories (no reference counterpart: the reference keeps one
 * CSR per term, time_evolution.py:365-372). Every off-diagonal entry sits in one array of
 * nslots x R uint32 slots (slot m of row i at slots[m * stride + i]), one slot per (row, op):
 * an op moves one key field of a row by a fixed
 * amount (a hop the exciton position, a vib op the phonon count at the exciton), so the column
 * order of a row's neighbours under a term's ops is the same for every row of an open lattice.
 * Slot slot_off[t] + r of row i holds row i's entry of term t whose column has rank r among the
 * term's ops: the column in the low col_bits bits and, for a vib term, the pair's upper phonon
 * count m above them (hopping: 0); PACES_SLOT_EMPTY where row i has no such entry. Read in slot
 * order, a row's slots of term t are its CSR row of term t (paces_holstein_pairs_fill_coded),
 * columns ascending, and the entry values are the coded ones (paces_csr_coded_t). */
#define PACES_SLOT_EMPTY 0xffffffffu
#define PACES_MAX_SLOTS 8
typedef struct {
  int32_t nterms;
  int32_t nslots;                         /* slots per row: 4 or 8 (the ops, padded) */
  int32_t col_bits;                       /* column bits of a slot; the code sits above them */
  int32_t sqrt_len;                       /* entries of sqrt_table */
  int32_t slot_off[PACES_MAX_TERMS + 1];  /* term t's slots: [slot_off[t], slot_off[t + 1]) */
  int32_t vib[PACES_MAX_TERMS];           /* 1: term t's slots carry codes (a vib term) */
  int32_t stride;                         /* rows: slot m of row i is slots[m * stride + i] */
  const uint32_t* slots;                  /* device, nslots x stride */
  const double* term_value;               /* device, ntraj x nterms */
  const double* sqrt_table;               /* device, sqrt_len */
} paces_slots_t;
/* The host fields of S (every field but the three pointers) for model h and R rows (stride R):
 * the ops' slots, padded with slots that stay empty to 4 or 8 per row. Returns
 * PACES_E_UNSUPPORTED when h has no slot form: not batched, a periodic lattice (its wrap-around
 * hops give row-dependent column orders), more than PACES_MAX_SLOTS ops, more than
 * PACES_CODE_TABLE codes, or R beyond what col_bits can index. */
int paces_holstein_slot_layout(const paces_holstein_t* h, int64_t R, paces_slots_t* S);
/* The slot matrix of the pairs: arguments as paces_holstein_pairs_fill_split, with prior null for
 * one image list (paces_holstein_pairs_fill_coded). Writes all nslots x R slots (stride R, nslots
 * from paces_holstein_slot_layout); no count pass, no workspace. */
int paces_holstein_pairs_slots(int64_t R, int64_t nbasis, const int64_t* run_offsets,
                               const int32_t* cand_pos, const uint8_t* in_basis,
                               const int32_t* img_src, const uint8_t* img_nup,
                               const paces_prior_images_t* prior, const paces_holstein_t* h,
                               uint32_t* slots, void* stream);

/* Pairs of explicit images into a slot matrix of R rows (stride R): image e of run k (op k; run
 * offsets on the host, ops + 1 entries, the first 0) of source img_src[e], where src_pos[s] is
 * the position of source s among the R rows and img_pos[img_map[e]] (img_pos[e] when img_map is
 * null) that of image e; writes slot rank[k] of the source's row and slot rank[k ^ 1] of the
 * image's row as paces_holstein_pairs_slots does. A position outside [0, R) skips the pair. */
int paces_holstein_slots_place_explicit(int64_t R, const int64_t* run_offsets,
                                        const int32_t* src_pos, const int32_t* img_pos,
                                        const int32_t* img_map, const int32_t* img_src,
                                        const uint8_t* img_nup, const paces_holstein_t* h,
                                        uint32_t* slots, void* stream);

/* ---- Incremental adaptation of the batched step (no reference counterpart: the reference
 * rebuilds each step's basis by expansions and a sort of every candidate, hamiltonian.py:383-492,
 * time_evolution.py:735-807). A row of the previous expanded basis B_e has every neighbour in the
 * previous basis and lists it in the previous slot matrix, so the new basis is found by marking
 * flag bytes over the previous rows through their slots; only marked rows outside the previous
 * B_e (in_be = 0) need their images formed (paces_holstein_images_*), which may be rows of
 * neither basis. The entry points below are the steps of that construction; the result equals
 * the full one (rows, slots, carried-over amplitudes) bit for bit. */
/* Flag bytes over the previous rows hold bit 0x80 for a row of the previous B_e
 * (paces_batch_select_flags writes it). Row i is a source when flags[i] has a bit of src_mask
 * and none of skip_mask: every row a source with 0x80 names in its slots gets mark_bit; a source
 * without 0x80 gets expl[i] = 1 (expl[i] = 0 for every other row) and is counted into
 * *expl_count (device, added to). mark_bit must not overlap src_mask, skip_mask or 0x80. */
int paces_slots_mark(const uint32_t* slots, int64_t stride, int32_t nslots, int32_t col_bits,
                     int64_t R, uint8_t* flags, int32_t src_mask, int32_t skip_mask,
                     int32_t mark_bit, uint8_t* expl, int32_t* expl_count, void* stream);
/* out_idx = the indices i < n with flags[i] & mask, ascending; *out_count (device) = their
 * number */
size_t paces_flag_indices_workspace_bytes(int64_t n);
int paces_flag_indices(const uint8_t* flags, int64_t n, int32_t mask, int32_t* out_idx,
                       int32_t* out_count, void* ws, size_t ws_bytes, void* stream);
/* out[j] = rows[idx[j]] (W words per row) for j < m */
int paces_gather_rows(const uint64_t* rows, int32_t W, const int32_t* idx, int64_t m,
                      uint64_t* out, void* stream);
/* dst[dmap ? dmap[j] : j] = src[smap ? smap[j] : j] for j < m (no reference counterpart: the
 * incremental adaptation composes row positions through it instead of searching for them) */
int paces_move_i32(const int32_t* src, const int32_t* smap, const int32_t* dmap, int64_t m,
                   int32_t* dst, void* stream);
/* dst[dmap[j]] = v (0..255) for j < m (no reference counterpart) */
int paces_fill_u8_at(uint8_t* dst, const int32_t* dmap, int64_t m, int32_t v, void* stream);
/* flags[idx[j]] |= bit where hit[j]; miss[j] = !hit[j] (j < m) */
int paces_mark_hits(const int32_t* idx, const uint8_t* hit, int64_t m, uint8_t* flags,
                    int32_t bit, uint8_t* miss, void* stream);
/* Positions in the new basis: the rows i < R with flags[i] & 0x7f, in order, with the nY new rows
 * merged in at ins[j] = their lower_bound among the R rows (ascending). NP[i] (i <= R) = the
 * flagged rows before i + #{j : ins[j] <= i}, row i's position when flagged; ypos[j] = new row
 * j's position; *out_total (device) = the new basis's rows. */
size_t paces_warm_positions_workspace_bytes(int64_t R);
int paces_warm_positions(const uint8_t* flags, int64_t R, const int32_t* ins, int64_t nY,
                         int32_t* NP, int32_t* ypos, int32_t* out_total, void* ws,
                         size_t ws_bytes, void* stream);
/* Every flagged row i < R (flags & 0x7f) of the previous basis to position NP[i]: its row,
 * amplitude (the reassignment of time_evolution.py:781-797), diagonal (float when diag_f32,
 * else double), out_be = (flags & be_mask) != 0, and its slots: an entry whose pair has an
 * endpoint with flags & be_mask is kept with its column renumbered through NP, any other slot
 * is empty. out_x0 (optional): the amplitude rounded to complex64 (a mixed series' first term
 * input, for paces_batch_series_launch's prebegun mode). */
int paces_warm_emit(const uint64_t* rows, int32_t W, const double* psi, const void* diag,
                    int32_t diag_f32, int64_t R, const uint8_t* flags, int32_t be_mask,
                    const int32_t* NP, const uint32_t* slots, int64_t stride_old, int32_t nslots,
                    int32_t col_bits, uint64_t* out_rows, double* out_psi, void* out_diag,
                    uint8_t* out_be, uint32_t* out_slots, int64_t stride_new, void* out_x0,
                    void* stream);
/* The nY new rows Y to positions ypos: amplitude 0 (and 0 in out_x0, optional), diagonal ydiag,
 * out_be = ybe, empty slots */
int paces_warm_emit_new(const uint64_t* Y, int32_t W, int64_t nY, const int32_t* ypos,
                        const void* ydiag, int32_t diag_f32, const uint8_t* ybe, int32_t nslots,
                        uint64_t* out_rows, double* out_psi, void* out_diag, uint8_t* out_be,
                        uint32_t* out_slots, int64_t stride_new, void* out_x0, void* stream);

/* ---- D1/D2: per-term canonical CSR of the Holstein H on the new basis (generate_mel +
 * _add_hc_and_generate_coo + coo_matrix().tocsr(): hamiltonian.py:383-547,
 * time_evolution.py:365-372). An entry (i, j) of term t exists when rows i, j are linked by t and
 * at least one of them is in the pre-expansion basis (in_basis[i] != 0). Two phases: _count
 * writes nnz per term to term_nnz (device, nterms entries); _fill writes indptr (R+1 each),
 * indices and data of every term. Columns ascend within a row. */
size_t paces_holstein_csr_workspace_bytes(int64_t R, const paces_holstein_t* h);
int paces_holstein_csr_count(const uint64_t* rows, int64_t R, const uint8_t* in_basis,
                             const paces_layout_t* lay, const paces_holstein_t* h,
                             int32_t* term_nnz, void* ws, size_t ws_bytes, void* stream);
int paces_holstein_csr_fill(const uint64_t* rows, int64_t R, const paces_layout_t* lay,
                            const paces_holstein_t* h, int32_t* const* indptr,
                            int32_t* const* indices, double* const* data, void* ws,
                            size_t ws_bytes, void* stream);

/* ---- A6: diagonal eps + hbar_omega*sum(n) (+ delta_eps*pos) (generate_mel_diag,
 * paces/models/holstein.py:280-287; sum_occ_pte_kernel, _holstein_kernels.py:183-210). */
int paces_holstein_diag(const uint64_t* rows, int64_t R, const paces_layout_t* lay,
                        const paces_holstein_t* h, double* diag, void* stream);
/* The same diagonal rounded to FP32 (for the mixed-precision series; no reference counterpart). */
int paces_holstein_diag32(const uint64_t* rows, int64_t R, const paces_layout_t* lay,
                          const paces_holstein_t* h, float* diag, void* stream);
/* paces_merge_unique that also writes the diagonal of every output row, as
 * paces_holstein_diag (diag, FP64) or paces_holstein_diag32 (diag32, FP32) would: exactly one of
 * them, each with room for sum(len) values (the output count is not known before the call). The
 * new basis's rows and diagonal come from one pass (no reference counterpart: the reference
 * forms the rows with cupy_unique, hamiltonian.py:434-452, then the diagonal,
 * holstein.py:280-287). Same workspace as paces_merge_unique; lay->W must equal W. */
int paces_merge_unique_diag(const paces_runs_t* runs, int32_t W, uint64_t* out_rows,
                            int32_t* out_count, uint8_t* out_from_run0, int32_t* cand_pos,
                            const paces_layout_t* lay, const paces_holstein_t* h, double* diag,
                            float* diag32, void* ws, size_t ws_bytes, void* stream);

/* ---- Per-position sums behind V_hop and V_vib_coupl (calculate_hopping / calculate_coupling,
 * paces/models/holstein.py:325-365, which apply _calculate_partitioned_sum, 380-396, to
 * conj(psi) * (M psi)): out[2s], out[2s+1] = sum over rows i in [seg_off[s], seg_off[s+1]) of
 * conj(x_i) * y_i, complex128 interleaved. One workgroup per segment, fixed reduction order. */
int paces_segment_vdot(const double* x, const double* y, const int32_t* seg_off, int32_t nseg,
                       double* out, void* stream);
/* <mu(t) mu(0)> (calculate_mu_to_0, paces/models/holstein.py:366-377): out[0..1] = the sum of
 * x over the rows whose bits after the first lead_bits (the position field) are all zero; with
 * rows sorted, that is the first row of a position's segment when it is the phonon vacuum. */
int paces_vacuum_sum(const uint64_t* rows, int32_t W, int32_t lead_bits, const double* x,
                     const int32_t* seg_off, int32_t nseg, double* out, void* stream);

/* ---- H1: observables n_exc[N], n_pho[N], sum |psi|^2 (calculate_n_exc / calculate_n_pho,
 * paces/models/holstein.py:311-322; calc_bath_n_b_kernel, _holstein_kernels.py:215-255).
 * out has 2N+1 doubles. Deterministic (fixed reduction order). */
size_t paces_holstein_observables_workspace_bytes(int64_t R, const paces_holstein_t* h);
int paces_holstein_observables(const uint64_t* rows, const double* psi, int64_t R,
                               const paces_layout_t* lay, const paces_holstein_t* h, double* out,
                               void* ws, size_t ws_bytes, void* stream);

/* ---- G1: reassignment (paces/core/time_evolution.py:781-797): new_psi[i] = old_psi[j] where
 * old_rows[j] == new_rows[i], else 0. Both row sets sorted. */
size_t paces_reassign_workspace_bytes(int64_t R);
int paces_reassign(const uint64_t* old_rows, int64_t M, const double* old_psi,
                   const uint64_t* new_rows, int64_t R, int32_t W, double* new_psi, void* ws,
                   size_t ws_bytes, void* stream);
/* No counterpart in the reference: paces_reassign for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It takes
 * no workspace: its per-tile bounds live in a static device array of
 * PACES_REASSIGN_NOALLOC_SCRATCH_BYTES (16 bytes per 256 new rows, so R up to
 * PACES_REASSIGN_NOALLOC_MAX_ROWS), held by each GPU the library runs on until the process ends.
 * R beyond that returns PACES_E_SCRATCH. It calls no hipMalloc or hipFree. Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (different streams, or host threads)
 * race and must be serialised by the caller. Results equal paces_reassign's bit for bit. */
#define PACES_REASSIGN_NOALLOC_SCRATCH_BYTES (4u << 20)
#define PACES_REASSIGN_NOALLOC_MAX_ROWS (((int64_t)PACES_REASSIGN_NOALLOC_SCRATCH_BYTES / 16 - 1) * 256)
int paces_reassign_noalloc(const uint64_t* old_rows, int64_t M, const double* old_psi,
                           const uint64_t* new_rows, int64_t R, int32_t W, double* new_psi,
                           void* stream);

/* ---- E1/E2: H.x and the Taylor series (total_ham / _simple_taylor,
 * paces/core/time_evolution.py:382-400, 482-512). FP64 parity order: per row
 * y = d*x, then each term's entries summed left to right from zero and added, no FMA. */
typedef struct {
  int32_t nterms;
  const int32_t* indptr[PACES_MAX_TERMS];
  const int32_t* indices[PACES_MAX_TERMS];
  const double* data[PACES_MAX_TERMS];
} paces_csr_set_t;

/* ---- The scaling-and-squaring propagator (expm_multiply_simple, paces/aux/_scipy/
 * _expm_multiply.py:106-202, 395-450; onenormest, paces/aux/_scipy/_onenormest.py:11-103,
 * 202-345). The series itself runs on paces_taylor_term with a shifted diagonal; these are the
 * shift, the exact 1-norm and the block 1-norm estimator's vector operations. A "pair" is two
 * real columns held as one complex128 vector (column 0 = real parts, column 1 = imaginary
 * parts). Reductions have a fixed order: results depend only on the inputs. */
size_t paces_pair_workspace_bytes(void);
/* out[0] = sum of x (mu = trace(A) / n, _expm_multiply.py:150-151) */
int paces_sum_f64(const double* x, int64_t n, double* out, void* ws, size_t ws_bytes,
                  void* stream);
/* out_i = x_i - m (the diagonal of A - mu I, _expm_multiply.py:152) */
int paces_shift_f64(const double* x, int64_t n, double m, double* out, void* stream);
/* out[0] = max_i |diag_i - m| + sum over terms of |row i's values|: the exact 1-norm of the
 * symmetric H - m I (_exact_1_norm, _expm_multiply.py:26-33) */
int paces_csr_shifted_onenorm(int64_t n, const double* diag, double m, const paces_csr_set_t* h,
                              double* out, void* ws, size_t ws_bytes, void* stream);
/* out[0..1] = sum |column 0|, sum |column 1| (_sum_abs_axis0, _onenormest.py:149-158) */
int paces_pair_abs_sums(const double* y, int64_t n, double* out, void* ws, size_t ws_bytes,
                        void* stream);
/* out[0..3] = a0.b0, a0.b1, a1.b0, a1.b1 (vectors_are_parallel, _onenormest.py:167-174) */
int paces_pair_dots(const double* a, const double* b, int64_t n, double* out, void* ws,
                    size_t ws_bytes, void* stream);
/* s = sign(y) per column, sign(0) = 1 (sign_round_up, _onenormest.py:128-142, real case) */
int paces_pair_signs(const double* y, int64_t n, double* s, void* stream);
/* h_i = max(|z_i0|, |z_i1|) (_max_abs_axis1, _onenormest.py:145-146) */
int paces_pair_rowmax(const double* z, int64_t n, double* h, void* stream);
/* *out = the index of the first entry of h after (after_value, after_index) in the order of
 * decreasing value, then decreasing index (all entries when after_index < 0), or -1: repeated,
 * it lists h as argsort(h)[::-1] does (_onenormest.py:312) */
int paces_argmax_after(const double* h, int64_t n, double after_value, int64_t after_index,
                       int64_t* out, void* ws, size_t ws_bytes, void* stream);
/* column col of x = +-scale from a hash of (seed, i) (resample_column, _onenormest.py:198-199;
 * the reference draws from CuPy's generator) */
int paces_pair_random_column(double* x, int64_t n, int32_t col, uint64_t seed, double scale,
                             void* stream);
/* x = (re, im) everywhere, then x[i0] column 0 and x[i1] column 1 = 1 when i0 / i1 >= 0
 * (starting matrix and unit vectors, _onenormest.py:259-280, 324-325) */
int paces_pair_fill(double* x, int64_t n, double re, double im, int64_t i0, int64_t i1,
                    void* stream);


int paces_spmv(int64_t R, const double* diag, const paces_csr_set_t* H, const double* x,
               double* y, void* stream);

/* Taylor state (device, 64 bytes): see paces_taylor_state_t. */
typedef struct {
  unsigned long long max_term2; /* bit pattern of max |term|^2 of the current term */
  unsigned long long max_res2;  /* bit pattern of max |res|^2 after adding it */
  double c1;                    /* inf-norm of the previous term */
  double c1_plus_c2;
  double ratio;
  int32_t done;
  int32_t j_ind;
  double rn; /* the last exact ||res||_inf (the mixed series' even terms divide by it) */
  int32_t pad[2];
} paces_taylor_state_t;

/* Workspace of per-workgroup partial maxima shared by _begin and _term. */
size_t paces_taylor_workspace_bytes(int64_t R);
/* res = v; state.c1 = ||v||_inf; state.done = 0. */
int paces_taylor_begin(int64_t R, const double* v, double* res, paces_taylor_state_t* state,
                       void* ws, size_t ws_bytes, void* stream);
/* One series term j: x_out = (-i*c) * (H x_in); res += x_out; then the convergence test
 * (c1 + c2) / ||res||_inf <= tol. No-op once state.done != 0. */
int paces_taylor_term(int64_t R, const double* diag, const paces_csr_set_t* H, const double* x_in,
                      double* x_out, double* res, double c, int32_t j, double tol,
                      paces_taylor_state_t* state, void* ws, size_t ws_bytes, void* stream);

/* ---- N4 launch shapes (no reference counterpart): the Taylor-term entry points below take the
 * workgroup size (128, 256, 512 or 1024) and rows per thread (1, 2 or 4) the autotuner picked;
 * a null cfg is 256 x 1, the shape of the entry points without the suffix. Every shape gives
 * bit-identical results. */
typedef struct {
  int32_t block;
  int32_t rows_per_thread;
} paces_launch_cfg_t;
int paces_taylor_term_tuned(int64_t R, const double* diag, const paces_csr_set_t* H,
                            const double* x_in, double* x_out, double* res, double c, int32_t j,
                            double tol, paces_taylor_state_t* state, void* ws, size_t ws_bytes,
                            const paces_launch_cfg_t* cfg, void* stream);

/* ---- N2 mixed precision: complex64 series terms with FP32 matrix values and FP32 accumulation,
 * summed into the complex128 result in FP64 (no reference counterpart; the FP64 path above is the
 * reference's arithmetic). Same state and stopping test; tol is the caller's (2^-24 suits FP32
 * terms). */
typedef struct {
  int32_t nterms;
  const int32_t* indptr[PACES_MAX_TERMS];
  const int32_t* indices[PACES_MAX_TERMS];
  const float* data[PACES_MAX_TERMS];
} paces_csr32_set_t;

int paces_convert_f64_to_f32(int64_t n, const double* in, float* out, void* stream);
/* res = v; x0 = (complex64) v; state.c1 = ||v||_inf */
int paces_taylor_begin_mixed(int64_t R, const double* v, double* res, float* x0,
                             paces_taylor_state_t* state, void* ws, size_t ws_bytes, void* stream);
/* x_out = (-i*c) * (H32 x_in) in complex64, summed into res in complex128 on odd terms j only:
 * an odd term adds its input and its output, (res + x_in) + x_out, the same roundings in the same
 * order as adding every term. An even term neither reads nor writes res, and its stopping test
 * divides by the last exact ||res||_inf (state.rn) instead of ||res + x_out||_inf. Even terms must
 * write x_out to one buffer (x_even); after the last term call paces_taylor_finish_mixed. */
int paces_taylor_term_mixed(int64_t R, const float* diag, const paces_csr32_set_t* H,
                            const float* x_in, float* x_out, double* res, float c, int32_t j,
                            double tol, paces_taylor_state_t* state, void* ws, size_t ws_bytes,
                            void* stream);
int paces_taylor_term_mixed_tuned(int64_t R, const float* diag, const paces_csr32_set_t* H,
                                  const float* x_in, float* x_out, double* res, float c,
                                  int32_t j, double tol, paces_taylor_state_t* state, void* ws,
                                  size_t ws_bytes, const paces_launch_cfg_t* cfg, void* stream);
/* res += x_even when the series' last term (state.j_ind, read on the device) was even */
int paces_taylor_finish_mixed(int64_t R, const float* x_even, double* res,
                              const paces_taylor_state_t* state, void* stream);

/* ---- N1 batched trajectories (no reference counterpart). Field 0 of every row is the
 * trajectory id (ceil(log2 ntraj) bits), so trajectory t owns the contiguous segment
 * [seg_off[t], seg_off[t+1]) of the sorted basis and H is block diagonal. Per trajectory:
 * the top-k select (the rule of paces_select_topk with k_per_traj[t]), the observables and the
 * Taylor stopping test; the series runs until every trajectory has stopped. */
typedef struct {
  double c1, c1_plus_c2, ratio;
  int32_t done, j_ind;
  double rn; /* the last exact ||res||_inf (begin: ||v||_inf; then every term of the FP64 form
              * and every odd term of the mixed form); the mixed form's even terms do not read
              * res and divide by it */
} paces_batch_taylor_t;
typedef struct {
  int32_t notdone; /* trajectories still running after the last term */
  int32_t pad[3];
} paces_batch_taylor_hdr_t;

/* seg_off (device, ntraj+1) from sorted rows */
int paces_batch_segments(const uint64_t* rows, int64_t R, int32_t W, int32_t ntraj,
                         int32_t* seg_off, void* stream);
size_t paces_batch_select_workspace_bytes(int64_t R, int32_t ntraj);
/* out_rows needs capacity R; *out_count (device) = sum_t min(k_t, R_t) */
int paces_batch_select_topk(const double* psi, const uint64_t* rows, int64_t R, int32_t W,
                            const int32_t* seg_off, int32_t ntraj, const int64_t* k_per_traj,
                            uint64_t* out_rows, int32_t* out_count, void* ws, size_t ws_bytes,
                            void* stream);
/* The same with the select keys already formed (keys: the bits of |psi|^2 per row, as
 * paces_holstein_observables_batched_finish writes them; psi then unread, may be null) */
int paces_batch_select_topk_keys(const double* psi, const uint64_t* keys, const uint64_t* rows,
                                 int64_t R, int32_t W, const int32_t* seg_off, int32_t ntraj,
                                 const int64_t* k_per_traj, uint64_t* out_rows,
                                 int32_t* out_count, void* ws, size_t ws_bytes, void* stream);
/* The same selection as one flag byte per row: keep[i] = 1 for a selected row, else 0 (every
 * row written), plus 0x80 where in_be (optional) is nonzero; *out_count (device) = the selected
 * rows. Same workspace; keys as above (optional: null forms them from psi). */
int paces_batch_select_flags(const double* psi, const uint64_t* keys, int64_t R,
                             const int32_t* seg_off, int32_t ntraj, const int64_t* k_per_traj,
                             const uint8_t* in_be, uint8_t* keep, int32_t* out_count, void* ws,
                             size_t ws_bytes, void* stream);
/* out: ntraj blocks of 2*nchain+1 doubles (n_exc, n_pho, sum |psi|^2); h->batched must be set */
size_t paces_holstein_observables_batched_workspace_bytes(int64_t R, int32_t ntraj,
                                                          int32_t nsites);
int paces_holstein_observables_batched(const uint64_t* rows, const double* psi, int64_t R,
                                       const int32_t* seg_off, const paces_layout_t* lay,
                                       const paces_holstein_t* h, double* out, void* ws,
                                       size_t ws_bytes, void* stream);
/* The same at the end of a series (no reference counterpart): with ts (the series' per-trajectory
 * states) and terms (a host array of nterms device buffers, as paces_batch_taylor_finish takes
 * them), psi first gets the series' finish in place, the work of paces_batch_taylor_finish
 * (mixed: psi holds v + t_0 + ... + t_(k0-1) and gets t_k0 .. t_j added, j = ts[t].j_ind, so a
 * finish launched while some series still run is completed by a second call with k0 = the terms
 * the first one saw; FP64 (k0 = 0): psi holds res and gets terms[0] where j is even); keys
 * (optional, R entries) receives each row's select key (the bits of |psi|^2) for
 * paces_batch_select_flags / paces_batch_select_topk_keys; vmax (optional, ntraj entries) each
 * trajectory's largest key (0 for an empty one). */
int paces_holstein_observables_batched_finish(uint64_t const* rows, double* psi, int64_t R,
                                              const int32_t* seg_off, const paces_layout_t* lay,
                                              const paces_holstein_t* h,
                                              const void* const* terms, int32_t nterms,
                                              int32_t k0, int32_t mixed,
                                              const paces_batch_taylor_t* ts, uint64_t* keys,
                                              uint64_t* vmax, double* out, void* ws,
                                              size_t ws_bytes, void* stream);
size_t paces_batch_taylor_workspace_bytes(int64_t R, int32_t ntraj);
/* res = v (not for mixed != 0: res is unused and may be null); x0 = v as complex64 (mixed) or
 * complex128; ts[t].c1 = ||v_t||_inf */
int paces_batch_taylor_begin(int64_t R, const double* v, double* res, void* x0, int32_t mixed,
                             const int32_t* seg_off, int32_t ntraj, paces_batch_taylor_t* ts,
                             void* ws, size_t ws_bytes, void* stream);
/* N4: the same with a launch shape (see paces_launch_cfg_t); begin and every term of one series
 * must use the same shape */
int paces_batch_taylor_begin_tuned(int64_t R, const double* v, double* res, void* x0,
                                   int32_t mixed, const int32_t* seg_off, int32_t ntraj,
                                   paces_batch_taylor_t* ts, void* ws, size_t ws_bytes,
                                   const paces_launch_cfg_t* cfg, void* stream);
/* one term for the trajectories still running; exactly one of H64 (FP64 values, complex128
 * terms, double diag) and H32 (FP32 values, complex64 terms, float diag) is non-null.
 * FP64: res is stored on odd terms j only (an even term's sum is formed for the stopping test
 * and completed by the next term, with the same roundings); even terms must write their x_out
 * to one buffer (x_even). Mixed (complex64 terms): no term reads or writes res (it may be
 * null), each term j writes x_out to a buffer of its own (term j + 1 reads it as x_in), and the
 * stopping test divides by ||v||_inf. After the last term call paces_batch_taylor_finish. */
int paces_batch_taylor_term(int64_t R, const void* diag, const paces_csr_set_t* H64,
                            const paces_csr32_set_t* H32, const void* x_in, void* x_out,
                            double* res, double c, int32_t j, double tol, const int32_t* seg_off,
                            int32_t ntraj, paces_batch_taylor_t* ts,
                            paces_batch_taylor_hdr_t* hdr, void* ws, size_t ws_bytes,
                            void* stream);
int paces_batch_taylor_term_tuned(int64_t R, const void* diag, const paces_csr_set_t* H64,
                                  const paces_csr32_set_t* H32, const void* x_in, void* x_out,
                                  double* res, double c, int32_t j, double tol,
                                  const int32_t* seg_off, int32_t ntraj, paces_batch_taylor_t* ts,
                                  paces_batch_taylor_hdr_t* hdr, void* ws, size_t ws_bytes,
                                  const paces_launch_cfg_t* cfg, void* stream);
/* Coded values (no reference counterpart). In the rows of trajectory r, every entry of a hopping
 * term t has the value v = term_value[r * nterms + t], and every entry of a vib term the value
 * sqrt_table[m] * v, m the upper phonon count of its pair: the values paces_holstein_pairs_fill
 * writes. A coded CSR stores m as one byte per vib entry (code[t]; null for a hopping term) in
 * place of the values, and the Taylor term forms each trajectory's few values in LDS. */
#define PACES_CODE_TABLE 64 /* codes 0 .. PACES_CODE_TABLE - 1: sqrt_len is at most this */
typedef struct {
  int32_t nterms;
  int32_t sqrt_len;                      /* entries of sqrt_table */
  const int32_t* indptr[PACES_MAX_TERMS];
  const int32_t* indices[PACES_MAX_TERMS];
  const uint8_t* code[PACES_MAX_TERMS];  /* null for a hopping term; else one byte per entry */
  const double* term_value;              /* device, ntraj x nterms */
  const double* sqrt_table;              /* device, sqrt_len */
} paces_csr_coded_t;
/* paces_batch_taylor_term_tuned with a coded CSR (complex64 terms when mixed != 0, with a float
 * diag; else complex128 and a double diag); the same results bit for bit. */
int paces_batch_taylor_term_coded(int64_t R, const void* diag, int32_t mixed,
                                  const paces_csr_coded_t* H, const void* x_in, void* x_out,
                                  double* res, double c, int32_t j, double tol,
                                  const int32_t* seg_off, int32_t ntraj, paces_batch_taylor_t* ts,
                                  paces_batch_taylor_hdr_t* hdr, void* ws, size_t ws_bytes,
                                  const paces_launch_cfg_t* cfg, void* stream);
/* paces_batch_taylor_term_coded with the slot matrix H (paces_slots_t): each row's slots are
 * loaded in one pass and its neighbours gathered in the next; the same results bit for bit */
int paces_batch_taylor_term_slots(int64_t R, const void* diag, int32_t mixed,
                                  const paces_slots_t* H, const void* x_in, void* x_out,
                                  double* res, double c, int32_t j, double tol,
                                  const int32_t* seg_off, int32_t ntraj, paces_batch_taylor_t* ts,
                                  paces_batch_taylor_hdr_t* hdr, void* ws, size_t ws_bytes,
                                  const paces_launch_cfg_t* cfg, void* stream);
/* after the series (terms and/or a replayed paces_batch_series_launch), with terms a host
 * array of nterms (1 .. PACES_SERIES_MAX_TERMS) device buffers. FP64: adds to res the last term
 * of every trajectory whose last term index ts[t].j_ind is even, from terms[0] (the buffer every
 * even term wrote). Mixed: res holds v and becomes v + t_0 + ... + t_j, j = ts[t].j_ind, added
 * in that order (terms[k] is term k's buffer; nterms must cover every trajectory's last term).
 * Same launch shape and workspace as the series. */
int paces_batch_taylor_finish(int64_t R, const void* const* terms, int32_t nterms, int32_t mixed,
                              double* res, const int32_t* seg_off, int32_t ntraj,
                              const paces_batch_taylor_t* ts, void* ws, size_t ws_bytes,
                              const paces_launch_cfg_t* cfg, void* stream);

/* ---- N1 the batched series as one graph-capturable launch sequence (no reference
 * counterpart). Every argument lives in a device-resident descriptor that the caller rewrites
 * before each step, so a HIP graph captured once from paces_batch_series_launch is replayed at
 * every step. Grids are sized for the capacity R_cap (>= the step's rows); workgroups past the
 * step's segments exit. The launch runs begin and terms 0 .. nterms-1 (FP64: term j reads x[2]
 * for j == 0, else x[(j-1) % 2], and writes x[j % 2], so x_even = x[0]; mixed: term j reads x[2]
 * for j == 0, else xt[j-1], and writes xt[j], and res is unused); terms of stopped
 * trajectories do nothing, and
 * a series not finished after nterms continues with paces_batch_taylor_term_tuned from term
 * nterms with the same state. Unlike paces_batch_taylor_begin, the launch does not copy v into
 * res: terms 0 and 1 read v as the partial sum and term 1 stores res. tol must be below 1 (no
 * series then stops at term 0) and nterms at least 2. */
#define PACES_SERIES_MAX_TERMS 64
typedef struct {
  int64_t R;                              /* rows this step (informational; grids use R_cap) */
  const void* diag;                       /* float (mixed) or double */
  int32_t nterms_h;                       /* Hamiltonian terms in the arrays below */
  int32_t pad0;
  const int32_t* indptr[PACES_MAX_TERMS];
  const int32_t* indices[PACES_MAX_TERMS];
  const void* data[PACES_MAX_TERMS];      /* float (mixed) or double values */
  const double* v;                        /* start vector, complex128 */
  double* res;                            /* result, complex128 */
  void* x[3];                             /* series buffers (complex64 or complex128) */
  const int32_t* seg_off;                 /* ntraj + 1 */
  paces_batch_taylor_t* ts;               /* ntraj */
  paces_batch_taylor_hdr_t* hdr;
  void* ws;                               /* paces_batch_taylor_workspace_bytes(R_cap, ntraj) */
  double tol;
  double c[PACES_SERIES_MAX_TERMS];       /* term coefficients fl(dt / (j + 1)) */
  /* coded != 0: the Hamiltonian is the coded CSR below (paces_csr_coded_t; data[] unused) */
  int32_t coded;
  int32_t sqrt_len;
  const uint8_t* code[PACES_MAX_TERMS];
  const double* term_value;
  const double* sqrt_table;
  /* slots != null (with coded != 0): the Hamiltonian is the slot matrix described by the fields
   * below, sqrt_len, term_value and sqrt_table (paces_slots_t); indptr, indices and code unused */
  const uint32_t* slots;
  int32_t nslots;
  int32_t col_bits;
  int32_t slot_off[PACES_MAX_TERMS + 1];
  int32_t slot_vib[PACES_MAX_TERMS];
  int32_t slot_stride;
  void* xt[PACES_SERIES_MAX_TERMS];       /* mixed: term j's output buffer (complex64) */
  /* prebegun launches: each trajectory's largest |v|^2 as the bits of a double (ntraj), as
   * paces_holstein_observables_batched_finish's vmax gives it */
  const uint64_t* vmax;
} paces_batch_series_t;
size_t paces_batch_series_desc_bytes(void); /* sizeof(paces_batch_series_t), for bindings */
/* prebegun != 0: x[2] already holds v as the series' term input (complex64 or complex128) and
 * vmax its per-trajectory largest |v|^2: the begin pass over v is replaced by the stopping
 * state's initialisation from vmax (the same state) */
int paces_batch_series_launch(const paces_batch_series_t* desc, int64_t R_cap, int32_t ntraj,
                              int32_t mixed, int32_t nslots, int32_t nterms, int32_t prebegun,
                              const paces_launch_cfg_t* cfg, void* stream);

/* ---- N3 candidate scorer (no reference counterpart): 64 inputs -> 3 x 256 GELU -> 1 logit,
 * BF16 matrix operands with FP32 accumulation on the matrix cores. params is the FP32 master
 * vector (PACES_MLP_NPARAM floats):
 *   Wt1 [256 x 64] | b1 [256] | Wt2 [256 x 256] | b2 | Wt3 [256 x 256] | b3 | w4 [256] | b4 [1]
 * (Wt output-major); wbf16 holds BF16 copies of Wt1 | Wt2 | Wt3 (PACES_MLP_NBF16 values).
 * Inputs X are BF16 rows of PACES_MLP_IN values. */
#define PACES_MLP_IN 64
#define PACES_MLP_HIDDEN 256
#define PACES_MLP_NPARAM 148481
#define PACES_MLP_NBF16 147456

/* wbf16 = bf16(Wt1 | Wt2 | Wt3) of params */
int paces_mlp_weights_bf16(const float* params, uint16_t* wbf16, void* stream);
size_t paces_mlp_train_workspace_bytes(int64_t B);
/* Gradient of the weighted mean binary cross-entropy (1/B) sum_b w_b (softplus(z_b) - y_b z_b),
 * w_b = pos_weight for y_b = 1, else 1, over B rows X[idx[b]] with labels labels[idx[b]] (idx may
 * be null: rows 0..B-1). grads receives PACES_MLP_NPARAM floats; *loss (optional, device) the
 * loss. Deterministic (fixed reduction order). */
int paces_mlp_grad(const uint16_t* X, const int32_t* idx, const uint8_t* labels, int64_t B,
                   float pos_weight, const float* params, const uint16_t* wbf16, float* grads,
                   float* loss, void* ws, size_t ws_bytes, void* stream);
/* One Adam step (step counted from 1) on params with moments m, v; refreshes wbf16. */
int paces_mlp_adam(float* params, const float* grads, float* m, float* v, int32_t step, float lr,
                   float beta1, float beta2, float eps, uint16_t* wbf16, void* stream);
/* prob[i] = sigmoid(z_i) and/or logit[i] = z_i for n rows of X (all layers fused). */
int paces_mlp_infer(const uint16_t* X, int64_t n, const float* params, const uint16_t* wbf16,
                    float* prob, float* logit, void* stream);
/* the same with tile_rows (64 or 128; 0 = default) rows per workgroup (N4 tuning knob) */
int paces_mlp_infer_tiled(const uint16_t* X, int64_t n, const float* params,
                          const uint16_t* wbf16, float* prob, float* logit, int32_t tile_rows,
                          void* stream);

/* ---- N3 candidate features and pruning (no reference counterpart).
 * paces_holstein_features writes one row of PACES_MLP_IN bf16 features per union row (R rows,
 * amplitudes psi and diagonal diag on the union): the phonon window of +-16 sites around the
 * exciton over (dim - 1) (-1 outside an open lattice), log10 of the largest neighbour |psi|
 * through the image pairs (nbasis, host run_offsets, cand_pos and img_src as for
 * paces_holstein_pairs_count), the scaled diagonal, log10 |psi| of the row, the vibronic and
 * hopping values and pos / nchain. Single trajectories only. */
size_t paces_holstein_features_workspace_bytes(int64_t R);
int paces_holstein_features(const uint64_t* rows, int64_t R, const double* psi,
                            const double* diag, int64_t nbasis, const int64_t* run_offsets,
                            const int32_t* cand_pos, const int32_t* img_src,
                            const paces_layout_t* lay, const paces_holstein_t* h, uint16_t* X,
                            void* ws, size_t ws_bytes, void* stream);
/* The same features for the added rows of a union only: the rows with in_basis[i] == 0 (in_basis
 * the 0/1 flags of the rows the expansion started from), in order. X holds (R - sum in_basis)
 * rows and row_index[j] = the union row of X's row j. The diagonal is computed from each row.
 * Replaces a full-union paces_holstein_features followed by a gather of the added rows. */
size_t paces_holstein_features_added_workspace_bytes(int64_t R);
int paces_holstein_features_added(const uint64_t* rows, int64_t R, const double* psi,
                                  int64_t nbasis, const int64_t* run_offsets,
                                  const int32_t* cand_pos, const int32_t* img_src,
                                  const uint8_t* in_basis, const paces_layout_t* lay,
                                  const paces_holstein_t* h, uint16_t* X, int32_t* row_index,
                                  void* ws, size_t ws_bytes, void* stream);
/* keep[i] = 1 when prob[i] >= threshold or i is one of forced_idx; else 2 when sample_bits >= 0
 * and the low sample_bits bits of a hash of (i, seed) are zero (a sample of the pruned rows kept
 * to measure their weight); else 0. */
int paces_prune_plan(const float* prob, int64_t n, float threshold, const int32_t* forced_idx,
                     int64_t nforced, int32_t sample_bits, uint64_t seed, uint8_t* keep,
                     void* stream);
/* The plan of paces_prune_plan from the scores of the added rows only (prob[j] for the union
 * row row_index[j], j < m, as paces_holstein_features_added writes them): keep = in_basis
 * (0/1) for the n union rows, then the decision of each scored row at its union position, with
 * the same sample hash of (union row, seed). Equals paces_prune_plan on the scores scattered to
 * the union with the in_basis rows forced. */
int paces_prune_plan_rows(const float* prob, const int32_t* row_index, int64_t m,
                          const uint8_t* in_basis, int64_t n, float threshold,
                          int32_t sample_bits, uint64_t seed, uint8_t* keep, void* stream);
/* out[idx[j]] = src[j] for j < m; other entries are untouched. */
int paces_scatter_u8(const int32_t* idx, int64_t m, const uint8_t* src, uint8_t* out,
                     void* stream);
/* out[idx[j]] = src[idx[j]] (or 1 when src is null) for j < m; other entries are untouched. */
int paces_mark_indices(const int32_t* idx, int64_t m, const uint8_t* src, uint8_t* out,
                       void* stream);
/* Rows with keep != 0 in order, with their psi (optional), in_basis flag (optional) and
 * out_sample = (keep == 2); out_cand_pos[g] = the kept index of cand_pos[g], or -1 when that row
 * was dropped (the pairs kernels then skip the pair). *out_count (device) = kept rows. */
size_t paces_compact_workspace_bytes(int64_t n);
int paces_compact_kept(const uint8_t* keep, int64_t n, const uint64_t* rows, int32_t W,
                       const double* psi, const uint8_t* in_basis, const int32_t* cand_pos,
                       int64_t C, uint64_t* out_rows, double* out_psi, uint8_t* out_in_basis,
                       uint8_t* out_sample, int32_t* out_cand_pos, int32_t* out_count, void* ws,
                       size_t ws_bytes, void* stream);

/* ---- deterministic vector reductions (norms, vdot) and Lanczos helpers (E3,
 * paces/core/time_evolution.py:515-549). out is a device double[2] (re, im) or double[1]. */
size_t paces_reduce_workspace_bytes(int64_t R);
int paces_norm2(const double* x, int64_t R, double* out, void* ws, size_t ws_bytes, void* stream);
/* sum of |x_i|^2 over the rows with mask[i] != 0 (N3: weight of the sampled pruned rows) */
int paces_norm2_masked(const double* x, const uint8_t* mask, int64_t R, double* out, void* ws,
                       size_t ws_bytes, void* stream);
int paces_vdot(const double* x, const double* y, int64_t R, double* out, void* ws,
               size_t ws_bytes, void* stream);
/* y = a*x + b*y with complex scalars given on the host; with b == 0, y is only written */
int paces_axpby(int64_t R, double ar, double ai, const double* x, double br, double bi, double* y,
                void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PACES_HIP_H */
