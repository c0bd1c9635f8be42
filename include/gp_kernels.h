/*
 * Internal C-ABI between the host C++ of libgpu_presolver_amd.so and its
 * hand-written HIP kernels (gfx950).  The reference launches its CUDA
 * kernels directly with <<<>>> from the rule files; this library keeps all
 * device code behind these extern "C" launchers, each taking plain pointers,
 * sizes and the stream to run on.  Every launcher is asynchronous on
 * `stream` and returns hipSuccess or the launch error.
 *
 * Kernel-to-reference map (reference paths under cpp/src/presolve/):
 *   gpk_row_activity_lu     gpu_presolve_kernels.cu:18-66
 *   gpk_col_max_abs         gpu_presolve_kernels.cu:68-82
 *   gpk_scan_*              cub::DeviceScan::InclusiveSum call sites (gpu_presolve.cu:67-81)
 *   gpk_compact_*           gpu_presolve.cu:297-531, 1881-1978 (compact_csr_by_masks_gpu)
 *   gpk_radix_sort_*        thrust::sort / sort_by_key call sites
 *   gpk_transpose_*         cusparseCsr2cscEx2 (gpu_presolve.cu:1804-1879)
 *   gpk_csr_stats           gpu_presolve.cu:249-276
 */
#ifndef GP_KERNELS_H
#define GP_KERNELS_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- fills and small helpers ------------------------------------------- */
hipError_t gpk_fill_u8(uint8_t* dst, uint8_t v, int64_t n, hipStream_t stream);
hipError_t gpk_fill_i32(int32_t* dst, int32_t v, int64_t n, hipStream_t stream);
hipError_t gpk_fill_f64(double* dst, double v, int64_t n, hipStream_t stream);
hipError_t gpk_iota_i32(int32_t* dst, int32_t n, hipStream_t stream);

/* ---- scans --------------------------------------------------------------
 * Exclusive prefix sum of n int32 values into out[0..n]; out[n] is the total.
 * `scratch` holds gpk_scan_scratch_ints(n) ints.  In-place (out == in) is allowed. */
int64_t gpk_scan_scratch_ints(int64_t n);
hipError_t gpk_scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* scratch,
                                  hipStream_t stream);
/* Exclusive scan of (mask[i] != 0): out[0..n], out[n] = popcount. */
hipError_t gpk_scan_mask_u8(const uint8_t* mask, int32_t* out, int64_t n, int32_t* scratch,
                            hipStream_t stream);
/* ---- single-pass scans (decoupled look-back) ----------------------------
 * One launch per scan: each 2048-item tile takes a ticket (dynamic tile id, in
 * dispatch order), publishes its aggregate, walks back over its predecessors'
 * published aggregates / inclusive prefixes, and writes its outputs.
 * Status words: [63:34] epoch, [33:32] 1 = aggregate, 2 = inclusive prefix,
 * [31:0] value; words of older epochs read as "not yet published", so the
 * status array never needs clearing.  A spin that exceeds its limit sets
 * *error (host-mapped memory) and the kernel finishes without waiting. */
typedef struct gpk_lookback {
  uint64_t* status; /* >= tiles words, used only by look-back kernels */
  int32_t* ticket;  /* one zeroed word per launch */
  int32_t* error;   /* host-mapped flag, set on a spin-limit exit */
  uint32_t epoch;   /* 1 .. 2^30 - 1, unique per launch */
} gpk_lookback;
int64_t gpk_lookback_tiles(int64_t n);
/* out[0..n] = exclusive prefix of in (out[n] = total); out == in allowed. */
hipError_t gpk_scan_lb_i32(const int32_t* in, int32_t* out, int64_t n, gpk_lookback lb, hipStream_t stream);
/* Same over (mask[i] != 0). */
hipError_t gpk_scan_lb_mask(const uint8_t* mask, int32_t* out, int64_t n, gpk_lookback lb, hipStream_t stream);
/* map[i] = mask[i] ? (number of set mask entries before i) : -1; *count = total. */
hipError_t gpk_mask_to_map(const uint8_t* mask, int32_t* map, int64_t n, int32_t* count, gpk_lookback lb,
                           hipStream_t stream);
/* Rank form of an old-to-new map, one word per 64 entries: bit b of word w
 * is set when entry 64w + b is kept; base is the new index of the word's
 * first kept entry.  gpk_build_rank writes ceil(n / 64) words. */
typedef struct __attribute__((aligned(16))) gpk_rank_word {
  uint64_t bits;
  int32_t base;
  int32_t pad;
} gpk_rank_word;
hipError_t gpk_build_rank(const int32_t* map, int32_t n, gpk_rank_word* out, hipStream_t stream);
/* The three keep maps of a plan application in one launch (three look-back
 * chains): row_map from keep_row (m), col_map from keep_col (n), and
 * fixed_map from the removed columns with a fixed finite interval
 * (!keep_col && |u - l| <= 1e-9; the mask is stored in fixed_mask);
 * map[i] = rank among the set entries or -1; counts[0..2] = the three totals;
 * row_rank / col_rank, when given, receive the maps' rank words (as
 * gpk_build_rank).
 * The look-back workspace covers gpk_apply_maps_tiles(m, n) tiles. */
typedef struct gpk_apply_maps_args {
  int32_t m, n;
  const uint8_t* keep_row;
  const uint8_t* keep_col;
  const double* l;
  const double* u;
  int32_t* row_map;
  int32_t* col_map;
  uint8_t* fixed_mask;
  int32_t* fixed_map;
  int32_t* counts;
  gpk_rank_word* row_rank; /* optional: rank words of row_map ((m + 63) / 64) */
  gpk_rank_word* col_rank; /* optional: rank words of col_map ((n + 63) / 64) */
} gpk_apply_maps_args;
int64_t gpk_apply_maps_tiles(int32_t m, int32_t n);
hipError_t gpk_apply_maps(gpk_apply_maps_args args, gpk_lookback lb, hipStream_t stream);

/* Stable compaction of a CSR by an old-to-new row map and the rank form of
 * the column map, over tiles of 256 old rows (gpk_compact_tiles(m_old)):
 *   gpk_compact_map    jn[p] = new column of entry p, or -1 when its row or
 *                      column is dropped; out_ptr[new row] = kept count;
 *                      tile_cnt[t] = kept entries of tile t;
 *   gpk_scan_lb_i32(tile_cnt, tile_cnt, tiles, ...) makes the tile counts
 *                      exclusive offsets (tile_cnt[tiles] = kept entries);
 *   gpk_compact_write  turns out_ptr into the new row pointers
 *                      (out_ptr[m_new] = kept entries) and writes the kept
 *                      (column, value) pairs.
 * jn holds one int32 per entry of A, tile_cnt tiles + 1.  col_rank NULL
 * means no column is dropped (every kept entry keeps its column index). */
int32_t gpk_compact_tiles(int32_t m_old);
hipError_t gpk_compact_map(const int32_t* ptr, const int32_t* idx, int32_t m_old, const int32_t* row_map,
                           const gpk_rank_word* col_rank, int32_t* jn, int32_t* out_ptr, int32_t* tile_cnt,
                           hipStream_t stream);
hipError_t gpk_compact_write(const int32_t* ptr, const double* val, int32_t m_old, const int32_t* row_map,
                             const int32_t* jn, const int32_t* tile_off, int32_t* out_ptr, int32_t m_new,
                             int32_t* out_idx, double* out_val, hipStream_t stream);
/* gpk_compact_map that also adds the kept entry count to *kept_total (one
 * integer atomic per tile; the caller zeroes it), so the count is known
 * before any scan. */
hipError_t gpk_compact_map_total(const int32_t* ptr, const int32_t* idx, int32_t m_old, const int32_t* row_map,
                                 const gpk_rank_word* col_rank, int32_t* jn, int32_t* out_ptr, int32_t* tile_cnt,
                                 int32_t* kept_total, hipStream_t stream);
/* A CSR and its transpose compacted with the same two maps (A by row_map
 * and the column rank words, AT by the column map and the row rank words)
 * in one map launch and one write launch.  tile_cnt holds
 * gpk_compact_tiles(a.m_old) + gpk_compact_tiles(t.m_old) + 1 words: A's tile
 * counts, then AT's.  The write pass takes either those raw counts
 * (from_counts, allowed when each side has at most 256 tiles)
 * or one exclusive scan of them (gpk_scan_lb_i32 over both sides' tiles).
 * kept_total (optional, zeroed by the caller) receives A's kept count. */
typedef struct gpk_compact_side {
  const int32_t* ptr;
  const int32_t* idx;
  const double* val;
  int32_t m_old;
  int32_t m_new;
  const int32_t* row_map;
  const gpk_rank_word* col_rank;
  int32_t* jn;
  int32_t* out_ptr;
  int32_t* out_idx;
  double* out_val;
} gpk_compact_side;
hipError_t gpk_compact_map_pair(gpk_compact_side a, gpk_compact_side t, int32_t* tile_cnt, int32_t* kept_total,
                                hipStream_t stream);
hipError_t gpk_compact_write_pair(gpk_compact_side a, gpk_compact_side t, const int32_t* tile_off,
                                  int32_t from_counts, hipStream_t stream);
/* The same without the scan, for at most 256 tiles (m_old <= 65,536): it
 * takes the raw tile counts of gpk_compact_map and derives the offsets. */
hipError_t gpk_compact_write_counts(const int32_t* ptr, const double* val, int32_t m_old, const int32_t* row_map,
                                    const int32_t* jn, const int32_t* tile_cnt, int32_t* out_ptr, int32_t m_new,
                                    int32_t* out_idx, double* out_val, hipStream_t stream);

/* Up to GPK_COPY_MAX copies (src != NULL) or zero fills (src == NULL) of
 * device byte ranges in one launch (block row y serves segment y; the unit
 * is the widest of 16/8/4/1 bytes that divides both addresses and the size).
 * No reference counterpart: it batches what the reference issues as one
 * cudaMemcpyAsync / cudaMemsetAsync per buffer (e.g. the plan's vector copies,
 * gpu_presolve.cu:197-247). */
#define GPK_COPY_MAX 12
typedef struct gpk_copy_seg {
  void* dst;
  const void* src;
  int64_t bytes;
} gpk_copy_seg;
hipError_t gpk_copy_many(const gpk_copy_seg* segs, int32_t count, hipStream_t stream);

/* Host mailbox (no reference counterpart): one workgroup copies `words`
 * (<= 28) 32-bit words from src (may be NULL when words = 0) to mail[4..],
 * makes them visible system-wide, then stores seq to mail[0].  mail is the
 * device address of host-mapped, coherent pinned memory. */
hipError_t gpk_mail(const uint32_t* src, int32_t words, uint32_t* mail, uint32_t seq, hipStream_t stream);

/* old_to_new[i] = mask[i] ? prefix[i] : -1. */
hipError_t gpk_old_to_new(const uint8_t* mask, const int32_t* prefix, int32_t* old_to_new, int32_t n,
                          hipStream_t stream);

/* ---- activity primitives ----------------------------------------------- */
/* Row activity (reference gpu_presolve_kernels.cu:18-111) over bounds packed
 * as lu[2j] = l[j], lu[2j+1] = u[j] (16-byte aligned, written by
 * gpk_pack_lu), so each entry needs one 16-byte gather; 256 rows per block,
 * each row summed in CSR order. */
hipError_t gpk_pack_lu(const double* l, const double* u, int32_t n, double* lu, hipStream_t stream);
hipError_t gpk_row_activity_lu(const int32_t* row_ptr, const int32_t* col_idx, const double* val, int32_t m,
                               const double* lu, double ztol, double* min_fin, double* max_fin, int32_t* neg_inf,
                               int32_t* pos_inf, hipStream_t stream);
hipError_t gpk_col_max_abs(const int32_t* at_ptr, const double* at_val, int32_t n, double* out,
                           hipStream_t stream);
/* Row activity with the (l, u) pairs staged in a fixed file-scope device
 * array (k_noalloc.hip) instead of a pool block: no allocation.  n at most
 * GP_ROW_ACTIVITY_NOALLOC_MAX_COLS (hipErrorInvalidValue beyond).  Calls that
 * overlap on one GPU share that array and race. */
hipError_t gpk_row_activity_noalloc(const int32_t* row_ptr, const int32_t* col_idx, const double* val, int32_t m,
                                    int32_t n, const double* l, const double* u, double ztol, double* min_fin,
                                    double* max_fin, int32_t* neg_inf, int32_t* pos_inf, hipStream_t stream);
/* Row stats of a CSR (counts include explicit zeros; singleton = first entry). */
hipError_t gpk_csr_stats(const int32_t* ptr, const int32_t* idx, const double* val, int32_t rows,
                         int32_t* nnz_out, uint8_t* empty_mask, uint8_t* singleton_mask,
                         int32_t* singleton_idx, double* singleton_val, hipStream_t stream);

/* ---- compaction helpers ---------------------------------------------- */
/* The gathers of a plan application in one launch (gpu_presolve.cu:2021-2147):
 * c, l, u and the column reduced-to-original map through col_map (n); AL, AU
 * and the row map through row_map (m).  Null r2o outputs are skipped. */
typedef struct gpk_apply_gathers_args {
  int32_t m, n;
  const int32_t* col_map;
  const double* c_in;
  const double* l_in;
  const double* u_in;
  const int32_t* col_r2o_in;
  double* c_out;
  double* l_out;
  double* u_out;
  int32_t* col_r2o_out;
  const int32_t* row_map;
  const double* al_in;
  const double* au_in;
  const int32_t* row_r2o_in;
  double* al_out;
  double* au_out;
  int32_t* row_r2o_out;
} gpk_apply_gathers_args;
hipError_t gpk_apply_gathers(gpk_apply_gathers_args args, hipStream_t stream);
/* dst[map[i]] = src[i] for map[i] >= 0, for up to 5 vectors of the same length. */
hipError_t gpk_gather5(const int32_t* map, int32_t n_old, const double* s0, double* d0,
                       const double* s1, double* d1, const double* s2, double* d2,
                       const double* s3, double* d3, const double* s4, double* d4,
                       hipStream_t stream);
hipError_t gpk_gather_i32(const int32_t* map, int32_t n_old, const int32_t* src, int32_t* dst,
                          hipStream_t stream);

/* ---- radix sort -----------------------------------------------------------
 * Stable LSD radix sort of (key, value) pairs, 8-bit digits over the low
 * `key_bits` bits.  Buffers *_alt have the same size; the sorted result is
 * left in keys/vals.  scratch holds gpk_radix_scratch_ints(n) ints. */
int64_t gpk_radix_scratch_ints(int64_t n);
hipError_t gpk_radix_sort_u64_i32(uint64_t* keys, int32_t* vals, uint64_t* keys_alt,
                                  int32_t* vals_alt, int32_t n, int32_t key_bits, int32_t* scratch,
                                  hipStream_t stream);
hipError_t gpk_radix_sort_u32_i32(uint32_t* keys, int32_t* vals, uint32_t* keys_alt,
                                  int32_t* vals_alt, int32_t n, int32_t key_bits, int32_t* scratch,
                                  hipStream_t stream);

/* ---- transpose -------------------------------------------------------------
 * Expands the row index of every entry (row_of[p] = i), for use as the sort
 * payload of a transpose, and the final gather of the transpose. */
/* Transpose building blocks: stable sort of the column indices of A (read
 * in place) carrying the entry positions, key_bits wide, through ka/va and
 * kb/vb (*result_in_b tells which pair holds the result); row pointers of
 * the transpose (n + 1) from the sorted keys. */
hipError_t gpk_transpose_sort(const int32_t* col, int32_t nnz, int32_t key_bits, uint32_t* ka, int32_t* va,
                              uint32_t* kb, int32_t* vb, int32_t* scratch, int32_t* result_in_b, hipStream_t stream);
hipError_t gpk_col_bounds(const uint32_t* sorted_keys, int32_t nnz, int32_t n, int32_t* ptr, hipStream_t stream);
/* The row and value of every entry of A, packed per entry (16 bytes, 16-byte
 * aligned array of nnz), then the transpose's columns and values gathered
 * through the sorted positions: at_idx[k] = rv[perm[k]].row, at_val[k] =
 * rv[perm[k]].val. */
typedef struct gpk_row_val {
  int32_t row;
  int32_t pad;
  double val;
} gpk_row_val;
hipError_t gpk_expand_rows(const int32_t* ptr, const double* val, int32_t m, gpk_row_val* rv, hipStream_t stream);
hipError_t gpk_transpose_gather(const int32_t* perm, const gpk_row_val* rv, int32_t nnz, int32_t* at_idx,
                                double* at_val, hipStream_t stream);
/* Input check of a CSR matrix, one thread per entry over tiles of 256 rows:
 * flags[0] bad row pointers, flags[1] a column index outside [0, cols),
 * flags[2] two equal neighbours in a row, flags[3] a row not sorted by column
 * (flags are set to 1, never cleared).  A row that is not sorted can repeat a
 * column at a distance: gpk_csr_check_rows (one thread per row, every pair of
 * an unsorted row) sets flags[2] for those as well. */
hipError_t gpk_csr_check(const int32_t* ptr, const int32_t* idx, int32_t rows, int32_t cols, int32_t nnz,
                         int32_t* flags, hipStream_t stream);
hipError_t gpk_csr_check_rows(const int32_t* ptr, const int32_t* idx, int32_t rows, int32_t cols, int32_t nnz,
                              int32_t* flags, hipStream_t stream);
/* flag[0] = 1 if some row of the CSR is not sorted by column (row pointers
 * trusted). */
hipError_t gpk_rows_unsorted(const int32_t* ptr, const int32_t* idx, int32_t rows, int32_t* flag,
                             hipStream_t stream);
/* In-place sort of every row of a CSR by column (values follow). */
hipError_t gpk_sort_rows_in_place(const int32_t* ptr, int32_t* idx, double* val, int32_t rows,
                                  hipStream_t stream);


/* ======================================================================
 * Rule kernels.  Views of the CSR and of the plan vectors are passed by
 * value.  Reference kernels restated are cited per launcher group.
 * ====================================================================== */
typedef struct gpk_csr {
  int32_t rows, cols, nnz;
  const int32_t* ptr;
  const int32_t* idx;
  const double* val;
} gpk_csr;

typedef struct gpk_plan {
  uint8_t* keep_row;
  uint8_t* keep_col;
  double* c;
  double* l;
  double* u;
  double* AL;
  double* AU;
} gpk_plan;

/* Fresh phase plan (gpu_presolve.cu:197-247): keep masks set to 1 and
 * copies of the working LP's c, l, u (n) and AL, AU (m), in one launch. */
hipError_t gpk_plan_init(int32_t m, int32_t n, gpk_plan plan, const double* c, const double* l, const double* u,
                         const double* AL, const double* AU, hipStream_t stream);

/* Fresh phase plan and phase statistics in one launch: gpk_plan_init plus
 * gpk_csr_stats of A (row counts, singleton rows) and of AT (column counts,
 * empty and singleton columns), gpu_presolve.cu:197-276. */
typedef struct gpk_phase_stats {
  int32_t* row_nnz;
  uint8_t* singleton_row_mask;
  int32_t* singleton_row_col;
  double* singleton_row_val;
  int32_t* col_nnz;
  uint8_t* empty_col_mask;
  uint8_t* column_singleton_mask;
  int32_t* column_singleton_row;
  double* column_singleton_val;
} gpk_phase_stats;
hipError_t gpk_phase_init(int32_t m, int32_t n, gpk_plan plan, const double* c, const double* l, const double* u,
                          const double* AL, const double* AU, gpk_csr A, gpk_csr AT, gpk_phase_stats stats,
                          hipStream_t stream);

/* Serving runtime (no reference counterpart).  gpk_synthetic_point: seeded
 * reduced-space point standing in for a solver (x in [l, u] on the 2^-10
 * grid, y and z uniform in [-1, 1)); element draws depend only on (seed,
 * index).  gpk_checksum_f64: *acc += sum_i mix(bits(v[i]), salt + i) mod
 * 2^64, independent of the summation order. */
hipError_t gpk_synthetic_point(uint64_t seed, int32_t n, int32_t m, const double* l, const double* u, double* x,
                               double* y, double* z, hipStream_t stream);
hipError_t gpk_checksum_f64(const double* v, int32_t n, uint64_t salt, unsigned long long* acc, hipStream_t stream);

/* Seeded LP generator (k_lpgen.hip).  Steps: columns (class, c, l, u, x0);
 * row classes, kinds and raw entry counts (extra: zeroed m words); after an
 * exclusive scan of raw_count into raw_ptr, the raw entries (own, cursor:
 * zeroed m words), per-row column sort and the final row lengths len; after
 * a scan of len into ptr, the CSR and the row bounds. */
typedef struct gpk_lpgen_spec {
  uint64_t seed;
  int32_t m, n;
  int32_t len_lo, len_hi; /* general row length U[len_lo, len_hi] when mean_len <= 0 */
  double mean_len;        /* > 0: general row length max(min_len, Poisson(mean_len)) */
  int32_t min_len;
} gpk_lpgen_spec;
hipError_t gpk_lpgen_columns(gpk_lpgen_spec spec, uint8_t* col_class, double* c, double* l, double* u, double* x0,
                             hipStream_t stream);
hipError_t gpk_lpgen_row_counts(gpk_lpgen_spec spec, const uint8_t* col_class, uint8_t* row_class, uint8_t* kind,
                                int32_t* raw_count, int32_t* scol_row, int32_t* extra, hipStream_t stream);
hipError_t gpk_lpgen_fill(gpk_lpgen_spec spec, const uint8_t* col_class, const uint8_t* row_class,
                          const int32_t* raw_ptr, const int32_t* scol_row, int32_t* own, int32_t* cursor,
                          int32_t* raw_col, double* raw_val, int32_t* unique, int32_t* src, int32_t* len,
                          hipStream_t stream);
hipError_t gpk_lpgen_write(gpk_lpgen_spec spec, const uint8_t* row_class, const uint8_t* kind, const int32_t* raw_ptr,
                           const int32_t* raw_col, const double* raw_val, const int32_t* src, const int32_t* ptr,
                           int32_t* idx, double* val, const double* x0, double* AL, double* AU, hipStream_t stream);

/* Batched singleton closure (reference gpu_presolve.cu:98-247, 2506-2831).
 * gpk_closure_degrees: row_nnz (col_nnz) lose the live entries of columns
 *   (rows) that were kept in col_before (row_before) and are removed now;
 *   the before masks are then set to the current ones.
 * gpk_closure_singletons: singleton rows / columns from the live degrees,
 *   with the first live entry as the singleton.
 * gpk_closure_snapshot: seen.{keep_row, c, l, u, AL, AU} := plan's.
 * gpk_closure_candidates: cand[j] = live column j whose c/l/u differ from
 *   seen, or that lies in a row removed or re-bounded since seen. */
hipError_t gpk_closure_degrees(gpk_csr A, gpk_csr AT, const uint8_t* keep_row, const uint8_t* keep_col,
                               uint8_t* row_before, uint8_t* col_before, int32_t* row_nnz, int32_t* col_nnz,
                               hipStream_t stream);
hipError_t gpk_closure_singletons(gpk_csr A, gpk_csr AT, const uint8_t* keep_row, const uint8_t* keep_col,
                                  const int32_t* row_nnz, const int32_t* col_nnz, uint8_t* srow_mask,
                                  int32_t* srow_col, double* srow_val, uint8_t* scol_mask, int32_t* scol_row,
                                  double* scol_val, hipStream_t stream);
hipError_t gpk_closure_snapshot(int32_t m, int32_t n, gpk_plan plan, gpk_plan seen, hipStream_t stream);
hipError_t gpk_closure_candidates(gpk_csr A, gpk_plan plan, gpk_plan seen, uint8_t* cand, hipStream_t stream);

/* Generic row shift: for each row i (all rows, or only keep_row[i] != 0 when
 * keep_row is given) shift = sum over entries with col_flag[j] of a*col_val[j]
 * in CSR order, then AL[i] -= shift, AU[i] -= shift for every row.  `abs_min`
 * skips entries with |a| <= abs_min when >= 0. */
hipError_t gpk_row_shift(gpk_csr A, const uint8_t* keep_row, const uint8_t* col_flag, const double* col_val,
                         double abs_min, double* AL, double* AU, hipStream_t stream);
/* out[k] = vals[i] for the k-th selected i (mask != 0), using prefix = exclusive scan of mask. */
hipError_t gpk_pack_f64(const uint8_t* mask, const double* vals, const int32_t* prefix, int32_t n, double* out,
                        hipStream_t stream);
hipError_t gpk_pack_index(const uint8_t* mask, const int32_t* prefix, int32_t n, int32_t* out, hipStream_t stream);

/* rule_empty_rows.cu:143-184 */
hipError_t gpk_empty_rows(int32_t m, gpk_plan plan, const int32_t* row_nnz, double tol, int32_t apply,
                          int32_t* flags, hipStream_t stream);
/* rule_singleton_rows.cu:54-125 */
hipError_t gpk_singleton_rows_bounds(int32_t m, int32_t n, gpk_plan plan, const uint8_t* mask, const int32_t* col,
                                     const double* val, double zero_tol, double* cand_l, double* cand_u,
                                     uint8_t* row_remove, hipStream_t stream);
hipError_t gpk_singleton_rows_check(int32_t m, int32_t n, gpk_plan plan, const double* cand_l, const double* cand_u,
                                    const uint8_t* row_remove, double bound_tol, int32_t* flags, hipStream_t stream);
hipError_t gpk_singleton_rows_commit(int32_t m, int32_t n, gpk_plan plan, const double* cand_l,
                                     const double* cand_u, const uint8_t* row_remove, hipStream_t stream);
/* rule_infeasible_redundant_rows.cu:20-117 */
hipError_t gpk_redundant_rows_classify(gpk_csr A, gpk_plan plan, const int32_t* row_nnz, double tol,
                                       uint8_t* full, uint8_t* dlo, uint8_t* dhi, int32_t* flags,
                                       hipStream_t stream);
hipError_t gpk_redundant_rows_apply(int32_t m, gpk_plan plan, const uint8_t* full, const uint8_t* dlo,
                                    const uint8_t* dhi, hipStream_t stream);
/* rule_implied_variable_bounds.cu:339-866.  The walks along AT read, for
 * every entry, its row's activity summary and bounds: gpk_implied_pack_rows
 * first packs them into one 64-byte record per row (one cache line per
 * entry's row instead of eight gathers). */
typedef struct __attribute__((aligned(64))) gpk_implied_row {
  double smin, smax; /* finite parts of the row's min and max activity */
  double lo, hi;     /* the plan's AL, AU */
  int32_t ninf, pinf; /* infinite terms of the min and max activity */
  int32_t live;      /* kept and more than one entry */
  int32_t pad[5];
} gpk_implied_row;
hipError_t gpk_implied_pack_rows(int32_t m, gpk_plan plan, const int32_t* row_nnz, const double* rmin,
                                 const double* rmax, const int32_t* rni, const int32_t* rpi, gpk_implied_row* rows,
                                 hipStream_t stream);
hipError_t gpk_implied_candidates(gpk_csr AT, gpk_plan plan, const gpk_implied_row* rows, double zero_tol,
                                  double* cand_l, double* cand_u, hipStream_t stream);
hipError_t gpk_implied_finalize(int32_t n, gpk_plan plan, const double* ref_l, const double* ref_u,
                                const double* cand_l, const double* cand_u, const double* vmax,
                                int32_t preserve_free_zero_cost, double zero_tol, double feas_tol, double* fin_l,
                                double* fin_u, uint8_t* lch, uint8_t* uch, uint8_t* fixed, double* fixed_val,
                                int32_t* flags, hipStream_t stream);
hipError_t gpk_implied_support(gpk_csr AT, gpk_plan plan, const gpk_implied_row* rows, const uint8_t* lch,
                               const uint8_t* uch, const double* fin_l, const double* fin_u, double zero_tol,
                               double tol, double support_tol, int32_t* support_l, int32_t* support_u,
                               int32_t* rec_count, hipStream_t stream);
hipError_t gpk_implied_pack_tape(int32_t n, gpk_plan plan, const uint8_t* lch, const uint8_t* uch,
                                 const int32_t* support_l, const int32_t* support_u, const double* fin_l,
                                 const double* fin_u, const int32_t* prefix, int32_t* rec_col, int32_t* rec_row,
                                 double* rec_vals, hipStream_t stream);
hipError_t gpk_implied_apply(int32_t n, gpk_plan plan, const double* fin_l, const double* fin_u,
                             const uint8_t* fixed, hipStream_t stream);
hipError_t gpk_remove_empty_rows_after_fix(gpk_csr A, gpk_plan plan, double feas_tol, int32_t* flags,
                                           hipStream_t stream);
/* fixed_col_tape.cu:41-103 */
hipError_t gpk_fixed_col_counts(gpk_csr AT, const uint8_t* fixed, const uint8_t* excluded, const uint8_t* keep_row,
                                uint8_t* selected, int32_t* active_nnz, hipStream_t stream);
hipError_t gpk_fixed_col_pack(gpk_csr AT, const uint8_t* selected, const int32_t* sel_prefix,
                              const int32_t* nnz_prefix, const double* fixed_val, const double* c,
                              const uint8_t* keep_row, int32_t* out_col, double* out_v, double* out_c,
                              int32_t* out_start, int32_t* out_rows, double* out_coef, hipStream_t stream);
/* rule_duplicate_rows.cu:78-418.  gpk_dup_row_hash computes the hash the
 * reference picks (:484-497): sequential FNV over (length, column,
 * a / a_first) per row, or, when A.nnz > 32 A.rows, the XOR of per-entry
 * pair hashes mixed into the length seed (:116-178), one wave per row. */
hipError_t gpk_dup_row_hash(gpk_csr A, const uint8_t* keep_row, double coeff_tol, uint64_t* hash,
                            hipStream_t stream);
hipError_t gpk_dup_row_groups(gpk_csr A, gpk_plan plan, const uint64_t* sorted_hash, const int32_t* sorted_rows,
                              double coeff_tol, double interval_tol, uint8_t* row_delete, int32_t* merge_to,
                              int32_t* merge_step, double* rec_vals, int32_t* flags, hipStream_t stream);
/* keep_row[i] = 0 where row_delete[i]; *deleted set to 1 when some row is. */
hipError_t gpk_delete_rows(int32_t m, uint8_t* keep_row, const uint8_t* row_delete, int32_t* deleted,
                           hipStream_t stream);
/* Deleted duplicate rows (del[i] != 0, prefix = exclusive scan of del) packed
 * in ascending row order: row, merge_to, step and the 5 record values. */
hipError_t gpk_dup_row_pack(int32_t m, const uint8_t* del, const int32_t* prefix, const int32_t* merge_to,
                            const int32_t* step, const double* rec, int32_t* out_row, int32_t* out_to,
                            int32_t* out_step, double* out_rec, hipStream_t stream);

/* rule_infeasible_fixed_variables.cu:42-104, rule_empty_cols.cu:42-108
 * (*count: set to 1 when some column is fixed) */
hipError_t gpk_fixed_vars_candidates(int32_t n, gpk_plan plan, double bound_tol, uint8_t* fixed, double* v,
                                     double* contrib, int32_t* count, hipStream_t stream);
hipError_t gpk_empty_cols(int32_t n, gpk_plan plan, const uint8_t* empty_mask, double bound_tol, double zero_tol,
                          uint8_t* fixed, double* v, double* contrib, int32_t* flags, hipStream_t stream);
/* l = u = v and keep_col = 0 for fixed columns. */
hipError_t gpk_fix_cols_apply(int32_t n, gpk_plan plan, const uint8_t* fixed, const double* v, hipStream_t stream);

/* rule_column_singletons.cu:411-1215 */
typedef struct gpk_row_cache {
  double* fmin;
  double* fmax;
  int32_t* mn;
  int32_t* mp;
  int32_t* xn;
  int32_t* xp;
  uint8_t* invalid;
} gpk_row_cache;
/* flags[0] = 1: some live singleton column; flags[1]: some on an equality
 * row; flags[2]: some on another row. */
hipError_t gpk_cs_precheck(int32_t n, int32_t m, gpk_plan plan, const uint8_t* mask, const int32_t* srow,
                           double bound_tol, uint8_t* needed, int32_t* flags, hipStream_t stream);
hipError_t gpk_cs_activity_cache(gpk_csr A, gpk_plan plan, const uint8_t* needed, gpk_row_cache cache,
                                 hipStream_t stream);
hipError_t gpk_cs_owner(gpk_csr A, gpk_plan plan, const uint8_t* mask, const int32_t* srow, const double* sval,
                        gpk_row_cache cache, double tol, double zero_tol, int32_t eq_mode, int32_t* owner,
                        int32_t* flags, hipStream_t stream);
hipError_t gpk_cs_process(int32_t n, int32_t m, gpk_plan plan, const uint8_t* mask, const int32_t* srow,
                          const double* sval, const int32_t* owner, gpk_row_cache cache, double tol,
                          double zero_tol, int32_t eq_mode, double* AL_new, double* AU_new, uint8_t* row_del,
                          uint8_t* row_changed, uint8_t* col_del, double* side, double* contrib, int32_t* flags,
                          hipStream_t stream);
hipError_t gpk_cs_objective(gpk_csr AT, gpk_plan plan, const uint8_t* col_del, const int32_t* owner,
                            const double* sval, double* c_new, hipStream_t stream);
hipError_t gpk_cs_subcol_counts(gpk_csr A, gpk_plan plan, const int32_t* owner, const uint8_t* col_del,
                                uint8_t* selected, int32_t* support_count, hipStream_t stream);
hipError_t gpk_cs_subcol_pack(gpk_csr A, gpk_plan plan, const int32_t* owner, const uint8_t* row_del,
                              const double* side, const uint8_t* selected, const int32_t* sel_prefix,
                              const int32_t* sup_prefix, int32_t* out_i2, double* out_v6, int32_t* out_scol,
                              double* out_sval, hipStream_t stream);
hipError_t gpk_cs_apply(int32_t m, int32_t n, gpk_plan plan, const uint8_t* row_del, const uint8_t* col_del,
                        const double* AL_new, const double* AU_new, const double* c_new, hipStream_t stream);

/* rule_dual_fix.cu:174-384 */
hipError_t gpk_dual_fix_candidates(gpk_csr AT, gpk_plan plan, const uint8_t* candidate, double zero_tol,
                                   double bound_tol, int32_t max_degree, uint8_t* fixed, uint8_t* inf_fix,
                                   uint8_t* finite_fix, double* v, double* contrib, int32_t* flags,
                                   hipStream_t stream);
hipError_t gpk_dual_fix_rows(gpk_csr A, gpk_plan plan, const uint8_t* finite_fix, const uint8_t* inf_fix,
                             const double* v, double* shift, uint8_t* row_remove, int32_t* flags,
                             hipStream_t stream);
hipError_t gpk_dual_fix_apply(int32_t m, int32_t n, gpk_plan plan, const uint8_t* fixed, const double* v,
                              const double* shift, const uint8_t* row_remove, hipStream_t stream);

/* rule_duplicate_columns.cu:95-556 */
hipError_t gpk_dup_col_hash(gpk_csr AT, gpk_plan plan, double zero_tol, uint64_t* hash, hipStream_t stream);
hipError_t gpk_dup_col_guard(gpk_csr AT, gpk_plan plan, const uint64_t* sorted_hash, const int32_t* sorted_cols,
                             double zero_tol, double coeff_tol, double obj_tol, double max_ratio,
                             int32_t max_group, int32_t* count, hipStream_t stream);
hipError_t gpk_dup_col_groups(gpk_csr AT, gpk_plan plan, const uint64_t* sorted_hash, const int32_t* sorted_cols,
                              double zero_tol, double coeff_tol, double obj_tol, double max_ratio,
                              const int32_t* guard_count, int32_t guard_min_groups, uint8_t* col_del,
                              uint8_t* fixed, double* fixed_val, int32_t* merge_to, double* rec5, int32_t* flags,
                              hipStream_t stream);
hipError_t gpk_dup_col_apply(int32_t n, gpk_plan plan, const uint8_t* col_del, const uint8_t* fixed,
                             const double* fixed_val, double* contrib, hipStream_t stream);

/* rule_doubleton_equations.cu:143-1314 (batch path) */
typedef struct gpk_dbl {
  uint8_t* acceptable;
  uint8_t* active;
  uint8_t* selected;
  uint8_t* blocked;
  int32_t* cec;
  int32_t* ckc;
  double* cev;
  double* ckv;
  uint64_t* owner; /* per column: (round << 32) | (0xffffffff - row) of the round's lowest claiming row */
  int32_t* live_nnz;
} gpk_dbl;
typedef struct gpk_dbl_cols {
  uint8_t* is_elim;
  uint8_t* is_fixed_keep;
  double* alpha_of;
  double* beta_of;
  double* fixed_val_of;
  int32_t* row_of;
  int32_t* keep_of;
} gpk_dbl_cols;
hipError_t gpk_dbl_live_nnz(gpk_csr AT, gpk_plan plan, double zero_tol, int32_t* live_nnz, hipStream_t stream);
hipError_t gpk_dbl_candidates(gpk_csr A, gpk_csr AT, gpk_plan plan, gpk_dbl d, double zero_tol, double tol,
                              double ratio_tol, int32_t max_fill, int32_t live_fill, int32_t* counts,
                              hipStream_t stream);
/* One claim round of the matching (rule_doubleton_equations.cu:876-913): every
 * active row with both columns unblocked claims them for itself, the lowest
 * row index winning a column.  Owners carry their round (1, 2, ...; the
 * owner array is zeroed once before round 1), so no reset between rounds is
 * needed.  gpk_dbl_select then selects the rows that own both of their
 * columns in that round and blocks those columns; counts[1] += selected. */
hipError_t gpk_dbl_claim(int32_t m, gpk_dbl d, int32_t round, hipStream_t stream);
hipError_t gpk_dbl_select(int32_t m, gpk_dbl d, int32_t round, int32_t* counts, hipStream_t stream);
hipError_t gpk_dbl_apply(int32_t k, const int32_t* sel, gpk_plan plan, gpk_dbl d, gpk_dbl_cols cols, double tol,
                         uint8_t* subst, uint8_t* fixedp, double* osub, double* ofix, int32_t* rec_i3,
                         double* rec_v12, int32_t* flags, hipStream_t stream);
hipError_t gpk_dbl_delta_count(gpk_csr A, gpk_plan plan, gpk_dbl_cols cols, double zero_tol, int32_t* counts,
                               hipStream_t stream);
hipError_t gpk_dbl_shift(gpk_csr A, gpk_plan plan, gpk_dbl_cols cols, double zero_tol, int32_t include_subst,
                         hipStream_t stream);
hipError_t gpk_dbl_delta_fill(gpk_csr A, gpk_plan plan, gpk_dbl_cols cols, double zero_tol,
                              const int32_t* delta_ptr, int32_t* dcol, double* dval, hipStream_t stream);
hipError_t gpk_dbl_merge(gpk_csr A, const int32_t* delta_ptr, const int32_t* dcol, const double* dval,
                         double zero_tol, const int32_t* out_ptr, int32_t* counts, int32_t* out_idx, double* out_val,
                         hipStream_t stream);

/* rule_doubleton_equations.cu:1779-1909 (single-doubleton path) */
hipError_t gpk_dbl_scalar_rows(gpk_csr A, int32_t target, int32_t ec, int32_t kc, double alpha, double zero_tol,
                               const int32_t* out_ptr, int32_t* counts, int32_t* out_idx, double* out_val,
                               hipStream_t stream);
/* ec_len: entries of column ec in AT (the launch covers exactly those). */
hipError_t gpk_dbl_scalar_shift(gpk_csr AT, int32_t ec, int32_t ec_len, double beta, double* AL, double* AU,
                                hipStream_t stream);

/* rule_redundant_bounds.cu:31-262 */
hipError_t gpk_rb_candidates(gpk_csr A, gpk_csr AT, gpk_plan plan, const double* c_lp, int32_t one_sided,
                             double zero_tol, double tol, uint8_t* dlo, uint8_t* dhi, hipStream_t stream);
hipError_t gpk_rb_select(gpk_csr A, gpk_csr AT, const uint8_t* dlo, const uint8_t* dhi, uint64_t* row_key,
                         uint8_t* sel_lo, uint8_t* sel_hi, int32_t* flags, hipStream_t stream);
hipError_t gpk_rb_apply(int32_t n, gpk_plan plan, const uint8_t* sel_lo, const uint8_t* sel_hi,
                        hipStream_t stream);

/* Fail-closed screens of the structure-specific rules (k_screens.hip).
 * Counter words in `out` start at zero.
 *   covering:   out[0] != 0 when a row, column or entry breaks the unit set
 *               covering form (rule_covering_cost_dominance.cu:112-196),
 *               out[1] = columns whose cost exceeds their row count;
 *   bounded:    *selected = non-conflicting projectable columns
 *               (rule_bounded_two_row_projection.cu:236-335); cand: n bytes,
 *               row_hits: m zeroed words;
 *   mccormick:  the same for orphan McCormick columns
 *               (rule_orphan_mccormick_projection.cu:117-244); ends: 2n words;
 *   pair probe: kind 0 antipodal, 1 Linf: out[0] != 0 when a sampled pair
 *               fails, out[1] = sampled merge rows
 *               (rule_antipodal_components.cu:41-125, rule_linf_components.cu:108-220);
 *   structural: *out = 1 when rows 0..3 open an L1 split, outer pair or
 *               graph block (rule_structural_l1_substitution.cu:2214-2376). */
hipError_t gpk_screen_covering(const int32_t* ptr, const int32_t* idx, const double* val, const int32_t* at_ptr,
                               const double* at_val, const double* c, const double* AL, const double* AU,
                               const double* l, const double* u, int32_t m, int32_t n, int32_t nnz, int32_t* out,
                               hipStream_t stream);
hipError_t gpk_screen_bounded_two_row(const int32_t* ptr, const int32_t* idx, const double* val,
                                      const int32_t* at_ptr, const int32_t* at_idx, const double* c,
                                      const double* AL, const double* AU, const double* l, const double* u,
                                      int32_t n, int32_t max_row, double ztol, double tol, uint8_t* cand,
                                      int32_t* row_hits, int32_t* selected, hipStream_t stream);
hipError_t gpk_screen_mccormick(const int32_t* ptr, const int32_t* idx, const double* val, const int32_t* at_ptr,
                                const int32_t* at_idx, const double* c, const double* AL, const double* AU,
                                const double* l, const double* u, int32_t n, uint8_t* cand, int32_t* ends,
                                int32_t* row_hits, int32_t* selected, hipStream_t stream);
hipError_t gpk_screen_pair_probe(int32_t kind, const int32_t* ptr, const int32_t* idx, const double* val,
                                 const int32_t* at_ptr, const int32_t* at_idx, const double* at_val,
                                 const double* c, const double* AL, const double* AU, const double* l,
                                 const double* u, int32_t m, int32_t n, int32_t pairs, int32_t pair_probes,
                                 int32_t row_probes, int32_t* out, hipStream_t stream);
hipError_t gpk_screen_structural_prefix(const int32_t* ptr, const int32_t* idx, const double* val,
                                        const double* AL, const double* AU, int32_t m, int32_t* out,
                                        hipStream_t stream);

/* ======================================================================
 * On-device LP solver of the reduced LP (restarted PDHG; no counterpart in
 * the reference, which hands the reduced LP to an external GPU solver).
 * step[0] = tau (primal), step[1] = sigma (dual) are read from device memory
 * so a captured run of iterations can be replayed after the host changes
 * them.  Row kernels take a gpk_rows: the CSR matrix (A, or AT for the
 * columns) and its row blocks, built on the host from the row pointers by
 * gpk_pdhg_row_blocks: a block is one row (every row of more than
 * GPK_PDHG_LONG_ROW entries is alone, reduced by the whole workgroup) or up
 * to 256 consecutive shorter rows with at most GPK_PDHG_ENTRY_BUDGET entries
 * in all (staged in LDS).
 * ====================================================================== */
#define GPK_PDHG_LONG_ROW 256
#define GPK_PDHG_ENTRY_BUDGET 2048
typedef struct gpk_rows {
  gpk_csr M;
  const int32_t* blk; /* device: blocks + 1 row starts, blk[blocks] = M.rows */
  int32_t blocks;
} gpk_rows;
/* Host: the row blocks of a CSR from its host row pointers (rows + 1);
 * writes blocks + 1 starts to blk (room for rows + 1; NULL only counts) and
 * returns the block count. */
int32_t gpk_pdhg_row_blocks(const int32_t* ptr, int32_t rows, int32_t* blk);
/* `iterations` PDHG steps from z = (x, y): the PDHG point T(z) is
 * xn = proj_[l,u](x - tau (c - A^T y)), xbar = 2 xn - x, yn the projected
 * dual step on xbar.  halpern == 0 (averaged PDHG): z <- T(z) and x_acc,
 * y_acc accumulate the iterates (the averages of a restart period).
 * halpern != 0 (reflected Halpern PDHG): x_acc, y_acc <- T(z) and
 * z <- w (2 T(z) - z) + (1 - w) (x0, y0) with w = (k + 1) / (k + 2), k the
 * iterations since the restart: step[2] at the first of these iterations. */
hipError_t gpk_pdhg_iterations(gpk_rows A, gpk_rows AT, const double* c, const double* l, const double* u,
                               const double* AL, const double* AU, double* x, double* xbar, double* x_acc, double* y,
                               double* y_acc, const double* x0, const double* y0, const double* step,
                               int32_t halpern, int32_t iterations, hipStream_t stream);
/* out = a x + b y (y may be null: out = a x). */
hipError_t gpk_pdhg_axpby(double a, const double* x, double b, const double* y, double* out, int32_t n,
                          hipStream_t stream);
hipError_t gpk_pdhg_mul(const double* x, const double* d, double* out, int32_t n, hipStream_t stream);
hipError_t gpk_pdhg_spmv(gpk_rows M, const double* x, double* y, hipStream_t stream);
/* One Ruiz step: d[i] /= sqrt(max |M.val[p]| over row i) (rows with no
 * nonzero keep d). */
hipError_t gpk_pdhg_ruiz(gpk_rows M, double* d, hipStream_t stream);
/* out[p] = M.val[p] * drow[row(p)] * dcol[idx[p]]. */
hipError_t gpk_pdhg_scale_csr(gpk_rows M, const double* drow, const double* dcol, double* out, hipStream_t stream);
/* out = v * d (mul != 0) or v / d. */
hipError_t gpk_pdhg_scale_vec(const double* v, const double* d, double* out, int32_t n, int32_t mul,
                              hipStream_t stream);
/* v *= 1 / sqrt(sumsq[0]) when sumsq[0] > 0 (sumsq in device memory). */
hipError_t gpk_pdhg_normalize(double* v, const double* sumsq, int32_t n, hipStream_t stream);
/* Doubles of partial sums a *_terms launch over `blocks` workgroups needs
 * (row blocks for row/col terms, ceil(n / 256) for sq_terms). */
int32_t gpk_pdhg_partials(int32_t blocks);
/* KKT sums over the rows of A (out[0..3]: squared row violation of Ax, the
 * rows' dual-objective terms of y, squared bound magnitudes, squared y on
 * rows whose bound of y's sign is infinite) and over the columns (out[0..3]:
 * squared dual residual of c - A^T y against the reduced-cost cone of the
 * bounds, the columns' dual-objective terms, squared c, c^T x; z, if not
 * null, gets the projected reduced costs).  Block sums are added in block
 * order by one workgroup: the results do not vary run to run. */
hipError_t gpk_pdhg_row_terms(gpk_rows A, const double* AL, const double* AU, const double* x, const double* y,
                              double* partial, double* out, hipStream_t stream);
hipError_t gpk_pdhg_col_terms(gpk_rows AT, const double* c, const double* l, const double* u, const double* x,
                              const double* y, double* z, double* partial, double* out, hipStream_t stream);
/* out[0] = sum (a - b)^2 (b null: sum a^2), out[1] = sum a^2. */
hipError_t gpk_pdhg_sq_terms(const double* a, const double* b, int32_t n, double* partial, double* out,
                             hipStream_t stream);
/* p[i] = 0.5 + frac(i * golden ratio): the power iteration's start vector. */
hipError_t gpk_pdhg_seed(double* p, int32_t n, hipStream_t stream);
/* out = the projection of xs * d on [l, u]. */
hipError_t gpk_pdhg_unscale_x(const double* xs, const double* d, const double* l, const double* u, double* out,
                              int32_t n, hipStream_t stream);
/* x = the projection of 0 on [l, u]. */
hipError_t gpk_pdhg_project_zero(const double* l, const double* u, double* x, int32_t n, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GP_KERNELS_H */
