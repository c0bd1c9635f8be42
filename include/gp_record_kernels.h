/*
 * Internal C-ABI launchers for record maintenance and postsolve
 * (part of the same thin layer as gp_kernels.h).
 *
 *   gpk_removed_fixed_mask / gpk_pack_fixed_records
 *       append_removed_fixed_bound_records (reference gpu_presolve.cu:1005-1233)
 *   gpk_map_tape_indices
 *       append_globalized_postsolve_tape (reference gpu_presolve.cu:1363-1533)
 *   gpk_scatter / gpk_mark / gpk_restore_fixed
 *       reference gpu_postsolve.cu:143-190
 *   gpk_replay
 *       _kernel_replay_postsolve_tape (reference gpu_postsolve.cu:605-1046)
 *   gpk_protect_columns / gpk_project_duals
 *       reference gpu_postsolve.cu:1370-1435
 */
#ifndef GP_RECORD_KERNELS_H
#define GP_RECORD_KERNELS_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Device copy of a postsolve tape. */
typedef struct gp_tape_dev {
  int32_t count;
  const int32_t* types;
  const int32_t* index_starts;
  const int32_t* value_starts;
  const int32_t* indices;
  const double* vals;
} gp_tape_dev;

/* Device view of the original model for the dual recoveries (present=0: none). */
typedef struct gp_orig_dev {
  int32_t present;
  const int32_t* a_ptr;
  const int32_t* a_idx;
  const double* a_val;
  const int32_t* at_ptr;
  const int32_t* at_idx;
  const double* at_val;
  const double* c;
  const double* l;
  const double* u;
} gp_orig_dev;

hipError_t gpk_removed_fixed_mask(int32_t n, const uint8_t* keep_col, const double* l, const double* u,
                                  uint8_t* mask, hipStream_t stream);
hipError_t gpk_pack_fixed_records(int32_t n, const uint8_t* mask, const int32_t* prefix, const int32_t* red2org,
                                  const double* l, const double* u, int32_t* out_idx, double* out_val,
                                  hipStream_t stream);
hipError_t gpk_map_tape_indices(int32_t count, int32_t* idx, const uint8_t* kind, const int32_t* row_red2org,
                                int32_t m, const int32_t* col_red2org, int32_t n, hipStream_t stream);
hipError_t gpk_scatter(const double* src, const int32_t* red2org, int32_t n_red, double* dst, hipStream_t stream);
hipError_t gpk_mark(const int32_t* red2org, int32_t n_red, uint8_t* mask, hipStream_t stream);
hipError_t gpk_restore_fixed(const int32_t* idx, const double* val, int32_t count, double* x, hipStream_t stream);
hipError_t gpk_replay(gp_tape_dev tape, gp_orig_dev orig, double* x, double* y, double* z, uint8_t* zret,
                      int32_t begin, int32_t end, int32_t primal, int32_t dual, double tol, hipStream_t stream);
/* Dependency-level replay (the reverse tape order split into levels whose
 * records touch disjoint variables).  gpk_replay_list: the records of one
 * level, one per thread.  gpk_replay_levels: nlevels consecutive levels in one
 * workgroup, level lv being records[level_start[lv] .. level_start[lv+1]). */
hipError_t gpk_replay_list(gp_tape_dev tape, gp_orig_dev orig, double* x, double* y, double* z, uint8_t* zret,
                           const int32_t* records, int32_t count, int32_t primal, int32_t dual, double tol,
                           hipStream_t stream);
hipError_t gpk_replay_levels(gp_tape_dev tape, gp_orig_dev orig, double* x, double* y, double* z, uint8_t* zret,
                             const int32_t* records, const int32_t* level_start, int32_t nlevels, int32_t primal,
                             int32_t dual, double tol, hipStream_t stream);
/* len[i] = ptr[list[i] + 1] - ptr[list[i]]; gpk_seg_copy packs line
 * list[i]'s idx entries at out + off[i]. */
hipError_t gpk_seg_lengths(const int32_t* ptr, const int32_t* list, int32_t count, int32_t* len, hipStream_t stream);
hipError_t gpk_seg_copy(const int32_t* ptr, const int32_t* idx, const int32_t* list, int32_t count,
                        const int32_t* off, int32_t* out, hipStream_t stream);
hipError_t gpk_protect_columns(gp_tape_dev tape, int32_t n, uint8_t* prot, hipStream_t stream);
/* long_count: one zeroed device word; long_cols: n words of scratch (columns
 * with more than 64 entries are projected one wave per column). */
hipError_t gpk_project_duals(gp_orig_dev orig, const uint8_t* prot, const double* x, const double* y, double* z,
                             int32_t n, double tol, int32_t* long_count, int32_t* long_cols, hipStream_t stream);

/* Primal recovery steps of the structure-specific reductions
 * (gpu_postsolve.cu:44-141), in place on x (k_recovery.hip).  Splits:
 * x[t] = t + e, x[e] = (t - e) / rho; outer pairs: x[free] = x[bound];
 * linked slacks: x[slack] = factor * x[t]; max slacks (one per slack, entries
 * starts[s] .. starts[s+1]): x[slack] = max(lower, max factor * x[t]) with
 * lower the finite original lower bound or 0 (original_l may be null);
 * antipodal: the root values are cached first, then x[plus] = plus_lower +
 * root_plus value and x[minus] = minus_lower + root_minus value. */
hipError_t gpk_restore_splits(double* x, const int32_t* t_cols, const int32_t* e_cols, const double* rho, int32_t k,
                              hipStream_t stream);
hipError_t gpk_restore_outer_pairs(double* x, const int32_t* bound_cols, const int32_t* free_cols, int32_t k,
                                   hipStream_t stream);
hipError_t gpk_restore_linked_slacks(double* x, const int32_t* slack_cols, const int32_t* t_cols,
                                     const double* factors, int32_t k, hipStream_t stream);
hipError_t gpk_restore_max_slacks(double* x, const int32_t* slack_cols, const int32_t* starts, const int32_t* t_cols,
                                  const double* factors, int32_t slacks, const double* original_l,
                                  hipStream_t stream);
hipError_t gpk_restore_antipodal(double* x, const int32_t* plus_cols, const int32_t* minus_cols,
                                 const int32_t* root_plus, const int32_t* root_minus, const double* plus_lower,
                                 const double* minus_lower, double* plus_cache, double* minus_cache, int32_t k,
                                 hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GP_RECORD_KERNELS_H */
