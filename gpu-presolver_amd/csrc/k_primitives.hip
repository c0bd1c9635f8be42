// Primitive kernels of the presolve library: fills, scans, row activity,
// column max-abs, CSR statistics, mask compaction, LSD radix sort and the
// transpose built on it.  gfx950, wave64, 256-thread blocks.
#include <hip/hip_runtime.h>

#include <cmath>
#include <climits>
#include <cstdint>

#include "gp_device.hpp"
#include "gp_kernels.h"

using namespace gpdev;

namespace {

// ---------------------------------------------------------------- fills ----
template <class T>
__global__ void k_fill(T* dst, T v, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = v;
}

__global__ void k_iota(int32_t* dst, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = i;
}

// ---------------------------------------------------------------- scans ----
// Reduce-then-scan: tile sums, one-block scan of tile sums, tile downsweep.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048

template <class In>
__device__ inline int32_t scan_value(const In* in, int64_t i);
template <>
__device__ inline int32_t scan_value<int32_t>(const int32_t* in, int64_t i) { return in[i]; }
template <>
__device__ inline int32_t scan_value<uint8_t>(const uint8_t* in, int64_t i) { return in[i] != 0 ? 1 : 0; }

template <class In>
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const In* in, int64_t n, int32_t* tile_sums) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += scan_value(in, base + k);
  int32_t total;
  (void)block_exclusive_sum(s, &total);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One block scans `tiles` tile sums in place (exclusive) and stores the grand total in out[n].
__global__ __launch_bounds__(kBlock) void k_scan_tile_sums(int32_t* tile_sums, int64_t tiles, int32_t* out_total) {
  __shared__ int32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int64_t b = 0; b < tiles; b += kBlock) {
    const int64_t i = b + threadIdx.x;
    const int32_t v = i < tiles ? tile_sums[i] : 0;
    int32_t total;
    const int32_t ex = block_exclusive_sum(v, &total);
    const int32_t carry = s_carry;
    if (i < tiles) tile_sums[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) s_carry = carry + total;
    __syncthreads();
  }
  if (threadIdx.x == 0) *out_total = s_carry;
}

template <class In>
__global__ __launch_bounds__(kBlock) void k_scan_down(const In* in, int32_t* out, int64_t n, const int32_t* tile_prefix) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? scan_value(in, base + k) : 0;
    s += v[k];
  }
  int32_t total;
  int32_t run = block_exclusive_sum(s, &total) + tile_prefix[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

__global__ void k_old_to_new(const uint8_t* mask, const int32_t* prefix, int32_t* o2n, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o2n[i] = mask[i] != 0 ? prefix[i] : -1;
}

// -------------------------------------------------------- row activity ----
// One thread per row, 256 consecutive rows per block.  The block's CSR
// segment is streamed through LDS in windows: all threads cooperatively load
// coalesced (col, a) entries, then gather the packed (l, u) pairs (all loads
// of a window issued before any is used), form the min/max terms and stage
// them; each thread then adds its own row's terms in CSR order, so every row
// sum has exactly the sequential rounding of the reference kernel
// (gpu_presolve_kernels.cu:18-66) for any row length.
constexpr int kActRows = kBlock;
constexpr int kActCap = 2048;
constexpr int kActItems = kActCap / kBlock;

// (l, u) interleaved, so an activity term needs one 16-byte gather.
__global__ void k_pack_lu(const double* __restrict__ l, const double* __restrict__ u, int32_t n,
                          double2* __restrict__ lu) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) lu[j] = make_double2(l[j], u[j]);
}

__global__ __launch_bounds__(kBlock) void k_row_activity_lu(const int32_t* __restrict__ row_ptr,
                                                            const int32_t* __restrict__ col_idx,
                                                            const double* __restrict__ val, int32_t m,
                                                            const double2* __restrict__ lu, double ztol,
                                                            double* __restrict__ min_fin,
                                                            double* __restrict__ max_fin,
                                                            int32_t* __restrict__ neg_inf,
                                                            int32_t* __restrict__ pos_inf) {
  // One term array (the min terms, then the max terms): 16 KB of LDS per
  // block, so eight blocks fit a CU and their load phases overlap.
  __shared__ double s_t[kActCap];
  __shared__ int32_t s_ptr[kActRows + 1];
  const int32_t r0 = blockIdx.x * kActRows;
  const int32_t rows = min(kActRows, m - r0);
  const int tid = threadIdx.x;
  for (int t = tid; t <= rows; t += kBlock) s_ptr[t] = row_ptr[r0 + t];
  __syncthreads();
  const int32_t seg_b = s_ptr[0], seg_e = s_ptr[rows];
  const int32_t my_b = tid < rows ? s_ptr[tid] : 0;
  const int32_t my_e = tid < rows ? s_ptr[tid + 1] : 0;
  double mn = 0.0, mx = 0.0;
  int32_t ni = 0, pi = 0;
  for (int32_t w = seg_b; w < seg_e; w += kActCap) {
    const int32_t we = min(w + kActCap, seg_e);
    double a[kActItems];
    double2 b[kActItems];
    int32_t j[kActItems];
    // Every load is issued unconditionally (positions past the window read
    // its last entry and are discarded), so the compiler emits the 16 loads
    // back to back and then the 8 gathers, with no wait in between.
#pragma unroll
    for (int k = 0; k < kActItems; ++k) {
      const int32_t pc = min(w + tid + k * kBlock, we - 1);
      a[k] = val[pc];
      j[k] = col_idx[pc];
    }
#pragma unroll
    for (int k = 0; k < kActItems; ++k) {
      if (w + tid + k * kBlock >= we) a[k] = 0.0;
      b[k] = lu[fabs(a[k]) > ztol ? j[k] : 0];
    }
    const int32_t rb = max(my_b, w), re = min(my_e, we);
    double tmx[kActItems];  // the max terms wait in registers while the min terms are summed
#pragma unroll
    for (int k = 0; k < kActItems; ++k) {
      const int32_t p = w + tid + k * kBlock;
      const bool on = fabs(a[k]) > ztol;
      tmx[k] = on ? (a[k] >= 0.0 ? a[k] * b[k].y : a[k] * b[k].x) : 0.0;
      if (p < we) s_t[p - w] = on ? (a[k] >= 0.0 ? a[k] * b[k].x : a[k] * b[k].y) : 0.0;
    }
    __syncthreads();
    for (int32_t p = rb; p < re; ++p) {
      const double t = s_t[p - w];
      if (isfinite(t)) mn += t; else if (t < 0.0) ++ni;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kActItems; ++k) {
      const int32_t p = w + tid + k * kBlock;
      if (p < we) s_t[p - w] = tmx[k];
    }
    __syncthreads();
    for (int32_t p = rb; p < re; ++p) {
      const double t = s_t[p - w];
      if (isfinite(t)) mx += t; else if (t > 0.0) ++pi;
    }
    __syncthreads();
  }
  if (tid < rows) {
    min_fin[r0 + tid] = mn;
    max_fin[r0 + tid] = mx;
    neg_inf[r0 + tid] = ni;
    pos_inf[r0 + tid] = pi;
  }
}

__global__ void k_col_max_abs(const int32_t* at_ptr, const double* at_val, int32_t n, double* out) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double v = 0.0;
  for (int32_t p = at_ptr[j]; p < at_ptr[j + 1]; ++p) v = fmax(v, fabs(at_val[p]));
  out[j] = v;
}

__global__ void k_csr_stats(const int32_t* ptr, const int32_t* idx, const double* val, int32_t rows,
                            int32_t* nnz_out, uint8_t* empty_mask, uint8_t* singleton_mask,
                            int32_t* singleton_idx, double* singleton_val) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const int32_t b = ptr[i];
  const int32_t cnt = ptr[i + 1] - b;
  if (nnz_out) nnz_out[i] = cnt;
  if (empty_mask) empty_mask[i] = cnt == 0 ? 1 : 0;
  if (singleton_mask) singleton_mask[i] = cnt == 1 ? 1 : 0;
  if (singleton_idx) singleton_idx[i] = cnt == 1 ? idx[b] : -1;
  if (singleton_val) singleton_val[i] = cnt == 1 ? val[b] : 0.0;
}

// ---------------------------------------------------------- compaction ----
// A tile's entry segment is processed in chunks of kCompChunk entries,
// striped: entry c + 256k + t goes to thread t, so every load instruction of
// a wave reads 64 consecutive entries.
constexpr int kCompItems = 8;
constexpr int kCompChunk = kBlock * kCompItems;

__device__ inline int find_row(const int32_t* s_ptr, int rows, int32_t p) {
  int lo = 0, hi = rows;  // s_ptr[lo] <= p < s_ptr[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s_ptr[mid] <= p) lo = mid; else hi = mid;
  }
  return lo;
}

// ------------------------------------------------ decoupled look-back ----
constexpr uint64_t kLbSpinLimit = 1ull << 20;

__device__ inline uint64_t lb_word(uint32_t epoch, uint32_t flag, int32_t value) {
  return (static_cast<uint64_t>(epoch) << 34) | (static_cast<uint64_t>(flag) << 32) |
         static_cast<uint64_t>(static_cast<uint32_t>(value));
}

// Called by the first wave of the tile (all 64 lanes).  Publishes the tile's
// aggregate, then looks back over kLbWindow predecessors at a time, lane l
// watching kLbPerLane consecutive ones (nearest first): every lane waits for
// its predecessors' words, the nearest inclusive prefix in the window ends the
// walk, otherwise the window's aggregates are summed and the walk moves a
// window further back.  A wide window matters on this GPU: a status word
// written on one XCD is read through memory by the others, so each step
// costs a memory round trip, and the walk's depth is tiles / window.
// Publishes the tile's inclusive prefix and returns its exclusive prefix (in
// every lane).  Tile ids are tickets taken in dispatch order, so every
// predecessor is already running and each wait ends; the spin limit only
// guards a broken invariant (*error is set and the result is meaningless).
// The status words carry all the data a reader needs, so relaxed
// device-scope atomics suffice: acquire/release would add cache
// invalidation / write-back on every poll.
constexpr int kLbPerLane = 4;
constexpr int kLbWindow = 64 * kLbPerLane;

__device__ inline uint64_t lb_wait(const gpk_lookback& lb, int32_t j, int32_t first = 0) {
  if (j < first) return lb_word(lb.epoch, 2u, 0);  // before the chain's first tile: an inclusive prefix of 0
  uint64_t spins = 0;
  while (true) {
    const uint64_t w = __hip_atomic_load(lb.status + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((w >> 34) == lb.epoch && ((w >> 32) & 3u) != 0) return w;
    if (++spins > kLbSpinLimit) {
      __hip_atomic_store(lb.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      return lb_word(lb.epoch, 2u, 0);
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// `first`: the first tile of this tile's chain (several independent scans
// can share one launch and one status array, chain after chain in ticket
// order).
__device__ int32_t lb_exclusive(const gpk_lookback& lb, int32_t tile, int32_t aggregate, int32_t first = 0) {
  const int lane = lane_id();
  if (lane == 0)
    __hip_atomic_store(lb.status + tile, lb_word(lb.epoch, tile == first ? 2u : 1u, aggregate), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  if (tile == first) return 0;
  int32_t excl = 0;
  for (int32_t j0 = tile - 1;; j0 -= kLbWindow) {
    uint64_t w[kLbPerLane];
#pragma unroll
    for (int q = 0; q < kLbPerLane; ++q) w[q] = lb_wait(lb, j0 - lane * kLbPerLane - q, first);
    int q_incl = kLbPerLane;  // this lane's nearest inclusive prefix
#pragma unroll
    for (int q = kLbPerLane - 1; q >= 0; --q)
      if (((w[q] >> 32) & 3u) == 2u) q_incl = q;
    const uint64_t incl = __ballot(q_incl < kLbPerLane);
    const int first = incl ? __ffsll(static_cast<unsigned long long>(incl)) - 1 : 64;
    int32_t v = 0;
    if (lane <= first) {
      const int last_q = lane < first ? kLbPerLane - 1 : q_incl;
#pragma unroll
      for (int q = 0; q < kLbPerLane; ++q)
        if (q <= last_q) v += static_cast<int32_t>(static_cast<uint32_t>(w[q]));
    }
    excl += __shfl(wave_inclusive_sum(v), 63, 64);
    if (incl) break;
  }
  if (lane == 0)
    __hip_atomic_store(lb.status + tile, lb_word(lb.epoch, 2u, excl + aggregate), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

__device__ inline int32_t lb_ticket(const gpk_lookback& lb) {
  __shared__ int32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(lb.ticket, 1);
  __syncthreads();
  return s_tile;
}

// Exclusive scan, 2048 items per tile (8 consecutive per thread).  mode 0:
// out = prefix (out[n] = total); mode 1 (mask input): out = prefix or -1 for
// unset entries, *count = total.
template <class In, int kMode>
__global__ __launch_bounds__(kBlock) void k_scan_lb(const In* in, int32_t* out, int64_t n, int32_t* count,
                                                    gpk_lookback lb) {
  constexpr int kWaves = kBlock / 64;
  __shared__ int32_t s_wbase[kScanItems * kWaves];
  __shared__ int32_t s_excl;
  const int32_t tile = lb_ticket(lb);
  if (tile >= static_cast<int32_t>(gridDim.x)) {  // the ticket word was not zeroed
    if (threadIdx.x == 0) __hip_atomic_store(lb.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return;
  }
  // Striped items: element base + 256k + t belongs to thread t, so every load
  // and store instruction of a wave covers 64 consecutive elements.  Each
  // slice k is scanned across the wave with shuffles; the (slice, wave)
  // totals, in element order, give each wave-slice its offset in the tile.
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int64_t base = static_cast<int64_t>(tile) * kScanTile + tid;
  int32_t v[kScanItems], inc[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kBlock;
    v[k] = i < n ? scan_value(in, i) : 0;
  }
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    inc[k] = wave_inclusive_sum(v[k]);
    if (lane == 63) s_wbase[k * kWaves + wave] = inc[k];
  }
  __syncthreads();
  if (tid < 64) {
    const int32_t c = tid < kScanItems * kWaves ? s_wbase[tid] : 0;
    const int32_t ci = wave_inclusive_sum(c);
    if (tid < kScanItems * kWaves) s_wbase[tid] = ci - c;
    const int32_t total = __shfl(ci, 63, 64);
    const int32_t e = lb_exclusive(lb, tile, total);
    if (tid == 0) s_excl = e;
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tid == 0 && tile == tiles - 1) {
      if (kMode == 1) *count = e + total;
      else out[n] = e + total;
    }
  }
  __syncthreads();
  const int32_t excl = s_excl;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kBlock;
    const int32_t prefix = excl + s_wbase[k * kWaves + wave] + inc[k] - v[k];
    if (i < n) out[i] = kMode == 1 ? (v[k] ? prefix : -1) : prefix;
  }
}

// The three keep maps of one plan application in one launch
// (build_old_to_new for rows and columns, gpu_presolve.cu:878-934, and the
// map of removed columns with a fixed interval): three look-back chains in
// one status array, ticket order rows, then columns, then fixed columns.
// The fixed-column mask is computed from keep_col, l and u on the way and
// stored (removed, both bounds finite, |u - l| <= 1e-9).
__global__ __launch_bounds__(kBlock) void k_apply_maps(gpk_apply_maps_args a, int32_t row_tiles, int32_t col_tiles,
                                                       gpk_lookback lb) {
  constexpr int kWaves = kBlock / 64;
  __shared__ int32_t s_wbase[kScanItems * kWaves];
  __shared__ int32_t s_excl;
  const int32_t tile = lb_ticket(lb);
  if (tile >= static_cast<int32_t>(gridDim.x)) {  // the ticket word was not zeroed
    if (threadIdx.x == 0) __hip_atomic_store(lb.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return;
  }
  const int chain = tile < row_tiles ? 0 : (tile < row_tiles + col_tiles ? 1 : 2);
  const int32_t first = chain == 0 ? 0 : (chain == 1 ? row_tiles : row_tiles + col_tiles);
  const int64_t n = chain == 0 ? a.m : a.n;
  int32_t* out = chain == 0 ? a.row_map : (chain == 1 ? a.col_map : a.fixed_map);
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int64_t base = static_cast<int64_t>(tile - first) * kScanTile + tid;
  int32_t v[kScanItems], inc[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kBlock;
    int32_t x = 0;
    if (i < n) {
      if (chain == 0) {
        x = a.keep_row[i] != 0;
      } else if (chain == 1) {
        x = a.keep_col[i] != 0;
      } else {
        const double lo = a.l[i], hi = a.u[i];
        x = (!a.keep_col[i] && isfinite(lo) && isfinite(hi) && fabs(lo - hi) <= 1.0e-9) ? 1 : 0;
        a.fixed_mask[i] = static_cast<uint8_t>(x);
      }
    }
    v[k] = x;
  }
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    inc[k] = wave_inclusive_sum(v[k]);
    if (lane == 63) s_wbase[k * kWaves + wave] = inc[k];
  }
  __syncthreads();
  if (tid < 64) {
    const int32_t c = tid < kScanItems * kWaves ? s_wbase[tid] : 0;
    const int32_t ci = wave_inclusive_sum(c);
    if (tid < kScanItems * kWaves) s_wbase[tid] = ci - c;
    const int32_t total = __shfl(ci, 63, 64);
    const int32_t e = lb_exclusive(lb, tile, total, first);
    if (tid == 0) s_excl = e;
    const int32_t last = chain == 0 ? row_tiles - 1 : (chain == 1 ? row_tiles + col_tiles - 1 : gridDim.x - 1);
    if (tid == 0 && tile == last) a.counts[chain] = e + total;
  }
  __syncthreads();
  const int32_t excl = s_excl;
  gpk_rank_word* rank = chain == 0 ? a.row_rank : (chain == 1 ? a.col_rank : nullptr);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kBlock;
    const int32_t prefix = excl + s_wbase[k * kWaves + wave] + inc[k] - v[k];
    if (i < n) out[i] = v[k] ? prefix : -1;
    if (rank) {
      // A wave's 64 items of slice k are 64 consecutive, 64-aligned entries:
      // one rank word (kept bits and the new index of the first kept one).
      const uint64_t bits = __ballot(v[k] != 0);
      const int lead = bits ? __ffsll(static_cast<unsigned long long>(bits)) - 1 : 0;
      const int32_t first_new = __shfl(prefix, lead, 64);
      const int64_t i0 = i - lane;
      if (lane == 0 && i0 < n) {
        gpk_rank_word w;
        w.bits = bits;
        w.base = bits ? first_new : 0;
        w.pad = 0;
        rank[i0 >> 6] = w;
      }
    }
  }
}

// Small scans in one workgroup (no tickets, no look-back): 1024 threads walk
// the array in chunks of 8192 with a running carry, each chunk striped and
// scanned like a look-back tile.  Same outputs as k_scan_lb.
constexpr int kSmallBlock = 1024;
constexpr int64_t kSmallScanMax = 16384;

template <class In, int kMode>
__global__ __launch_bounds__(kSmallBlock) void k_scan_small(const In* in, int32_t* out, int64_t n, int32_t* count) {
  constexpr int kWaves = kSmallBlock / 64;
  constexpr int kSlots = kScanItems * kWaves;  // 128
  __shared__ int32_t s_w[kSlots];
  __shared__ int32_t s_total;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  int32_t carry = 0;
  for (int64_t c = 0; c < n; c += static_cast<int64_t>(kScanItems) * kSmallBlock) {
    int32_t v[kScanItems], inc[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const int64_t i = c + k * kSmallBlock + tid;
      v[k] = i < n ? scan_value(in, i) : 0;
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      inc[k] = wave_inclusive_sum(v[k]);
      if (lane == 63) s_w[k * kWaves + wave] = inc[k];
    }
    __syncthreads();
    if (tid < 64) {  // exclusive prefix of the 128 (slice, wave) totals, two per lane
      const int32_t a = s_w[2 * tid], b = s_w[2 * tid + 1];
      const int32_t pi = wave_inclusive_sum(a + b);
      s_w[2 * tid] = pi - a - b;
      s_w[2 * tid + 1] = pi - b;
      if (tid == 63) s_total = pi;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const int64_t i = c + k * kSmallBlock + tid;
      const int32_t prefix = carry + s_w[k * kWaves + wave] + inc[k] - v[k];
      if (i < n) out[i] = kMode == 1 ? (v[k] ? prefix : -1) : prefix;
    }
    carry += s_total;
    __syncthreads();  // s_w and s_total are rewritten by the next chunk
  }
  if (tid == 0) {
    if (kMode == 1) *count = carry;
    else out[n] = carry;
  }
}

// Compaction (gpu_presolve.cu:1881-1978): after the column map's rank
// words, three launches over tiles of 256 old rows: k_compact_map maps
// every entry once (row kept and column rank,
// one cache-resident rank gather per entry; with no rank words, when no
// column is dropped, the column keeps its index), stores the new column index or
// -1 per entry, each kept row's kept count at out_ptr[new row] and the
// tile's kept count; an exclusive scan of the tile counts (one small array,
// m/256 entries) gives every tile its output offset; k_compact_write turns
// its rows' counts into row pointers (a block scan plus the tile offset),
// then streams the mapped indices and the values and writes the kept
// entries, positioned by wave ballots and staged in LDS for coalesced stores.
constexpr int kCompRows = kBlock;

// Rank form of a keep map (map[j] = new index or -1), one 16-byte word per 64
// entries: bit b of word w is set when entry 64w + b is kept, and base is the
// new index of the word's first kept entry.  A lookup then reads one small,
// cache-resident word instead of gathering from the n-entry map.
__global__ void k_build_rank(const int32_t* __restrict__ map, int32_t n, gpk_rank_word* __restrict__ out) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t v = j < n ? map[j] : -1;
  const uint64_t bits = __ballot(v >= 0);
  const int first = bits ? __ffsll(static_cast<unsigned long long>(bits)) - 1 : 0;
  const int32_t base = __shfl(v, first, 64);
  if (lane_id() == 0 && (j >> 6) <= ((n - 1) >> 6)) {
    gpk_rank_word w;
    w.bits = bits;
    w.base = bits ? base : 0;
    w.pad = 0;
    out[j >> 6] = w;
  }
}

// One 16-byte load of the word, then the rank; -1 when entry j is dropped.
__device__ inline uint4 rank_word(const gpk_rank_word* __restrict__ rank, int32_t j) {
  return *reinterpret_cast<const uint4*>(rank + (j >> 6));
}
__device__ inline int32_t rank_of(uint4 w, int32_t j) {
  const uint64_t bits = static_cast<uint64_t>(w.x) | (static_cast<uint64_t>(w.y) << 32);
  const uint64_t bit = 1ull << (j & 63);
  const int32_t r = static_cast<int32_t>(w.z) + static_cast<int32_t>(__popcll(bits & (bit - 1)));
  return (bits & bit) ? r : -1;
}

// Per-row kept counts of one striped item per lane: the lanes of a wave that
// hold entries of the same row are consecutive, so the first lane of each
// run adds the run's kept count with one LDS atomic.
__device__ inline void comp_count_run(int row, bool valid, bool kept, int32_t* s_cnt) {
  const int lane = lane_id();
  const uint64_t kb = __ballot(kept);
  const int prev = __shfl_up(row, 1, 64);
  const bool head = valid && (lane == 0 || prev != row);
  const uint64_t hb = __ballot(head);
  if (!head) return;
  const uint64_t after = lane == 63 ? 0ull : (hb & (~0ull << (lane + 1)));
  const int next = after ? __ffsll(static_cast<unsigned long long>(after)) - 1 : 64;
  const uint64_t run = (next == 64 ? ~0ull : ((1ull << next) - 1)) & ~lanemask_lt(lane);
  const int c = static_cast<int>(__popcll(kb & run));
  if (c) atomicAdd(&s_cnt[row], c);
}

// The map pass of one 256-row tile (tile_cnt[tile]: its kept count).
__device__ __forceinline__ void compact_map_tile(int32_t tile, const int32_t* __restrict__ ptr,
                                                 const int32_t* __restrict__ idx, int32_t m_old,
                                                 const int32_t* __restrict__ row_map,
                                                 const gpk_rank_word* __restrict__ col_rank,
                                                 int32_t* __restrict__ jn_out, int32_t* __restrict__ out_ptr,
                                                 int32_t* __restrict__ tile_cnt, int32_t* __restrict__ kept_total) {
  __shared__ int32_t s_cnt[kCompRows];
  __shared__ int32_t s_wsum[kBlock / 64];
  __shared__ uint8_t s_keep[kCompRows];
  __shared__ uint8_t s_row[kCompChunk];  // tile-local row of each entry of the chunk
  static_assert(kCompRows <= 256, "tile-local rows are stored in one byte");
  const int32_t r0 = tile * kCompRows;
  const int rows = min(kCompRows, m_old - r0);
  const int tid = threadIdx.x;
  const int32_t my_new_row = tid < rows ? row_map[r0 + tid] : -1;
  const int32_t my_b = tid < rows ? ptr[r0 + tid] : 0;
  const int32_t my_e = tid < rows ? ptr[r0 + tid + 1] : 0;
  s_cnt[tid] = 0;
  s_keep[tid] = my_new_row >= 0 ? 1 : 0;
  const int32_t seg_b = ptr[r0], seg_e = ptr[r0 + rows];
  for (int32_t c = seg_b; c < seg_e; c += kCompChunk) {
    // Each thread labels its row's entries of the chunk; then rows and
    // liveness come from LDS, and every load goes out at once (a register a
    // load is still writing is never assigned).
    for (int32_t q = max(my_b, c); q < min(my_e, c + kCompChunk); ++q) s_row[q - c] = static_cast<uint8_t>(tid);
    __syncthreads();
    int rr[kCompItems];
    bool live[kCompItems];
#pragma unroll
    for (int k = 0; k < kCompItems; ++k) {
      const int32_t p = c + k * kBlock + tid;
      rr[k] = p < seg_e ? static_cast<int>(s_row[p - c]) : -1;
      live[k] = rr[k] >= 0 && s_keep[rr[k]] != 0;
    }
    int32_t cj[kCompItems];
#pragma unroll
    for (int k = 0; k < kCompItems; ++k) cj[k] = idx[min(c + k * kBlock + tid, seg_e - 1)];
    uint4 w[kCompItems];
    if (col_rank != nullptr) {
#pragma unroll
      for (int k = 0; k < kCompItems; ++k) w[k] = rank_word(col_rank, live[k] ? cj[k] : 0);
    }
#pragma unroll
    for (int k = 0; k < kCompItems; ++k) {
      const int32_t p = c + k * kBlock + tid;
      const int32_t jn = !live[k] ? -1 : col_rank != nullptr ? rank_of(w[k], cj[k]) : cj[k];
      if (p < seg_e) jn_out[p] = jn;
      comp_count_run(rr[k], rr[k] >= 0, jn >= 0, s_cnt);
    }
    __syncthreads();  // s_row is relabelled for the next chunk
  }
  const int32_t mine = my_new_row >= 0 ? s_cnt[tid] : 0;
  if (my_new_row >= 0) out_ptr[my_new_row] = mine;
  const int32_t inc = wave_inclusive_sum(mine);
  if (lane_id() == 63) s_wsum[tid >> 6] = inc;
  __syncthreads();
  if (tid == 0) {
    int32_t t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += s_wsum[w];
    tile_cnt[tile] = t;
    if (kept_total != nullptr && t != 0) atomicAdd(kept_total, t);  // integer: order-free
  }
}

__global__ __launch_bounds__(kBlock) void k_compact_map(const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ idx, int32_t m_old,
                                                        const int32_t* __restrict__ row_map,
                                                        const gpk_rank_word* __restrict__ col_rank,
                                                        int32_t* __restrict__ jn_out, int32_t* __restrict__ out_ptr,
                                                        int32_t* __restrict__ tile_cnt,
                                                        int32_t* __restrict__ kept_total) {
  compact_map_tile(static_cast<int32_t>(blockIdx.x), ptr, idx, m_old, row_map, col_rank, jn_out, out_ptr, tile_cnt,
                   kept_total);
}

// A and its transpose compacted with the same maps in one launch: blocks
// [0, tiles_a) map A's tiles, the rest AT's; the tile counts go to one array
// (A's, then AT's), so one scan serves both.
__global__ __launch_bounds__(kBlock) void k_compact_map_pair(gpk_compact_side a, gpk_compact_side t,
                                                             int32_t tiles_a, int32_t* __restrict__ tile_cnt,
                                                             int32_t* __restrict__ kept_total) {
  const int32_t b = static_cast<int32_t>(blockIdx.x);
  if (b < tiles_a)
    compact_map_tile(b, a.ptr, a.idx, a.m_old, a.row_map, a.col_rank, a.jn, a.out_ptr, tile_cnt, kept_total);
  else
    compact_map_tile(b - tiles_a, t.ptr, t.idx, t.m_old, t.row_map, t.col_rank, t.jn, t.out_ptr,
                     tile_cnt + tiles_a, nullptr);
}

// The write pass of tile `tile` of `ntiles`.  seg: the tiles' raw kept
// counts (from_counts; ntiles <= kBlock) or an exclusive scan of them whose
// first word is subtracted (a segment of a longer scan).
__device__ __forceinline__ void compact_write_tile(int32_t tile, int32_t ntiles, bool from_counts,
                                                   const int32_t* __restrict__ ptr, const double* __restrict__ val,
                                                   int32_t m_old, const int32_t* __restrict__ row_map,
                                                   const int32_t* __restrict__ jn_in,
                                                   const int32_t* __restrict__ seg, int32_t* __restrict__ out_ptr,
                                                   int32_t m_new, int32_t* __restrict__ out_idx,
                                                   double* __restrict__ out_val) {
  constexpr int kWaves = kBlock / 64;
  __shared__ int32_t s_oidx[kCompChunk];
  __shared__ double s_oval[kCompChunk];
  __shared__ int32_t s_wcnt[kCompItems * kWaves];
  __shared__ int32_t s_wsum[kWaves];
  __shared__ int32_t s_total;
  const int32_t r0 = tile * kCompRows;
  const int rows = min(kCompRows, m_old - r0);
  const int tid = threadIdx.x;
  const int lane = lane_id(), wave = tid >> 6;
  const int32_t my_new_row = tid < rows ? row_map[r0 + tid] : -1;
  int32_t base;
  if (from_counts) {
    // Few tiles (ntiles <= kBlock): seg holds the raw per-tile kept counts
    // and each block sums those of the tiles before it (and the last block
    // all of them) instead of waiting for a scan launch.
    const int32_t t = tid < ntiles ? seg[tid] : 0;
    const int32_t before = wave_inclusive_sum(tid < tile ? t : 0);
    const int32_t all = wave_inclusive_sum(t);
    if (lane == 63) {
      s_wsum[wave] = before;
      s_wcnt[wave] = all;  // s_wcnt is free until the chunk loop
    }
    __syncthreads();
    int32_t b = 0, a = 0;
    for (int w = 0; w < kWaves; ++w) {
      b += s_wsum[w];
      a += s_wcnt[w];
    }
    base = b;
    if (tile == ntiles - 1 && tid == 0) out_ptr[m_new] = a;
    __syncthreads();  // s_wsum is reused below
  } else {
    const int32_t s0 = seg[0];
    base = seg[tile] - s0;
    if (tile == ntiles - 1 && tid == 0) out_ptr[m_new] = seg[ntiles] - s0;
  }
  {  // row pointers of the tile's kept rows: tile offset + block exclusive scan of the counts
    const int32_t cnt = my_new_row >= 0 ? out_ptr[my_new_row] : 0;
    const int32_t inc = wave_inclusive_sum(cnt);
    if (lane == 63) s_wsum[wave] = inc;
    __syncthreads();
    int32_t before = 0, all = 0;
    for (int w = 0; w < kWaves; ++w) {
      before += w < wave ? s_wsum[w] : 0;
      all += s_wsum[w];
    }
    if (my_new_row >= 0) out_ptr[my_new_row] = base + before + inc - cnt;
    if (all == 0) return;  // no kept entry in the tile (the same in every thread)
  }
  const int32_t seg_b = ptr[r0], seg_e = ptr[r0 + rows];
  const uint64_t below = lanemask_lt(lane);
  for (int32_t c = seg_b; c < seg_e; c += kCompChunk) {
    int32_t jn[kCompItems];
    double v[kCompItems];
#pragma unroll
    for (int k = 0; k < kCompItems; ++k) {
      const int32_t pc = min(c + k * kBlock + tid, seg_e - 1);
      jn[k] = jn_in[pc];
      v[k] = val[pc];
    }
    uint64_t ball[kCompItems];
#pragma unroll
    for (int k = 0; k < kCompItems; ++k) {
      const bool keep = c + k * kBlock + tid < seg_e && jn[k] >= 0;
      ball[k] = __ballot(keep);
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < kCompItems; ++k) s_wcnt[k * kWaves + wave] = static_cast<int32_t>(__popcll(ball[k]));
    }
    __syncthreads();
    if (tid < 64) {  // exclusive prefix over (slice, wave) in entry order
      const int32_t cnt = tid < kCompItems * kWaves ? s_wcnt[tid] : 0;
      const int32_t inc = wave_inclusive_sum(cnt);
      if (tid < kCompItems * kWaves) s_wcnt[tid] = inc - cnt;
      if (tid == 63) s_total = inc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kCompItems; ++k) {
      if ((ball[k] >> lane) & 1ull) {
        const int32_t pos = s_wcnt[k * kWaves + wave] + static_cast<int32_t>(__popcll(ball[k] & below));
        s_oidx[pos] = jn[k];
        s_oval[pos] = v[k];
      }
    }
    __syncthreads();
    const int32_t total = s_total;
    for (int t = tid; t < total; t += kBlock) {
      out_idx[base + t] = s_oidx[t];
      out_val[base + t] = s_oval[t];
    }
    base += total;
    __syncthreads();
  }
}

template <bool kOffsetsFromCounts>
__global__ __launch_bounds__(kBlock) void k_compact_write(const int32_t* __restrict__ ptr,
                                                          const double* __restrict__ val, int32_t m_old,
                                                          const int32_t* __restrict__ row_map,
                                                          const int32_t* __restrict__ jn_in,
                                                          const int32_t* __restrict__ tile_off,
                                                          int32_t* __restrict__ out_ptr, int32_t m_new,
                                                          int32_t* __restrict__ out_idx,
                                                          double* __restrict__ out_val) {
  compact_write_tile(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x), kOffsetsFromCounts, ptr, val,
                     m_old, row_map, jn_in, tile_off, out_ptr, m_new, out_idx, out_val);
}

// Both write passes of a pair map (k_compact_map_pair) in one launch.
__global__ __launch_bounds__(kBlock) void k_compact_write_pair(gpk_compact_side a, gpk_compact_side t,
                                                               int32_t tiles_a, int32_t tiles_t,
                                                               const int32_t* __restrict__ tile_off,
                                                               int32_t from_counts) {
  const int32_t b = static_cast<int32_t>(blockIdx.x);
  if (b < tiles_a)
    compact_write_tile(b, tiles_a, from_counts != 0, a.ptr, a.val, a.m_old, a.row_map, a.jn, tile_off, a.out_ptr,
                       a.m_new, a.out_idx, a.out_val);
  else
    compact_write_tile(b - tiles_a, tiles_t, from_counts != 0, t.ptr, t.val, t.m_old, t.row_map, t.jn,
                       tile_off + tiles_a, t.out_ptr, t.m_new, t.out_idx, t.out_val);
}

__global__ void k_mail(const uint32_t* __restrict__ src, int32_t words, uint32_t* mail, uint32_t seq) {
  const int t = threadIdx.x;
  if (t < words) mail[4 + t] = src[t];
  __threadfence_system();
  __syncthreads();
  if (t == 0) *reinterpret_cast<volatile uint32_t*>(mail) = seq;
}

struct CopyList {
  gpk_copy_seg seg[GPK_COPY_MAX];
};

template <class U>
__device__ inline void copy_units(const gpk_copy_seg& g) {
  const int64_t units = g.bytes / static_cast<int64_t>(sizeof(U));
  U* d = static_cast<U*>(g.dst);
  const U* src = static_cast<const U*>(g.src);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < units; i += stride)
    d[i] = src ? src[i] : U{};
}

__global__ __launch_bounds__(kBlock) void k_copy_many(CopyList list) {
  const gpk_copy_seg g = list.seg[blockIdx.y];
  const uint64_t a = reinterpret_cast<uint64_t>(g.dst) | reinterpret_cast<uint64_t>(g.src) |
                     static_cast<uint64_t>(g.bytes);
  if ((a & 15u) == 0) copy_units<uint4>(g);
  else if ((a & 7u) == 0) copy_units<uint2>(g);
  else if ((a & 3u) == 0) copy_units<uint32_t>(g);
  else copy_units<uint8_t>(g);
}

// The vector gathers of one plan application in one launch: c, l, u and
// the column reduced-to-original map through col_map; AL, AU and the row
// map through row_map (dst[map[i]] = src[i] for map[i] >= 0; a null
// destination is skipped).
__global__ void k_apply_gathers(gpk_apply_gathers_args a) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n) {
    const int32_t j = a.col_map[i];
    if (j >= 0) {
      a.c_out[j] = a.c_in[i];
      a.l_out[j] = a.l_in[i];
      a.u_out[j] = a.u_in[i];
      if (a.col_r2o_out) a.col_r2o_out[j] = a.col_r2o_in[i];
    }
  }
  if (i < a.m) {
    const int32_t r = a.row_map[i];
    if (r >= 0) {
      a.al_out[r] = a.al_in[i];
      a.au_out[r] = a.au_in[i];
      if (a.row_r2o_out) a.row_r2o_out[r] = a.row_r2o_in[i];
    }
  }
}

__global__ void k_gather5(const int32_t* map, int32_t n_old, const double* s0, double* d0, const double* s1,
                          double* d1, const double* s2, double* d2, const double* s3, double* d3,
                          const double* s4, double* d4) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_old) return;
  const int32_t t = map[i];
  if (t < 0) return;
  if (s0) d0[t] = s0[i];
  if (s1) d1[t] = s1[i];
  if (s2) d2[t] = s2[i];
  if (s3) d3[t] = s3[i];
  if (s4) d4[t] = s4[i];
}

__global__ void k_gather_i32(const int32_t* map, int32_t n_old, const int32_t* src, int32_t* dst) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_old) return;
  const int32_t t = map[i];
  if (t >= 0) dst[t] = src[i];
}

// ------------------------------------------------------------ radix sort ----
// Stable LSD radix sort, 8-bit digits.  Tile = 2048 keys; wave w of a block
// owns the contiguous keys [w*512, w*512+512) of the tile, read 64 at a time.
// Per 64-key step the lanes holding the same digit are found with eight
// ballots, ranked by lane order, and appended to a per-wave digit counter in
// LDS; digit offsets across waves and tiles come from the global histogram
// scan, so equal digits keep their input order.  The ranked tile is staged
// in LDS in digit order and stored run by run.
constexpr int kRsTile = 2048;
constexpr int kRsWaveItems = 512;
constexpr int kRsWaves = kBlock / 64;
// Up to this many tiles, each scatter block reads the whole histogram
// (256 x tiles words) instead of waiting for a scan launch.
constexpr int kRsHistTilesMax = 64;

template <class K>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const K* __restrict__ keys, int32_t n, int shift,
                                                      int32_t tiles, int32_t* __restrict__ hist) {
  __shared__ int32_t s_h[256];
  s_h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kRsTile;
  for (int k = 0; k < kRsTile / kBlock; ++k) {
    const int64_t e = base + k * kBlock + threadIdx.x;
    if (e < n) atomicAdd(&s_h[static_cast<int>((keys[e] >> shift) & 255u)], 1);
  }
  __syncthreads();
  hist[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x] = s_h[threadIdx.x];
}

// kFromHist: offs is the raw histogram (digit-major, tiles per digit) and
// every block derives its own digit offsets from it (for few tiles, this
// saves the separate scan launch); otherwise offs is the scanned histogram.
template <class K, bool kFromHist>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const K* __restrict__ kin, const int32_t* __restrict__ vin,
                                                         K* __restrict__ kout, int32_t* __restrict__ vout,
                                                         int32_t n, int shift, int32_t tiles,
                                                         const int32_t* __restrict__ offs) {
  __shared__ int32_t s_cnt[kRsWaves][256];
  __shared__ int32_t s_off[kRsWaves][256];
  __shared__ int32_t s_base[256];
  __shared__ int32_t s_start[256];
  __shared__ int32_t s_wsum[kRsWaves];
  __shared__ K s_key[kRsTile];
  __shared__ int32_t s_val[kRsTile];
  const int tid = threadIdx.x;
  for (int i = tid; i < kRsWaves * 256; i += kBlock) s_cnt[i >> 8][i & 255] = 0;
  if (kFromHist) {
    static_assert(kBlock == 256, "one thread per digit");
    int32_t total = 0, before = 0;
    for (int32_t t = 0; t < tiles; ++t) {
      const int32_t h = offs[static_cast<int64_t>(tid) * tiles + t];
      total += h;
      before += t < static_cast<int32_t>(blockIdx.x) ? h : 0;
    }
    const int32_t inc = wave_inclusive_sum(total);
    if ((tid & 63) == 63) s_wsum[tid >> 6] = inc;
    __syncthreads();
    int32_t wbase = 0;
    for (int w = 0; w < (tid >> 6); ++w) wbase += s_wsum[w];
    s_base[tid] = wbase + inc - total + before;
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const uint64_t lt = lanemask_lt(lane);
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kRsTile + wave * kRsWaveItems;
  constexpr int kSteps = kRsWaveItems / 64;
  K kk[kSteps];
  int32_t vv[kSteps];
  int dig[kSteps];
  int rank[kSteps];
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    const int64_t e = base + k * 64 + lane;
    const bool valid = e < n;
    int d = 0;
    if (valid) {
      kk[k] = kin[e];
      vv[k] = vin ? vin[e] : static_cast<int32_t>(e);  // no values: the input positions
      d = static_cast<int>((kk[k] >> shift) & 255u);
    }
    dig[k] = d;
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t bal = __ballot(bit);
      match &= bit ? bal : ~bal;
    }
    rank[k] = 0;
    if (valid) {
      const int r = __popcll(match & lt);
      const int prev = s_cnt[wave][d];
      rank[k] = prev + r;
      if (r == 0) s_cnt[wave][d] = prev + __popcll(match);
    }
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // Thread `tid` is digit tid: offsets of the waves' runs inside the tile's
  // run of the digit, and the digit's run start in the tile's digit-sorted
  // order.
  int32_t tile_cnt;
  {
    int32_t run = 0;
    for (int w = 0; w < kRsWaves; ++w) {
      s_off[w][tid] = run;
      run += s_cnt[w][tid];
    }
    s_start[tid] = block_exclusive_sum(run, &tile_cnt);
    if (!kFromHist) s_base[tid] = offs[static_cast<int64_t>(tid) * tiles + blockIdx.x];
  }
  __syncthreads();
  // The tile is staged in LDS in digit order, then written out by
  // consecutive threads to consecutive positions of each digit's run, so
  // the stores of a digit's run are contiguous.
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    const int64_t e = base + k * 64 + lane;
    if (e < n) {
      const int d = dig[k];
      const int32_t local = s_start[d] + s_off[wave][d] + rank[k];
      s_key[local] = kk[k];
      s_val[local] = vv[k];
    }
  }
  __syncthreads();
  for (int i = tid; i < tile_cnt; i += kBlock) {
    const K key = s_key[i];
    const int d = static_cast<int>((key >> shift) & 255u);
    const int32_t pos = s_base[d] + (i - s_start[d]);
    kout[pos] = key;
    vout[pos] = s_val[i];
  }
}

// ------------------------------------------------------------- transpose ----
// (row, value) of every entry of A packed into 16 bytes, so that the
// transpose's gather reads one pair per entry instead of two scattered
// words: one thread per entry over tiles of 256 rows (an entry finds its row
// by a binary search of the tile's row pointers in LDS), so reads and writes
// stay coalesced whatever the row lengths.
__global__ __launch_bounds__(kBlock) void k_expand_rows(const int32_t* __restrict__ ptr,
                                                        const double* __restrict__ val, int32_t m,
                                                        gpk_row_val* __restrict__ rv) {
  __shared__ int32_t s_ptr[kBlock + 1];
  const int32_t r0 = blockIdx.x * kBlock;
  const int nr = min(kBlock, m - r0);
  for (int t = threadIdx.x; t <= nr; t += kBlock) s_ptr[t] = ptr[r0 + t];
  __syncthreads();
  for (int32_t p = s_ptr[0] + threadIdx.x; p < s_ptr[nr]; p += kBlock) {
    const double v = val[p];
    reinterpret_cast<int4*>(rv)[p] =
        make_int4(r0 + find_row(s_ptr, nr, p), 0, __double2loint(v), __double2hiint(v));
  }
}

// Row pointers of the transpose from its sorted column keys: position p
// starts every column in (key[p-1], key[p]] (p = nnz closes the columns up
// to n), so each pointer is written once and no counter is shared.
__global__ void k_col_bounds(const uint32_t* keys, int32_t nnz, int32_t n, int32_t* ptr) {
  const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > nnz) return;
  const int32_t hi = p < nnz ? static_cast<int32_t>(keys[p]) : n;
  const int32_t lo = p > 0 ? static_cast<int32_t>(keys[p - 1]) : -1;
  for (int32_t j = lo + 1; j <= hi; ++j) ptr[j] = p;
}

__global__ void k_transpose_gather(const int32_t* __restrict__ perm, const gpk_row_val* __restrict__ rv,
                                   int32_t nnz, int32_t* __restrict__ at_idx, double* __restrict__ at_val) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int4 w = reinterpret_cast<const int4*>(rv)[perm[k]];  // one 16-byte load
  at_idx[k] = w.x;
  at_val[k] = __hiloint2double(w.w, w.z);
}

// CSR checks over tiles of 256 rows with one thread per entry, so a long
// row (a popular column of AT) is shared by the whole tile instead of walked
// by one thread.  The tile's row pointers go to LDS; every entry finds its
// row by a binary search of them and is compared with the entry before it in
// the same row.  kFull: flags[0] bad row pointers (not monotone, not from 0
// to nnz: the tile then reads no entry), flags[1] a column outside [0, cols),
// flags[2] two equal neighbours, flags[3] a descending pair.  Otherwise the
// pointers are trusted and only a descending pair is flagged, in flags[0].
template <bool kFull>
__global__ __launch_bounds__(kBlock) void k_csr_tiles(const int32_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ idx, int32_t rows, int32_t cols,
                                                      int32_t nnz, int32_t* flags) {
  __shared__ int32_t s_ptr[kBlock + 1];
  __shared__ int32_t s_bad;
  const int32_t r0 = blockIdx.x * kBlock;
  const int nr = min(kBlock, rows - r0);
  const int tid = threadIdx.x;
  if (tid == 0) s_bad = 0;
  for (int t = tid; t <= nr; t += kBlock) s_ptr[t] = ptr[r0 + t];
  __syncthreads();
  if (kFull && tid < nr) {
    const int32_t b = s_ptr[tid], e = s_ptr[tid + 1];
    if ((r0 + tid == 0 && b != 0) || (r0 + tid == rows - 1 && e != nnz) || b > e || b < 0 || e > nnz) s_bad = 1;
  }
  __syncthreads();
  if (s_bad != 0) {
    if (tid == 0) __hip_atomic_store(flags, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  bool range = false, equal = false, desc = false;
  for (int32_t p = s_ptr[0] + tid; p < s_ptr[nr]; p += kBlock) {
    const int row = find_row(s_ptr, nr, p);
    const int32_t j = idx[p];
    if (kFull) range = range || j < 0 || j >= cols;
    if (p > s_ptr[row]) {
      const int32_t q = idx[p - 1];
      equal = equal || j == q;
      desc = desc || j < q;
    }
  }
  if (kFull) {
    if (range) wave_set_flag(flags + 1);
    if (equal) wave_set_flag(flags + 2);
    if (desc) wave_set_flag(flags + 3);
  } else if (desc) {
    wave_set_flag(flags);
  }
}

// Input check of a CSR matrix, one row per thread (for matrices whose rows
// arrive unsorted: a column can repeat at a distance): flags[0] row pointers
// that are not monotone or do not start at 0 / end at nnz, flags[1] a column
// index outside [0, cols), flags[2] a column repeated within a row, flags[3]
// a row whose columns are not ascending.  Sorted rows are checked for repeats
// by neighbours; an unsorted row compares every pair of its entries.
__global__ void k_csr_check(const int32_t* ptr, const int32_t* idx, int32_t rows, int32_t cols, int32_t nnz,
                            int32_t* flags) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const int32_t b = ptr[i], e = ptr[i + 1];
  if ((i == 0 && b != 0) || (i == rows - 1 && e != nnz) || b > e || b < 0 || e > nnz) {
    wave_set_flag(flags);
    return;
  }
  bool range = false, repeat = false, unsorted = false;
  int32_t prev = -1;
  for (int32_t p = b; p < e; ++p) {
    const int32_t j = idx[p];
    range = range || j < 0 || j >= cols;
    if (p > b) {
      repeat = repeat || j == prev;
      unsorted = unsorted || j < prev;
    }
    prev = j;
  }
  if (unsorted && !repeat) {
    for (int32_t p = b + 1; p < e && !repeat; ++p) {
      const int32_t j = idx[p];
      for (int32_t q = b; q < p; ++q)
        if (idx[q] == j) {
          repeat = true;
          break;
        }
    }
  }
  if (range) wave_set_flag(flags + 1);
  if (repeat) wave_set_flag(flags + 2);
  if (unsorted) wave_set_flag(flags + 3);
}

// Stable insertion sort of one row per thread (used only for inputs whose
// rows arrive unsorted).
__global__ void k_sort_rows(const int32_t* ptr, int32_t* idx, double* val, int32_t rows) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const int32_t b = ptr[i], e = ptr[i + 1];
  for (int32_t p = b + 1; p < e; ++p) {
    const int32_t key = idx[p];
    const double v = val[p];
    int32_t q = p - 1;
    while (q >= b && idx[q] > key) {
      idx[q + 1] = idx[q];
      val[q + 1] = val[q];
      --q;
    }
    idx[q + 1] = key;
    val[q + 1] = v;
  }
}

// Sorts up to kSortSmallMax (key, value) pairs in one workgroup: a bitonic
// network in LDS over (key, input position), so the order equals that of the
// stable radix sort; the values follow through the positions.
constexpr int kSortSmallMax = 4096;
constexpr int kSortSmallBlock = 1024;

template <class K>
__global__ __launch_bounds__(kSortSmallBlock) void k_sort_small(K* __restrict__ keys, int32_t* __restrict__ vals,
                                                               int32_t n) {
  __shared__ K s_k[kSortSmallMax];
  __shared__ int32_t s_v[kSortSmallMax];
  __shared__ int16_t s_p[kSortSmallMax];
  static_assert(kSortSmallMax <= 32768, "positions are stored in 16 bits");
  const int tid = threadIdx.x;
  int len = 1;
  while (len < n) len <<= 1;
  for (int i = tid; i < len; i += kSortSmallBlock) {
    s_k[i] = i < n ? keys[i] : static_cast<K>(~static_cast<K>(0));
    s_p[i] = static_cast<int16_t>(i);
    if (i < n) s_v[i] = vals[i];
  }
  __syncthreads();
  for (int size = 2; size <= len; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = tid; t < (len >> 1); t += kSortSmallBlock) {
        const int i = 2 * t - (t & (stride - 1));  // lower index of the pair
        const int j = i + stride;
        const bool up = (i & size) == 0;
        const K ki = s_k[i], kj = s_k[j];
        const int pi = s_p[i], pj = s_p[j];
        const bool gt = ki > kj || (ki == kj && pi > pj);
        if (gt == up) {
          s_k[i] = kj;
          s_k[j] = ki;
          s_p[i] = static_cast<int16_t>(pj);
          s_p[j] = static_cast<int16_t>(pi);
        }
      }
      // A stage with stride <= 64 only touches its wave's own 128-element
      // segments (pair t of thread t lies in [128 (t / 64), +128)), so between
      // two such stages the wave's LDS order suffices; any wider stage, before
      // or after, meets the other waves at the workgroup barrier.
      const int next = stride > 1 ? stride >> 1 : (size < len ? size : kSortSmallBlock);
      if (stride > 64 || next > 64) {
        __syncthreads();
      } else {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  for (int i = tid; i < n; i += kSortSmallBlock) {
    keys[i] = s_k[i];
    vals[i] = s_v[s_p[i]];
  }
}

template <class In>
hipError_t scan_impl(const In* in, int32_t* out, int64_t n, int32_t* scratch, hipStream_t s) {
  if (n <= 0) {
    return hipMemsetAsync(out, 0, sizeof(int32_t), s);
  }
  if (n <= kSmallScanMax) {
    hipLaunchKernelGGL((k_scan_small<In, 0>), dim3(1), dim3(kSmallBlock), 0, s, in, out, n, nullptr);
    return hipGetLastError();
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_tiles<In>, dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, s, in, n, scratch);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kBlock), 0, s, scratch, tiles, out + n);
  hipLaunchKernelGGL(k_scan_down<In>, dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, s, in, out, n,
                     static_cast<const int32_t*>(scratch));
  return hipGetLastError();
}

template <class K>
hipError_t radix_impl(K* keys, int32_t* vals, K* keys_alt, int32_t* vals_alt, int32_t n, int32_t key_bits,
                      int32_t* scratch, hipStream_t s) {
  if (n <= 1) return hipSuccess;
  if (n <= kSortSmallMax) {
    hipLaunchKernelGGL(k_sort_small<K>, dim3(1), dim3(kSortSmallBlock), 0, s, keys, vals, n);
    return hipGetLastError();
  }
  const int32_t tiles = (n + kRsTile - 1) / kRsTile;
  int32_t* hist = scratch;
  int32_t* scan_scratch = scratch + static_cast<int64_t>(256) * tiles + 1;
  const int passes = key_bits <= 0 ? 1 : (key_bits + 7) / 8;
  const bool from_hist = tiles <= kRsHistTilesMax;
  K* ka = keys;
  int32_t* va = vals;
  K* kb = keys_alt;
  int32_t* vb = vals_alt;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = pass * 8;
    hipLaunchKernelGGL(k_radix_hist<K>, dim3(tiles), dim3(kBlock), 0, s, ka, n, shift, tiles, hist);
    if (from_hist) {
      hipLaunchKernelGGL((k_radix_scatter<K, true>), dim3(tiles), dim3(kBlock), 0, s, ka, va, kb, vb, n, shift,
                         tiles, static_cast<const int32_t*>(hist));
    } else {
      hipError_t e = scan_impl<int32_t>(hist, hist, static_cast<int64_t>(256) * tiles, scan_scratch, s);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((k_radix_scatter<K, false>), dim3(tiles), dim3(kBlock), 0, s, ka, va, kb, vb, n, shift,
                         tiles, static_cast<const int32_t*>(hist));
    }
    K* tk = ka; ka = kb; kb = tk;
    int32_t* tv = va; va = vb; vb = tv;
  }
  if (ka != keys) {
    hipError_t e = hipMemcpyAsync(keys, ka, sizeof(K) * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(vals, va, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

inline unsigned grid_fill(int64_t n) {
  const unsigned g = grid_for(n);
  return g > 65535u ? 65535u : g;
}

}  // namespace

extern "C" {

hipError_t gpk_fill_u8(uint8_t* dst, uint8_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<uint8_t>, dim3(grid_fill(n)), dim3(kBlock), 0, s, dst, v, n);
  return hipGetLastError();
}
hipError_t gpk_fill_i32(int32_t* dst, int32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<int32_t>, dim3(grid_fill(n)), dim3(kBlock), 0, s, dst, v, n);
  return hipGetLastError();
}
hipError_t gpk_fill_f64(double* dst, double v, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_fill(n)), dim3(kBlock), 0, s, dst, v, n);
  return hipGetLastError();
}
hipError_t gpk_iota_i32(int32_t* dst, int32_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, n);
  return hipGetLastError();
}

int64_t gpk_scan_scratch_ints(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

hipError_t gpk_scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* scratch, hipStream_t s) {
  return scan_impl<int32_t>(in, out, n, scratch, s);
}

hipError_t gpk_scan_mask_u8(const uint8_t* mask, int32_t* out, int64_t n, int32_t* scratch, hipStream_t s) {
  return scan_impl<uint8_t>(mask, out, n, scratch, s);
}

int64_t gpk_lookback_tiles(int64_t n) { return n > 0 ? (n + kScanTile - 1) / kScanTile : 1; }

hipError_t gpk_scan_lb_i32(const int32_t* in, int32_t* out, int64_t n, gpk_lookback lb, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n <= kSmallScanMax) {
    hipLaunchKernelGGL((k_scan_small<int32_t, 0>), dim3(1), dim3(kSmallBlock), 0, s, in, out, n, nullptr);
    return hipGetLastError();
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL((k_scan_lb<int32_t, 0>), dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, s, in, out, n,
                     nullptr, lb);
  return hipGetLastError();
}

hipError_t gpk_scan_lb_mask(const uint8_t* mask, int32_t* out, int64_t n, gpk_lookback lb, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n <= kSmallScanMax) {
    hipLaunchKernelGGL((k_scan_small<uint8_t, 0>), dim3(1), dim3(kSmallBlock), 0, s, mask, out, n, nullptr);
    return hipGetLastError();
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL((k_scan_lb<uint8_t, 0>), dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, s, mask, out, n,
                     nullptr, lb);
  return hipGetLastError();
}

int64_t gpk_apply_maps_tiles(int32_t m, int32_t n) {
  const int64_t rt = m > 0 ? (m + kScanTile - 1) / kScanTile : 0;
  const int64_t ct = n > 0 ? (n + kScanTile - 1) / kScanTile : 0;
  return rt + 2 * ct;
}

hipError_t gpk_apply_maps(gpk_apply_maps_args args, gpk_lookback lb, hipStream_t s) {
  const int64_t rt = args.m > 0 ? (args.m + kScanTile - 1) / kScanTile : 0;
  const int64_t ct = args.n > 0 ? (args.n + kScanTile - 1) / kScanTile : 0;
  const int64_t tiles = rt + 2 * ct;
  if (tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply_maps, dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, s, args,
                     static_cast<int32_t>(rt), static_cast<int32_t>(ct), lb);
  return hipGetLastError();
}

hipError_t gpk_mask_to_map(const uint8_t* mask, int32_t* map, int64_t n, int32_t* count, gpk_lookback lb,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n <= kSmallScanMax) {
    hipLaunchKernelGGL((k_scan_small<uint8_t, 1>), dim3(1), dim3(kSmallBlock), 0, s, mask, map, n, count);
    return hipGetLastError();
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL((k_scan_lb<uint8_t, 1>), dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, s, mask, map, n,
                     count, lb);
  return hipGetLastError();
}



hipError_t gpk_build_rank(const int32_t* map, int32_t n, gpk_rank_word* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_build_rank, dim3(grid_for(n)), dim3(kBlock), 0, s, map, n, out);
  return hipGetLastError();
}

int32_t gpk_compact_tiles(int32_t m_old) { return m_old > 0 ? static_cast<int32_t>(grid_for(m_old, kCompRows)) : 0; }

hipError_t gpk_compact_map(const int32_t* ptr, const int32_t* idx, int32_t m_old, const int32_t* row_map,
                           const gpk_rank_word* col_rank, int32_t* jn, int32_t* out_ptr, int32_t* tile_cnt,
                           hipStream_t s) {
  return gpk_compact_map_total(ptr, idx, m_old, row_map, col_rank, jn, out_ptr, tile_cnt, nullptr, s);
}

hipError_t gpk_compact_map_total(const int32_t* ptr, const int32_t* idx, int32_t m_old, const int32_t* row_map,
                                 const gpk_rank_word* col_rank, int32_t* jn, int32_t* out_ptr, int32_t* tile_cnt,
                                 int32_t* kept_total, hipStream_t s) {
  if (m_old <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_compact_map, dim3(grid_for(m_old, kCompRows)), dim3(kBlock), 0, s, ptr, idx, m_old, row_map,
                     col_rank, jn, out_ptr, tile_cnt, kept_total);
  return hipGetLastError();
}

hipError_t gpk_compact_write(const int32_t* ptr, const double* val, int32_t m_old, const int32_t* row_map,
                             const int32_t* jn, const int32_t* tile_off, int32_t* out_ptr, int32_t m_new,
                             int32_t* out_idx, double* out_val, hipStream_t s) {
  if (m_old <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_compact_write<false>, dim3(grid_for(m_old, kCompRows)), dim3(kBlock), 0, s, ptr, val, m_old,
                     row_map, jn, tile_off, out_ptr, m_new, out_idx, out_val);
  return hipGetLastError();
}

hipError_t gpk_compact_map_pair(gpk_compact_side a, gpk_compact_side t, int32_t* tile_cnt, int32_t* kept_total,
                                hipStream_t s) {
  if (a.m_old < 0 || t.m_old < 0) return hipErrorInvalidValue;
  const int32_t ta = a.m_old > 0 ? static_cast<int32_t>(grid_for(a.m_old, kCompRows)) : 0;
  const int32_t tt = t.m_old > 0 ? static_cast<int32_t>(grid_for(t.m_old, kCompRows)) : 0;
  if (ta + tt == 0) return hipSuccess;
  hipLaunchKernelGGL(k_compact_map_pair, dim3(ta + tt), dim3(kBlock), 0, s, a, t, ta, tile_cnt, kept_total);
  return hipGetLastError();
}

hipError_t gpk_compact_write_pair(gpk_compact_side a, gpk_compact_side t, const int32_t* tile_off, int32_t from_counts,
                                  hipStream_t s) {
  if (a.m_old < 0 || t.m_old < 0) return hipErrorInvalidValue;
  const int32_t ta = a.m_old > 0 ? static_cast<int32_t>(grid_for(a.m_old, kCompRows)) : 0;
  const int32_t tt = t.m_old > 0 ? static_cast<int32_t>(grid_for(t.m_old, kCompRows)) : 0;
  if (ta + tt == 0) return hipSuccess;
  if (from_counts && (ta > kBlock || tt > kBlock)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_compact_write_pair, dim3(ta + tt), dim3(kBlock), 0, s, a, t, ta, tt, tile_off, from_counts);
  return hipGetLastError();
}

hipError_t gpk_compact_write_counts(const int32_t* ptr, const double* val, int32_t m_old, const int32_t* row_map,
                                    const int32_t* jn, const int32_t* tile_cnt, int32_t* out_ptr, int32_t m_new,
                                    int32_t* out_idx, double* out_val, hipStream_t s) {
  if (m_old <= 0) return hipSuccess;
  const unsigned tiles = grid_for(m_old, kCompRows);
  if (tiles > static_cast<unsigned>(kBlock)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_compact_write<true>, dim3(tiles), dim3(kBlock), 0, s, ptr, val, m_old, row_map, jn, tile_cnt,
                     out_ptr, m_new, out_idx, out_val);
  return hipGetLastError();
}

hipError_t gpk_old_to_new(const uint8_t* mask, const int32_t* prefix, int32_t* o2n, int32_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_old_to_new, dim3(grid_for(n)), dim3(kBlock), 0, s, mask, prefix, o2n, n);
  return hipGetLastError();
}

hipError_t gpk_pack_lu(const double* l, const double* u, int32_t n, double* lu, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_lu, dim3(grid_for(n)), dim3(kBlock), 0, s, l, u, n, reinterpret_cast<double2*>(lu));
  return hipGetLastError();
}

hipError_t gpk_row_activity_lu(const int32_t* row_ptr, const int32_t* col_idx, const double* val, int32_t m,
                               const double* lu, double ztol, double* min_fin, double* max_fin, int32_t* neg_inf,
                               int32_t* pos_inf, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_row_activity_lu, dim3(grid_for(m, kActRows)), dim3(kBlock), 0, s, row_ptr, col_idx,
                     val, m, reinterpret_cast<const double2*>(lu), ztol, min_fin, max_fin, neg_inf, pos_inf);
  return hipGetLastError();
}

hipError_t gpk_col_max_abs(const int32_t* at_ptr, const double* at_val, int32_t n, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_col_max_abs, dim3(grid_for(n)), dim3(kBlock), 0, s, at_ptr, at_val, n, out);
  return hipGetLastError();
}

hipError_t gpk_csr_stats(const int32_t* ptr, const int32_t* idx, const double* val, int32_t rows, int32_t* nnz_out,
                         uint8_t* empty_mask, uint8_t* singleton_mask, int32_t* singleton_idx,
                         double* singleton_val, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_stats, dim3(grid_for(rows)), dim3(kBlock), 0, s, ptr, idx, val, rows, nnz_out,
                     empty_mask, singleton_mask, singleton_idx, singleton_val);
  return hipGetLastError();
}

hipError_t gpk_mail(const uint32_t* src, int32_t words, uint32_t* mail, uint32_t seq, hipStream_t s) {
  if (words < 0 || words > 28 || (words > 0 && src == nullptr) || mail == nullptr) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mail, dim3(1), dim3(64), 0, s, src, words, mail, seq);
  return hipGetLastError();
}

hipError_t gpk_copy_many(const gpk_copy_seg* segs, int32_t count, hipStream_t s) {
  if (count < 0 || count > GPK_COPY_MAX) return hipErrorInvalidValue;
  CopyList list{};
  int32_t used = 0;
  int64_t most = 0;
  for (int32_t k = 0; k < count; ++k) {
    if (segs[k].bytes <= 0) continue;
    list.seg[used++] = segs[k];
    most = segs[k].bytes > most ? segs[k].bytes : most;
  }
  if (used == 0) return hipSuccess;
  const int64_t per_block = static_cast<int64_t>(kBlock) * 16;
  const int64_t blocks = (most + per_block - 1) / per_block;
  hipLaunchKernelGGL(k_copy_many, dim3(static_cast<unsigned>(blocks < 1024 ? blocks : 1024), used), dim3(kBlock), 0,
                     s, list);
  return hipGetLastError();
}

hipError_t gpk_apply_gathers(gpk_apply_gathers_args args, hipStream_t s) {
  const int32_t len = args.m > args.n ? args.m : args.n;
  if (len <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply_gathers, dim3(grid_for(len)), dim3(kBlock), 0, s, args);
  return hipGetLastError();
}

hipError_t gpk_gather5(const int32_t* map, int32_t n_old, const double* s0, double* d0, const double* s1, double* d1,
                       const double* s2, double* d2, const double* s3, double* d3, const double* s4, double* d4,
                       hipStream_t s) {
  if (n_old <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather5, dim3(grid_for(n_old)), dim3(kBlock), 0, s, map, n_old, s0, d0, s1, d1, s2, d2, s3,
                     d3, s4, d4);
  return hipGetLastError();
}

hipError_t gpk_gather_i32(const int32_t* map, int32_t n_old, const int32_t* src, int32_t* dst, hipStream_t s) {
  if (n_old <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_i32, dim3(grid_for(n_old)), dim3(kBlock), 0, s, map, n_old, src, dst);
  return hipGetLastError();
}

int64_t gpk_radix_scratch_ints(int64_t n) {
  const int64_t tiles = (n + kRsTile - 1) / kRsTile;
  return 256 * tiles + 1 + gpk_scan_scratch_ints(256 * tiles) + 1;
}

hipError_t gpk_radix_sort_u64_i32(uint64_t* keys, int32_t* vals, uint64_t* keys_alt, int32_t* vals_alt, int32_t n,
                                  int32_t key_bits, int32_t* scratch, hipStream_t s) {
  return radix_impl<uint64_t>(keys, vals, keys_alt, vals_alt, n, key_bits, scratch, s);
}

hipError_t gpk_radix_sort_u32_i32(uint32_t* keys, int32_t* vals, uint32_t* keys_alt, int32_t* vals_alt, int32_t n,
                                  int32_t key_bits, int32_t* scratch, hipStream_t s) {
  return radix_impl<uint32_t>(keys, vals, keys_alt, vals_alt, n, key_bits, scratch, s);
}

hipError_t gpk_transpose_sort(const int32_t* col, int32_t nnz, int32_t key_bits, uint32_t* ka, int32_t* va,
                              uint32_t* kb, int32_t* vb, int32_t* scratch, int32_t* result_in_b, hipStream_t s) {
  *result_in_b = 0;
  if (nnz <= 0) return hipSuccess;
  if (nnz <= kSortSmallMax) {
    hipError_t e = hipMemcpyAsync(ka, col, sizeof(uint32_t) * nnz, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_iota, dim3(grid_for(nnz)), dim3(kBlock), 0, s, va, nnz);
    hipLaunchKernelGGL(k_sort_small<uint32_t>, dim3(1), dim3(kSortSmallBlock), 0, s, ka, va, nnz);
    return hipGetLastError();
  }
  // Stable LSD passes; the first reads the column indices in place with the
  // entry positions as values.
  const int32_t tiles = (nnz + kRsTile - 1) / kRsTile;
  int32_t* hist = scratch;
  int32_t* scan_scratch = scratch + static_cast<int64_t>(256) * tiles + 1;
  const int passes = key_bits <= 0 ? 1 : (key_bits + 7) / 8;
  const bool from_hist = tiles <= kRsHistTilesMax;
  const uint32_t* kin = reinterpret_cast<const uint32_t*>(col);
  const int32_t* vin = nullptr;
  uint32_t* kout = ka;
  int32_t* vout = va;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = pass * 8;
    hipLaunchKernelGGL(k_radix_hist<uint32_t>, dim3(tiles), dim3(kBlock), 0, s, kin, nnz, shift, tiles, hist);
    if (from_hist) {
      hipLaunchKernelGGL((k_radix_scatter<uint32_t, true>), dim3(tiles), dim3(kBlock), 0, s, kin, vin, kout, vout, nnz,
                         shift, tiles, static_cast<const int32_t*>(hist));
    } else {
      hipError_t e = scan_impl<int32_t>(hist, hist, static_cast<int64_t>(256) * tiles, scan_scratch, s);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((k_radix_scatter<uint32_t, false>), dim3(tiles), dim3(kBlock), 0, s, kin, vin, kout, vout,
                         nnz, shift, tiles, static_cast<const int32_t*>(hist));
    }
    kin = kout;
    vin = vout;
    const bool to_b = kout == ka;
    kout = to_b ? kb : ka;
    vout = to_b ? vb : va;
  }
  *result_in_b = kin == kb ? 1 : 0;
  return hipGetLastError();
}

hipError_t gpk_col_bounds(const uint32_t* sorted_keys, int32_t nnz, int32_t n, int32_t* ptr, hipStream_t s) {
  hipLaunchKernelGGL(k_col_bounds, dim3(grid_for(static_cast<int64_t>(nnz) + 1)), dim3(kBlock), 0, s, sorted_keys,
                     nnz, n, ptr);
  return hipGetLastError();
}

hipError_t gpk_expand_rows(const int32_t* ptr, const double* val, int32_t m, gpk_row_val* rv, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_expand_rows, dim3(grid_for(m)), dim3(kBlock), 0, s, ptr, val, m, rv);
  return hipGetLastError();
}

hipError_t gpk_transpose_gather(const int32_t* perm, const gpk_row_val* rv, int32_t nnz, int32_t* at_idx,
                                double* at_val, hipStream_t s) {
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_transpose_gather, dim3(grid_for(nnz)), dim3(kBlock), 0, s, perm, rv, nnz, at_idx, at_val);
  return hipGetLastError();
}

hipError_t gpk_rows_unsorted(const int32_t* ptr, const int32_t* idx, int32_t rows, int32_t* flag, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_tiles<false>, dim3(grid_for(rows)), dim3(kBlock), 0, s, ptr, idx, rows, 0, 0, flag);
  return hipGetLastError();
}

hipError_t gpk_csr_check(const int32_t* ptr, const int32_t* idx, int32_t rows, int32_t cols, int32_t nnz,
                         int32_t* flags, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_tiles<true>, dim3(grid_for(rows)), dim3(kBlock), 0, s, ptr, idx, rows, cols, nnz, flags);
  return hipGetLastError();
}

hipError_t gpk_csr_check_rows(const int32_t* ptr, const int32_t* idx, int32_t rows, int32_t cols, int32_t nnz,
                              int32_t* flags, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_check, dim3(grid_for(rows)), dim3(kBlock), 0, s, ptr, idx, rows, cols, nnz, flags);
  return hipGetLastError();
}

hipError_t gpk_sort_rows_in_place(const int32_t* ptr, int32_t* idx, double* val, int32_t rows, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_rows, dim3(grid_for(rows)), dim3(kBlock), 0, s, ptr, idx, val, rows);
  return hipGetLastError();
}

}  // extern "C"
