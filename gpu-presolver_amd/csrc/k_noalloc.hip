// Allocation-free row activity: the interleaved (l, u) pairs that
// gp_row_activity stages in a pool block live here in a fixed file-scope
// device array instead, for callers that cannot allocate (code under HIP
// graph capture, serving loops that forbid hipMalloc).  The kernels are the
// ones gp_row_activity runs (k_pack_lu, k_row_activity_lu in
// k_primitives.hip); only the staging buffer differs.
//
// Each GPU holds its own copy of the array (32 MiB) until the process ends.
// Host code never passes the array's name to a kernel or a copy: the name is
// a host-side stand-in, not a device address.  The device address is looked
// up once per device, with that device current, by hipGetSymbolAddress and
// kept in g_scratch[device].  Every call on one GPU shares that GPU's array
// and overwrites it, so calls that overlap on one GPU (two streams, two
// threads) race; calls ordered on one stream do not.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "gp_kernels.h"
#include "gp_presolve_capi.h"

namespace {

constexpr int64_t kNoAllocMaxCols = int64_t{1} << 21;  // 2,097,152 columns
static_assert(kNoAllocMaxCols == GP_ROW_ACTIVITY_NOALLOC_MAX_COLS, "public limit and array size agree");
constexpr int kMaxDevices = 64;

// (l_j, u_j) at [2j], [2j + 1].
__device__ __align__(256) double scratch[2 * kNoAllocMaxCols];

std::once_flag g_once[kMaxDevices];
void* g_scratch[kMaxDevices];
hipError_t g_lookup[kMaxDevices];

// Device address of the calling device's copy of `scratch` (looked up on the
// first call on that device, with it current).
hipError_t scratch_ptr(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::call_once(g_once[dev], [dev] {
    void* p = nullptr;
    g_lookup[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch[dev] = g_lookup[dev] == hipSuccess ? p : nullptr;
  });
  if (g_lookup[dev] != hipSuccess) return g_lookup[dev];
  *out = static_cast<double*>(g_scratch[dev]);
  return hipSuccess;
}

}  // namespace

extern "C" {

hipError_t gpk_row_activity_noalloc(const int32_t* row_ptr, const int32_t* col_idx, const double* val, int32_t m,
                                    int32_t n, const double* l, const double* u, double ztol, double* min_fin,
                                    double* max_fin, int32_t* neg_inf, int32_t* pos_inf, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  if (n < 0 || n > kNoAllocMaxCols) return hipErrorInvalidValue;
  double* lu = nullptr;
  hipError_t e = scratch_ptr(&lu);
  if (e != hipSuccess) return e;
  e = gpk_pack_lu(l, u, n, lu, s);
  if (e != hipSuccess) return e;
  return gpk_row_activity_lu(row_ptr, col_idx, val, m, lu, ztol, min_fin, max_fin, neg_inf, pos_inf, s);
}

}  // extern "C"
