// Tool layer between the MPS reader and the presolve API: environment
// overrides of PresolveParams, host <-> device copies, the device LP owner
// and its upload, binary vector files, the reduced-LP export directory and
// the postsolve checker's figures.  Same names and behaviour as the
// reference's tool header (cpp/tools/gpu_presolver_tool_common.hpp:20-310)
// and the helpers of its export and checker tools
// (gpu_presolver_export_reduced.cpp:16-57,
// gpu_presolver_check_postsolve_solution.cpp:20-246); HIP in place of CUDA.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "gpu_presolver/presolve/gpu_presolve.hpp"
#include "gpu_presolver/presolve/presolve_structs.hpp"
#include "mps_reader/mpsreader.hpp"

namespace gpu_presolver::tools {

// "1/true/yes/on" and "0/false/no/off" (any case); anything else, or no
// value, gives the fallback.
bool parse_bool_env_value(const char* value, bool fallback);
std::int64_t parse_int64_env_value(const char* value, std::int64_t fallback);
double parse_double_env_value(const char* value, double fallback);
// GPUPRESOLVER_* variables (GPUPRESOLVER_MAX_ITERS, _SCHEDULER,
// _ENABLE_DOUBLETON_EQUATIONS, ...) applied to params.
void apply_presolve_env_overrides(presolve::PresolveParams& params);

// std::runtime_error("<context>: <HIP error string>") on failure.
void check_hip(hipError_t status, const char* context);

template <class T>
T* copy_to_device(const std::vector<T>& values) {
  if (values.empty()) return nullptr;
  T* device = nullptr;
  check_hip(hipMalloc(reinterpret_cast<void**>(&device), sizeof(T) * values.size()), "hipMalloc");
  const hipError_t st = hipMemcpy(device, values.data(), sizeof(T) * values.size(), hipMemcpyHostToDevice);
  if (st != hipSuccess) {
    (void)hipFree(device);
    check_hip(st, "hipMemcpy H2D");
  }
  return device;
}

template <class T>
std::vector<T> copy_to_host(const T* device, std::size_t size) {
  std::vector<T> values(size);
  if (size > 0) check_hip(hipMemcpy(values.data(), device, sizeof(T) * size, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  return values;
}

// A device LP whose arrays this object hipFree()s; move-only.
struct DeviceLpOwner {
  presolve::LPInfoGpu lp{};
  DeviceLpOwner() = default;
  DeviceLpOwner(const DeviceLpOwner&) = delete;
  DeviceLpOwner& operator=(const DeviceLpOwner&) = delete;
  DeviceLpOwner(DeviceLpOwner&& other) noexcept : lp(std::exchange(other.lp, {})) {}
  DeviceLpOwner& operator=(DeviceLpOwner&& other) noexcept {
    if (this != &other) {
      release();
      lp = std::exchange(other.lp, {});
    }
    return *this;
  }
  ~DeviceLpOwner() { release(); }

 private:
  void release() noexcept;
};

// The reader's LP on the device: A from its CSR, AT from its CSC (rows
// sorted), c, AL, AU, l, u and the objective constant.
DeviceLpOwner upload_lp(const mps_reader::LpModel& model);
// "-" reads standard input.
mps_reader::MpsModel read_mps_file(const std::string& path);

// Flushes and closes; throws when any write failed.
void close_output_file(std::ofstream& out, const std::filesystem::path& path);

template <class T>
std::streamsize binary_vector_bytes(std::size_t count) {
  if (count > static_cast<std::size_t>(std::numeric_limits<std::streamsize>::max()) / sizeof(T))
    throw std::length_error("binary vector is too large");
  return static_cast<std::streamsize>(count * sizeof(T));
}

// Raw little-endian values, no header.
template <class T>
void write_binary_vector(const std::filesystem::path& path, const std::vector<T>& values) {
  const std::streamsize bytes = binary_vector_bytes<T>(values.size());
  std::ofstream out(path, std::ios::binary);
  if (!out.is_open()) throw std::runtime_error("failed to open " + path.string());
  if (!values.empty()) out.write(reinterpret_cast<const char*>(values.data()), bytes);
  close_output_file(out, path);
}

// Exactly `count` values: a file of any other size is an error.
template <class T>
std::vector<T> read_binary_vector(const std::filesystem::path& path, std::size_t count) {
  const std::streamsize bytes = binary_vector_bytes<T>(count);
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in.is_open()) throw std::runtime_error("failed to open " + path.string());
  if (in.tellg() != std::streampos(bytes)) throw std::runtime_error("unexpected binary file size: " + path.string());
  in.seekg(0);
  std::vector<T> values(count);
  if (count > 0) {
    in.read(reinterpret_cast<char*>(values.data()), bytes);
    if (!in) throw std::runtime_error("failed to read " + path.string());
  }
  if (!in || in.peek() != std::char_traits<char>::eof() || in.bad())
    throw std::runtime_error("failed to read complete file " + path.string());
  return values;
}

// Reduced-LP export directory (format gpu_presolver_reduced_lp_v1):
// meta.txt (model, status, dimensions, obj_constant, objective_shift,
// iterations, doubles at max_digits10) and A_rowPtr_i32, A_colVal_i32,
// A_nzVal_f64, c_f64, l_f64, u_f64, AL_f64, AU_f64 (.bin).
void write_lp_export(const std::filesystem::path& out_dir, const std::string& model_name,
                     const presolve::GpuPresolveSummary& summary);

// 64-bit FNV-1a over raw bytes; vectors and strings are prefixed by their
// 64-bit length.
class Fnv1a64 {
 public:
  template <class T>
  void add_scalar(const T& value) {
    static_assert(std::is_trivially_copyable_v<T>, "raw bytes");
    add_bytes(&value, sizeof(value));
  }
  template <class T>
  void add_vector(const std::vector<T>& values) {
    add_scalar(static_cast<std::uint64_t>(values.size()));
    if (!values.empty()) add_bytes(values.data(), values.size() * sizeof(T));
  }
  void add_string(const std::string& value) {
    add_scalar(static_cast<std::uint64_t>(value.size()));
    add_bytes(value.data(), value.size());
  }
  std::uint64_t value() const { return hash_; }

 private:
  void add_bytes(const void* data, std::size_t size);
  std::uint64_t hash_ = 14695981039346656037ull;
};

template <class T>
std::uint64_t raw_vector_hash(const std::vector<T>& values) {
  Fnv1a64 h;
  h.add_vector(values);
  return h.value();
}
// Every field of the record, in declaration order.
std::uint64_t record_hash(const presolve::PresolveRecordGpu& record);

struct PrimalFeasibility {
  double normalized = 0.0;  // max(row_l2, col_l2) / (1 + ||finite row bounds||_2)
  double row_l2 = 0.0;
  double row_max = 0.0;
  double col_l2 = 0.0;
  double col_max = 0.0;
};
// Violations of AL <= A x <= AU and l <= x <= u (hypot accumulation);
// throws on a non-finite x, activity or bound norm.
PrimalFeasibility compute_primal_feasibility(const mps_reader::LpModel& model, const std::vector<double>& x);

}  // namespace gpu_presolver::tools
