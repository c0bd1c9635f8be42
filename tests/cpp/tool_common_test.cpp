// The tool layer (include/gpu_presolver/tools.hpp) as a caller sees it:
// binary vector files (round trip, wrong sizes, missing files, failed
// writes) on the host, and with the argument "gpu" the move-only device LP
// owner of an uploaded MPS model.  Cases of the reference's
// cpp/tests/test_tool_common.cpp:11-90.  Prints "ok <case>"; exits 1 at the
// first failure.
#include <unistd.h>

#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "gpu_presolver/tools.hpp"

namespace tools = gpu_presolver::tools;
namespace fs = std::filesystem;

static_assert(!std::is_copy_constructible_v<tools::DeviceLpOwner>);
static_assert(!std::is_copy_assignable_v<tools::DeviceLpOwner>);
static_assert(std::is_nothrow_move_constructible_v<tools::DeviceLpOwner>);
static_assert(std::is_nothrow_move_assignable_v<tools::DeviceLpOwner>);

namespace {

void require(bool ok, const char* what) {
  if (!ok) throw std::runtime_error(what);
}

template <class F>
void require_error(F f, const char* what) {
  bool caught = false;
  try {
    f();
  } catch (const std::exception&) {
    caught = true;
  }
  require(caught, what);
}

void write_text(const fs::path& p, const std::string& text) {
  std::ofstream out(p);
  out << text;
  tools::close_output_file(out, p);
}

fs::path g_dir;

void binary_io() {
  const fs::path bin = g_dir / "vector.bin";
  const std::vector<double> v = {1.0, -2.0};
  tools::write_binary_vector(bin, v);
  require(tools::read_binary_vector<double>(bin, 2) == v, "round trip");
  for (std::size_t count : {0, 1, 3})
    require_error([&] { tools::read_binary_vector<double>(bin, count); }, "wrong count accepted");
  tools::write_binary_vector<double>(bin, {});
  require(tools::read_binary_vector<double>(bin, 0).empty(), "empty round trip");
  write_text(bin, "x");
  require_error([&] { tools::read_binary_vector<double>(bin, 0); }, "odd file accepted (0)");
  require_error([&] { tools::read_binary_vector<double>(bin, 1); }, "odd file accepted (1)");
  require_error([&] { tools::read_binary_vector<double>(g_dir / "missing", 0); }, "missing file accepted");
}

void failed_writes() {
  if (!fs::exists("/dev/full")) return;
  require_error([] { tools::write_binary_vector<double>("/dev/full", {1.0}); }, "full device write accepted");
  require_error([] { write_text("/dev/full", "metadata"); }, "full device text accepted");
}

void env_values() {
  require(tools::parse_bool_env_value("YES", false) && !tools::parse_bool_env_value("off", true), "bool words");
  require(tools::parse_bool_env_value("maybe", true) && !tools::parse_bool_env_value(nullptr, false), "bool fallback");
  require(tools::parse_int64_env_value("42", 0) == 42 && tools::parse_int64_env_value("4x", 7) == 7, "int64");
  require(tools::parse_double_env_value("0.5", 0) == 0.5 && tools::parse_double_env_value("inf", 3) == 3, "double");
}

void owner_moves() {
  write_text(g_dir / "toy.mps",
             "NAME toy\nROWS\n N OBJ\n G DEMAND\nCOLUMNS\n    X OBJ 1 DEMAND 1\n    Y OBJ 2 DEMAND 1\nRHS\n"
             "    RHS1 DEMAND 3\nBOUNDS\n UP BND1 X 2\n UP BND1 Y 4\nENDATA\n");
  const auto model = tools::read_mps_file((g_dir / "toy.mps").string());
  {
    auto owner = tools::upload_lp(model.lp);
    tools::DeviceLpOwner moved(std::move(owner));
    require(owner.lp.c == nullptr && owner.lp.A.rowPtr == nullptr, "move construction kept ownership");
    auto assigned = tools::upload_lp(model.lp);
    assigned = std::move(moved);
    require(moved.lp.c == nullptr && moved.lp.AT.nzVal == nullptr, "move assignment kept ownership");
    require(tools::copy_to_host(assigned.lp.c, 2) == model.lp.objective(), "moved device storage");
  }
  tools::check_hip(hipDeviceSynchronize(), "owner lifetime");
}

}  // namespace

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::string(argv[1]) == "gpu";
  g_dir = fs::temp_directory_path() / ("gp_tool_common_" + std::to_string(::getpid()));
  fs::create_directories(g_dir);
  std::vector<std::pair<const char*, std::function<void()>>> cases = {
      {"binary_io", binary_io}, {"failed_writes", failed_writes}, {"env_values", env_values}};
  if (gpu) cases.push_back({"owner_moves", owner_moves});
  int rc = 0;
  for (const auto& [name, fn] : cases) {
    try {
      fn();
      std::printf("ok %s\n", name);
    } catch (const std::exception& e) {
      std::printf("FAIL %s: %s\n", name, e.what());
      rc = 1;
      break;
    }
  }
  fs::remove_all(g_dir);
  return rc;
}
