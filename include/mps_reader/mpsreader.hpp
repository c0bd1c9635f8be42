// Continuous-MPS reader (reference interface cpp/mps_reader/include/
// mps_reader/mpsreader.hpp:8-15).  Free format is tried first, then the
// fixed column layout; an input neither accepts throws
// std::invalid_argument with the free-format error.  Minimisation only;
// integer markers and bounds, quadratic sections, non-finite numbers and
// sums that overflow are rejected.
#pragma once

#include <iosfwd>
#include <string>

#include "mps_reader/lp_model.hpp"

namespace mps_reader {

struct MpsModel {
  std::string name;
  LpModel lp;
};

MpsModel read_mps(std::istream& input);

}  // namespace mps_reader
