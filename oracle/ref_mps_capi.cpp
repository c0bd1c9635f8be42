// Test infrastructure only: a C entry point over the reference's own MPS
// reader (cpp/mps_reader, plain C++17), compiled from its sources where they
// lie by `make ref` into oracle/_ref/ (never shipped, never copied).  The
// tests read the same text with both readers and compare every array bit
// for bit.  The reference reader's result is exposed in the layout of
// gp_mps_view (include/gp_presolve_capi.h).
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "mps_reader/mpsreader.hpp"

namespace {

struct RefModel {
  mps_reader::MpsModel m;
  std::vector<int> csr_ptr, csr_idx, csc_ptr, csc_idx;
};

}  // namespace

extern "C" {

struct ref_mps_view {
  const char* name;
  int rows, cols, nnz;
  const int* csr_row_ptr;
  const int* csr_col_idx;
  const double* csr_val;
  const int* csc_col_ptr;
  const int* csc_row_idx;
  const double* csc_val;
  const double* c;
  const double* row_lower;
  const double* row_upper;
  const double* col_lower;
  const double* col_upper;
  double obj_constant;
};

// 0: read, *out set; 1: rejected, the message in err.
int ref_mps_parse(const char* text, size_t length, void** out, char* err, size_t err_len) {
  *out = nullptr;
  try {
    std::istringstream in(std::string(text, length));
    auto* r = new RefModel{mps_reader::read_mps(in), {}, {}, {}, {}};
    r->csr_ptr = r->m.lp.csr().row_ptr();
    r->csr_idx = r->m.lp.csr().col_idx();
    r->csc_ptr = r->m.lp.csc().col_ptr();
    r->csc_idx = r->m.lp.csc().row_idx();
    *out = r;
    return 0;
  } catch (const std::exception& e) {
    if (err && err_len) {
      std::strncpy(err, e.what(), err_len - 1);
      err[err_len - 1] = '\0';
    }
    return 1;
  }
}

void ref_mps_get_view(void* h, ref_mps_view* v) {
  const RefModel* r = static_cast<const RefModel*>(h);
  const auto& lp = r->m.lp;
  v->name = r->m.name.c_str();
  v->rows = lp.num_rows();
  v->cols = lp.num_cols();
  v->nnz = lp.csr().nnz();
  v->csr_row_ptr = r->csr_ptr.data();
  v->csr_col_idx = r->csr_idx.data();
  v->csr_val = lp.csr().values().data();
  v->csc_col_ptr = r->csc_ptr.data();
  v->csc_row_idx = r->csc_idx.data();
  v->csc_val = lp.csc().values().data();
  v->c = lp.objective().data();
  v->row_lower = lp.row_lower().data();
  v->row_upper = lp.row_upper().data();
  v->col_lower = lp.col_lower().data();
  v->col_upper = lp.col_upper().data();
  v->obj_constant = lp.obj_constant();
}

void ref_mps_free(void* h) { delete static_cast<RefModel*>(h); }

}  // extern "C"
