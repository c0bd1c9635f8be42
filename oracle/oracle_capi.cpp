// C entry points of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.
// They mirror the gp_* functions of include/gp_presolve_capi.h with host
// pointers instead of device pointers, so a test can run the same case
// through both and compare.
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

#include "oracle_core.hpp"

namespace {

thread_local std::string g_err;

orc::CsrV csr_of(const gp_csr& c) {
  return orc::CsrV{c.rows, c.cols, c.nnz, c.row_ptr, c.col_idx, c.val};
}

orc::LpV lp_of(const gp_lp& lp) {
  orc::LpV v;
  v.A = csr_of(lp.A);
  v.AT = csr_of(lp.AT);
  v.c = lp.c;
  v.AL = lp.AL;
  v.AU = lp.AU;
  v.l = lp.l;
  v.u = lp.u;
  v.obj_constant = lp.obj_constant;
  return v;
}

template <class T>
T* dup_array(const std::vector<T>& v) {
  T* p = static_cast<T*>(std::malloc(sizeof(T) * (v.empty() ? 1 : v.size())));
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

void fill_tape_view(const orc::Tape& t, gp_record_view* out) {
  out->tape_count = static_cast<int32_t>(t.types.size());
  out->tape_types = t.types.data();
  out->tape_index_starts = t.index_starts.data();
  out->tape_value_starts = t.value_starts.data();
  out->tape_dual_modes = t.dual_modes.data();
  out->tape_index_count = static_cast<int32_t>(t.indices.size());
  out->tape_indices = t.indices.data();
  out->tape_value_count = static_cast<int32_t>(t.vals.size());
  out->tape_vals = t.vals.data();
}

}  // namespace

struct orc_tape {
  orc::Tape tape;
};

struct orc_result {
  orc::Result res;
};

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void orc_params_default(gp_params* p) { orc::default_params(p); }

void orc_free(void* p) { std::free(p); }

// gpu_presolve.cu:2151-2158
int32_t orc_has_good_nnz_progress(int32_t before, int32_t after, double ratio) {
  return orc::has_good_nnz_progress(before, after, ratio) ? 1 : 0;
}

void orc_set_batched_closure_min_nnz(int32_t nnz) { orc::set_batched_closure_min_nnz(nnz); }
int32_t orc_batched_closure_min_nnz(void) { return orc::batched_closure_min_nnz(); }
int64_t orc_batched_closure_runs(void) { return orc::batched_closure_runs(); }

// rules/rule_duplicate_rows.cu:78-178 (the hash the rule sorts rows by).
int orc_dup_row_hashes(const gp_csr* A, const uint8_t* keep_row, double ctol, uint64_t* out) {
  orc::duplicate_row_hashes(csr_of(*A), keep_row, ctol, out);
  return 0;
}

int orc_row_activity(const gp_csr* A, const double* l, const double* u, double ztol, double* mn, double* mx,
                     int32_t* ni, int32_t* pi) {
  orc::row_activity(csr_of(*A), l, u, ztol, mn, mx, ni, pi);
  return 0;
}

int orc_col_max_abs(const gp_csr* AT, double* out) {
  orc::col_max_abs(csr_of(*AT), out);
  return 0;
}

// Outputs are malloc'ed; release with orc_free.
int orc_compact_csr(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col, gp_csr* out) {
  orc::Csr c = orc::compact(csr_of(*A), keep_row, keep_col);
  out->rows = c.rows;
  out->cols = c.cols;
  out->nnz = c.nnz();
  out->row_ptr = dup_array(c.ptr);
  out->col_idx = dup_array(c.idx);
  out->val = dup_array(c.val);
  return 0;
}

int orc_transpose_csr(const gp_csr* A, gp_csr* out) {
  orc::Csr c = orc::transpose(csr_of(*A));
  out->rows = c.rows;
  out->cols = c.cols;
  out->nnz = c.nnz();
  out->row_ptr = dup_array(c.ptr);
  out->col_idx = dup_array(c.idx);
  out->val = dup_array(c.val);
  return 0;
}

int orc_rule_apply(int32_t rule, gp_plan_view* pv, const gp_lp* lp, const gp_stats_view* sv,
                   const gp_params* prm, const uint8_t* cand, orc_tape** tape) {
  try {
    orc::Plan p;
    p.keep_row = pv->keep_row;
    p.keep_col = pv->keep_col;
    p.c = pv->c;
    p.l = pv->l;
    p.u = pv->u;
    p.AL = pv->AL;
    p.AU = pv->AU;
    p.obj_delta = pv->obj_constant_delta;
    p.has_change = pv->has_change != 0;
    p.has_row_action = pv->has_row_action != 0;
    p.has_col_action = pv->has_col_action != 0;
    p.has_infeasible = pv->has_infeasible != 0;
    p.has_unbounded = pv->has_unbounded != 0;
    orc::Stats s;
    if (sv) {
      s.row_nnz = sv->row_nnz;
      s.singleton_row_mask = sv->singleton_row_mask;
      s.singleton_row_col = sv->singleton_row_col;
      s.singleton_row_val = sv->singleton_row_val;
      s.col_nnz = sv->col_nnz;
      s.empty_col_mask = sv->empty_col_mask;
      s.column_singleton_mask = sv->column_singleton_mask;
      s.column_singleton_row = sv->column_singleton_row;
      s.column_singleton_val = sv->column_singleton_val;
    }
    const orc::LpV v = lp_of(*lp);
    switch (rule) {
      case GP_RULE_EMPTY_ROWS: orc::rule_empty_rows(p, v, s, *prm); break;
      case GP_RULE_SINGLETON_ROWS: orc::rule_singleton_rows(p, v, s, *prm); break;
      case GP_RULE_INFEASIBLE_REDUNDANT_ROWS: orc::rule_infeasible_redundant_rows(p, v, s, *prm); break;
      case GP_RULE_IMPLIED_VARIABLE_BOUNDS: orc::rule_implied_variable_bounds(p, v, s, *prm); break;
      case GP_RULE_DUPLICATE_ROWS: orc::rule_duplicate_rows(p, v, s, *prm); break;
      case GP_RULE_INFEASIBLE_FIXED_VARIABLES: orc::rule_infeasible_fixed_variables(p, v, *prm); break;
      case GP_RULE_EMPTY_COLS: orc::rule_empty_cols(p, v, s, *prm); break;
      case GP_RULE_COLUMN_SINGLETONS_DUAL_INFER: orc::rule_column_singletons(p, v, s, *prm, false, false); break;
      case GP_RULE_COLUMN_SINGLETONS_EQ: orc::rule_column_singletons(p, v, s, *prm, true, false); break;
      case GP_RULE_DOUBLETON_EQUATIONS: orc::rule_doubleton_equations(p, v, s, *prm); break;
      case GP_RULE_DUAL_FIX: orc::rule_dual_fix(p, v, s, *prm, cand); break;
      case GP_RULE_DUPLICATE_COLUMNS: orc::rule_duplicate_columns(p, v, s, *prm); break;
      case GP_RULE_REDUNDANT_BOUNDS: orc::rule_redundant_bounds(p, v, s, *prm); break;
      default: g_err = "unknown rule"; return 1;
    }
    pv->obj_constant_delta = p.obj_delta;
    pv->has_change = p.has_change;
    pv->has_row_action = p.has_row_action;
    pv->has_col_action = p.has_col_action;
    pv->has_infeasible = p.has_infeasible;
    pv->has_unbounded = p.has_unbounded;
    pv->has_new_A = p.has_new_A;
    if (p.has_new_A) {
      pv->new_A.rows = p.new_A.rows;
      pv->new_A.cols = p.new_A.cols;
      pv->new_A.nnz = p.new_A.nnz();
      pv->new_A.row_ptr = dup_array(p.new_A.ptr);
      pv->new_A.col_idx = dup_array(p.new_A.idx);
      pv->new_A.val = dup_array(p.new_A.val);
    }
    if (tape) {
      if (!*tape) *tape = new orc_tape();
      orc::Tape& t = (*tape)->tape;
      const int32_t bi = static_cast<int32_t>(t.indices.size());
      const int32_t bv = static_cast<int32_t>(t.vals.size());
      t.types.insert(t.types.end(), p.tape.types.begin(), p.tape.types.end());
      t.dual_modes.insert(t.dual_modes.end(), p.tape.dual_modes.begin(), p.tape.dual_modes.end());
      t.indices.insert(t.indices.end(), p.tape.indices.begin(), p.tape.indices.end());
      t.vals.insert(t.vals.end(), p.tape.vals.begin(), p.tape.vals.end());
      for (size_t k = 1; k < p.tape.index_starts.size(); ++k) {
        t.index_starts.push_back(bi + p.tape.index_starts[k]);
        t.value_starts.push_back(bv + p.tape.value_starts[k]);
      }
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 3;
  }
}

int orc_tape_view(const orc_tape* t, gp_record_view* out) {
  std::memset(out, 0, sizeof(*out));
  fill_tape_view(t->tape, out);
  return 0;
}

void orc_tape_free(orc_tape* t) { delete t; }

int orc_presolve(const gp_lp* lp, const gp_params* prm, int32_t keep, orc_result** out) {
  try {
    auto* r = new orc_result();
    r->res = orc::presolve(lp_of(*lp), *prm, keep != 0);
    *out = r;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 3;
  }
}

int orc_summary(const orc_result* r, gp_summary* s) {
  const orc::Result& x = r->res;
  s->original_rows = x.original_rows;
  s->original_cols = x.original_cols;
  s->reduced_rows = x.reduced_rows;
  s->reduced_cols = x.reduced_cols;
  s->reduced_nnz = x.reduced_nnz;
  s->iterations = x.iterations;
  s->elapsed_seconds = x.elapsed_seconds;
  s->obj_constant_delta = x.obj_constant_delta;
  s->has_infeasible = x.has_infeasible;
  s->has_unbounded = x.has_unbounded;
  return 0;
}

int orc_record(const orc_result* r, gp_record_view* v) {
  const orc::Record& R = r->res.record;
  std::memset(v, 0, sizeof(*v));
  v->m0 = R.m0;
  v->n0 = R.n0;
  v->m1 = R.m1;
  v->n1 = R.n1;
  v->row_red2org = R.row_red2org.data();
  v->col_red2org = R.col_red2org.data();
  v->fixed_count = static_cast<int32_t>(R.fixed_idx.size());
  v->fixed_idx = R.fixed_idx.data();
  v->fixed_val = R.fixed_val.data();
  v->obj_constant_old = R.obj_constant_old;
  v->obj_constant_new = R.obj_constant_new;
  fill_tape_view(R.tape, v);
  return 0;
}

int orc_reduced_lp(const orc_result* r, gp_lp* out) {
  if (!r->res.has_reduced) { g_err = "reduced LP not kept"; return 1; }
  const orc::OwnedLp& lp = r->res.reduced;
  auto fill = [](const orc::Csr& c, gp_csr* o) {
    o->rows = c.rows;
    o->cols = c.cols;
    o->nnz = c.nnz();
    o->row_ptr = const_cast<int32_t*>(c.ptr.data());
    o->col_idx = const_cast<int32_t*>(c.idx.data());
    o->val = const_cast<double*>(c.val.data());
  };
  fill(lp.A, &out->A);
  fill(lp.AT, &out->AT);
  out->c = const_cast<double*>(lp.c.data());
  out->AL = const_cast<double*>(lp.AL.data());
  out->AU = const_cast<double*>(lp.AU.data());
  out->l = const_cast<double*>(lp.l.data());
  out->u = const_cast<double*>(lp.u.data());
  out->obj_constant = lp.obj_constant;
  return 0;
}

void orc_result_free(orc_result* r) { delete r; }

int orc_postsolve(const gp_record_view* rv, const double* x_red, const double* y_red, const double* z_red,
                  const gp_lp* original, double* x, double* y, double* z) {
  try {
    orc::RecordV R{rv->m0, rv->n0, rv->m1, rv->n1, rv->row_red2org, rv->col_red2org, rv->fixed_count,
                   rv->fixed_idx, rv->fixed_val, rv->tape_count, rv->tape_types, rv->tape_index_starts,
                   rv->tape_value_starts, rv->tape_indices, rv->tape_vals};
    orc::LpV ov;
    if (original) ov = lp_of(*original);
    orc::postsolve(R, x_red, y_red, z_red, original ? &ov : nullptr, x, y, z);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 3;
  }
}

}  // extern "C"
