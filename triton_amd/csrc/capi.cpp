// capi.cpp -- C ABI for loaded cases, writers and error reporting.
#include <algorithm>
#include <cstring>
#include <exception>
#include <string>

#include "../../include/triton_amd.h"
#include "host/case.hpp"

struct tr_case {
  tr::Case c;
};

namespace tr {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
}  // namespace tr

extern "C" {

const char* tr_last_error(void) { return tr::g_last_error.c_str(); }

const char* tr_version(void) { return "triton_amd 0.1.0 (gfx950)"; }

int tr_case_load(const char* cfg_path, const char* base_dir, tr_case** out) {
  if (!cfg_path || !out) {
    tr::set_error("tr_case_load: NULL argument");
    return TR_ERR_INVALID;
  }
  *out = nullptr;
  try {
    auto* c = new tr_case();
    c->c = tr::load_case(cfg_path, base_dir ? base_dir : "");
    *out = c;
    return TR_OK;
  } catch (const tr::InputError& e) {
    tr::set_error(e.what());
    return TR_ERR_INPUT;
  } catch (const std::bad_alloc&) {
    tr::set_error("tr_case_load: out of host memory");
    return TR_ERR_NOMEM;
  } catch (const std::exception& e) {
    tr::set_error(std::string("tr_case_load: ") + e.what());
    return TR_ERR_INPUT;
  }
}

int tr_case_view_get(tr_case* cp, tr_case_view* v) {
  if (!cp || !v) {
    tr::set_error("tr_case_view_get: NULL argument");
    return TR_ERR_INVALID;
  }
  tr::Case& c = cp->c;
  std::memset(v, 0, sizeof(*v));
  v->rows = c.rows;
  v->cols = c.cols;
  v->dx = c.dx;
  v->xll = c.dem_hdr.xll;
  v->yll = c.dem_hdr.yll;
  v->courant = c.cfg.courant;
  v->hextra = c.cfg.hextra;
  v->h = c.h.data();
  v->qx = c.qx.data();
  v->qy = c.qy.data();
  v->dem = c.dem.data();
  v->nman2 = c.nman2.data();
  v->maxh = c.maxh.data();
  v->runoff_id = c.num_runoffs > 0 ? c.runoff_id.data() : nullptr;
  v->num_runoffs = c.num_runoffs;
  v->roff_nrows = c.roff_nrows;
  v->roff_row0 = c.roff_row0;
  v->roff_time = c.roff_time.empty() ? nullptr : c.roff_time.data();
  v->roff_rate = c.roff_rate.empty() ? nullptr : c.roff_rate.data();
  v->num_src = int32_t(c.src_pos.size());
  v->hyg_nrows = c.hyg_nrows;
  v->hyg_time = c.hyg_time.empty() ? nullptr : c.hyg_time.data();
  v->hyg_val = c.hyg_val.empty() ? nullptr : c.hyg_val.data();
  v->src_pos = c.src_pos.empty() ? nullptr : c.src_pos.data();
  v->num_extbc_cells = int32_t(c.bc_index.size());
  v->bc_index = c.bc_index.empty() ? nullptr : c.bc_index.data();
  v->bc_type = c.bc_type.empty() ? nullptr : c.bc_type.data();
  v->bc_start = c.bc_start.empty() ? nullptr : c.bc_start.data();
  v->bc_nrows = c.bc_nrows.empty() ? nullptr : c.bc_nrows.data();
  v->bc_nvars = int32_t(c.bc_var1.size());
  v->bc_var1 = c.bc_var1.empty() ? nullptr : c.bc_var1.data();
  v->bc_var2 = c.bc_var2.empty() ? nullptr : c.bc_var2.data();
  v->open_boundaries = c.cfg.open_boundaries;
  v->sim_start = c.cfg.sim_start_time;
  v->sim_duration = c.cfg.sim_duration;
  v->print_interval = c.cfg.print_interval;
  v->print_observation = c.cfg.print_observation;
  v->init_dt = c.init_dt;
  v->time_series_flag = c.cfg.time_series_flag;
  v->checkpoint_id = c.cfg.checkpoint_id;
  v->it_print = c.cfg.it_print;
  v->it_count = c.cfg.it_count;
  v->num_obs = c.cfg.time_series_flag ? int32_t(c.obs_pos.size()) : 0;
  v->obs_pos = v->num_obs > 0 ? c.obs_pos.data() : nullptr;
  return TR_OK;
}

int32_t tr_case_config_value(const tr_case* c, const char* key, char* buf, int32_t buflen) {
  if (!c || !key || buflen < 0 || (buflen > 0 && !buf)) {
    tr::set_error("tr_case_config_value: bad arguments");
    return -2;
  }
  auto it = c->c.cfg.raw.find(key);
  if (it == c->c.cfg.raw.end()) return -1;
  const std::string& v = it->second;
  if (buflen > 0) {
    const size_t n = std::min(v.size(), size_t(buflen - 1));
    std::memcpy(buf, v.data(), n);
    buf[n] = '\0';
  }
  return int32_t(v.size());
}

void tr_case_free(tr_case* c) { delete c; }

int tr_write_output_ascii(const char* path, int32_t rows, int32_t cols, const double* a) {
  if (!path || !a || rows < 3 || cols < 3) {
    tr::set_error("tr_write_output_ascii: bad arguments");
    return TR_ERR_INVALID;
  }
  try {
    tr::write_output_ascii(path, rows, cols, a);
  } catch (const std::exception& e) {
    tr::set_error(e.what());
    return TR_ERR_INPUT;
  }
  return TR_OK;
}

int tr_write_output_binary(const char* path, int32_t rows, int32_t cols, const double* a) {
  if (!path || !a || rows < 3 || cols < 3) {
    tr::set_error("tr_write_output_binary: bad arguments");
    return TR_ERR_INVALID;
  }
  try {
    tr::write_output_binary(path, rows, cols, a);
  } catch (const std::exception& e) {
    tr::set_error(e.what());
    return TR_ERR_INPUT;
  }
  return TR_OK;
}

}  // extern "C"
