// serve.cpp -- continuous-batching request loop over one tr_batch (tr_serve_*).
//
// Requests (cases run to their sim_duration in the FP32 form) wait in arrival
// order; each scheduling round queues chunk_steps batched launches for the
// running requests, admits waiting requests first fit while their device
// bytes fit the HBM budget (their host-side preparation overlapping the
// queued chunk; they step from the next round), waits, then retires the
// finished ones (final status and, optionally, state downloaded; member
// removed, memory freed).  New requests
// thus join while others are mid-run, and a request's result is the result it
// has alone: a member's arithmetic does not depend on the other members of
// its launches.  The reference has no serving layer; it runs ensemble
// members as separate MPI jobs (src/Ensify.h:43-72).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <deque>
#include <filesystem>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "../../../include/triton_amd.h"

namespace tr {
void set_error(const std::string& msg);
int check_member_view(const tr_case_view* v);
}
const char* tr_last_error(void);

namespace {

struct PrintRec {
  int32_t print_id = 0;
  double simtime = 0.0;
  std::vector<double> h, qx, qy, maxh;
};

struct Request {
  int32_t id = 0;
  tr_case_view view{};  // caller-owned arrays, valid until admission
  tr_serve_request info{};
  std::vector<double> h, qx, qy;  // final state (keep_state)
  int64_t rows = 0, cols = 0;
  // print products
  std::string outdir;
  int32_t out_format = 0;
  bool keep_prints = false;
  bool track_maxh = false;
  int32_t last_print = 0;
  std::vector<PrintRec> prints;
  bool wants_prints() const { return keep_prints || !outdir.empty(); }
};

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// numpy's default percentile: linear interpolation between order statistics.
double percentile(std::vector<double> v, double q) {
  if (v.empty()) return std::numeric_limits<double>::quiet_NaN();
  std::sort(v.begin(), v.end());
  const double pos = q / 100.0 * double(v.size() - 1);
  const size_t lo = size_t(std::floor(pos));
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (pos - double(lo)) * (v[hi] - v[lo]);
}

}  // namespace

struct tr_serve {
  tr_serve_options opt{};
  tr_batch* batch = nullptr;
  std::map<int32_t, Request> req;  // every request not yet released
  std::deque<int32_t> waiting;     // arrival order
  std::map<int32_t, int32_t> running;  // batch member -> request id
  int32_t next_id = 0;
  tr_serve_stats stats{};
  std::vector<double> latencies;
  bool rearm = false;  // a member stopped at a print: resume it before the next chunk
};

namespace {

int fail(int code, const std::string& msg) {
  tr::set_error(msg);
  return code;
}

// End a request that will not finish: state TR_REQ_FAILED (code = the TR_*
// error) or TR_REQ_CANCELLED; not counted in the completion statistics.
void end_request(tr_serve* s, Request& r, int32_t state, int code) {
  r.info.state = state;
  r.info.error = code;
  r.info.member = -1;
  r.info.done = now_s();
  r.view = tr_case_view{};
  if (state == TR_REQ_FAILED) s->stats.failed++;
  else s->stats.cancelled++;
}

// The reference's raster names (outfile_pattern "%s/%s/%s_%02d_%02d" + ".out").
std::string raster_path(const std::string& dir, const char* sub, const char* name, int pid) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "/%s/%s_%02d_%02d.out", sub, name, pid, 0);
  return dir + buf;
}

// Capture the prints of running requests whose member stopped at a print
// boundary since the last round: write the rasters and/or keep the state.
int capture_prints(tr_serve* s) {
  for (auto& kv : s->running) {
    Request& r = s->req.at(kv.second);
    if (!r.wants_prints()) continue;
    tr_sim_status st{};
    int rc = tr_batch_poll(s->batch, kv.first, &st);
    if (rc) return rc;
    if (st.print_id <= r.last_print) continue;
    PrintRec p;
    p.print_id = st.print_id;
    p.simtime = st.simtime;
    const size_t n = size_t(r.rows * r.cols);
    p.h.resize(n);
    p.qx.resize(n);
    p.qy.resize(n);
    if (r.track_maxh) p.maxh.resize(n);
    if ((rc = tr_batch_download(s->batch, kv.first, p.h.data(), p.qx.data(), p.qy.data(),
                                r.track_maxh ? p.maxh.data() : nullptr)))
      return rc;
    if (!r.outdir.empty()) {
      const struct {
        const char* name;
        const std::vector<double>* a;
      } fields[] = {{"H", &p.h}, {"QX", &p.qx}, {"QY", &p.qy}, {"MH", &p.maxh}};
      for (const auto& f : fields) {
        if (f.a->empty()) continue;
        for (int asc = 0; asc < (r.out_format == 0 ? 2 : 1); asc++) {
          const char* sub = asc ? "asc" : "bin";
          std::error_code ec;
          std::filesystem::create_directories(r.outdir + "/" + sub, ec);
          const std::string path = raster_path(r.outdir, sub, f.name, p.print_id);
          rc = asc ? tr_write_output_ascii(path.c_str(), int32_t(r.rows), int32_t(r.cols), f.a->data())
                   : tr_write_output_binary(path.c_str(), int32_t(r.rows), int32_t(r.cols), f.a->data());
          if (rc) return rc;
        }
      }
    }
    r.last_print = p.print_id;
    s->rearm = true;
    if (r.keep_prints) r.prints.push_back(std::move(p));
  }
  return TR_OK;
}

// Retire every running request whose member has finished.
int retire(tr_serve* s) {
  for (auto it = s->running.begin(); it != s->running.end();) {
    const int32_t mid = it->first;
    tr_sim_status st{};
    int rc = tr_batch_poll(s->batch, mid, &st);
    if (rc) return rc;
    if (!st.finished) {
      ++it;
      continue;
    }
    Request& r = s->req.at(it->second);
    if ((rc = tr_batch_status(s->batch, mid, &r.info.status))) return rc;
    if (s->opt.keep_state) {
      const size_t n = size_t(r.rows * r.cols);
      r.h.resize(n);
      r.qx.resize(n);
      r.qy.resize(n);
      if ((rc = tr_batch_download(s->batch, mid, r.h.data(), r.qx.data(), r.qy.data(), nullptr))) return rc;
    }
    if ((rc = tr_batch_remove(s->batch, mid))) return rc;
    r.info.state = TR_REQ_FINISHED;
    r.info.member = -1;
    r.info.done = now_s();
    s->stats.completed++;
    s->stats.cell_steps += r.info.cells * r.info.status.steps;
    s->latencies.push_back(r.info.done - r.info.arrival);
    it = s->running.erase(it);
  }
  return TR_OK;
}

// Admit waiting requests by the tr_serve_admit policy.
int admit(tr_serve* s) {
  const int32_t n = int32_t(s->waiting.size());
  if (n == 0) return TR_OK;
  std::vector<int64_t> bytes(static_cast<size_t>(n));
  std::vector<int32_t> skipped(static_cast<size_t>(n));
  std::vector<uint8_t> take(size_t(n), 0);
  for (int32_t k = 0; k < n; k++) {
    const Request& r = s->req.at(s->waiting[size_t(k)]);
    bytes[size_t(k)] = r.info.bytes;
    skipped[size_t(k)] = r.info.skipped;
  }
  int rc = tr_serve_admit(bytes.data(), skipped.data(), n, tr_batch_device_bytes(s->batch),
                          s->opt.hbm_budget_bytes, int32_t(s->running.size()), s->opt.max_members,
                          s->opt.starve_rounds, take.data());
  if (rc) return rc;
  // A request whose add fails does not hold up the others: out of device
  // memory it waits for the next round (and fails when nothing else runs, so
  // nothing will free memory for it); any other error fails it.
  std::deque<int32_t> keep;
  for (int32_t k = 0; k < n; k++) {
    Request& r = s->req.at(s->waiting[size_t(k)]);
    r.info.skipped = skipped[size_t(k)];
    if (!take[size_t(k)]) {
      keep.push_back(r.id);
      continue;
    }
    int32_t mid = -1;
    const int arc = tr_batch_add(s->batch, &r.view, &mid);
    if (arc == TR_ERR_NOMEM && !s->running.empty()) {
      r.info.skipped++;
      keep.push_back(r.id);
      continue;
    }
    if (arc) {
      end_request(s, r, TR_REQ_FAILED, arc);
      continue;
    }
    if (r.wants_prints()) {
      int prc = tr_batch_set_print_halt(s->batch, mid, 1);
      tr_sim_status st0{};
      if (!prc) prc = tr_batch_poll(s->batch, mid, &st0);
      if (prc) return prc;
      r.last_print = st0.print_id;
    }
    r.info.state = TR_REQ_RUNNING;
    r.info.member = mid;
    r.info.admitted = now_s();
    r.view = tr_case_view{};  // the batch holds its own copy now
    s->running[mid] = r.id;
  }
  s->waiting.swap(keep);
  s->stats.max_members = std::max(s->stats.max_members, int32_t(s->running.size()));
  s->stats.peak_bytes = std::max(s->stats.peak_bytes, tr_batch_device_bytes(s->batch));
  return TR_OK;
}

}  // namespace

extern "C" {

double tr_serve_clock(void) { return now_s(); }

int tr_serve_admit(const int64_t* bytes, int32_t* skipped, int32_t n_waiting, int64_t held, int64_t budget,
                   int32_t running, int32_t max_members, int32_t starve_rounds, uint8_t* admit) {
  if (n_waiting < 0 || (n_waiting > 0 && (!bytes || !skipped || !admit)))
    return fail(TR_ERR_INVALID, "tr_serve_admit: bad argument");
  const int32_t cap = max_members > 0 ? max_members : 256;
  bool blocked = false;
  for (int32_t k = 0; k < n_waiting; k++) {
    const bool fits = (budget <= 0 || held + bytes[k] <= budget) && running < cap;
    admit[k] = 0;
    if (blocked || !fits) {
      if (!fits) {
        skipped[k]++;
        blocked = blocked || skipped[k] > starve_rounds;
      }
      continue;
    }
    admit[k] = 1;
    held += bytes[k];
    running++;
  }
  return TR_OK;
}

int tr_serve_create(const tr_serve_options* opt, tr_serve** out) {
  if (!opt || !out) return fail(TR_ERR_INVALID, "tr_serve_create: NULL argument");
  if (opt->chunk_steps < 1 || opt->hbm_budget_bytes < 0 || opt->max_members < 0 || opt->starve_rounds < 0)
    return fail(TR_ERR_INVALID, "tr_serve_create: chunk_steps must be >= 1 and the limits >= 0");
  tr_serve* s = new tr_serve;
  s->opt = *opt;
  if (s->opt.max_members == 0) s->opt.max_members = 256;
  int rc = tr_batch_create(opt->hbm_budget_bytes, opt->waves_per_simd, &s->batch);
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return TR_OK;
}

int tr_serve_submit(tr_serve* s, const tr_case_view* v, double arrival, int32_t* request_id) {
  return tr_serve_submit_ex(s, v, arrival, nullptr, request_id);
}

int tr_serve_submit_ex(tr_serve* s, const tr_case_view* v, double arrival, const tr_serve_request_options* ro,
                       int32_t* request_id) {
  if (!s || !v || !request_id) return fail(TR_ERR_INVALID, "tr_serve_submit: NULL argument");
  if (ro && (ro->output_format < 0 || ro->output_format > 1))
    return fail(TR_ERR_INVALID, "tr_serve_submit: output_format must be 0 (ASC) or 1 (BIN)");
  // a request that could never be admitted is refused here, not at admission
  if (tr::check_member_view(v)) return fail(TR_ERR_INVALID, std::string("tr_serve_submit: ") + tr_last_error());
  const int64_t need = tr_batch_member_bytes(v);
  if (need < 0) return fail(TR_ERR_INVALID, "tr_serve_submit: the case cannot run in the batched FP32 form");
  if (s->opt.hbm_budget_bytes > 0 && need > s->opt.hbm_budget_bytes)
    return fail(TR_ERR_NOMEM, "tr_serve_submit: the request needs " + std::to_string(need) +
                                  " B, more than the whole HBM budget " +
                                  std::to_string(s->opt.hbm_budget_bytes) + " B");
  Request r;
  r.id = s->next_id++;
  r.view = *v;
  r.rows = v->rows;
  r.cols = v->cols;
  r.info.state = TR_REQ_WAITING;
  r.info.member = -1;
  r.info.bytes = need;
  r.info.cells = int64_t(v->rows - 2) * int64_t(v->cols - 2);
  r.info.arrival = arrival >= 0.0 ? arrival : now_s();
  r.info.admitted = -1.0;
  r.info.done = -1.0;
  if (ro) {
    if (ro->output_dir) r.outdir = ro->output_dir;
    r.out_format = ro->output_format;
    r.keep_prints = ro->keep_prints != 0;
  }
  r.track_maxh = v->maxh != nullptr;
  s->waiting.push_back(r.id);
  s->req.emplace(r.id, std::move(r));
  s->stats.submitted++;
  *request_id = s->next_id - 1;
  return TR_OK;
}

int tr_serve_step(tr_serve* s, int32_t* ran) {
  if (!s) return fail(TR_ERR_INVALID, "tr_serve_step: NULL argument");
  if (ran) *ran = 0;
  int rc = TR_OK;
  // With nothing running, admit first so the round has work; otherwise admit
  // once the chunk is queued, so the host side of an admission (converting
  // and uploading the request, about 175 ms for 4087^2) overlaps the chunk.
  if (s->running.empty() && (rc = admit(s))) return rc;
  if (s->running.empty()) return TR_OK;
  const double t0 = now_s();
  int32_t active = 0;
  // Members run to their horizon with no step limit (as tr_batch_add leaves
  // them), so a round re-arms the batch only to resume members that stopped
  // at a print the last round captured.
  if (s->rearm) {
    if ((rc = tr_batch_arm(s->batch, 0.0, -1))) return rc;
    s->rearm = false;
  }
  if ((rc = tr_batch_enqueue(s->batch, s->opt.chunk_steps))) return rc;
  if ((rc = admit(s))) return rc;
  if ((rc = tr_batch_sync(s->batch, &active))) return rc;
  if ((rc = capture_prints(s))) return rc;
  if ((rc = retire(s))) return rc;
  s->stats.busy_s += now_s() - t0;
  s->stats.rounds++;
  if (ran) *ran = 1;
  return TR_OK;
}

int tr_serve_drain(tr_serve* s) {
  if (!s) return fail(TR_ERR_INVALID, "tr_serve_drain: NULL argument");
  while (!s->waiting.empty() || !s->running.empty()) {
    int32_t ran = 0;
    int rc = tr_serve_step(s, &ran);
    if (rc) return rc;
    if (!ran && !s->waiting.empty())
      return fail(TR_ERR_STATE, "tr_serve_drain: waiting requests cannot be admitted");
  }
  return TR_OK;
}

int tr_serve_poll(const tr_serve* s, int32_t request_id, tr_serve_request* out) {
  if (!s || !out) return fail(TR_ERR_INVALID, "tr_serve_poll: NULL argument");
  auto it = s->req.find(request_id);
  if (it == s->req.end()) return fail(TR_ERR_INVALID, "tr_serve_poll: unknown or released request");
  *out = it->second.info;
  return TR_OK;
}

int tr_serve_result(const tr_serve* s, int32_t request_id, double* h, double* qx, double* qy) {
  if (!s || !h || !qx || !qy) return fail(TR_ERR_INVALID, "tr_serve_result: NULL argument");
  auto it = s->req.find(request_id);
  if (it == s->req.end()) return fail(TR_ERR_INVALID, "tr_serve_result: unknown or released request");
  const Request& r = it->second;
  if (r.info.state != TR_REQ_FINISHED) return fail(TR_ERR_STATE, "tr_serve_result: request not finished");
  if (r.h.empty()) return fail(TR_ERR_STATE, "tr_serve_result: the server does not keep states");
  const size_t bytes = r.h.size() * sizeof(double);
  std::memcpy(h, r.h.data(), bytes);
  std::memcpy(qx, r.qx.data(), bytes);
  std::memcpy(qy, r.qy.data(), bytes);
  return TR_OK;
}

int tr_serve_release(tr_serve* s, int32_t request_id) {
  if (!s) return fail(TR_ERR_INVALID, "tr_serve_release: NULL argument");
  auto it = s->req.find(request_id);
  if (it == s->req.end()) return fail(TR_ERR_INVALID, "tr_serve_release: unknown or released request");
  const int32_t st = it->second.info.state;
  if (st == TR_REQ_WAITING || st == TR_REQ_RUNNING)
    return fail(TR_ERR_STATE, "tr_serve_release: request still waiting or running (tr_serve_cancel ends it)");
  s->req.erase(it);
  return TR_OK;
}

int tr_serve_print_count(const tr_serve* s, int32_t request_id, int32_t* n) {
  if (!s || !n) return fail(TR_ERR_INVALID, "tr_serve_print_count: NULL argument");
  auto it = s->req.find(request_id);
  if (it == s->req.end()) return fail(TR_ERR_INVALID, "tr_serve_print_count: unknown or released request");
  *n = int32_t(it->second.prints.size());
  return TR_OK;
}

int tr_serve_print(const tr_serve* s, int32_t request_id, int32_t k, int32_t* print_id, double* simtime,
                   double* h, double* qx, double* qy, double* maxh) {
  if (!s) return fail(TR_ERR_INVALID, "tr_serve_print: NULL argument");
  auto it = s->req.find(request_id);
  if (it == s->req.end()) return fail(TR_ERR_INVALID, "tr_serve_print: unknown or released request");
  const Request& r = it->second;
  if (k < 0 || k >= int32_t(r.prints.size())) return fail(TR_ERR_INVALID, "tr_serve_print: no such print");
  const PrintRec& p = r.prints[size_t(k)];
  if (print_id) *print_id = p.print_id;
  if (simtime) *simtime = p.simtime;
  const size_t bytes = p.h.size() * sizeof(double);
  if (h) std::memcpy(h, p.h.data(), bytes);
  if (qx) std::memcpy(qx, p.qx.data(), bytes);
  if (qy) std::memcpy(qy, p.qy.data(), bytes);
  if (maxh && !p.maxh.empty()) std::memcpy(maxh, p.maxh.data(), bytes);
  return TR_OK;
}

int tr_serve_cancel(tr_serve* s, int32_t request_id) {
  if (!s) return fail(TR_ERR_INVALID, "tr_serve_cancel: NULL argument");
  auto it = s->req.find(request_id);
  if (it == s->req.end()) return fail(TR_ERR_INVALID, "tr_serve_cancel: unknown or released request");
  Request& r = it->second;
  if (r.info.state == TR_REQ_WAITING) {
    s->waiting.erase(std::find(s->waiting.begin(), s->waiting.end(), request_id));
  } else if (r.info.state == TR_REQ_RUNNING) {
    const int32_t mid = r.info.member;
    int rc = tr_batch_poll(s->batch, mid, &r.info.status);
    if (!rc) rc = tr_batch_remove(s->batch, mid);
    if (rc) return rc;
    s->running.erase(mid);
  } else {
    return fail(TR_ERR_STATE, "tr_serve_cancel: request already ended");
  }
  end_request(s, r, TR_REQ_CANCELLED, TR_OK);
  return TR_OK;
}

int tr_serve_stats_get(const tr_serve* s, tr_serve_stats* out) {
  if (!s || !out) return fail(TR_ERR_INVALID, "tr_serve_stats_get: NULL argument");
  *out = s->stats;
  out->waiting = int64_t(s->waiting.size());
  out->running = int64_t(s->running.size());
  out->p50_latency_s = percentile(s->latencies, 50.0);
  out->p99_latency_s = percentile(s->latencies, 99.0);
  out->max_latency_s = s->latencies.empty() ? std::numeric_limits<double>::quiet_NaN()
                                             : *std::max_element(s->latencies.begin(), s->latencies.end());
  return TR_OK;
}

void tr_serve_destroy(tr_serve* s) {
  if (!s) return;
  tr_batch_destroy(s->batch);
  delete s;
}

}  // extern "C"
