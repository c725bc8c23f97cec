// abl_serve.cpp -- continuous batching of simulation runs on one GPU.
//
// Requests (a program, its parameters, a timestep count, the generator
// state its initialisation starts from) arrive at any time through
// abl_server_submit.  Two threads serve them:
//   issuer    the only thread that drives the device.  It admits queued
//             requests in arrival order while the estimated device footprint
//             of the admitted runs stays within the HBM budget and their
//             count within `max_active`, creates each admitted run's
//             simulation, uploads the initial state the initialiser built,
//             advances every ready run `chunk_steps` timesteps at a time on
//             that run's own stream (a launch queue, never a wait), so runs
//             overlap on the device, and downloads and releases each finished
//             run;
//   starter   builds each created run's initial state on the host (the
//             model's initialisation with the run's generator state; the
//             simulation holds the records back under hip.defer_upload, so
//             this thread makes no device call; the generator is
//             process-wide, so initialisations are serialised here).
// The costly host part of starting a run therefore never holds up issuing
// work to the runs already on the device, and every device call of the
// server comes from one thread.  A run's state, error text, timestamps and
// results are published together under the server's mutex, which is also
// how every reader takes them.  Each run is computed exactly as it would be
// alone: runs share no device state, and agent order inside a run depends
// only on that run's positions and ids.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/abl_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

// Public states (abl_hip.h) plus the hand-over state, reported as ACTIVE.
enum { RUN_QUEUED = 0, RUN_ACTIVE = 1, RUN_DONE = 2, RUN_FAILED = 3, RUN_STARTING = 4 };

struct Run {
  uint64_t id = 0;
  std::string program;
  int precision = 0;
  std::vector<std::string> names;
  std::vector<double> values;
  std::vector<std::pair<std::string, std::string>> config;
  int timesteps = 0;
  uint64_t rng[2] = {0xdeadbeefULL, 0xbeefdeadULL};
  size_t footprint = 0;
  int state = RUN_QUEUED;  // guarded by abl_server::mu
  std::string error;       // guarded by abl_server::mu
  abl_sim *sim = nullptr;  // owned by the thread that holds the run
  bool init_ok = false;    // starter -> issuer
  std::string init_error;  // starter -> issuer
  int issued = 0;          // timesteps handed to the device so far (issuer)
  size_t agents = 0;
  Clock::time_point t_submit, t_admit, t_done;
  std::vector<std::vector<char>> results;  // per agent type, generated-struct records
  std::vector<size_t> counts;
  std::string log;
};

}  // namespace

struct abl_server {
  std::mutex mu;
  std::condition_variable cv;       // waiters: a run ended
  std::condition_variable work_cv;  // issuer, starter: work arrived
  std::deque<Run *> queue;          // submitted, not admitted
  std::deque<Run *> to_init;        // created, waiting for the starter
  std::vector<Run *> inited;        // initial state built, back to the issuer
  std::map<uint64_t, Run *> runs;
  std::thread issuer, starter;
  bool stop = false;
  size_t budget = 0, used = 0, peak = 0;
  int max_active = 8;
  int admitted = 0;  // runs holding budget (starting or ready)
  int chunk = 20;
  uint64_t next_id = 1;
  double agent_steps = 0.0;
  Clock::time_point t_first_admit;
  Clock::time_point t_last_done;
  bool any_admitted = false;
};

namespace {

thread_local std::string g_serve_error;

int serve_fail(int code, const std::string &msg) {
  g_serve_error = msg;
  return code;
}

// End a run (caller holds s->mu): publish state, error and time together,
// release its budget, wake its waiters.
void end_run_locked(abl_server *s, Run *r, int state, const std::string &error) {
  r->error = error;
  r->t_done = Clock::now();
  r->state = state;
  if (state == RUN_DONE) {
    s->agent_steps += (double) r->agents * (double) r->timesteps;
    s->t_last_done = r->t_done;
  }
  s->used -= r->footprint;
  s->admitted -= 1;
  s->cv.notify_all();
  s->work_cv.notify_all();
}

// A failed run's message; its device memory is released (issuer).
std::string drop_sim(Run *r, const std::string &what) {
  std::string msg = what + ": " + abl_hip_last_error();
  if (r->sim) {
    abl_hip_destroy(r->sim);
    r->sim = nullptr;
  }
  return msg;
}

// The run's simulation with its parameters and configuration, holding
// uploads back on the host (issuer).
bool create_run(Run *r, std::string *err) {
  if (abl_hip_create(r->program.c_str(), r->precision, &r->sim)) {
    *err = drop_sim(r, "create");
    return false;
  }
  for (size_t k = 0; k < r->names.size(); k++) {
    if (abl_hip_set_param(r->sim, r->names[k].c_str(), r->values[k])) {
      *err = drop_sim(r, "parameter " + r->names[k]);
      return false;
    }
  }
  for (const auto &kv : r->config) {
    if (abl_hip_set_config(r->sim, kv.first.c_str(), kv.second.c_str())) {
      *err = drop_sim(r, "config " + kv.first);
      return false;
    }
  }
  if (abl_hip_set_config(r->sim, "hip.defer_upload", "1")) {
    *err = drop_sim(r, "config hip.defer_upload");
    return false;
  }
  return true;
}

// The initial state, on the host only (starter).
void init_run(Run *r) {
  abl_rng_seed(r->rng[0], r->rng[1]);
  r->init_ok = abl_hip_init_agents(r->sim) == ABL_OK;
  if (!r->init_ok) r->init_error = std::string("init: ") + abl_hip_last_error();
}

// Upload the initial state the starter built (issuer).
bool upload_run(Run *r, std::string *err) {
  if (!r->init_ok) {
    *err = r->init_error;
    drop_sim(r, "init");
    return false;
  }
  if (abl_hip_set_config(r->sim, "hip.defer_upload", "0")) {
    *err = drop_sim(r, "upload");
    return false;
  }
  r->agents = 0;
  for (int t = 0; t < abl_hip_agent_type_count(r->sim); t++) {
    size_t n = 0;
    abl_hip_count(r->sim, t, &n);
    r->agents += n;
  }
  return true;
}

// Download the results and release the device memory (issuer).
bool finish_run(Run *r, std::string *err) {
  if (abl_hip_synchronize(r->sim)) {
    *err = drop_sim(r, "synchronize");
    return false;
  }
  const int nt = abl_hip_agent_type_count(r->sim);
  r->results.resize((size_t) nt);
  r->counts.resize((size_t) nt);
  for (int t = 0; t < nt; t++) {
    size_t n = 0;
    if (abl_hip_count(r->sim, t, &n)) {
      *err = drop_sim(r, "count");
      return false;
    }
    const size_t rec = abl_hip_agent_size(r->sim, t);
    r->results[(size_t) t].resize(std::max<size_t>(n, 1) * rec);
    if (abl_hip_download(r->sim, t, r->results[(size_t) t].data(), n, &r->counts[(size_t) t])) {
      *err = drop_sim(r, "download");
      return false;
    }
  }
  size_t len = 0;
  abl_hip_log_text(r->sim, nullptr, 0, &len);
  if (len) {
    r->log.resize(len + 1);
    abl_hip_log_text(r->sim, &r->log[0], len + 1, &len);
    r->log.resize(len);
  }
  abl_hip_destroy(r->sim);
  r->sim = nullptr;
  return true;
}

// Admission (caller holds s->mu): in arrival order while memory and slots
// allow; a run that can never fit fails at once.  Admitted runs are handed
// back to the issuer for creation.
void admit_locked(abl_server *s, std::vector<Run *> *admitted) {
  while (!s->queue.empty() && s->admitted < s->max_active) {
    Run *r = s->queue.front();
    if (s->used + r->footprint > s->budget) {
      if (s->admitted == 0 && r->footprint > s->budget) {
        s->queue.pop_front();
        s->used += r->footprint;  // end_run_locked releases it
        s->admitted += 1;
        end_run_locked(s, r, RUN_FAILED, "run needs more device memory than the server's budget");
        continue;
      }
      break;
    }
    s->queue.pop_front();
    s->used += r->footprint;
    s->peak = std::max(s->peak, s->used);
    s->admitted += 1;
    r->state = RUN_STARTING;
    r->t_admit = Clock::now();
    if (!s->any_admitted) {
      s->any_admitted = true;
      s->t_first_admit = r->t_admit;
    }
    admitted->push_back(r);
  }
}

void starter_loop(abl_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->work_cv.wait(lk, [&] { return s->stop || !s->to_init.empty(); });
    if (s->stop) return;
    Run *r = s->to_init.front();
    s->to_init.pop_front();
    lk.unlock();
    init_run(r);
    lk.lock();
    s->inited.push_back(r);
    s->work_cv.notify_all();
  }
}

// The issuer: admission and creation, uploads of initialised runs, then a
// pass over the started runs: each run whose stream is idle gets its next
// chunk of timesteps, or is downloaded and released.
void issuer_loop(abl_server *s) {
  std::vector<Run *> ready;  // started runs, held by this thread only
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    std::vector<Run *> fresh, uploaded;
    admit_locked(s, &fresh);
    uploaded.swap(s->inited);
    lk.unlock();

    bool progressed = !fresh.empty() || !uploaded.empty();
    std::vector<Run *> created;
    std::vector<std::pair<Run *, std::string>> ended;  // (run, error: empty = done)
    for (Run *r : fresh) {
      std::string err;
      if (create_run(r, &err)) created.push_back(r);
      else ended.push_back({r, err.empty() ? "create failed" : err});
    }
    for (Run *r : uploaded) {
      std::string err;
      if (upload_run(r, &err)) {
        lk.lock();
        r->state = RUN_ACTIVE;
        lk.unlock();
        ready.push_back(r);
      } else {
        ended.push_back({r, err.empty() ? "init failed" : err});
      }
    }
    std::vector<Run *> keep;
    for (Run *r : ready) {
      int done = 0;
      if (abl_hip_query(r->sim, &done)) {
        ended.push_back({r, drop_sim(r, "query")});
        continue;
      }
      if (!done) {
        keep.push_back(r);
        continue;
      }
      if (r->issued >= r->timesteps) {
        std::string err;
        ended.push_back({r, finish_run(r, &err) ? std::string() : (err.empty() ? "finish failed" : err)});
        progressed = true;
        continue;
      }
      const int k = std::min(s->chunk, r->timesteps - r->issued);
      if (abl_hip_simulate(r->sim, k)) {
        ended.push_back({r, drop_sim(r, "simulate")});
        continue;
      }
      r->issued += k;
      progressed = true;
      keep.push_back(r);
    }
    ready.swap(keep);

    lk.lock();
    for (Run *r : created) s->to_init.push_back(r);
    if (!created.empty()) s->work_cv.notify_all();
    for (auto &e : ended) end_run_locked(s, e.first, e.second.empty() ? RUN_DONE : RUN_FAILED, e.second);
    if (!progressed) s->work_cv.wait_for(lk, std::chrono::microseconds(200));
  }
  lk.unlock();
  // stopping: release the runs this thread still holds (device calls stay
  // on this thread)
  for (Run *r : ready) drop_sim(r, "stopped");
}

double seconds(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

int public_state(int st) { return st == RUN_STARTING ? RUN_ACTIVE : st; }

}  // namespace

extern "C" const char *abl_server_last_error(void) { return g_serve_error.c_str(); }

extern "C" int abl_server_create(size_t hbm_budget, int max_active, int chunk_steps,
                                 abl_server **out) {
  if (!out || max_active < 1 || chunk_steps < 1) return serve_fail(ABL_ERR_ARG, "bad argument");
  abl_server *s = new abl_server();
  s->budget = hbm_budget;
  s->max_active = max_active;
  s->chunk = chunk_steps;
  s->issuer = std::thread(issuer_loop, s);
  s->starter = std::thread(starter_loop, s);
  *out = s;
  return ABL_OK;
}

extern "C" int abl_server_submit(abl_server *s, const abl_run_request *req, uint64_t *run_id) {
  if (!s || !req || !req->program || req->timesteps < 0 || req->num_params < 0 ||
      (req->num_params && (!req->param_names || !req->param_values)) || req->num_config < 0 ||
      (req->num_config && (!req->config_keys || !req->config_values)))
    return serve_fail(ABL_ERR_ARG, "bad request");
  Run *r = new Run();
  r->program = req->program;
  r->precision = req->precision;
  for (int k = 0; k < req->num_params; k++) {
    r->names.push_back(req->param_names[k]);
    r->values.push_back(req->param_values[k]);
  }
  for (int k = 0; k < req->num_config; k++) {
    if (!req->config_keys[k] || !req->config_values[k]) {
      delete r;
      return serve_fail(ABL_ERR_ARG, "bad request");
    }
    r->config.push_back({req->config_keys[k], req->config_values[k]});
  }
  r->timesteps = req->timesteps;
  r->rng[0] = req->rng_state[0];
  r->rng[1] = req->rng_state[1];
  size_t fp = 0;
  if (abl_hip_estimate_footprint(req->program, req->precision, req->num_params, req->param_names,
                                 req->param_values, req->agents_hint, req->cells_hint, &fp)) {
    std::string m = abl_hip_last_error();
    delete r;
    return serve_fail(ABL_ERR_ARG, m);
  }
  r->footprint = fp;
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> g(s->mu);
  r->id = s->next_id++;
  s->runs[r->id] = r;
  s->queue.push_back(r);
  s->work_cv.notify_all();
  if (run_id) *run_id = r->id;
  return ABL_OK;
}

static Run *find_run(abl_server *s, uint64_t id) {
  auto it = s->runs.find(id);
  return it == s->runs.end() ? nullptr : it->second;
}

extern "C" int abl_server_status(abl_server *s, uint64_t id, int *state) {
  if (!s || !state) return serve_fail(ABL_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> g(s->mu);
  Run *r = find_run(s, id);
  if (!r) return serve_fail(ABL_ERR_ARG, "unknown run");
  *state = public_state(r->state);
  return ABL_OK;
}

extern "C" int abl_server_wait(abl_server *s, uint64_t id) {
  if (!s) return serve_fail(ABL_ERR_ARG, "bad argument");
  std::unique_lock<std::mutex> lk(s->mu);
  Run *r = find_run(s, id);
  if (!r) return serve_fail(ABL_ERR_ARG, "unknown run");
  s->cv.wait(lk, [&] { return r->state == RUN_DONE || r->state == RUN_FAILED; });
  if (r->state == RUN_FAILED) return serve_fail(ABL_ERR_STATE, r->error);
  return ABL_OK;
}

extern "C" int abl_server_result(abl_server *s, uint64_t id, int agent_type, void *aos,
                                 size_t capacity, size_t *count) {
  if (!s) return serve_fail(ABL_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> g(s->mu);
  Run *r = find_run(s, id);
  if (!r) return serve_fail(ABL_ERR_ARG, "unknown run");
  if (r->state != RUN_DONE) return serve_fail(ABL_ERR_STATE, "run not finished");
  if (agent_type < 0 || agent_type >= (int) r->counts.size())
    return serve_fail(ABL_ERR_ARG, "agent type out of range");
  const size_t n = r->counts[(size_t) agent_type];
  if (count) *count = n;
  if (!aos) return ABL_OK;
  if (capacity < n) return serve_fail(ABL_ERR_ARG, "buffer too small");
  const size_t rec = n ? r->results[(size_t) agent_type].size() / std::max<size_t>(n, 1) : 0;
  std::memcpy(aos, r->results[(size_t) agent_type].data(), n * rec);
  return ABL_OK;
}

extern "C" int abl_server_run_log(abl_server *s, uint64_t id, char *buf, size_t capacity,
                                  size_t *len) {
  if (!s) return serve_fail(ABL_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> g(s->mu);
  Run *r = find_run(s, id);
  if (!r || r->state != RUN_DONE) return serve_fail(ABL_ERR_STATE, "run not finished");
  if (len) *len = r->log.size();
  if (buf && capacity) {
    size_t m = std::min(capacity - 1, r->log.size());
    std::memcpy(buf, r->log.data(), m);
    buf[m] = 0;
  }
  return ABL_OK;
}

extern "C" int abl_server_run_times(abl_server *s, uint64_t id, double out[3]) {
  if (!s || !out) return serve_fail(ABL_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> g(s->mu);
  Run *r = find_run(s, id);
  if (!r) return serve_fail(ABL_ERR_ARG, "unknown run");
  if (r->state != RUN_DONE) return serve_fail(ABL_ERR_STATE, "run not finished");
  out[0] = seconds(r->t_submit, r->t_admit);
  out[1] = seconds(r->t_admit, r->t_done);
  out[2] = seconds(r->t_submit, r->t_done);
  return ABL_OK;
}

// out: [0] completed runs, [1] p50 latency s, [2] p99 latency s, [3] agent-steps
// completed, [4] seconds from first admission to last completion, [5] peak
// admitted footprint in bytes.
extern "C" int abl_server_stats(abl_server *s, double out[6]) {
  if (!s || !out) return serve_fail(ABL_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> g(s->mu);
  std::vector<double> lat;
  for (auto &kv : s->runs) {
    if (kv.second->state == RUN_DONE) lat.push_back(seconds(kv.second->t_submit, kv.second->t_done));
  }
  std::sort(lat.begin(), lat.end());
  auto pct = [&](double p) {
    if (lat.empty()) return 0.0;
    size_t k = (size_t) std::ceil(p * (double) lat.size()) - 1;
    return lat[std::min(k, lat.size() - 1)];
  };
  out[0] = (double) lat.size();
  out[1] = pct(0.50);
  out[2] = pct(0.99);
  out[3] = s->agent_steps;
  out[4] = s->any_admitted ? seconds(s->t_first_admit, s->t_last_done) : 0.0;
  out[5] = (double) s->peak;
  return ABL_OK;
}

extern "C" int abl_server_destroy(abl_server *s) {
  if (!s) return serve_fail(ABL_ERR_ARG, "bad argument");
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv.notify_all();
    s->work_cv.notify_all();
  }
  s->issuer.join();
  s->starter.join();
  // runs the issuer had not taken back from the starter (their simulations
  // hold no device work: both threads have ended)
  for (auto &kv : s->runs) {
    if (kv.second->sim) abl_hip_destroy(kv.second->sim);
    delete kv.second;
  }
  delete s;
  return ABL_OK;
}
