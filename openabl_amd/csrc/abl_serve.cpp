// abl_serve.cpp -- continuous batching of simulation runs on one GPU.
//
// Requests (a program, its parameters, a timestep count, the generator
// state its initialisation starts from) arrive at any time through
// abl_server_submit.  A worker thread admits queued requests while the
// estimated device footprint of the active runs stays within the HBM budget
// and the active count within `max_active`, builds each admitted run's initial
// state on the host (the generator is process-wide, so initialisations are
// serialised here), and then advances every active run `chunk_steps`
// timesteps at a time on that run's own stream.  Work of different runs
// overlaps on the device; a run that finishes is downloaded, its memory is
// released, and the next queued request can take its place.  Each run is
// computed exactly as it would be alone: runs share no device state, and
// agent order inside a run depends only on that run's positions and ids.
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

enum { RUN_QUEUED = 0, RUN_ACTIVE = 1, RUN_DONE = 2, RUN_FAILED = 3 };

struct Run {
  uint64_t id = 0;
  std::string program;
  int precision = 0;
  std::vector<std::string> names;
  std::vector<double> values;
  int timesteps = 0;
  uint64_t rng[2] = {0xdeadbeefULL, 0xbeefdeadULL};
  size_t footprint = 0;
  std::atomic<int> state{RUN_QUEUED};
  std::string error;
  abl_sim *sim = nullptr;
  int issued = 0;  // timesteps handed to the device so far
  size_t agents = 0;
  Clock::time_point t_submit, t_admit, t_done;
  std::vector<std::vector<char>> results;  // per agent type, generated-struct records
  std::vector<size_t> counts;
  std::string log;
};

}  // namespace

struct abl_server {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Run *> queue;
  std::vector<Run *> active;
  std::map<uint64_t, Run *> runs;
  std::thread worker;
  bool stop = false;
  size_t budget = 0, used = 0, peak = 0;
  int max_active = 8;
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

void fail_run(Run *r, const std::string &what) {
  r->state = RUN_FAILED;
  r->error = what + ": " + abl_hip_last_error();
  if (r->sim) {
    abl_hip_destroy(r->sim);
    r->sim = nullptr;
  }
  r->t_done = Clock::now();
}

// Build the run's simulation and initial state (worker thread only).
bool start_run(Run *r) {
  if (abl_hip_create(r->program.c_str(), r->precision, &r->sim)) {
    fail_run(r, "create");
    return false;
  }
  for (size_t k = 0; k < r->names.size(); k++) {
    if (abl_hip_set_param(r->sim, r->names[k].c_str(), r->values[k])) {
      fail_run(r, "parameter " + r->names[k]);
      return false;
    }
  }
  abl_rng_seed(r->rng[0], r->rng[1]);
  if (abl_hip_init_agents(r->sim)) {
    fail_run(r, "init");
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

bool finish_run(Run *r) {
  const int nt = abl_hip_agent_type_count(r->sim);
  r->results.resize((size_t) nt);
  r->counts.resize((size_t) nt);
  for (int t = 0; t < nt; t++) {
    size_t n = 0;
    if (abl_hip_count(r->sim, t, &n)) {
      fail_run(r, "count");
      return false;
    }
    const size_t rec = abl_hip_agent_size(r->sim, t);
    r->results[(size_t) t].resize(std::max<size_t>(n, 1) * rec);
    if (abl_hip_download(r->sim, t, r->results[(size_t) t].data(), n, &r->counts[(size_t) t])) {
      fail_run(r, "download");
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
  r->state = RUN_DONE;
  r->t_done = Clock::now();
  return true;
}

void worker_loop(abl_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    // admission: in arrival order while memory and slots allow
    std::vector<Run *> admitted;
    bool rejected = false;
    while (!s->queue.empty() && (int) s->active.size() + (int) admitted.size() < s->max_active) {
      Run *r = s->queue.front();
      if (s->used + r->footprint > s->budget) {
        if (s->active.empty() && admitted.empty() && r->footprint > s->budget) {
          s->queue.pop_front();  // can never fit
          r->state = RUN_FAILED;
          r->error = "run needs more device memory than the server's budget";
          r->t_done = Clock::now();
          rejected = true;
          continue;
        }
        break;
      }
      s->queue.pop_front();
      s->used += r->footprint;
      s->peak = std::max(s->peak, s->used);
      r->state = RUN_ACTIVE;
      r->t_admit = Clock::now();
      if (!s->any_admitted) {
        s->any_admitted = true;
        s->t_first_admit = r->t_admit;
      }
      admitted.push_back(r);
    }
    // a waiter may already be blocked on a run rejected here
    if (rejected) s->cv.notify_all();
    std::vector<Run *> act = s->active;
    lk.unlock();

    bool progressed = !admitted.empty();
    std::vector<Run *> still;
    std::vector<Run *> ended;
    for (Run *r : admitted) {
      // a run that failed to start frees its budget and wakes its waiter
      if (start_run(r)) act.push_back(r);
      else ended.push_back(r);
    }
    for (Run *r : act) {
      if (r->state != RUN_ACTIVE) {
        ended.push_back(r);
        continue;
      }
      int done = 0;
      if (abl_hip_query(r->sim, &done)) {
        fail_run(r, "query");
        ended.push_back(r);
        continue;
      }
      if (!done) {
        still.push_back(r);
        continue;
      }
      if (r->issued >= r->timesteps) {
        finish_run(r);
        ended.push_back(r);
        progressed = true;
        continue;
      }
      const int k = std::min(s->chunk, r->timesteps - r->issued);
      if (abl_hip_simulate(r->sim, k)) {
        fail_run(r, "simulate");
        ended.push_back(r);
        continue;
      }
      r->issued += k;
      progressed = true;
      still.push_back(r);
    }

    lk.lock();
    s->active = still;
    for (Run *r : ended) {
      s->used -= r->footprint;
      if (r->state == RUN_DONE) {
        s->agent_steps += (double) r->agents * (double) r->timesteps;
        s->t_last_done = r->t_done;
      }
    }
    if (!ended.empty()) s->cv.notify_all();
    if (!progressed) s->cv.wait_for(lk, std::chrono::microseconds(200));
  }
}

double seconds(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

}  // namespace

extern "C" const char *abl_server_last_error(void) { return g_serve_error.c_str(); }

extern "C" int abl_server_create(size_t hbm_budget, int max_active, int chunk_steps,
                                 abl_server **out) {
  if (!out || max_active < 1 || chunk_steps < 1) return serve_fail(ABL_ERR_ARG, "bad argument");
  abl_server *s = new abl_server();
  s->budget = hbm_budget;
  s->max_active = max_active;
  s->chunk = chunk_steps;
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return ABL_OK;
}

extern "C" int abl_server_submit(abl_server *s, const abl_run_request *req, uint64_t *run_id) {
  if (!s || !req || !req->program || req->timesteps < 0 || (req->num_params && !req->param_names))
    return serve_fail(ABL_ERR_ARG, "bad request");
  Run *r = new Run();
  r->program = req->program;
  r->precision = req->precision;
  for (int k = 0; k < req->num_params; k++) {
    r->names.push_back(req->param_names[k]);
    r->values.push_back(req->param_values[k]);
  }
  r->timesteps = req->timesteps;
  r->rng[0] = req->rng_state[0];
  r->rng[1] = req->rng_state[1];
  size_t fp = 0;
  if (abl_hip_estimate_footprint(req->program, req->precision, req->agents_hint, req->cells_hint,
                                 &fp)) {
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
  s->cv.notify_all();
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
  *state = r->state;
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
  }
  s->worker.join();
  for (auto &kv : s->runs) {
    if (kv.second->sim) abl_hip_destroy(kv.second->sim);
    delete kv.second;
  }
  delete s;
  return ABL_OK;
}
