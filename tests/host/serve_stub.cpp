// serve_stub.cpp -- the serving engine (openabl_amd/csrc/abl_serve.cpp) on the
// CPU, against a stand-in for the runtime's C ABI whose "device" is a
// per-simulation completion time: abl_hip_simulate queues work that ends a
// random 0-2 ms later, abl_hip_query reports whether it has, and
// abl_hip_synchronize sleeps until it has.  Every call also sleeps a little at
// random, so the issuer, starter and finisher threads interleave in many
// ways.  main() streams requests in from two client threads and a poller
// (status/stats/run_times while runs end), waits for every run with a time
// limit, and checks that each run ended exactly once with its own results,
// that every call which would reach the device came from one server thread
// (or from the caller's thread after the server stopped), and that the
// initialisation ran with its uploads held back on the host.
// Built with -fsanitize=thread by tests/test_serving_host.py.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/abl_hip.h"

using Clock = std::chrono::steady_clock;

namespace {
thread_local std::string g_err;
std::mutex g_rng_mu;
uint64_t g_rng[2] = {1, 2};
std::atomic<int> g_live{0};
thread_local std::mt19937 g_jit{std::random_device{}()};

std::thread::id g_main;
std::mutex g_dev_mu;
std::thread::id g_dev;  // the one server thread that calls the device

// A call that would reach the device.
void device_call(const char *what) {
  const auto me = std::this_thread::get_id();
  if (me == g_main) return;
  std::lock_guard<std::mutex> g(g_dev_mu);
  if (g_dev == std::thread::id()) g_dev = me;
  if (g_dev != me) {
    std::fprintf(stderr, "%s: device call from a second server thread\n", what);
    std::abort();
  }
}

void jitter(int max_us) {
  std::uniform_int_distribution<int> d(0, max_us);
  std::this_thread::sleep_for(std::chrono::microseconds(d(g_jit)));
}
}  // namespace

struct abl_sim {
  std::mutex mu;  // the stand-in's own state; a real stream is one object too
  Clock::time_point busy_until = Clock::now();
  uint64_t seed = 0, pending_seed = 0;
  bool defer = false, has_pending = false;
  int steps = 0;
  double n = 0;
  std::atomic<int> users{0};  // calls in flight: must never exceed 1
};

struct Guard {  // one thread at a time per simulation, as the engine promises
  abl_sim *s;
  explicit Guard(abl_sim *x) : s(x) {
    if (s->users.fetch_add(1) != 0) {
      std::fprintf(stderr, "two threads in one simulation\n");
      std::abort();
    }
  }
  ~Guard() { s->users.fetch_sub(1); }
};

extern "C" {
const char *abl_hip_last_error(void) { return g_err.c_str(); }
void abl_rng_seed(uint64_t s0, uint64_t s1) {
  std::lock_guard<std::mutex> g(g_rng_mu);
  g_rng[0] = s0;
  g_rng[1] = s1;
}
int abl_hip_create(const char *program, int, abl_sim **out) {
  device_call("create");
  jitter(300);
  if (std::strcmp(program, "fail_create") == 0) {
    g_err = "no such program";
    return ABL_ERR_ARG;
  }
  *out = new abl_sim();
  g_live++;
  return ABL_OK;
}
int abl_hip_destroy(abl_sim *s) {
  device_call("destroy");
  { Guard g(s); jitter(300); }
  delete s;
  g_live--;
  return ABL_OK;
}
int abl_hip_set_param(abl_sim *s, const char *name, double v) {
  Guard g(s);
  if (std::strcmp(name, "num_agents") == 0) s->n = v;
  return ABL_OK;
}
int abl_hip_set_config(abl_sim *s, const char *key, const char *value) {
  Guard g(s);
  if (std::strcmp(key, "hip.defer_upload") == 0) {
    if (std::strcmp(value, "1") == 0) {
      s->defer = true;
    } else {
      device_call("upload");
      s->defer = false;
      if (s->has_pending) s->seed = s->pending_seed;
      s->has_pending = false;
      jitter(300);
    }
  }
  return ABL_OK;
}
int abl_hip_agent_type_count(abl_sim *s) { Guard g(s); return 1; }
size_t abl_hip_agent_size(abl_sim *s, int) { Guard g(s); return 16; }
int abl_hip_init_agents(abl_sim *s) {
  Guard g(s);
  if (!s->defer) {
    std::fprintf(stderr, "initialisation would upload to the device\n");
    std::abort();
  }
  {
    std::lock_guard<std::mutex> l(g_rng_mu);
    s->pending_seed = g_rng[0] ^ (g_rng[1] << 1);
  }
  s->has_pending = true;
  jitter(1000);
  return ABL_OK;
}
int abl_hip_count(abl_sim *s, int, size_t *n) {
  Guard g(s);
  *n = (size_t) s->n;
  return ABL_OK;
}
int abl_hip_simulate(abl_sim *s, int k) {
  Guard g(s);
  device_call("abl_hip_simulate");
  std::lock_guard<std::mutex> l(s->mu);
  std::uniform_int_distribution<int> d(0, 2000);
  const auto start = std::max(Clock::now(), s->busy_until);
  s->busy_until = start + std::chrono::microseconds(d(g_jit));
  s->steps += k;
  return ABL_OK;
}
int abl_hip_query(abl_sim *s, int *done) {
  Guard g(s);
  device_call("abl_hip_query");
  std::lock_guard<std::mutex> l(s->mu);
  *done = Clock::now() >= s->busy_until;
  return ABL_OK;
}
int abl_hip_synchronize(abl_sim *s) {
  Guard g(s);
  device_call("abl_hip_synchronize");
  Clock::time_point t;
  {
    std::lock_guard<std::mutex> l(s->mu);
    t = s->busy_until;
  }
  std::this_thread::sleep_until(t);
  return ABL_OK;
}
int abl_hip_download(abl_sim *s, int, void *aos, size_t cap, size_t *count) {
  Guard g(s);
  device_call("abl_hip_download");
  jitter(500);
  const size_t n = (size_t) s->n;
  if (count) *count = n;
  if (cap < n) return ABL_ERR_ARG;
  // each record: the run's generator seed and its timestep count
  for (size_t i = 0; i < n; i++) {
    uint64_t rec[2] = {s->seed, (uint64_t) s->steps};
    std::memcpy((char *) aos + 16 * i, rec, 16);
  }
  return ABL_OK;
}
int abl_hip_log_text(abl_sim *s, char *buf, size_t cap, size_t *len) {
  Guard g(s);
  const std::string t = "steps," + std::to_string(s->steps) + "\n";
  if (len) *len = t.size();
  if (buf && cap) {
    const size_t m = std::min(cap - 1, t.size());
    std::memcpy(buf, t.data(), m);
    buf[m] = 0;
  }
  return ABL_OK;
}
int abl_hip_estimate_footprint(const char *, int, int num_params, const char *const *names,
                               const double *values, size_t, size_t, size_t *bytes) {
  double n = 1000;
  for (int k = 0; k < num_params; k++)
    if (std::strcmp(names[k], "num_agents") == 0) n = values[k];
  *bytes = (size_t) (n * 100.0);
  return ABL_OK;
}
}  // extern "C"

int main(int argc, char **argv) {
  const int rounds = argc > 1 ? std::atoi(argv[1]) : 20;
  const int per_client = argc > 2 ? std::atoi(argv[2]) : 40;
  g_main = std::this_thread::get_id();
  for (int round = 0; round < rounds; round++) {
    {
      std::lock_guard<std::mutex> g(g_dev_mu);
      g_dev = std::thread::id();
    }
    abl_server *srv = nullptr;
    // a budget that admits a few runs at a time, so admission waits happen
    if (abl_server_create(400000, 1 + round % 8, 1 + round % 7, &srv)) return 2;
    struct Sub {
      uint64_t id;
      uint64_t rng0;
      int steps, n;
      bool expect_fail;
    };
    std::mutex subs_mu;
    std::vector<Sub> subs;
    std::atomic<bool> stop_poll{false};
    auto client = [&](int c) {
      std::mt19937 r(1000 * round + c);
      for (int k = 0; k < per_client; k++) {
        const int n = 1 + (int) (r() % 3000);
        const int steps = (int) (r() % 120);
        const bool bad = r() % 17 == 0;
        const bool huge = r() % 23 == 0;  // more than the whole budget: fails at once
        const char *names[] = {"num_agents"};
        const double vals[] = {huge ? 1e7 : (double) n};
        abl_run_request q{};
        q.program = bad ? "fail_create" : "stub";
        q.num_params = 1;
        q.param_names = names;
        q.param_values = vals;
        q.timesteps = steps;
        q.rng_state[0] = 7919ull * (uint64_t) (round * 100000 + c * 1000 + k) + 1;
        q.rng_state[1] = 3;
        uint64_t id = 0;
        if (abl_server_submit(srv, &q, &id)) std::abort();
        std::lock_guard<std::mutex> g(subs_mu);
        subs.push_back({id, q.rng_state[0], steps, huge ? 10000000 : n, bad || huge});
        std::this_thread::sleep_for(std::chrono::microseconds(r() % 400));
      }
    };
    std::thread poller([&] {
      while (!stop_poll.load()) {
        double st[6], t[3];
        abl_server_stats(srv, st);
        uint64_t id = 1 + (uint64_t) (std::rand() % (2 * per_client + 1));
        int state = 0;
        if (abl_server_status(srv, id, &state) == ABL_OK && state == ABL_RUN_DONE) {
          abl_server_run_times(srv, id, t);
          if (!(t[0] >= 0 && t[1] >= 0 && t[2] >= t[0])) {
            std::fprintf(stderr, "inconsistent run times\n");
            std::abort();
          }
        }
        std::this_thread::sleep_for(std::chrono::microseconds(50));
      }
    });
    std::thread c0(client, 0), c1(client, 1);
    c0.join();
    c1.join();
    const auto deadline = Clock::now() + std::chrono::seconds(60);
    for (const Sub &s : subs) {
      // the engine's own wait blocks; a watchdog turns a stuck run into an
      // error message naming it
      std::atomic<bool> done{false};
      std::thread dog([&] {
        while (!done.load()) {
          if (Clock::now() > deadline) {
            int st = -1;
            abl_server_status(srv, s.id, &st);
            std::fprintf(stderr, "round %d: run %llu never ended (state %d)\n", round,
                         (unsigned long long) s.id, st);
            std::_Exit(3);
          }
          std::this_thread::sleep_for(std::chrono::milliseconds(5));
        }
      });
      const int rc = abl_server_wait(srv, s.id);
      done = true;
      dog.join();
      if ((rc != ABL_OK) != s.expect_fail) {
        std::fprintf(stderr, "run %llu: wait %d, expected %s\n", (unsigned long long) s.id, rc,
                     s.expect_fail ? "failure" : "success");
        return 4;
      }
      if (rc != ABL_OK) continue;
      std::vector<char> buf((size_t) std::max(s.n, 1) * 16);
      size_t count = 0;
      if (abl_server_result(srv, s.id, 0, buf.data(), (size_t) s.n, &count) || count != (size_t) s.n)
        return 5;
      uint64_t rec[2];
      std::memcpy(rec, buf.data(), 16);
      if (rec[0] != (s.rng0 ^ (3ull << 1)) || rec[1] != (uint64_t) s.steps) {
        std::fprintf(stderr, "run %llu: results of another run\n", (unsigned long long) s.id);
        return 6;
      }
    }
    stop_poll = true;
    poller.join();
    double st[6];
    abl_server_stats(srv, st);
    abl_server_destroy(srv);
    if (g_live.load() != 0) {
      std::fprintf(stderr, "simulations left alive: %d\n", g_live.load());
      return 7;
    }
  }
  std::printf("ok\n");
  return 0;
}
