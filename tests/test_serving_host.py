"""Host-side serving logic, on the CPU.

* A run larger than the server's HBM budget is rejected at admission, and a
  waiter blocked on it is woken (repeated so the waiter is sometimes already
  blocked when the rejection happens).
* The engine's threads (tests/host/serve_stub.cpp): the engine
  (openabl_amd/csrc/abl_serve.cpp) against a stand-in for the runtime whose
  streams finish at random times, with random delays in every call.  Checked
  in every round of submissions from two client threads and a poller: every
  run ends exactly once (none is left waiting), with its own results and
  timestep count; oversized and failing runs fail cleanly; per-run times are
  consistent while runs end; every call that would reach the device comes
  from one server thread; initialisation holds its uploads back on the host;
  no simulation is left alive.  Once plainly and once under ThreadSanitizer
  (data races and lock misuse are errors).
"""
import os
import shutil
import subprocess
import threading

import pytest


def test_rejected_runs_wake_their_waiters():
    from openabl_amd import AblError
    from openabl_amd.serve import Server
    errors = []

    def body():
        try:
            for _ in range(100):
                srv = Server(hbm_budget=1 << 20, max_active=2, chunk_steps=10)
                rid = srv.submit("boids2d", "f64", {"num_agents": 100000, "agent_density": 0.065536}, 10)
                with pytest.raises(AblError):
                    srv.wait(rid)
                assert srv.status(rid) == 3
                srv.close()
        except BaseException as e:  # reported by the main thread
            errors.append(e)

    t = threading.Thread(target=body, daemon=True)
    t.start()
    t.join(60)
    assert not t.is_alive(), "a waiter was never woken"
    assert not errors, errors


REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(REPO, "tests", "host", "serve_stub.cpp"),
       os.path.join(REPO, "openabl_amd", "csrc", "abl_serve.cpp")]


def build(tmp_path, tsan):
    exe = str(tmp_path / ("serve_tsan" if tsan else "serve"))
    cmd = ["g++", "-O1" if tsan else "-O2", "-g", "-std=c++17", "-pthread", *SRC, "-o", exe]
    if tsan:
        cmd[1:1] = ["-fsanitize=thread", "-include",
                    os.path.join(REPO, "tests", "host", "tsan_condvar.h")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        if tsan:
            pytest.skip(f"no ThreadSanitizer build here: {r.stderr[-300:]}")
        raise AssertionError(r.stderr)
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.parametrize("tsan,rounds", [(False, 24), (True, 6)], ids=["plain", "tsan"])
def test_serving_threads_under_stress(tmp_path, tsan, rounds):
    exe = build(tmp_path, tsan)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([exe, str(rounds), "40"], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-3000:]
