"""Host-side serving logic that needs no device: a run larger than the
server's HBM budget is rejected at admission, and a waiter blocked on it is
woken (the worker notifies waiters of runs it rejects).  Repeated so the
waiter is sometimes already blocked when the rejection happens."""
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
