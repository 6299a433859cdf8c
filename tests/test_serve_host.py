"""Host logic of the serving loop, on the CPU: admission order, chunked
stepping and failure isolation, with the device region replaced by a stand-in
that counts steps (the real region runs in tests/test_gpu_batch_serve.py)."""
import numpy as np

from rocflupicl_amd import serve


class FakeRegion:
    """Stands in for rflu.Region: `idle` is true every other query, a
    request whose rid is in `fail_step` raises on its second chunk, one in
    `fail_wait` raises when its result is checked."""
    log = []

    def __init__(self, rid, fail_step, fail_wait):
        self.rid, self.steps, self.queries, self.closed = rid, 0, 0, False
        self.fail_step, self.fail_wait, self.chunks = fail_step, fail_wait, 0

    def idle(self):
        self.queries += 1
        return self.queries % 2 == 0

    def step_async(self, n):
        self.chunks += 1
        if self.rid in self.fail_step and self.chunks == 2:
            raise RuntimeError(f"EHIP: launch failed in request {self.rid}")
        self.steps += n

    def wait(self):
        if self.rid in self.fail_wait:
            raise RuntimeError("RFLU_CheckPositivity: negative positive-definite variables")

    def get_state(self):
        return np.full((5, 3), float(self.steps)), None

    def close(self):
        assert not self.closed
        self.closed = True
        FakeRegion.log.append(self.rid)


def run_trace(monkeypatch, fail_step=(), fail_wait=(), fail_build=()):
    made = {}

    class Srv(serve.Server):
        def _build(self, req):
            if req.rid in fail_build:
                raise RuntimeError("mesh: bad connectivity")
            made[req.rid] = FakeRegion(req.rid, set(fail_step), set(fail_wait))
            return made[req.rid]

    monkeypatch.setattr(serve.rflu, "mem_info", lambda: (1 << 40, 1 << 40))
    FakeRegion.log = []
    reqs = [serve.Request(i, 0.0005 * i, 7 + 3 * i, None, cells=100, particles=0) for i in range(6)]
    srv = Srv(max_active=3, chunk=4, builders=2, keep_state=True)
    results, wall = srv.run(reqs)
    return reqs, results, made


def test_every_request_gets_its_steps(monkeypatch):
    reqs, results, made = run_trace(monkeypatch)
    assert [r.rid for r in results] == [q.rid for q in reqs]
    for q, r in zip(reqs, results):
        assert r.error is None and made[q.rid].steps == q.steps
        assert np.all(r.cv == q.steps)
        assert r.done >= r.started >= r.admitted >= r.arrival - 1e-9
    assert sorted(FakeRegion.log) == list(range(6))  # every region freed once


def test_failures_stay_with_their_request(monkeypatch):
    reqs, results, made = run_trace(monkeypatch, fail_step=(1,), fail_wait=(3,), fail_build=(4,))
    err = {r.rid: r.error for r in results}
    assert "launch failed" in err[1]
    assert "negative" in err[3]
    assert "bad connectivity" in err[4]
    for rid in (0, 2, 5):
        assert err[rid] is None and made[rid].steps == reqs[rid].steps
    assert sorted(FakeRegion.log) == [0, 1, 2, 3, 5]  # failed regions are freed too
    s = serve.summarize(results, 1.0)
    assert s["failed"] == 3 and s["requests"] == 6
