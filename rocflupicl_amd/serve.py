"""Serving loop: continuous batching of shock / particle solve requests on one
GPU (no reference counterpart; Rocflu runs one job per MPI launch).

A request is a region to advance a number of steps: a mesh and its initial
state, optionally a particle cloud.  The server

* admits requests in arrival order while the number in flight is below
  ``max_active`` and the device's free memory, less a reserve, covers the
  request's estimated footprint (``bytes_per_cell`` x cells +
  ``bytes_per_particle`` x particles);
* sets admitted requests up on builder threads (the host mesh pipeline and
  the upload run in C++ with the GIL released), so set-up overlaps with the
  device work of the requests already running, and lets the autotuner pick
  the gradient kernel for each request's mesh (``tune``);
* advances every request in flight by ``chunk`` steps at a time with
  asynchronous graph replays on its own non-blocking stream, so the work of
  different requests overlaps on the device; a request whose stream is busy
  is skipped, not waited for;
* on completion checks AfterUpdateMP's conditions, hands the final state to
  ``on_done`` and frees the request's device memory at once, which lets the
  next request in;
* isolates failures: a request that cannot be set up, or whose steps fail
  (a launch error, NaN or negative pressure), is closed and reported in its
  own Result.error while the others keep running.

Each request's result is independent of the mix it ran in: regions share no
device state, and every kernel is deterministic, so a request gives the same
bits alone and under load (tests/test_gpu_batch_serve.py checks this).
"""
import concurrent.futures as cf
import dataclasses
import time

import numpy as np

from . import rflu


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float              # seconds after the start of the run
    steps: int
    make: object                # () -> (mesh, params, ic(cofg) -> cv0, particles or None, picl params or None)
    cells: int = 0              # for admission; filled from the mesh when 0
    particles: int = 0


@dataclasses.dataclass
class Result:
    rid: int
    cells: int
    particles: int
    steps: int
    arrival: float
    admitted: float
    started: float
    done: float
    cv: np.ndarray = None
    error: str = None

    @property
    def latency(self):
        return self.done - self.arrival


class Server:
    def __init__(self, max_active=8, chunk=10, builders=4, reserve_frac=0.05, bytes_per_cell=3000,
                 bytes_per_particle=800, keep_state=False, on_done=None, tune=True):
        self.max_active = max_active
        self.chunk = chunk
        self.builders = builders
        self.reserve_frac = reserve_frac
        self.bytes_per_cell = bytes_per_cell
        self.bytes_per_particle = bytes_per_particle
        self.keep_state = keep_state
        self.on_done = on_done
        self.tune = tune

    def _estimate(self, req):
        return self.bytes_per_cell * req.cells + self.bytes_per_particle * req.particles

    def _build(self, req):
        mesh, params, ic, parts, picl = req.make()
        reg = rflu.Region(mesh, **params)
        reg.set_state(ic(reg.mesh_arrays()["cofg"]))
        if parts is not None:
            y, pc = parts
            reg.picl_init(y, pc, **(picl or {}))
        if self.tune:
            reg.autotune(reps=2)
        return reg

    def run(self, requests):
        """Serve `requests` (arrival times relative to now); returns the
        Results in request order and the wall time of the run."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        results = {}
        building = {}      # future -> (req, admitted time, reserved bytes)
        active = []        # [req, region, steps left, started, reserved bytes]
        reserved = 0
        free0, total = rflu.mem_info()
        budget = free0 - self.reserve_frac * total
        t0 = time.perf_counter()
        pool = cf.ThreadPoolExecutor(max_workers=self.builders)
        try:
            while pending or building or active:
                now = time.perf_counter() - t0
                # admission, in arrival order (no overtaking)
                while pending and pending[0].arrival <= now and len(building) + len(active) < self.max_active:
                    req = pending[0]
                    need = self._estimate(req)
                    if reserved + need > budget and (building or active):
                        break
                    pending.pop(0)
                    reserved += need
                    building[pool.submit(self._build, req)] = (req, now, need)
                # set-ups that finished
                for fut in [f for f in building if f.done()]:
                    req, adm, need = building.pop(fut)
                    now = time.perf_counter() - t0
                    try:
                        reg = fut.result()
                    except Exception as e:  # a request that cannot be set up fails alone
                        reserved -= need
                        results[req.rid] = Result(req.rid, req.cells, req.particles, req.steps, req.arrival, adm,
                                                  now, now, error=str(e))
                        continue
                    active.append([req, reg, req.steps, now, need, adm])
                # advance every idle request by one chunk; finish the done ones
                progressed = False
                for slot in list(active):
                    req, reg, left, started, need, adm = slot
                    err, cv = None, None
                    try:
                        if not reg.idle():
                            continue
                        progressed = True
                        if left > 0:
                            n = min(self.chunk, left)
                            reg.step_async(n)
                            slot[2] = left - n
                            continue
                        reg.wait()
                        if self.keep_state or self.on_done:
                            cv = reg.get_state()[0]
                    except Exception as e:  # a request whose steps fail ends alone; the others go on
                        progressed = True
                        err, cv = str(e), None
                    done = time.perf_counter() - t0
                    res = Result(req.rid, req.cells, req.particles, req.steps, req.arrival, adm, started, done,
                                 cv if self.keep_state else None, err)
                    if self.on_done and cv is not None:
                        self.on_done(req, cv)
                    reg.close()
                    reserved -= need
                    active.remove(slot)
                    results[req.rid] = res
                if not progressed:
                    time.sleep(0.0005)
        finally:
            for slot in active:
                slot[1].close()
            pool.shutdown(wait=True)
            for fut in building:  # set-ups still in flight when a run is cut short
                if fut.done() and fut.exception() is None:
                    fut.result().close()
        wall = time.perf_counter() - t0
        return [results[r.rid] for r in sorted(requests, key=lambda r: r.rid)], wall


def summarize(results, wall):
    ok = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    cell_updates = sum(r.cells * r.steps for r in ok)
    return dict(requests=len(results), failed=len(results) - len(ok), wall_s=wall,
                p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                mean_latency_s=float(lat.mean()), mcell_updates_per_s=cell_updates / wall / 1e6,
                particle_updates_per_s=sum(r.particles * r.steps for r in ok) / wall)


def trace(n, rate, seed=5, cells=(10_000, 2_000_000), particles=(0, 500_000), mach=(1.2, 3.0),
          steps=(100, 2000), particle_share=0.5):
    """A serving trace: n requests with Poisson arrivals at `rate` per second;
    cell counts log-uniform in `cells` (a 5-tet box, or a 5:1:1 hex box when
    the request carries particles), particle counts uniform in `particles`
    for a `particle_share` of the requests (lognormal diameters, median
    115 um, sigma_ln 0.2, rho_p 2500), shock Mach number uniform in `mach`,
    step counts uniform in `steps`.  Second-order AUSM+ with the sign-safe
    Venkatakrishnan guard (limiterGuard = 1)."""
    from . import configs
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for i in range(n):
        nc = float(np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1]))))
        m = float(rng.uniform(*mach))
        st = int(rng.integers(steps[0], steps[1] + 1))
        with_p = particles[1] > 0 and rng.random() < particle_share
        npart = int(rng.integers(max(particles[0], 1), particles[1] + 1)) if with_p else 0
        sd = int(rng.integers(1 << 30))
        if npart:
            k = max(2, int(round((nc / 5.0) ** (1.0 / 3.0))))

            def make(k=k, m=m, npart=npart, sd=sd):
                return configs.curtain(nx=5 * k, ny=k, nz=k, npart=npart, mach=m, seed=sd, limiterGuard=1)
            cells_n = 5 * k * k * k
        else:
            k = max(2, int(round((nc / 5.0) ** (1.0 / 3.0))))

            def make(k=k, m=m, sd=sd):
                mesh, params, ic = configs.shock_tets(n=k, discr=4, mach=m, seed=sd, limiterGuard=1)
                return mesh, params, ic, None, None
            cells_n = 5 * k ** 3
        reqs.append(Request(i, float(t[i]), st, make, cells_n, npart))
    return reqs

