"""Batched ensemble and serving loop on the GPU.

A batch (rflu_batch_*) is the disjoint union of its instances in one handle;
every instance keeps its own dt and time, and its result must equal the
same region run alone, bit for bit.  The serving loop runs requests on
their own streams with asynchronous graph replays; each request's result
must equal the request run alone, bit for bit, whatever else is in flight.
"""
import numpy as np
import pytest

from rocflupicl_amd import configs, rflu, serve

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def alone(mesh, params, ic, steps, parts=None, picl=None):
    r = rflu.Region(mesh, **params)
    r.set_state(ic(r.mesh_arrays()["cofg"]))
    if parts is not None:
        r.picl_init(*parts, **picl)
    r.RFLU_TimeStepping(steps, use_graph=True)
    return r


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_instances_equal_regions_alone(prec):
    specs = [(6, 1.3, 11), (9, 2.6, 12), (4, 1.8, 13), (7, 3.0, 14)]
    cases = [configs.shock_tets(n=n, discr=4, mach=m, seed=sd, limiterGuard=1) for n, m, sd in specs]
    params = dict(cases[0][1], fluxPrecision=prec)
    b = rflu.Batch([c[0] for c in cases], **params)
    for k, (mesh, _, ic) in enumerate(cases):
        r = rflu.Region(mesh, **params)
        b.set_state(k, ic(r.mesh_arrays()["cofg"]))
        r.close()
    b.RFLU_TimeStepping(12, use_graph=True)
    for k, (mesh, _, ic) in enumerate(cases):
        a = alone(mesh, params, ic, 12)
        assert same_bits(b.get_state(k)[0], a.get_state()[0]), k
        assert same_bits(b.get_state(k)[1], a.get_state()[1]), k
        assert b.time(k) == a.time(), k
    # instances advance with their own dt
    dts = {b.time(k)[1] for k in range(len(cases))}
    assert len(dts) > 1


def test_batch_of_hexes_eager_equals_graph():
    cases = [configs.shock_hex(nx=10 + 3 * k, ny=4, nz=4, mach=1.5 + 0.5 * k, limiterGuard=1) for k in range(3)]
    params = cases[0][1]
    bs = []
    for graph in (True, False):
        b = rflu.Batch([c[0] for c in cases], **params)
        for k, (mesh, _, ic) in enumerate(cases):
            r = rflu.Region(mesh, **params)
            b.set_state(k, ic(r.mesh_arrays()["cofg"]))
            r.close()
        b.RFLU_TimeStepping(6, use_graph=graph)
        bs.append(b)
    for k in range(3):
        assert same_bits(bs[0].get_state(k)[0], bs[1].get_state(k)[0])


def test_served_results_equal_requests_alone():
    reqs = serve.trace(6, rate=200.0, seed=21, cells=(600, 6000), particles=(100, 600), steps=(7, 23),
                       particle_share=0.5)
    assert any(r.particles for r in reqs) and any(not r.particles for r in reqs)
    srv = serve.Server(max_active=3, chunk=4, builders=2, keep_state=True)
    results, wall = srv.run(reqs)
    assert all(r.error is None for r in results), [r.error for r in results]
    for req, res in zip(reqs, results):
        mesh, params, ic, parts, picl = req.make()
        a = alone(mesh, params, ic, req.steps, parts, picl)
        assert same_bits(res.cv, a.get_state()[0]), req.rid
    s = serve.summarize(results, wall)
    assert s["failed"] == 0 and s["p99_latency_s"] >= s["p50_latency_s"] > 0


def test_async_stepping_matches_blocking_steps():
    mesh, params, ic = configs.shock_tets(n=6, discr=1, limiterGuard=1)
    a = alone(mesh, params, ic, 9)
    r = rflu.Region(mesh, **params)
    r.set_state(ic(r.mesh_arrays()["cofg"]))
    for n in (4, 3, 2):
        r.step_async(n)
    r.wait()
    assert r.idle()
    assert same_bits(r.get_state()[0], a.get_state()[0])
    assert r.device_bytes() > 0
    free, total = rflu.mem_info()
    assert 0 < free <= total


def test_autotune_keeps_the_bits():
    mesh, params, ic = configs.shock_tets(n=10, discr=1, limiterGuard=1)
    a = alone(mesh, params, ic, 6)
    r = rflu.Region(mesh, **params)
    r.set_state(ic(r.mesh_arrays()["cofg"]))
    r.RFLU_TimeStepping(2, use_graph=True)
    v, ms = r.autotune(reps=3)
    assert v in (2, 3, 4, 24, 25, 26) and ms > 0.0
    r.RFLU_TimeStepping(4, use_graph=True)
    assert same_bits(r.get_state()[0], a.get_state()[0])
