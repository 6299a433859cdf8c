"""The allocation-free point location (picl_gpu_locate_static) against the
allocating one (picl_gpu_locate), through the C ABI: the same element ids on
two calls in a row, the second with fewer points, and at the size limit; a
call beyond the limit is refused with an error before any device work."""
import numpy as np
import pytest

from rocflupicl_amd import _lib, configs, rflu

pytestmark = pytest.mark.gpu


def region(jitter=0.15, nx=9, ny=6, nz=5):
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=nx, ny=ny, nz=nz, npart=40, jitter=jitter, seed=12)
    g = rflu.Region(mesh, **params)
    g.set_state(ic(g.mesh_arrays()["cofg"]))
    g.picl_init(y, pc, **picl)
    return g, np.array([nx, ny, nz]) * 0.0005


def points(n, hi, seed):
    rng = np.random.default_rng(seed)
    return hi * (rng.random((n, 3)) * 1.04 - 0.02)  # some outside the box: element -1


def test_static_locate_equals_allocating_locate():
    g, hi = region()
    before = g.device_bytes()
    for n, seed in ((50_000, 1), (1_234, 2)):  # the second call is smaller
        pts = points(n, hi, seed)
        a = g.picl_locate(pts, static=True)
        b = g.picl_locate(pts)
        assert np.array_equal(a, b), n
        assert (a >= 0).any() and (a < 0).any()
    assert g.device_bytes() == before


def test_static_locate_at_and_beyond_its_limit():
    g, hi = region(jitter=0.0, nx=5, ny=4, nz=4)
    n = _lib.PICL_LOCATE_STATIC_MAX
    pts = points(n, hi, 3)
    assert np.array_equal(g.picl_locate(pts, static=True), g.picl_locate(pts))
    with pytest.raises(_lib.GpuError, match="PICL_LOCATE_STATIC_MAX"):
        g.picl_locate(points(n + 1, hi, 4), static=True)
    small = points(7, hi, 5)  # still usable after the refusal
    assert np.array_equal(g.picl_locate(small, static=True), g.picl_locate(small))
