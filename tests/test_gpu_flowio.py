"""Restart through a Rocflu flow file on the GPU: writing the binary flow
file after n steps and resuming a new region from it gives the same bits as
running straight through (the binary format is exact); the ASCII file
resumes within its 16 significant digits."""
import numpy as np
import pytest

from rocflupicl_amd import configs, rflu

pytestmark = pytest.mark.gpu


def region(mesh, params, ic):
    r = rflu.Region(mesh, **params)
    r.set_state(ic(r.mesh_arrays()["cofg"]))
    return r


@pytest.mark.parametrize("binary", [True, False])
def test_resume_from_flow_file(tmp_path, binary):
    mesh, params, ic = configs.shock_tets(n=8, discr=1, limiterGuard=1)
    a = region(mesh, params, ic)
    a.RFLU_TimeStepping(7)
    b = region(mesh, params, ic)
    b.RFLU_TimeStepping(3)
    name = b.write_flow("shock", str(tmp_path), binary=binary)
    c = rflu.Region(mesh, **params)
    t = c.read_flow(name)
    if binary:
        assert t == b.time()[0]
    else:  # E23.16 keeps 16 significant digits
        assert t == pytest.approx(b.time()[0], rel=1e-15)
    c.RFLU_TimeStepping(4)
    ca, cc = a.get_state()[0], c.get_state()[0]
    if binary:
        assert np.array_equal(ca.view(np.uint64), cc.view(np.uint64))
        assert c.time() == a.time()
    else:
        err = (np.abs(cc - ca) / np.abs(ca).max(axis=0)).max()
        assert err < 1e-13, err
    assert (tmp_path / "shock.rin").exists()


def test_particle_file_of_a_coupled_run(tmp_path):
    from rocflupicl_amd import flowio
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=12, ny=4, nz=4, npart=300, jitter=0.05)
    r = region(mesh, params, ic)
    r.picl_init(y, pc, **picl)
    r.RFLU_TimeStepping(3)
    r.write_particles(tmp_path / "par00001.vtu", cycle=3)
    d = flowio.read_particle_vtu(tmp_path / "par00001.vtu")
    p = r.picl_particles()
    assert d["cycle"] == 3 and d["y"].shape == (300, 12)
    assert np.array_equal(d["y"], p["y"].astype(np.float32).astype(np.float64))
    assert np.array_equal(d["tags"][:, 2], p["elem"] + 1.0)
