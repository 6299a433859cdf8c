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
    assert t == b.time()[0]
    c.RFLU_TimeStepping(4)
    ca, cc = a.get_state()[0], c.get_state()[0]
    if binary:
        assert np.array_equal(ca.view(np.uint64), cc.view(np.uint64))
        assert c.time() == a.time()
    else:
        err = (np.abs(cc - ca) / np.abs(ca).max(axis=0)).max()
        assert err < 1e-13, err
    assert (tmp_path / "shock.rin").exists()
