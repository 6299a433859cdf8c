"""Pins of the oracle's optional ppiclF models: the quasi-steady drag laws of
ppiclf_user_QuasiSteady.f, the trilinear evaluation and the Gaussian-filter
projection.  CPU only.

The drag laws are checked against what the physics fixes, not against a
retyped formula: in the incompressible, dilute limit every law reduces to the
standard sphere drag curve (Clift and Gauvin's fit, which Parmar's Re law is),
the Gidaspow (Wen-Yu) law to the isolated-sphere drag of its Schiller-Naumann
coefficient, the modified Parmar law to the Osnes law at low Mach number (the
claim of ppiclf_user_QuasiSteady.f:138-140), and the Osnes law is continuous
across its Mach-number branches (the constants were tuned for that,
:304-322).  The Gaussian filter conserves the projected quantity, has the
stated full width at half maximum, and drops exactly the nodes whose weight
falls below alpha.
"""
import math

import numpy as np
import pytest

from rocflupicl_amd import configs

NODE2V = [0, 3, 4, 7, 1, 2, 5, 6]
MU, DP, RHO = 1.8e-5, 1.0e-4, 1.2


def clift_gauvin(re):
    return 24.0 / re * (1.0 + 0.15 * re ** 0.687) + 0.42 / (1.0 + 42500.0 * re ** -1.16)


def cd_of_beta(beta, v):
    # F = beta v = Cd (rho v^2 / 2) (pi d^2 / 4)
    return 8.0 * beta / (math.pi * RHO * v * DP * DP)


@pytest.mark.parametrize("qs,res,tol", [
    (1, [0.5, 10, 100, 500, 2000], 3e-3),   # Parmar
    (2, [10, 50, 100, 500], 1.5e-2),        # Osnes (Loth) at phi -> 0
    (3, [10, 50, 100, 500, 2000], 1.5e-2),  # modified Parmar at phi -> 0
])
def test_drag_laws_reduce_to_the_standard_curve(oracle_mod, qs, res, tol):
    for re in res:
        v = re * MU / (RHO * DP)
        phi = 0.0 if qs == 1 else 1e-4
        beta = oracle_mod.qs_beta(qs, 0.01, phi, re, MU, DP, RHO, v)
        assert abs(cd_of_beta(beta, v) / clift_gauvin(re) - 1.0) < tol, (qs, re)


def test_gidaspow_dilute_limit_is_the_isolated_sphere(oracle_mod):
    for re in (1.0, 100.0, 900.0):
        v = re * MU / (RHO * DP)
        beta = oracle_mod.qs_beta(4, 0.01, 1e-4, re, MU, DP, RHO, v)
        cd_sn = 24.0 / re * (1.0 + 0.15 * re ** 0.687)
        assert abs(beta / (math.pi / 8.0 * cd_sn * RHO * v * DP * DP) - 1.0) < 1e-3


def test_schiller_naumann_coefficient(oracle_mod):
    for re in (0.1, 1.0, 50.0, 800.0):
        beta = oracle_mod.qs_beta(5, 0.3, 0.2, re, MU, DP)
        assert beta / (3.0 * math.pi * MU * DP) == pytest.approx(1.0 + 0.15 * re ** 0.687, rel=1e-14)


def test_osnes_mach_branches_are_continuous(oracle_mod):
    for re, mps in ((300.0, (0.8, 1.0, 1.5)), (30.0, (1.0,))):
        for mp in mps:
            lo = oracle_mod.qs_beta(2, mp * (1 - 1e-12), 0.1, re, MU, DP)
            hi = oracle_mod.qs_beta(2, mp * (1 + 1e-12), 0.1, re, MU, DP)
            assert abs(lo - hi) / lo < 1e-9, (re, mp)


def test_modified_parmar_matches_osnes_at_low_mach(oracle_mod):
    for phi in (0.01, 0.1, 0.3):
        for re in (50.0, 200.0, 1000.0):
            a = oracle_mod.qs_beta(3, 0.01, phi, re, MU, DP)
            b = oracle_mod.qs_beta(2, 0.01, phi, re, MU, DP)
            assert abs(a / b - 1.0) < 5e-3, (phi, re)


def region(oracle_mod, n=16, h=0.5, y=None, pc=None, **pp):
    mesh, params, ic, (y0, pc0), picl = configs.curtain(nx=n, ny=n, nz=n, h=h, npart=4)
    picl.update(pp)
    o = oracle_mod.Region(mesh, **params)
    o.set_cv(ic(o.cells()[1]))
    o.picl_init(y if y is not None else y0, pc if pc is not None else pc0, **picl)
    return o, mesh


def particles(xyz, vp):
    y = np.zeros((len(xyz), 12))
    y[:, :3] = xyz
    y[:, 6] = 300.0
    pc = np.stack([np.full(len(xyz), 2500.0), np.full(len(xyz), 1e-3), np.asarray(vp, float),
                   np.ones(len(xyz))], axis=1)
    return y, pc


def test_gaussian_projection_conserves_the_deposit(oracle_mod):
    h = 0.5
    y, pc = particles([[4.1, 3.9, 4.2], [3.7, 4.35, 3.8]], [1e-3, 2e-3])
    o, _ = region(oracle_mod, h=h, y=y, pc=pc, projMode=1, filterWidth=2 * h, filterAlpha=1e-12)
    nodes = o.picl_nodes()
    vol = o.cells()[0]
    total = (vol * (0.125 * nodes[:, 0, :]).sum(axis=1)).sum()
    assert abs(total / pc[:, 2].sum() - 1.0) < 1e-4
    # piclVF is the same read-back
    assert abs((o.picl_cells()["phi"] * vol).sum() / pc[:, 2].sum() - 1.0) < 1e-4


def test_gaussian_filter_width_and_cutoff(oracle_mod):
    h, width, alpha = 0.5, 1.0, 0.2
    y, pc = particles([[4.0, 4.0, 4.0]], [1e-3])  # on a mesh vertex
    o, mesh = region(oracle_mod, h=h, y=y, pc=pc, projMode=1, filterWidth=width, filterAlpha=alpha)
    nodes = o.picl_nodes()[:, 0, :]
    xyz, c2v = mesh["xyz"], mesh["c2v"]
    pos = xyz[c2v[:, NODE2V]]  # (nCells, 8, 3)
    r2 = ((pos - y[0, :3]) ** 2).sum(axis=2)
    peak = nodes[r2 == 0.0]
    assert peak.size and np.all(peak == peak[0])
    ratio = nodes / peak[0]
    # half maximum one spacing away: the filter width is the Gaussian's FWHM
    at_h = ratio[np.isclose(r2, h * h)]
    assert at_h.size and np.allclose(at_h, 0.5, rtol=1e-13, atol=0)
    # weights below alpha are dropped, the others kept
    sigma = width / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    w = np.exp(-r2 / (2.0 * sigma * sigma))
    assert np.array_equal(nodes > 0.0, w >= alpha)
    assert np.allclose(ratio[nodes > 0], w[nodes > 0], rtol=1e-13, atol=0)


def test_trilinear_evaluation_returns_the_element_value(oracle_mod):
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=8, ny=4, nz=4, npart=200, jitter=0.1)
    o = oracle_mod.Region(mesh, **params)
    o.set_cv(ic(o.cells()[1]))
    o.picl_init(y, pc, interpMode=1, **{k: v for k, v in picl.items() if k != "interpMode"})
    o.step(1)
    fields = o.picl_cells()["fields"]
    elem = o.picl_particles()["elem"]
    from rocflupicl_amd import rflu
    rp = o.picl_rprop()[:, np.array(rflu.PICL_FIELD_RPROP) - 1]
    ref = fields[elem]
    scale = np.maximum(np.abs(ref), 1e-300)
    assert (np.abs(rp - ref) / scale).max() <= 8 * np.finfo(float).eps


# ---- heat transfer (HEATTRANSFER 1-4, ppiclf_user_HeatTransfer.f) ----
def _ht_stage(oracle_mod, ht, tp=250.0):
    """One coupled RK stage of a dilute curtain (phi 1 %) sitting in the
    post-shock stream; returns (particles, rprop, picl params) of that stage."""
    n, h = 12, 0.0005
    mesh, params, ic, (y, pc), picl = configs.curtain(nx=n, ny=4, nz=4, npart=60, jitter=0.05, seed=4, phi=0.01,
                                                      x_shock=0.9 * n * h, x_curtain=0.3 * n * h)
    y = y.copy()
    y[:, 6] = tp
    picl.update(rmuFlag=0, heatTransfer=ht)
    o = oracle_mod.Region(mesh, **params)
    o.set_cv(ic(o.cells()[1]))
    o.picl_init(y, pc, **picl)
    o.rk_stage(1, 1e-8)
    return o.picl_particles(), o.picl_rprop(), picl


@pytest.mark.parametrize("ht", [1, 2, 3, 4])
def test_heat_transfer_nusselt_laws(oracle_mod, ht):
    p, rp, picl = _ht_stage(oracle_mod, ht)
    y0 = p["y"][:, 6] - (8.0 / 15.0) * 1e-8 * p["ydot"][:, 6]  # stage-1 input temperature
    rhop, rhof, dp, volp, phip = rp[:, 0], rp[:, 1], rp[:, 2], rp[:, 3], rp[:, 4]
    u = rp[:, 5:8]
    cs, tg = rp[:, 8], rp[:, 23]
    vrel = np.linalg.norm(u - (p["y"][:, 3:6] - (8.0 / 15.0) * 1e-8 * p["ydot"][:, 3:6]), axis=1)
    mu = picl["rmuRef"]
    re = np.maximum(0.1, vrel * dp * rhof / mu)
    ma = vrel / cs
    vg = 1.0 - np.minimum(phip, 0.62)
    kappa = 1004.64 * mu / 0.70
    moving = vrel > 1e-3
    assert moving.sum() >= 10
    nu = p["ydot"][:, 6] * rhop * volp * picl["rcpPart"] / (math.pi * kappa * dp * (tg - y0))
    pr3 = 0.7 ** (1.0 / 3.0)
    expect = {1: np.full_like(re, 2.0),                                          # conduction, Nu = 2
              2: 2.0 + 0.6 * np.sqrt(re) * pr3,                                  # Ranz-Marshall
              3: (7 - 10 * vg + 5 * vg ** 2) * (1 + 0.7 * re ** 0.2 * pr3)
              + (1.33 - 2.4 * vg + 1.2 * vg ** 2) * re ** 0.7 * pr3,            # Gunn
              4: 2 * np.exp(-ma) / (1 + 17 * ma / re)
              + 0.495 * pr3 * re ** 0.55 * (1 + 0.5 * np.exp(-17 * ma / re)) / 1.5}[ht]  # Fox
    assert np.allclose(nu[moving], expect[moving], rtol=1e-9, atol=0)
    # a hotter particle gives heat to the gas, a colder one takes it
    assert np.all(np.sign(p["ydot"][moving, 6]) == np.sign(tg - y0)[moving])


def test_heat_transfer_feeds_the_gas_energy_back(oracle_mod):
    on, _, picl = _ht_stage(oracle_mod, 2)
    off, rp, _ = _ht_stage(oracle_mod, 0)
    rmass_therm = rp[:, 0] * rp[:, 3] * picl["rcpPart"]
    spl = rp[:, 21]
    d = on["ydotc"][:, 6] - off["ydotc"][:, 6]
    q = spl * rmass_therm * on["ydot"][:, 6]
    assert np.allclose(d, q, rtol=1e-9, atol=1e-12 * np.abs(q).max())
    assert np.array_equal(on["ydot"][:, 3:6], off["ydot"][:, 3:6])  # momentum untouched
