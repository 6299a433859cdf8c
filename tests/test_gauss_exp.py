"""The Gaussian deposit's exp (rocflupicl_amd/csrc/gauss_exp.h) against the
exact exponential.

The header is compiled for the host with the same operations the device runs
(fma, rint and ldexp are exact or correctly rounded on both), so its bits here
are the device's.  The weights ppiclF evaluates are exp(-r^2 / (2 sigma^2))
with r^2 up to the cut-off, i.e. arguments in [ln(alpha), 0]
(ppiclf_solve.f:3196-3246); the header is valid on [-700, 0].
"""
import ctypes
import os
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "rocflupicl_amd", "csrc", "gauss_exp.h")

SRC = r"""
#include "gauss_exp.h"
extern "C" void gx_run(int n, const double *x, double *y) {
  for (int i = 0; i < n; i++) y[i] = gx::exp_tab(x[i], gx::kTab);
}
"""


@pytest.fixture(scope="module")
def gx_exp(tmp_path_factory):
    d = tmp_path_factory.mktemp("gx")
    src = d / "gx.cpp"
    src.write_text(SRC)
    so = d / "gx.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                    "-I", os.path.dirname(HDR), str(src), "-o", str(so)], check=True)
    lib = ctypes.CDLL(str(so))

    def run(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib.gx_run(len(x), x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p))
        return y
    return run


def _ulps_to_exact(x, y):
    getcontext().prec = 40
    worst = 0.0
    for a, b in zip(x, y):
        t = Decimal(float(a)).exp()
        worst = max(worst, float(abs(Decimal(float(b)) - t) / Decimal(float(np.spacing(float(t))))))
    return worst


def test_within_one_ulp_of_the_exact_exponential_on_the_filter_range(gx_exp):
    rng = np.random.default_rng(11)
    # ln(alpha) for alpha = 1e-2 (the benchmark's) is -4.6; wider alphas reach further
    x = np.concatenate([rng.uniform(-4.7, 0.0, 3000), rng.uniform(-700.0, -4.7, 1000)])
    assert _ulps_to_exact(x, gx_exp(x)) < 1.0


def test_within_one_ulp_of_libm_on_many_arguments(gx_exp):
    rng = np.random.default_rng(12)
    x = rng.uniform(-700.0, 0.0, 200000)
    y, e = gx_exp(x), np.exp(x)
    assert np.max(np.abs(y - e) / np.spacing(e)) <= 1.0


def test_special_points(gx_exp):
    # exp(0) = 1, tiny arguments round to 1 or its predecessor, table nodes
    # x = -k ln2/32 land on 2^(-k/32)
    y = gx_exp(np.array([0.0, -0.0, -1e-300, -5e-17]))
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.0
    assert y[3] in (1.0, np.nextafter(1.0, 0.0))
    k = np.arange(0, 1000)
    x = -k * (np.log(2.0) / 32.0)
    assert np.max(np.abs(gx_exp(x) - np.exp(x)) / np.spacing(np.exp(x))) <= 1.0
    # the reduction's split constant: kL1 + kL2 is ln2/32 to 1e-32
    getcontext().prec = 50
    ln2_32 = Decimal("0.693147180559945309417232121458176568075500134360") / 32
    l1, l2 = float.fromhex("0x1.62e42fefa3800p-6"), float.fromhex("0x1.ef35793c76730p-50")
    assert abs(Decimal(l1) + Decimal(l2) - ln2_32) < Decimal("1e-32")


def test_monotone_on_the_filter_range(gx_exp):
    # a weight never increases with distance
    x = np.linspace(-4.7, 0.0, 400001)
    assert np.all(np.diff(gx_exp(x)) >= 0.0)
