"""Writes tests/golden/briney_am.json: inputs and outputs of the reference's
Briney added-mass functions (libpicl/user_files/Added_Mass/
ppiclf_user_AddedMass_functions.f and _subroutines.f), compiled from the
reference tree by `make -C oracle ref` into oracle/_ref/libppiclf_am_ref.so.

Run in a checkout where the reference tree exists:
    python tests/golden/make_briney_am.py
The vectors pin the oracle's restatement (tests/test_picl_briney_pins.py)
where the compiled reference is not at hand.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libppiclf_am_ref.so")


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    R = C.CDLL(LIB)
    D = C.c_double

    def call(name, *a):
        f = getattr(R, name + "_")
        f.restype = D
        return f(*[C.byref(D(float(x))) for x in a])

    out = {"source": "reference Added_Mass/*.f compiled with amdflang -O2 -ffp-contract=off", "curves": [],
           "corrections": [], "rdf": [], "resistance_pair": []}
    for name in ("b11_11", "b11_22", "b12_11", "b12_22"):
        for d in np.linspace(2.0, 24.0, 23):
            for al in (0.0, 0.1, 0.25, 0.4):
                out["curves"].append([name, float(d), al, call(name, d, al)])
    rad = 5.75e-5
    for name in ("ia_analytical", "ii_analytical"):
        for m in (3.49, 4.0, 6.0, 8.0, 12.0, 20.0, 40.0):
            for al in (0.05, 0.2, 0.4):
                out["corrections"].append([name, m * rad, rad, al, call(name, m * rad, rad, al)])
    for name in ("ia_numerical", "ii_numerical"):
        for m in (2.0, 2.5, 3.0, 3.4):
            for al in (0.1, 0.3):
                out["corrections"].append([name, m * rad, rad, al, call(name, m * rad, rad, al)])
    for r in np.linspace(0.9, 4.0, 32):
        for al in (0.05, 0.2, 0.4):
            out["rdf"].append([float(r), al, call("rdf_analytical", r, al)])
    rng = np.random.default_rng(11)
    f = R.resistance_pair_
    for _ in range(24):
        x, y, z = (rng.normal(size=3) * 3e-4).tolist()
        al = float(rng.uniform(0.0, 0.4))
        Rm = np.zeros((6, 6), order="F")
        f(*[C.byref(D(v)) for v in (x, y, z, al, rad)], Rm.ctypes.data_as(C.c_void_p))
        out["resistance_pair"].append([x, y, z, al, rad, Rm.tolist()])
    path = os.path.join(HERE, "briney_am.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
