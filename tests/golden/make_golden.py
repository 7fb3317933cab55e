"""Write the committed golden fixtures from the reference's own compiled code.

Run in the build container, where /root/reference exists:
    make -C oracle ref && python tests/golden/make_golden.py
It loads oracle/_ref/libdeltaref.so -- the reference's basis.cpp, convolve.cpp
and catalog.cpp compiled where they lie -- and stores inputs and outputs as
small .npz files next to this script.  The GPU tests read only the .npz files.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def bn_fixture():
    rng = np.random.default_rng(20240611)
    cases = []
    for (h, w, beta, n_max, radius) in ((41, 53, 1.3, 4, 7), (50, 60, 1.7, 3, 6),
                                        (40, 52, 1.0, 6, 15), (33, 35, 2.2, 2, 10)):
        img = (1000 + 50 * rng.standard_normal((h, w))).astype(np.float32)
        img[rng.integers(0, h), :] += 4e4  # a bright streak
        bn = oracle.ref_basis_convolve(img, beta, n_max, radius)
        b1d = np.zeros((n_max + 1) * (2 * radius + 1))
        oracle.ref().ref_basis1d(beta, n_max, radius, b1d)
        cases.append(dict(img=img, beta=beta, n_max=n_max, radius=radius, bn=bn,
                          b1d=b1d.reshape(n_max + 1, -1)))
    out = {}
    for i, c in enumerate(cases):
        for k, v in c.items():
            out[f"c{i}_{k}"] = np.asarray(v)
    out["n_cases"] = np.asarray(len(cases))
    np.savez_compressed(os.path.join(HERE, "basis_convolve_ref.npz"), **out)


def main():
    if oracle.ref() is None:
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    bn_fixture()
    if hasattr(oracle, "ccl_fixture"):
        oracle.ccl_fixture(HERE)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
