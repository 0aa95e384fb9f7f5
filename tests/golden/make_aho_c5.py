"""Writes aho_c5.json: FP64 transfer-matrix values of the discretised primitive action
for the anharmonic oscillator V = x^2/2 + 0.1 x^4 at N = 4096 slices, beta = 40
(eps = 0.009765625): the virial energy <x^2 + 0.3 x^4>, <x^2>, the two-time
correlator <x_0 x_tau> for tau = 0..512 and the gap E1 - E0 of the transfer matrix.

Grid: [-7, 7] with 1400 points (h = 0.01, a tenth of the kernel width sqrt(eps));
the same quantities on 1000 points are stored too, as a convergence check.
Run:  python tests/golden/make_aho_c5.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

LAM, N, BETA, LAGS = 0.1, 4096, 40.0, 512


def V(x):
    return 0.5 * x * x + LAM * x ** 4


def values(M):
    x, p = O.transfer_matrix_density(V, N, BETA, L=7.0, M=M)
    x2, _, cor = O.transfer_matrix_1d(V, N, BETA, L=7.0, M=M)
    return {"E_virial": float(np.sum(p * (x * x + 3 * LAM * x ** 4))), "x2": x2,
            "correlator": [float(c) for c in cor[: LAGS + 1]],
            "gap": O.transfer_matrix_gap(V, N, BETA, L=7.0, M=M)}


def main():
    out = {"system": "V = x^2/2 + 0.1 x^4, m = hbar = 1", "N": N, "beta": BETA, "eps": BETA / N,
           "grid": {"L": 7.0, "M": 1400}, "values": values(1400), "values_M1000": values(1000)}
    with open(os.path.join(HERE, "aho_c5.json"), "w") as f:
        json.dump(out, f, indent=1)
    v = out["values"]
    print("E_vir %.12f  <x^2> %.12f  gap %.8f" % (v["E_virial"], v["x2"], v["gap"]))


if __name__ == "__main__":
    main()
