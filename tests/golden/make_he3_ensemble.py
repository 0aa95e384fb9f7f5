"""Writes he3_ensemble_fp64.json: the FP64 CPU reference of the He3 workload.

Three He atoms (LM2M2, Cartesian DOF 9, mass 4.0026 u), N = 1024 slices,
beta = 10 1/K, bisection level S = 6, plain Metropolis bisection in the
reference's work-item order (oracle, FP64).  The trimer is unbound in free
space at this temperature (it slowly dissociates), so there is no stationary
long-run value to quote; what is fixed is the ensemble mean after T sweeps
from a given initial ensemble.  The initial ensemble is the one the benchmark
uses: Lm2m2Cluster(3).initialPositions(walkers, numpy.random.default_rng(0),
side=3.5, sigma=3.0), each path constant in imaginary time; this file holds
all 16384 walkers.  For each T in SWEEPS the virial energy and the mean
pair distance (slice means) are averaged over walkers, with the standard
error of that mean; for T in PAIRED the per-walker values are kept as well
(he3_ensemble_fp64_walkers.npz, float32), so a run from the same initial
paths can be compared walker by walker.

Run:  python tests/golden/make_he3_ensemble.py     (about 16 minutes on 8 cores)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402
from feynsimul_amd.physical_systems import Lm2m2Cluster  # noqa: E402

N, S, BETA, WALKERS, SEED = 1024, 6, 10.0, 16384, 0
SWEEPS = (32, 128, 256, 512, 736)
PAIRED = (128, 736)  # sweeps whose per-walker values are kept


def main():
    system = Lm2m2Cluster(3)
    lib = O.build(system, N, BETA, operators=(system.energyOp, system.meanPairDistanceOp), double=True,
                  native=True)
    pos = system.initialPositions(WALKERS, np.random.default_rng(SEED), side=3.5, sigma=3.0)
    paths = np.ascontiguousarray(np.repeat(pos[:, :, None], N, axis=2).reshape(WALKERS, 9 * N))
    seeds = None
    rows, done, keep = [], 0, {}
    t0 = time.time()
    for T in SWEEPS:
        cfg = O.make_cfg(N, S, 9, WALKERS, BETA, 1, T - done, nOps=2)
        if seeds is None:
            seeds = O.xorshift_rows(20261203, 0, cfg.nThreads + 1)
        out = O.run_reference(lib, cfg, paths, seeds=seeds)
        done = T
        per = out["opMeans"].reshape(WALKERS, -1, 2).mean(axis=1)  # threads of a walker -> walker
        rows.append({"sweeps": T, "E_K": float(per[:, 0].mean()),
                     "E_se_K": float(per[:, 0].std(ddof=1) / np.sqrt(WALKERS)),
                     "pair_distance_A": float(per[:, 1].mean()),
                     "pair_distance_se_A": float(per[:, 1].std(ddof=1) / np.sqrt(WALKERS))})
        if T in PAIRED:
            keep["E_K_T%d" % T] = per[:, 0].astype(np.float32)
            keep["pair_distance_A_T%d" % T] = per[:, 1].astype(np.float32)
        print(rows[-1], "%.0f s" % (time.time() - t0), flush=True)
    out = {"system": "LM2M2 He3, Cartesian, mass 4.0026 u", "N": N, "S": S, "beta_per_K": BETA,
           "walkers": WALKERS, "initial": "Lm2m2Cluster(3).initialPositions(walkers, default_rng(0), side=3.5, "
                                          "sigma=3.0), constant paths", "precision": "FP64",
           "order": "reference work-item order (ParallelizePath, one walker per work-group)", "rows": rows}
    out["paired_sweeps"] = list(PAIRED)
    with open(os.path.join(HERE, "he3_ensemble_fp64.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "he3_ensemble_fp64_walkers.npz"), **keep)


if __name__ == "__main__":
    main()
