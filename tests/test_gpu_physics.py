"""Estimators of the GPU samplers against FP64 closed forms.

Harmonic oscillator, primitive action: <x^2>_N = E_vir = (1/N) sum_k 1/a_k with
a_k = (2 - 2 cos(2 pi k / N)) / eps + eps; the correlator and the density are
fixed by the same modes (oracle.ho_discrete_*).  Walkers are independent
chains, so the standard error is the spread of per-walker means / sqrt(W).
"""
import numpy as np
import pytest

from oracle import oracle as O
from feynsimul_amd.kernel_args import KernelArgs
from feynsimul_amd.kernel import PIMCKernel
from feynsimul_amd.physical_systems import HarmOsc, AnharmOsc

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case] " + msg, flush=True)


def _ka(system, W, N, S, beta, opRuns, steps, sampler, Wg=1, corr=False, bins=False, seed=2024, lags=None,
        ranlux=False, ops=None):
    return KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=S, enableOperator=True,
                      enableCorrelator=corr, enableBisection=True, enablePathShift=False,
                      enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                      enableGlobalOldPath=False, enableBins=bins, enableDouble=False,
                      enableRanlux=ranlux, luxuaryFactor=1 if ranlux else None,
                      ranluxIntMax=2 ** 31 - 1 if ranlux else None,
                      metroStepsPerOperatorRun=steps, operatorRuns=opRuns, nbrOfWalkersPerWorkGroup=Wg,
                      operators=ops or (system.energyOp,), correlators=("x1",) if corr else None,
                      xMin=-3.5, xMax=3.5, binResolutionPerDOF=70, seed=seed, sampler=sampler,
                      correlatorLags=lags)


def test_ho_wave_energy_c1(gpu):
    """1D HO, N=256, beta=10, 4096 walkers, S=4: 2000 burn-in sweeps, measure every 32 sweeps."""
    _case("HO wave C1 energy")
    system = HarmOsc()
    W, N, S, beta = 4096, 256, 4, 10.0
    rng = np.random.default_rng(0)
    burn = PIMCKernel(_ka(system, W, N, S, beta, 1, 2000, "wave"))
    burn.setPaths(rng.uniform(-1, 1, (W, N)))
    burn.run()
    k = PIMCKernel(_ka(system, W, N, S, beta, 100, 32, "wave", seed=99))
    k.setPaths(burn.getPaths())
    k.run()
    per = k.getOperators()[0]
    E, se = per.mean(), per.std() / np.sqrt(W)
    exact = O.ho_discrete_x2(N, beta)
    assert abs(exact - 0.499950082044) < 1e-10
    print("E = %.6f +- %.6f, exact %.9f, AR %.3f" % (E, se, exact, k.getAcceptanceRate()))
    assert abs(E - exact) < 4 * se
    assert se < 2e-3


def test_ho_exact_tutorial_run(gpu):
    """The tutorial run (doc/tutorial/index.rst:31-74): N=128, beta=11, S=6, 896 walkers,
    4 per work-group, 150 x 40 steps; the reference printed 0.499803373124."""
    _case("HO exact tutorial")
    system = HarmOsc()
    W, N, S, beta = 896, 128, 6, 11.0
    k = PIMCKernel(_ka(system, W, N, S, beta, 150, 40, "exact", Wg=4))
    k.run()  # paths start at zero, as in the reference
    k2 = PIMCKernel(_ka(system, W, N, S, beta, 150, 40, "exact", Wg=4, seed=5))
    k2.setPaths(k.getPaths())
    k2.run()
    per = k2.getOperators()[0]
    E, se = per.mean(), per.std() / np.sqrt(W)
    exact = O.ho_discrete_x2(N, beta)
    print("E = %.6f +- %.6f, exact %.9f, AR %.3f" % (E, se, exact, k2.getAcceptanceRate()))
    assert abs(E - exact) < 4 * se
    # the reference's own printed value is a single draw from the same distribution
    assert abs(0.499803373124 - exact) < 6 * se


def test_ho_correlator_and_density(gpu):
    _case("HO wave correlator + bins")
    system = HarmOsc()
    W, N, S, beta = 1024, 128, 4, 8.0
    rng = np.random.default_rng(1)
    burn = PIMCKernel(_ka(system, W, N, S, beta, 1, 1000, "wave"))
    burn.setPaths(rng.uniform(-1, 1, (W, N)))
    burn.run()
    k = PIMCKernel(_ka(system, W, N, S, beta, 50, 16, "wave", corr=True, bins=True, seed=7, lags=32))
    k.setPaths(burn.getPaths())
    k.run()
    mean, se = k.getCorrelator()
    exact = O.ho_discrete_correlator(N, beta)[:32]
    z = np.abs(mean[0] - exact) / se[0]
    print("correlator max z %.2f" % z.max())
    assert z.max() < 5
    dens = k.getBinCounts()
    x2 = O.ho_discrete_x2(N, beta)
    edges = np.linspace(-3.5, 3.5, 71)
    from math import erf, sqrt
    cdf = np.array([0.5 * (1 + erf(e / sqrt(2 * x2))) for e in edges])
    p_exact = np.diff(cdf) / (edges[1] - edges[0])
    assert np.abs(dens - p_exact).max() < 0.02


def test_aho_energy_vs_transfer_matrix(gpu):
    _case("AHO wave energy vs transfer matrix")
    system = AnharmOsc(0.1)
    W, N, S, beta = 2048, 128, 4, 8.0
    rng = np.random.default_rng(2)
    burn = PIMCKernel(_ka(system, W, N, S, beta, 1, 1000, "wave"))
    burn.setPaths(rng.uniform(-1, 1, (W, N)))
    burn.run()
    k = PIMCKernel(_ka(system, W, N, S, beta, 50, 16, "wave", seed=8, ops=(system.positionSquaredOp,)))
    k.setPaths(burn.getPaths())
    k.run()
    per = k.getOperators()[0]
    x2, se = per.mean(), per.std() / np.sqrt(W)
    ref, _, _ = O.transfer_matrix_1d(system.V, N, beta)
    print("<x^2> = %.6f +- %.6f, transfer matrix %.6f" % (x2, se, ref))
    assert abs(x2 - ref) < 4 * se


@pytest.mark.parametrize("ranlux", [False, True])
def test_aho_correlator_and_gap(gpu, ranlux):
    """Anharmonic oscillator x^2/2 + 0.1 x^4: two-time correlator <x(0) x(tau)> in FP64 Kahan bins
    against the FP64 transfer matrix of the discretised action, with xorshift128 and with RANLUX
    (luxury 2), and the gap from the correlator's decay."""
    _case("AHO correlator ranlux=%s" % ranlux)
    system = AnharmOsc(0.1)
    W, N, S, beta, lags = 4096, 512, 4, 20.0, 96
    rng = np.random.default_rng(3)
    burn = PIMCKernel(_ka(system, W, N, S, beta, 1, 600, "wave", ranlux=ranlux))
    if ranlux:
        burn.initRanlux()
        burn.groupInitRanlux()
    burn.setPaths(rng.uniform(-1, 1, (W, N)))
    burn.run()
    k = PIMCKernel(_ka(system, W, N, S, beta, 40, 8, "wave", corr=True, seed=13, lags=lags, ranlux=ranlux,
                       ops=(system.positionSquaredOp,)))
    if ranlux:
        k.initRanlux()
        k.groupInitRanlux()
    k.setPaths(burn.getPaths())
    k.run()
    mean, se = k.getCorrelator()
    x2, _, cor = O.transfer_matrix_1d(system.V, N, beta, L=7.0, M=1000)
    z = np.abs(mean[0] - cor[:lags]) / se[0]
    print("max z %.2f, C(0) %.5f vs %.5f" % (z.max(), mean[0][0], cor[0]))
    assert z.max() < 5.5
    # C(0) is the slice mean of x^2 over the same measurements
    assert abs(mean[0][0] - k.getOperators()[0].mean()) < 1e-5 * cor[0]
    # effective gap -d ln C / d tau over a window where the first odd state dominates
    gap = O.transfer_matrix_gap(system.V, N, beta, L=7.0, M=1000)
    eps = beta / N
    t1, t2 = 30, 60
    g = -np.log(mean[0][t2] / mean[0][t1]) / ((t2 - t1) * eps)
    g_tm = -np.log(cor[t2] / cor[t1]) / ((t2 - t1) * eps)
    print("gap from correlator %.4f, same window of the transfer matrix %.4f, eigen gap %.4f" % (g, g_tm, gap))
    assert abs(g - g_tm) < 0.03 * g_tm
    assert abs(g_tm - gap) < 0.05 * gap


def test_aho_full_size_ranlux_launch(gpu):
    """The correlator/RNG stress workload at full size: 1D AHO, N=4096, beta=40, 65536 chains, RANLUX
    luxury 2, 512 correlator lags.  Checked through size-independent identities: C(0) equals the
    slice mean of x^2, C is bounded by C(0), streams advance, acceptance is a probability."""
    import torch
    _case("AHO full size RANLUX")
    system = AnharmOsc(0.1)
    W, N, S, beta, lags = 65536, 4096, 4, 40.0, 512
    k = PIMCKernel(_ka(system, W, N, S, beta, 2, 2, "wave", corr=True, seed=21, lags=lags, ranlux=True,
                       ops=(system.positionSquaredOp,)))
    k.initRanlux()
    k.groupInitRanlux()
    st0 = k._rl[:1024].clone()
    rng = np.random.default_rng(4)
    k.setPaths(rng.uniform(-1, 1, (W, N)).astype(np.float32))
    k.run()
    mean, se = k.getCorrelator()
    assert np.isfinite(mean).all() and np.isfinite(se).all()
    x2 = k.getOperators()[0]
    assert abs(mean[0][0] - x2.mean()) < 1e-5 * abs(x2.mean())
    assert (np.abs(mean[0]) <= mean[0][0] * (1 + 1e-6)).all()
    assert 0.0 < k.getAcceptanceRate() < 1.0
    assert not torch.equal(st0, k._rl[:1024])
    del k
    torch.cuda.empty_cache()


def test_ho_energy_with_single_node_and_path_shift_moves(gpu):
    """No bisection: single-node (alpha 0.6) and path-shift (PSAlpha 0.3) moves alone, one
    work-item per walker (kernel.c:640-705), must sample <x^2> of the discretised HO."""
    _case("HO single-node + path-shift energy")
    system = HarmOsc()
    W, N, beta = 4096, 32, 4.0

    def ka(opRuns, seed):
        return KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=None, enableOperator=True,
                          enableCorrelator=False, enableBisection=False, enablePathShift=True,
                          enableSingleNodeMove=True, enableParallelizePath=False, enableGlobalPath=False,
                          enableGlobalOldPath=False, enableBins=False, enableDouble=False, enableRanlux=False,
                          metroStepsPerOperatorRun=N, operatorRuns=opRuns, nbrOfWalkersPerWorkGroup=1,
                          operators=(system.energyOp,), alpha=0.6, PSAlpha=0.3, seed=seed)
    burn = PIMCKernel(ka(300, 31))
    burn.setPaths(np.random.default_rng(6).uniform(-1, 1, (W, N)))
    burn.run()
    k = PIMCKernel(ka(300, 32))
    k.setPaths(burn.getPaths())
    k.run()
    per = k.getOperators()[0]
    E, se = per.mean(), per.std() / np.sqrt(W)
    exact = O.ho_discrete_x2(N, beta)
    print("E = %.6f +- %.6f, exact %.9f, AR %.3f" % (E, se, exact, k.getAcceptanceRate() / 2))
    assert abs(E - exact) < 4 * se


def test_jacobi_cluster_compiles_evaluates_and_samples(gpu):
    """Lm2m2_cluster(3, 0.1 K) (Jacobi, DOF 6): the compiled potential and virial operator
    match the FP64 evaluation of the same strings, and the wave sampler runs on it."""
    import math
    from feynsimul_amd.physical_systems import Lm2m2_cluster
    from feynsimul_amd.physical_systems.lm2m2 import lm2m2_kelvin
    from feynsimul_amd.physical_systems.lm2m2_cluster import jacobi_matrix
    _case("Jacobi He3")
    system = Lm2m2_cluster(3, 0.1)
    W, N, S = 256, 256, 5
    k = PIMCKernel(KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=system.beta, S=S, enableOperator=True,
                              enableCorrelator=False, enableBisection=True, enablePathShift=False,
                              enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                              enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=8,
                              operatorRuns=4, nbrOfWalkersPerWorkGroup=1,
                              operators=(system.energyOp, system.meanSquaredRadiusOp), seed=3))
    rng = np.random.default_rng(0)
    J = jacobi_matrix(3, system.m)
    r = rng.normal(0, 3.0, (64, 3, 3))
    x = np.einsum("ij,wjk->wik", J, r)[:, :2, :].reshape(64, 6)
    pot, ops = k.evalSystem(x)
    for w in range(64):
        d = [np.linalg.norm(r[w, a] - r[w, b]) for a, b in ((0, 1), (0, 2), (1, 2))]
        V = sum(lm2m2_kelvin(v) for v in d)
        assert abs(pot[w] - V) < 2e-4 * max(1.0, abs(V))
        assert abs(ops[w, 1] - sum(v * v for v in d)) < 1e-4 * sum(v * v for v in d)
    # straight paths at a compact configuration, then sampling
    base = np.einsum("ij,wjk->wik", J, rng.normal(0, 1.5, (W, 3, 3)))[:, :2, :].reshape(W, 6)
    k.setPaths(np.repeat(base[:, :, None], N, axis=2).reshape(W, 6 * N))
    k.run()
    ar = k.getAcceptanceRate()
    E = k.getOperators()[0].mean()
    print("AR %.3f, E %.4f K" % (ar, E))
    assert 0.05 < ar < 0.99 and np.isfinite(E)


def test_he3_ensemble_vs_fp64_oracle(gpu):
    """He3 (LM2M2, N = 1024, beta = 10/K, S = 6): the wave sampler in FP32 over 16384 chains against
    the FP64 CPU run of the reference-order sampler over the same 16384 initial paths
    (tests/golden/he3_ensemble_fp64.json).  The trimer dissociates slowly at this temperature, so the
    comparison is of the ensemble means after T sweeps, T = 32 .. 736: virial energy and mean pair
    distance, each within 4 combined standard errors, and walker by walker at T = 128 and 736 (both
    sides start from the same 16384 initial paths)."""
    import json
    import os
    from feynsimul_amd.physical_systems import Lm2m2Cluster
    _case("He3 ensemble vs FP64 oracle")
    with open(os.path.join(os.path.dirname(__file__), "golden", "he3_ensemble_fp64.json")) as f:
        g = json.load(f)
    system = Lm2m2Cluster(3)
    W, N, S, beta = 16384, g["N"], g["S"], g["beta_per_K"]
    pos = system.initialPositions(W, np.random.default_rng(0), side=3.5, sigma=3.0)
    assert np.array_equal(pos[:4], system.initialPositions(4, np.random.default_rng(0), side=3.5, sigma=3.0))
    ka = KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=S, enableOperator=True,
                    enableCorrelator=False, enableBisection=True, enablePathShift=False,
                    enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                    enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=32, operatorRuns=1,
                    nbrOfWalkersPerWorkGroup=1, operators=(system.energyOp, system.meanPairDistanceOp),
                    seed=99, sampler="wave")
    k = PIMCKernel(ka)
    k.setPaths(np.repeat(pos[:, :, None], N, axis=2).reshape(W, 9 * N))
    last = g["rows"][-1]["sweeps"]
    k.setSeriesCapacity(last // 32)
    for _ in range(last // 32):
        k.run(wait=False)
    ser = k.getSeries()
    # the same 16384 initial paths on both sides: walker-by-walker differences remove the spread of the
    # initial ensemble from the error
    walkers = np.load(os.path.join(os.path.dirname(__file__), "golden", "he3_ensemble_fp64_walkers.npz"))
    worst = 0.0
    for T in g["paired_sweeps"]:
        v = ser[T // 32 - 1].astype(np.float64)
        for col, key in ((0, "E_K"), (1, "pair_distance_A")):
            d = v[:, col] - walkers["%s_T%d" % (key, T)]
            z = abs(d.mean()) / (d.std(ddof=1) / np.sqrt(W))
            print("T=%d %s paired: mean difference %.5f +- %.5f, z %.2f" % (T, key, d.mean(),
                                                                          d.std(ddof=1) / np.sqrt(W), z))
            worst = max(worst, z)
    for row in g["rows"]:
        v = ser[row["sweeps"] // 32 - 1].astype(np.float64)
        for col, key in ((0, "E_K"), (1, "pair_distance_A")):
            m, se = v[:, col].mean(), v[:, col].std(ddof=1) / np.sqrt(W)
            z = abs(m - row[key]) / np.hypot(se, row[key.replace("_K", "_se_K").replace("_A", "_se_A")])
            print("T=%d %s: GPU %.5f +- %.5f, FP64 oracle %.5f, z %.2f" % (row["sweeps"], key, m, se, row[key], z))
            worst = max(worst, z)
    assert worst < 4.0
