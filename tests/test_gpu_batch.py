"""Batched persistent launches (several systems in one kernel) on the GPU.

A member of a batch runs the same specialisation of the wave sampler as a
kernel of its own, on its own buffers and streams, so one batched launch must
reproduce the separate launches: identical stream states, the same
acceptance counts and the same paths to rounding.
"""
import numpy as np
import pytest

from feynsimul_amd.kernel_args import KernelArgs
from feynsimul_amd.kernel import PIMCKernel
from feynsimul_amd.batch import BatchedPIMC
from feynsimul_amd.physical_systems import HarmOsc, Lm2m2Cluster, AnharmOsc

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case] " + msg, flush=True)


def _ka(system, W, N, S, beta, opRuns, steps, seed, storage=None):
    return KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=S, enableOperator=True,
                      enableCorrelator=False, enableBisection=True, enablePathShift=False,
                      enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                      enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=steps,
                      operatorRuns=opRuns, nbrOfWalkersPerWorkGroup=1, operators=(system.energyOp,),
                      seed=seed, sampler="wave", pathStorage=storage)


def _paths(system, W, N, rng):
    if isinstance(system, Lm2m2Cluster):
        pos = system.initialPositions(W, rng, sigma=1.0)
        return np.repeat(pos[:, :, None], N, axis=2).reshape(W, system.DOF * N)
    return rng.uniform(-1, 1, (W, system.DOF * N))


MEMBERS = [(HarmOsc(), 40, 256, 4, 10.0), (Lm2m2Cluster(3), 24, 256, 6, 10.0),
           (AnharmOsc(0.1), 16, 2048, 5, 20.0), (Lm2m2Cluster(4), 8, 512, 5, 10.0)]


@pytest.mark.parametrize("order", ["cost", "given"])
def test_batch_matches_separate_kernels(gpu, order):
    _case("batch vs separate, order " + order)
    rng = np.random.default_rng(0)
    starts = [_paths(s, W, N, rng) for s, W, N, S, b in MEMBERS]
    kas = [_ka(s, W, N, S, b, 2, 3, seed=100 + i) for i, (s, W, N, S, b) in enumerate(MEMBERS)]
    batch = BatchedPIMC(kas, order=order)
    for k, p in zip(batch.kernels, starts):
        k.setPaths(p)
    batch.run()
    for i, (s, W, N, S, b) in enumerate(MEMBERS):
        single = PIMCKernel(_ka(s, W, N, S, b, 2, 3, seed=100 + i))
        single.setPaths(starts[i])
        single.run()
        kb = batch.kernels[i]
        np.testing.assert_array_equal(kb._seeds.cpu().numpy(), single._seeds.cpu().numpy())
        acc_b = kb._accepts.cpu().numpy()
        acc_s = single._accepts.cpu().numpy()
        same = acc_b == acc_s
        assert same.mean() >= 0.9, "member %d: acceptance differs on %d walkers" % (i, (~same).sum())
        err = np.abs(kb.getPaths()[same] - single.getPaths()[same]).max()
        assert err < 1e-3, "member %d path error %g" % (i, err)
        np.testing.assert_allclose(kb.getOperators()[0][same], single.getOperators()[0][same], rtol=1e-4, atol=1e-6)


def test_graph_replay_matches_plain_runs(gpu):
    _case("batch graph replay")
    rng = np.random.default_rng(1)
    starts = [_paths(s, W, N, rng) for s, W, N, S, b in MEMBERS[:2]]
    kas = [_ka(s, W, N, S, b, 1, 2, seed=7 + i) for i, (s, W, N, S, b) in enumerate(MEMBERS[:2])]
    plain = BatchedPIMC(kas)
    graphed = BatchedPIMC(kas)
    for k1, k2, p in zip(plain.kernels, graphed.kernels, starts):
        k1.setPaths(p)
        k2.setPaths(p)
        k2.setSeriesCapacity(3)
    g = graphed.captureGraph(3)
    for rep in range(2):
        for _ in range(3):
            plain.run()
        graphed.replay(g)
        # the program's own eager work between replays
        for k in graphed.kernels:
            assert k.getSeries().shape[0] == 3
            assert np.isfinite(k.getSeries()).all()
        for k1, k2 in zip(plain.kernels, graphed.kernels):
            np.testing.assert_array_equal(k1._seeds.cpu().numpy(), k2._seeds.cpu().numpy())
            same = k1._accepts.cpu().numpy() == k2._accepts.cpu().numpy()
            assert same.mean() >= 0.9
            assert np.abs(k1.getPaths()[same] - k2.getPaths()[same]).max() < 1e-3


def test_ragged_lm2m2_batch_full_size(gpu):
    """The ragged workload at full size: LM2M2 clusters n in {3,4,6,8} x N in {256,512,1024,2048},
    2048 chains each, one persistent launch.  Size-independent checks: every walker ran (streams
    advanced), acceptance is a probability, energies are finite, and the batched launch covers all
    32768 chains."""
    import torch
    _case("ragged LM2M2 batch 16 x 2048")
    rng = np.random.default_rng(2)
    kas, starts = [], []
    for n in (3, 4, 6, 8):
        sysobj = Lm2m2Cluster(n)
        for N in (256, 512, 1024, 2048):
            kas.append(_ka(sysobj, 2048, N, 6, 10.0, 1, 2, seed=1000 + n * 10000 + N))
            starts.append(_paths(sysobj, 2048, N, rng).astype(np.float32))
    batch = BatchedPIMC(kas)
    seeds0 = [k._seeds.clone() for k in batch.kernels]
    for k, p in zip(batch.kernels, starts):
        k.setPaths(p)
    batch.run()
    upd = batch.sliceUpdatesPerLaunch()
    print("ragged batch: %.3f ms for %d slice-updates -> %.3g /s" % (batch.getRunTime() * 1e3, upd,
                                                                     upd / batch.getRunTime()))
    for k, s0 in zip(batch.kernels, seeds0):
        s1 = k._seeds
        # every lane stream of every walker advanced
        lanes_changed = (s1[: k._nbrOfWalkers * 64] != s0[: k._nbrOfWalkers * 64]).any(dim=1)
        assert bool(lanes_changed.all())
        ar = k.getAcceptanceRate()
        assert 0.0 < ar < 1.0
        assert np.isfinite(k.getOperators()).all()
    torch.cuda.empty_cache()


def test_occupancy_cap_keeps_results_and_autotune_kernel(gpu):
    """The occupancy cap only pads the launch's LDS request: paths, operators and streams are bit-identical
    with and without it; autotune_kernel times every (waves per workgroup, cap) pair."""
    from feynsimul_amd.autotune import autotune_kernel
    _case("occupancy cap + autotune_kernel")
    system = AnharmOsc(0.1)
    W, N, S = 512, 1024, 4
    p0 = np.random.default_rng(12).uniform(-1, 1, (W, N))
    out = []
    for cap in (0, 3):
        k = PIMCKernel(_ka(system, W, N, S, 20.0, 2, 8, seed=5))
        k.setPaths(p0)
        k.setWavesPerSimd(cap)
        k.run()
        out.append((k.getPaths(), k.getOperators(), k._seeds.cpu().numpy()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])
    best, trials = autotune_kernel(_ka(system, W, N, S, 20.0, 1, 8, seed=5), initialPaths=p0, wavesPerBlock=(2, 4),
                                   wavesPerSimd=(0, 4), launches=1, use_cache=False)
    assert len(trials) == 4 and best in trials
    assert all(t["ms_per_launch"] > 0 for t in trials)


def test_autotune_levels_follows_the_acceptance_window(gpu):
    """autotune_levels measures every feasible S per member and keeps the largest whose acceptance is at
    least 20 % (else the best): a stiff oscillator gets a smaller S than a soft one."""
    from feynsimul_amd.autotune import autotune_levels
    _case("autotune_levels")
    kas = [_ka(HarmOsc(), 64, 256, 6, 10.0, 1, 8, seed=1), _ka(HarmOsc(k=400.0), 64, 256, 6, 10.0, 1, 8, seed=2),
           _ka(Lm2m2Cluster(3), 32, 128, 5, 10.0, 1, 8, seed=3)]
    rng = np.random.default_rng(4)
    starts = [_paths(ka.system, ka.nbrOfWalkers, ka.N, rng) for ka in kas]
    levels, table = autotune_levels(kas, initialPaths=starts, levels=(2, 3, 4, 5, 6), warmupSweeps=16, sweeps=8)
    print(levels, table)
    assert sorted(table[2]) == [2, 3, 4, 5, 6]  # N = 128 allows S <= 6 (64 lanes)
    for t, S in enumerate(levels):
        assert S in table[t]
        ok = [s for s, a in table[t].items() if a >= 0.2]
        assert S == (max(ok) if ok else max(table[t], key=table[t].get))
    assert levels[1] < levels[0]
