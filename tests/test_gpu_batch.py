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


def _same_state(k1, k2, what):
    """Bit equality of two kernels' streams, acceptance counts, paths and operator means."""
    for name in ("_seeds", "_rl", "_grl", "_accepts"):
        a, b = getattr(k1, name), getattr(k2, name)
        if a is not None:
            np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy(), err_msg="%s: %s" % (what, name))
    np.testing.assert_array_equal(k1.getPaths(), k2.getPaths(), err_msg=what + ": paths")
    np.testing.assert_array_equal(k1.getOperators(), k2.getOperators(), err_msg=what + ": operators")


@pytest.mark.parametrize("order", ["cost", "given"])
def test_batch_matches_separate_kernels(gpu, order):
    """A member of a batch runs its type's specialisation of the same device code as a kernel of its
    own, on its own buffers: the results are bit-identical."""
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
        _same_state(batch.kernels[i], single, "member %d" % i)


def test_graph_replay_matches_plain_runs(gpu):
    """Replays of a captured graph equal the same number of plain launches bit for bit, with the
    program's own eager work between replays."""
    _case("batch graph replay")
    rng = np.random.default_rng(1)
    starts = [_paths(s, W, N, rng) for s, W, N, S, b in MEMBERS[:2]]
    kas = [_ka(s, W, N, S, b, 1, 2, seed=7 + i) for i, (s, W, N, S, b) in enumerate(MEMBERS[:2])]
    plain = BatchedPIMC(kas)
    graphed = BatchedPIMC(kas)
    for k1, k2, p in zip(plain.kernels, graphed.kernels, starts):
        k1.setPaths(p)
        k2.setPaths(p)
        k1.setSeriesCapacity(3)
        k2.setSeriesCapacity(3)
    g = graphed.captureGraph(3)
    for rep in range(2):
        plain.resetSeries()
        for _ in range(3):
            plain.run()
        graphed.replay(g)
        for i, (k1, k2) in enumerate(zip(plain.kernels, graphed.kernels)):
            assert k2.getSeries().shape[0] == 3
            np.testing.assert_array_equal(k1.getSeries(), k2.getSeries())
            _same_state(k1, k2, "replay %d member %d" % (rep, i))


def test_stale_graph_is_refused_and_recapture_works(gpu):
    """A graph keeps the dispatch tables of its capture: after setSeriesCapacity (a new series buffer)
    or setOrder, replay refuses it; a new capture replays correctly, and a plain run with another
    order in between does not change what the graph dispatches."""
    _case("stale graph")
    rng = np.random.default_rng(3)
    starts = [_paths(s, W, N, rng) for s, W, N, S, b in MEMBERS[:3]]
    kas = [_ka(s, W, N, S, b, 1, 2, seed=30 + i) for i, (s, W, N, S, b) in enumerate(MEMBERS[:3])]
    plain = BatchedPIMC(kas)
    graphed = BatchedPIMC(kas)
    for k1, k2, p in zip(plain.kernels, graphed.kernels, starts):
        k1.setPaths(p)
        k2.setPaths(p)
        k2.setSeriesCapacity(2)
    g = graphed.captureGraph(1)
    graphed.kernels[1].setSeriesCapacity(4)
    with pytest.raises(Exception, match="capture it again"):
        graphed.replay(g)
    g = graphed.captureGraph(1)
    order0 = list(graphed.order)
    graphed.setOrder(order0[::-1])
    with pytest.raises(Exception, match="capture it again"):
        graphed.replay(g)
    # a plain run with the reversed order uploads other tables; the graph keeps its own
    graphed.run()
    plain.run()
    graphed.setOrder(order0)
    graphed.replay(g)
    plain.run()
    for i, (k1, k2) in enumerate(zip(plain.kernels, graphed.kernels)):
        _same_state(k1, k2, "member %d" % i)


def _oracle_member(k, p0, seed):
    """One launch of a member on the CPU oracle (wave ordering) from the member's initial paths and
    the same seeds; returns the oracle's results and final paths."""
    from oracle import oracle as O
    from feynsimul_amd.kernel import _GROUP_OFFSET
    s = k._system
    W, N = k._nbrOfWalkers, k._N
    lib = O.build(s, N, k._beta, operators=k._operators)
    cfg = O.make_cfg(N, k._S, s.DOF, W, k._beta, int(k._operatorRuns), int(k._metroStepsPerOperatorRun),
                     ranlux=bool(k._enableRanlux), nskip=O.nskip_of_lux(int(k._luxuaryFactor or 0)),
                     nOps=len(k._operators))
    paths = np.ascontiguousarray(p0.astype(np.float32))
    if k._enableRanlux:
        lst = O.ranlux_init(lib, O.ranlux_seeds(seed, 0, W * 64))
        gst = O.ranlux_init(lib, O.ranlux_seeds(seed, _GROUP_OFFSET, W))
        ref = O.run_wave(lib, cfg, paths, lane_rl=lst, group_rl=gst)
        np.testing.assert_array_equal(k._rl.cpu().numpy().view(np.uint8).reshape(W * 64, 112),
                                      lst.view(np.uint8).reshape(W * 64, 112))
        np.testing.assert_array_equal(k._grl.cpu().numpy().view(np.uint8).reshape(W, 112),
                                      gst.view(np.uint8).reshape(W, 112))
    else:
        lane = O.xorshift_rows(seed, 0, W * 64)
        grp = O.xorshift_rows(seed, W * 64, W)
        ref = O.run_wave(lib, cfg, paths, lane_seeds=lane, group_seeds=grp)
        dev = k._seeds.cpu().numpy().view(np.uint32)
        np.testing.assert_array_equal(dev[: W * 64], lane)
        np.testing.assert_array_equal(dev[W * 64:], grp)
    return ref, paths


# LM2M2 clusters of 3, 4, 6 and 8 atoms (DOF 9 to 24) beside 1D oscillators, ragged N and S
ORACLE_MEMBERS = [(HarmOsc(), 32, 256, 4, 10.0), (Lm2m2Cluster(3), 16, 256, 6, 10.0),
                  (AnharmOsc(0.1), 16, 1024, 5, 20.0), (Lm2m2Cluster(4), 16, 512, 5, 10.0),
                  (Lm2m2Cluster(6), 16, 128, 4, 10.0), (Lm2m2Cluster(8), 16, 256, 6, 10.0)]


@pytest.mark.parametrize("ranlux", [False, True], ids=["xorshift", "ranlux-lux2"])
def test_batch_members_match_oracle(gpu, ranlux):
    """Every member of one batched launch against the CPU oracle's restatement of the wave sampler
    on the same initial paths and seeds: stream states bit-identical, acceptance counts on at least
    90 % of walkers (a Metropolis decision whose margin is below the rounding difference of gcc and
    hiprtc may flip), paths to 2e-4 (1D) / 2e-3 (LM2M2) and operator means to 1e-3 relative on the
    walkers whose decisions agree."""
    _case("batch members vs oracle, ranlux=%d" % ranlux)
    rng = np.random.default_rng(21)
    members = ORACLE_MEMBERS if not ranlux else ORACLE_MEMBERS[1:4]
    starts = [_paths(s, W, N, rng) for s, W, N, S, b in members]
    kas = []
    for i, (s, W, N, S, b) in enumerate(members):
        ka = _ka(s, W, N, S, b, 2, 2, seed=500 + i)
        ka.fastMath = False
        if ranlux:
            ka.enableRanlux, ka.luxuaryFactor, ka.ranluxIntMax = True, 1, 2 ** 31 - 1
        kas.append(ka)
    batch = BatchedPIMC(kas)
    for k, p in zip(batch.kernels, starts):
        if ranlux:
            k.initRanlux()
            k.groupInitRanlux()
        k.setPaths(p)
    batch.run()
    for i, (k, p0) in enumerate(zip(batch.kernels, starts)):
        ref, paths = _oracle_member(k, p0, 500 + i)
        W = k._nbrOfWalkers
        acc = k._accepts.cpu().numpy().view(np.uint32)
        same = acc == ref["accepts"]
        assert same.mean() >= 0.9, "member %d: acceptance differs on %d of %d walkers" % (i, (~same).sum(), W)
        tol = 2e-3 if isinstance(k._system, Lm2m2Cluster) else 2e-4
        err = np.abs(k.getPaths()[same] - paths[same]).max()
        assert err < tol, "member %d: path error %g" % (i, err)
        ops = k.getOperators()[0]
        rel = np.abs(ops[same] - ref["opMeans"][same]) / (np.abs(ref["opMeans"][same]) + 1e-3)
        assert rel.max() < 1e-3, "member %d: operator mean error %g" % (i, rel.max())


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
    with and without it; autotune_kernel times every (waves per workgroup, cap) pair and the reference-order
    sampler at one work-item per segment and per chain (the lanes-per-chain choice)."""
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
                                   wavesPerSimd=(0, 4), launches=1, use_cache=False, lanesPerChain=(64, "segments", 1))
    assert len(trials) == 6 and best in trials
    assert all(t["ms_per_launch"] > 0 for t in trials)
    # lanes per chain: a wavefront (the wave sampler), one work-item per segment (N / 2^S = 64), one per chain
    assert sorted({t["lanesPerChain"] for t in trials}) == [1, 64]
    assert [t["sampler"] for t in trials].count("exact") == 2
    print("lanes per chain: %s" % [(t["lanesPerChain"], t["sampler"], round(t["ms_per_launch"], 3)) for t in trials])


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
