"""Edge shapes of both samplers against the oracle: one walker, walker counts that do not fill a
workgroup, paths shorter than a wavefront (down to N = 2), the deepest bisection a path allows, and a
cluster with 18 coordinates per slice.  Same checks as the parity tests: bit-identical streams,
acceptance counts, paths and operator means."""
import numpy as np
import pytest

from oracle import oracle as O
from feynsimul_amd.kernel_args import KernelArgs
from feynsimul_amd.kernel import PIMCKernel
from feynsimul_amd.physical_systems import HarmOsc, Lm2m2Cluster, AnharmOsc

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case] " + msg, flush=True)


def _system(name):
    return {"ho": HarmOsc(), "aho": AnharmOsc(0.1), "lm2m2-6": Lm2m2Cluster(6)}[name]


def _paths(system, W, N, rng):
    if isinstance(system, Lm2m2Cluster):
        pos = system.initialPositions(W, rng, sigma=0.3)
        p = np.repeat(pos[:, :, None], N, axis=2) + rng.normal(0, 0.02, (W, system.DOF, N))
        return p.reshape(W, system.DOF * N)
    return rng.uniform(-1.0, 1.0, (W, system.DOF * N))


def _ka(system, W, N, S, beta, sampler, parallel=True, opRuns=2, steps=3):
    return KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=S, enableOperator=True,
                      enableCorrelator=False, enableBisection=True, enablePathShift=False,
                      enableSingleNodeMove=False, enableParallelizePath=parallel, enableGlobalPath=False,
                      enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=steps,
                      operatorRuns=opRuns, nbrOfWalkersPerWorkGroup=1, operators=(system.energyOp,), seed=11,
                      sampler=sampler, fastMath=False)


WAVE = [
    # name, W, N, S, beta
    ("ho", 1, 64, 3, 5.0),
    ("ho", 5, 8, 1, 2.0),
    ("ho", 3, 2, 1, 1.0),
    ("ho", 7, 64, 6, 5.0),
    ("aho", 9, 32, 5, 4.0),
    ("lm2m2-6", 3, 64, 6, 10.0),
]


@pytest.mark.parametrize("case", WAVE, ids=lambda c: "%s-W%d-N%d-S%d" % c[:4])
def test_wave_edges_match_oracle(gpu, case):
    name, W, N, S, beta = case
    _case("wave edge " + str(case))
    system = _system(name)
    p0 = _paths(system, W, N, np.random.default_rng(31))
    k = PIMCKernel(_ka(system, W, N, S, beta, "wave"))
    k.setPaths(p0)
    k.run()
    lib = O.build(system, N, beta, operators=(system.energyOp,))
    cfg = O.make_cfg(N, S, system.DOF, W, beta, 2, 3, nOps=1)
    paths = np.ascontiguousarray(p0.astype(np.float32))
    lane = O.xorshift_rows(11, 0, W * 64)
    grp = O.xorshift_rows(11, W * 64, W)
    ref = O.run_wave(lib, cfg, paths, lane_seeds=lane, group_seeds=grp)
    dev = k._seeds.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(dev[: W * 64], lane)
    np.testing.assert_array_equal(dev[W * 64:], grp)
    acc = k._accepts.cpu().numpy().view(np.uint32)
    same = acc == ref["accepts"]
    assert same.sum() >= max(1, int(0.8 * W)), "acceptance counts differ on %d of %d" % ((~same).sum(), W)
    tol = 2e-3 if name.startswith("lm2m2") else 2e-4
    err = np.abs(k.getPaths()[same] - paths[same]).max()
    assert err < tol, "path error %g" % err
    ops = k.getOperators()[0]
    rel = np.abs(ops[same] - ref["opMeans"][same]) / (np.abs(ref["opMeans"][same]) + 1e-3)
    assert rel.max() < 1e-3, "operator mean error %g" % rel.max()


EXACT = [
    # name, W, N, S, beta, parallel
    ("ho", 1, 4, 2, 2.0, True),
    ("ho", 3, 8, 3, 2.0, True),
    ("ho", 5, 16, 2, 3.0, False),
]


@pytest.mark.parametrize("case", EXACT, ids=lambda c: "%s-W%d-N%d-S%d-par%d" % (c[:4] + (c[5],)))
def test_exact_edges_match_oracle(gpu, case):
    name, W, N, S, beta, par = case
    _case("exact edge " + str(case))
    system = _system(name)
    p0 = _paths(system, W, N, np.random.default_rng(37))
    k = PIMCKernel(_ka(system, W, N, S, beta, "exact", parallel=par))
    k.setPaths(p0)
    k.run()
    lib = O.build(system, N, beta, operators=(system.energyOp,))
    cfg = O.make_cfg(N, S, system.DOF, W, beta, 2, 3, parallelizePath=par, nOps=1)
    paths = np.ascontiguousarray(p0.astype(np.float32))
    seeds = O.xorshift_rows(11, 0, cfg.nThreads + 1)
    ref = O.run_reference(lib, cfg, paths, seeds=seeds)
    np.testing.assert_array_equal(k._seeds.cpu().numpy().view(np.uint32), seeds)
    acc = k._accepts.cpu().numpy().view(np.uint32)
    assert (acc == ref["accepts"]).mean() >= 0.8
    T = cfg.nThreads // W
    same_w = (acc == ref["accepts"]).reshape(W, T).all(axis=1)
    assert same_w.any()
    err = np.abs(k.getPaths()[same_w] - paths[same_w]).max()
    assert err < 2e-4, "path error %g" % err


@pytest.mark.parametrize("sampler", ["exact", "wave"])
def test_limits_and_verbose_dump(gpu, sampler, tmp_path, monkeypatch):
    """getMaxKernelWGSize reads the compiled kernel's work-group limit, exceedsLimits checks the launch
    against it and the device (kernel.py:732-775), and verbose=True writes the generated source and its
    preprocessed form (kernel.py:351-364)."""
    _case("limits and verbose dump, " + sampler)
    monkeypatch.chdir(tmp_path)
    system = HarmOsc()
    ka = _ka(system, 8, 1024, 1, 10.0, sampler)   # exact: 512 work-items per work-group
    k = PIMCKernel(ka, verbose=True)
    info = k._launchInfo()
    assert k.getMaxKernelWGSize() == info.max_threads_per_block >= info.threads_per_block
    assert info.threads_per_block == (512 if sampler == "exact" else 64 * (info.threads_per_block // 64))
    assert 0 < info.lds_limit and info.lds_bytes <= info.lds_limit
    assert k.exceedsLimits() is False
    assert (tmp_path / "cleaned.hip").read_text() == k._engine.source
    pre = (tmp_path / "preprocessed.hip").read_text()
    assert "potential" in pre and "#define" not in pre


def test_calls_under_another_current_stream(gpu):
    """The kernel's and the learned proposal's torch work on their device buffers runs on the engine's
    stream whatever stream the caller has made current: the same calls made under a second current stream
    give the same paths, operator means and proposal acceptance bit for bit."""
    import torch
    from feynsimul_amd.learned import LearnedProposal
    _case("calls under another current stream")
    system, W, N = AnharmOsc(0.1), 256, 128
    p0 = _paths(system, W, N, np.random.default_rng(5))
    out = []
    for other in (False, True):
        k = PIMCKernel(_ka(system, W, N, 4, 4.0, "wave", opRuns=3, steps=8))
        lp = LearnedProposal(k, batch=2048, seed=3)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s) if other else torch.cuda.stream(torch.cuda.current_stream()):
            k.setPaths(p0)
            k.run()
            lp.train(2)
            ar = lp.sweep(2)
            lp.run(2, 1)
            out.append((k.getPaths(), k.getOperators(), ar, lp.lastAcceptance))
        del lp, k
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2:] == out[1][2:]


def test_potential_cache_kept_across_launches(gpu):
    """Memory storage keeps the potential of every slice in HBM; consecutive launches on unchanged paths reuse
    it (FS_RUN_POT_CACHE_VALID) instead of recomputing it at the start of each launch.  The results equal
    those of launches that recompute it, bit for bit, and setPaths invalidates it (new paths are not scored
    with the old paths' potentials)."""
    _case("potential cache across launches")
    system = Lm2m2Cluster(3)
    W, N = 64, 128
    rng = np.random.default_rng(8)
    p0, p1 = _paths(system, W, N, rng), _paths(system, W, N, rng)
    out = []
    for keep in (True, False):
        k = PIMCKernel(KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=5.0, S=6, enableOperator=True,
                                  enableCorrelator=False, enableBisection=True, enablePathShift=False,
                                  enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                                  enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=4,
                                  operatorRuns=2, nbrOfWalkersPerWorkGroup=1, operators=(system.energyOp,), seed=21,
                                  sampler="wave", pathStorage="memory", potentialCache=True))
        assert k._potCache is not None
        k.setPaths(p0)
        got = []
        for paths in (None, None, None, p1, None):
            if paths is not None:
                k.setPaths(paths)
            if not keep:
                k._potValid = False
            k.run()
            got.append((k.getPaths(), k.getOperators()))
        out.append(got)
        del k
    for (pa, oa), (pb, ob) in zip(*out):
        assert np.array_equal(pa, pb)
        assert np.array_equal(oa, ob)
