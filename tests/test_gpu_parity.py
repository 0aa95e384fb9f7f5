"""One launch of each GPU sampler vs the CPU oracle on the same inputs.

exact sampler  <-> oracle.run_reference: the reference's work-item geometry,
                   stream rows and draw order (kernel.c:514-837).
wave sampler   <-> oracle.run_wave: the engine's slice-per-lane ordering.

Random-number consumption is fixed per move, so the final stream states must
be bit-identical.  Paths, acceptance counts and operator means must agree to
FP32 tolerance; a Metropolis decision whose margin is below the rounding
difference of the two compilers may flip, so a small fraction of walkers is
allowed to differ and is excluded from the path comparison.
"""
import numpy as np
import pytest

from oracle import oracle as O
from feynsimul_amd.kernel_args import KernelArgs
from feynsimul_amd.kernel import PIMCKernel, _GROUP_OFFSET
from feynsimul_amd.physical_systems import HarmOsc, Lm2m2Cluster, AnharmOsc

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case] " + msg, flush=True)


def _system(name):
    if name == "ho":
        return HarmOsc()
    if name == "aho":
        return AnharmOsc(0.1)
    if name.startswith("lm2m2x"):
        return Lm2m2Cluster(int(name[6:]))
    return Lm2m2Cluster(3)


def _initial_paths(system, W, N, rng):
    if isinstance(system, Lm2m2Cluster):
        pos = system.initialPositions(W, rng, sigma=0.3)
        p = np.repeat(pos[:, :, None], N, axis=2) + rng.normal(0, 0.02, (W, system.DOF, N))
        return p.reshape(W, system.DOF * N)
    return rng.uniform(-1.0, 1.0, (W, system.DOF * N))


def _ka(system, W, N, S, beta, opRuns, steps, parallel, ranlux, double, sampler, Wg=1, seed=11, storage=None):
    return KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=S, enableOperator=True,
                      enableCorrelator=False, enableBisection=True, enablePathShift=False,
                      enableSingleNodeMove=False, enableParallelizePath=parallel, enableGlobalPath=False,
                      enableGlobalOldPath=False, enableBins=False, enableDouble=double,
                      enableRanlux=ranlux, luxuaryFactor=1 if ranlux else None,
                      ranluxIntMax=2 ** 31 - 1 if ranlux else None, metroStepsPerOperatorRun=steps,
                      operatorRuns=opRuns, nbrOfWalkersPerWorkGroup=Wg, operators=(system.energyOp,),
                      seed=seed, sampler=sampler, fastMath=False, pathStorage=storage)


def _compare(gp, rp, acc, racc, walker_of_thread, tol, min_match=0.95):
    same_t = acc == racc
    assert same_t.mean() >= min_match, "acceptance counts differ on %d of %d" % ((~same_t).sum(), len(same_t))
    bad_w = np.unique(walker_of_thread[~same_t])
    ok = np.ones(gp.shape[0], bool)
    ok[bad_w] = False
    err = np.abs(gp[ok] - rp[ok]).max()
    assert err < tol, "path error %g" % err
    return ok


CASES_EXACT = [
    # name, W, N, S, beta, parallel, ranlux, double, Wg
    ("ho", 32, 64, 3, 5.0, True, False, False, 1),
    ("ho", 32, 64, 3, 5.0, True, False, False, 4),
    ("ho", 64, 32, 2, 4.0, False, False, False, 8),
    ("ho", 16, 64, 4, 5.0, True, True, False, 1),
    ("ho", 16, 64, 3, 5.0, True, False, True, 1),
    ("lm2m2", 8, 128, 4, 10.0, True, False, False, 1),
]


@pytest.mark.parametrize("case", CASES_EXACT, ids=lambda c: "%s-W%d-N%d-S%d-par%d-rl%d-f64%d-Wg%d" % (
    c[0], c[1], c[2], c[3], c[5], c[6], c[7], c[8]))
def test_exact_sampler_matches_reference_order(gpu, case):
    name, W, N, S, beta, par, rl, dbl, Wg = case
    _case("exact " + str(case))
    system = _system(name)
    rng = np.random.default_rng(3)
    p0 = _initial_paths(system, W, N, rng)
    opRuns, steps = 3, 2
    k = PIMCKernel(_ka(system, W, N, S, beta, opRuns, steps, par, rl, dbl, "exact", Wg))
    if rl:
        k.initRanlux()
        k.groupInitRanlux()
    k.setPaths(p0)
    k.run()
    ft = np.float64 if dbl else np.float32
    lib = O.build(system, N, beta, operators=(system.energyOp,), double=dbl)
    cfg = O.make_cfg(N, S, system.DOF, W, beta, opRuns, steps, walkersPerGroup=Wg, parallelizePath=par,
                     ranlux=rl, nskip=O.nskip_of_lux(2), nOps=1)
    paths = np.ascontiguousarray(p0.astype(ft))
    T = cfg.nThreads
    tid, grow = k._thread_layout()
    walker_of_thread = np.empty(T, int)
    walker_of_thread[tid.reshape(-1)] = np.repeat(np.arange(W), tid.shape[1])
    if rl:
        st = O.ranlux_init(lib, O.ranlux_seeds(11, 0, T))
        row_of_thread = np.empty(T, np.int64)
        row_of_thread[tid.reshape(-1)] = np.repeat(grow, tid.shape[1])
        gseeds = np.array([O.splitmix64_pair((11 + _GROUP_OFFSET + int(r)) & ((1 << 64) - 1))[0]
                           for r in row_of_thread], dtype=np.uint64)
        gst = O.ranlux_init(lib, gseeds)
        ref = O.run_reference(lib, cfg, paths, rl=st, grl=gst)
        np.testing.assert_array_equal(k._rl.cpu().numpy().view(np.uint8).reshape(T, 112),
                                      st.view(np.uint8).reshape(T, 112))
    else:
        seeds = O.xorshift_rows(11, 0, T + 1)
        ref = O.run_reference(lib, cfg, paths, seeds=seeds)
        np.testing.assert_array_equal(k._seeds.cpu().numpy().view(np.uint32), seeds)
    acc = k._accepts.cpu().numpy().view(np.uint32)
    tol = 2e-4 if name == "ho" else 2e-3
    ok = _compare(k.getPaths(), paths, acc, ref["accepts"], walker_of_thread, tol)
    # operator means: the reference layout is per work-item
    opm = k._operatorValues.cpu().numpy()
    ok_t = ok[walker_of_thread]
    rel = np.abs(opm[ok_t] - ref["opMeans"][ok_t]) / (np.abs(ref["opMeans"][ok_t]) + 1e-3)
    assert rel.max() < 1e-3, "operator mean error %g" % rel.max()


CASES_MOVES = [
    # name, moves, ranlux, W, N, beta, Wg
    ("ho", "shift", False, 64, 32, 4.0, 1),
    ("ho", "single", False, 64, 32, 4.0, 1),
    ("ho", "all", False, 64, 32, 4.0, 4),
    ("ho", "all", True, 32, 32, 4.0, 1),
    ("lm2m2", "all", False, 16, 64, 10.0, 1),
]


@pytest.mark.parametrize("case", CASES_MOVES, ids=lambda c: "%s-%s-rl%d-Wg%d" % (c[0], c[1], c[2], c[6]))
def test_exact_sampler_moves_match_reference_order(gpu, case):
    """Path-shift and single-node moves (kernel.c:407-473, 640-705), one work-item per walker,
    against the oracle's reference-order loop: streams bit-identical, decisions and paths as
    in the exact-sampler test."""
    name, moves, rl, W, N, beta, Wg = case
    _case("exact moves " + str(case))
    system = _system(name)
    S = 2
    rng = np.random.default_rng(5)
    p0 = _initial_paths(system, W, N, rng)
    opRuns, steps = 3, 2 * N
    shift, single, bis = moves in ("shift", "all"), moves in ("single", "all"), moves == "all"
    alpha, psalpha = (0.5, 0.3) if name == "ho" else (0.05, 0.05)
    ka = KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=S, enableOperator=True,
                    enableCorrelator=False, enableBisection=bis, enablePathShift=shift,
                    enableSingleNodeMove=single, enableParallelizePath=False, enableGlobalPath=False,
                    enableGlobalOldPath=False, enableBins=False, enableRanlux=rl,
                    luxuaryFactor=1 if rl else None, ranluxIntMax=2 ** 31 - 1 if rl else None,
                    metroStepsPerOperatorRun=steps, operatorRuns=opRuns, nbrOfWalkersPerWorkGroup=Wg,
                    operators=(system.energyOp,), alpha=alpha if single else None,
                    PSAlpha=psalpha if shift else None, seed=11, sampler="exact", fastMath=False)
    k = PIMCKernel(ka)
    if rl:
        k.initRanlux()
        k.groupInitRanlux()
    k.setPaths(p0)
    k.run()
    lib = O.build(system, N, beta, operators=(system.energyOp,))
    cfg = O.make_cfg(N, S, system.DOF, W, beta, opRuns, steps, walkersPerGroup=Wg, parallelizePath=False,
                     ranlux=rl, nskip=O.nskip_of_lux(2), nOps=1, bisection=bis, pathShift=shift,
                     singleNode=single, alpha=alpha, psAlpha=psalpha)
    paths = np.ascontiguousarray(p0.astype(np.float32))
    T = cfg.nThreads
    if rl:
        st = O.ranlux_init(lib, O.ranlux_seeds(11, 0, T))
        gseeds = np.array([O.splitmix64_pair((11 + _GROUP_OFFSET + w // Wg) & ((1 << 64) - 1))[0]
                           for w in range(W)], dtype=np.uint64)
        gst = O.ranlux_init(lib, gseeds)
        ref = O.run_reference(lib, cfg, paths, rl=st, grl=gst)
        np.testing.assert_array_equal(k._rl.cpu().numpy().view(np.uint8).reshape(T, 112),
                                      st.view(np.uint8).reshape(T, 112))
    else:
        seeds = O.xorshift_rows(11, 0, T + 1)
        ref = O.run_reference(lib, cfg, paths, seeds=seeds)
        np.testing.assert_array_equal(k._seeds.cpu().numpy().view(np.uint32), seeds)
    acc = k._accepts.cpu().numpy().view(np.uint32)
    moves_per_step = int(shift) + int(single) + int(bis)
    print("acceptance %.3f" % (ref["accepts"].sum() / (W * opRuns * steps * moves_per_step)))
    assert 0.05 < ref["accepts"].sum() / (W * opRuns * steps * moves_per_step) < 0.98
    tol = 2e-4 if name == "ho" else 2e-3
    ok = _compare(k.getPaths(), paths, acc, ref["accepts"], np.arange(W), tol, min_match=0.9)
    opm = k._operatorValues.cpu().numpy()
    rel = np.abs(opm[ok] - ref["opMeans"][ok]) / (np.abs(ref["opMeans"][ok]) + 1e-3)
    assert rel.max() < 1e-3, "operator mean error %g" % rel.max()


CASES_WAVE = [
    # name, W, N, S, beta, ranlux, double, path storage
    ("ho", 64, 256, 4, 10.0, False, False, "registers"),
    ("ho", 64, 256, 4, 10.0, False, False, "memory"),
    ("ho", 32, 32, 3, 4.0, False, False, None),
    ("ho", 32, 128, 6, 10.0, True, False, "registers"),
    ("ho", 32, 128, 6, 10.0, True, False, "memory"),
    ("aho", 32, 256, 5, 10.0, False, True, None),
    ("aho", 16, 2048, 4, 20.0, False, False, None),
    ("lm2m2", 16, 256, 6, 10.0, False, False, "registers"),
    ("lm2m2", 16, 1024, 6, 10.0, False, False, "registers"),
    ("lm2m2", 16, 1024, 6, 10.0, False, False, "memory"),
    # several segments per 64-slice block with the potential cache; 2^S = N < 64; RANLUX with ten draws per block
    ("lm2m2", 16, 512, 4, 10.0, False, False, "memory"),
    ("ho", 32, 32, 5, 4.0, False, False, None),
    ("lm2m2", 16, 256, 5, 10.0, True, False, "memory"),
    # RANLUX in FP64 (two float4 draws per four doubles): two and ten draws per block
    ("aho", 32, 256, 5, 10.0, True, True, None),
    ("lm2m2", 16, 256, 5, 10.0, True, True, "memory"),
    # 4- and 8-atom clusters (DOF 12 and 24, 6 and 28 pairs)
    ("lm2m2x4", 16, 512, 6, 10.0, False, False, "memory"),
    ("lm2m2x8", 16, 256, 5, 10.0, False, False, "memory"),
]


@pytest.mark.parametrize("case", CASES_WAVE, ids=lambda c: "%s-W%d-N%d-S%d-rl%d-f64%d-%s" % (
    c[0], c[1], c[2], c[3], c[5], c[6], c[7]))
def test_wave_sampler_matches_oracle(gpu, case):
    name, W, N, S, beta, rl, dbl, storage = case
    _case("wave " + str(case))
    system = _system(name)
    rng = np.random.default_rng(5)
    p0 = _initial_paths(system, W, N, rng)
    opRuns, steps = 3, 2
    k = PIMCKernel(_ka(system, W, N, S, beta, opRuns, steps, True, rl, dbl, "wave", storage=storage))
    if rl:
        k.initRanlux()
        k.groupInitRanlux()
    k.setPaths(p0)
    k.run()
    ft = np.float64 if dbl else np.float32
    lib = O.build(system, N, beta, operators=(system.energyOp,), double=dbl)
    cfg = O.make_cfg(N, S, system.DOF, W, beta, opRuns, steps, ranlux=rl, nskip=O.nskip_of_lux(2), nOps=1)
    paths = np.ascontiguousarray(p0.astype(ft))
    if rl:
        lst = O.ranlux_init(lib, O.ranlux_seeds(11, 0, W * 64))
        gst = O.ranlux_init(lib, O.ranlux_seeds(11, _GROUP_OFFSET, W))
        ref = O.run_wave(lib, cfg, paths, lane_rl=lst, group_rl=gst)
        np.testing.assert_array_equal(k._rl.cpu().numpy().view(np.uint8).reshape(W * 64, 112),
                                      lst.view(np.uint8).reshape(W * 64, 112))
        np.testing.assert_array_equal(k._grl.cpu().numpy().view(np.uint8).reshape(W, 112),
                                      gst.view(np.uint8).reshape(W, 112))
    else:
        lane = O.xorshift_rows(11, 0, W * 64)
        grp = O.xorshift_rows(11, W * 64, W)
        ref = O.run_wave(lib, cfg, paths, lane_seeds=lane, group_seeds=grp)
        dev = k._seeds.cpu().numpy().view(np.uint32)
        np.testing.assert_array_equal(dev[: W * 64], lane)
        np.testing.assert_array_equal(dev[W * 64:], grp)
    acc = k._accepts.cpu().numpy().view(np.uint32)
    tol = 2e-3 if name.startswith("lm2m2") else 2e-4
    ok = _compare(k.getPaths(), paths, acc, ref["accepts"], np.arange(W), tol, min_match=0.9)
    ops = k.getOperators()[0]
    rel = np.abs(ops[ok] - ref["opMeans"][ok]) / (np.abs(ref["opMeans"][ok]) + 1e-3)
    assert rel.max() < 1e-3, "operator mean error %g" % rel.max()


def test_system_section_matches_oracle(gpu):
    """The hiprtc-compiled potential and virial operator vs gcc-compiled strings and the FP64 formula."""
    from feynsimul_amd.physical_systems.lm2m2 import lm2m2_kelvin
    _case("eval_system lm2m2 3 atoms")
    system = Lm2m2Cluster(3)
    k = PIMCKernel(_ka(system, 8, 64, 3, 10.0, 1, 1, True, False, True, "wave"))
    rng = np.random.default_rng(1)
    r = rng.uniform(2.2, 12.0, 4000)
    x = np.zeros((4000, 9))
    x[:, 3] = r * system.massScale  # atom 2 on the x axis, atom 3 far away
    x[:, 6:] = 1e3
    pot, ops = k.evalSystem(x)
    ref = np.array([lm2m2_kelvin(v) for v in r])
    np.testing.assert_allclose(pot, ref, rtol=1e-9, atol=1e-9)
    lib = O.build(system, 64, 10.0, operators=(system.energyOp,), double=True)
    xo = np.ascontiguousarray(x)
    for i in range(0, 4000, 97):
        v = lib.fs_oracle_potential(xo[i].ctypes.data)
        assert abs(v - pot[i]) <= 1e-9 * max(1.0, abs(v))
        acc = np.zeros(2)
        lib.fs_oracle_operators(xo[i].ctypes.data, acc.ctypes.data)
        assert abs(acc[0] - ops[i, 0]) <= 1e-9 * max(1.0, abs(acc[0]))


@pytest.mark.parametrize("double", [False, True])
def test_lm2m2_core_on_gpu(gpu, double):
    """The hiprtc-compiled LM2M2 pair energy and virial term through the core (r = 0 up to 2 A), FP32
    with relaxed math and FP64: finite, with the reference's zero damped dispersion term at and below
    x = 0.0005 / r_m, against the FP64 formula."""
    from test_oracle import _core_points
    _case("eval_system lm2m2 core double=%d" % double)
    system = Lm2m2Cluster(3)
    ka = _ka(system, 8, 64, 3, 10.0, 1, 1, True, False, double, "wave")
    ka.fastMath = not double
    k = PIMCKernel(ka)
    pts, pot, vir = _core_points(system)
    got, ops = k.evalSystem(pts)
    assert np.isfinite(got).all() and np.isfinite(ops).all()
    np.testing.assert_allclose(got, pot, rtol=1e-9 if double else 1e-5)
    np.testing.assert_allclose(ops[:, 0], vir, rtol=1e-4)


@pytest.mark.parametrize("storage", ["memory", "registers"])
def test_wave_observable_histogram_matches_oracle(gpu, storage):
    """256-bin pair-distance histogram of the He3 cluster (the observables of every slice at every
    measurement) against the oracle's restatement of the same launch."""
    _case("wave pair-distance histogram " + storage)
    system = Lm2m2Cluster(3)
    W, N, S, beta = 16, 256, 6, 10.0
    obs = tuple(system.pairDistance(i, j) for i, j in system.pairs)
    rng = np.random.default_rng(9)
    p0 = _initial_paths(system, W, N, rng)
    ka = _ka(system, W, N, S, beta, 3, 2, True, False, False, "wave", storage=storage)
    ka.histogramObservables, ka.histogramRange, ka.histogramBins = obs, (0.0, 12.0), 256
    k = PIMCKernel(ka)
    k.setPaths(p0)
    k.run()
    lib = O.build(system, N, beta, operators=(system.energyOp,), observables=obs)
    cfg = O.make_cfg(N, S, system.DOF, W, beta, 3, 2, nOps=1, nHistObs=3, histBins=256, histRange=(0.0, 12.0))
    paths = np.ascontiguousarray(p0.astype(np.float32))
    ref = O.run_wave(lib, cfg, paths, lane_seeds=O.xorshift_rows(11, 0, W * 64),
                     group_seeds=O.xorshift_rows(11, W * 64, W))
    _, dens, counts = k.getObservableHistogram()
    assert counts.sum() == ref["obsHist"].sum() or abs(counts.sum() - ref["obsHist"].sum()) <= 3 * N
    # walkers whose Metropolis decisions flipped move a few slices; allow that many counts to move
    assert np.abs(counts - ref["obsHist"]).sum() <= 0.02 * counts.sum() + 6
    assert counts.sum() > 0.5 * W * N * 3 * 3


def test_wave_observable_histogram_recount_and_density(gpu):
    """With one measurement per launch the GPU's pair-distance histogram equals an independent numpy
    recount of the GPU's own final paths (bin for bin, up to values within a few FP32 ulps of an edge),
    and getObservableHistogram's density integrates to the in-range fraction of the values."""
    from test_oracle import _recount, _assert_histogram
    _case("wave pair-distance histogram recount")
    system = Lm2m2Cluster(3)
    W, N, S, beta, bins, lo, hi = 32, 512, 6, 10.0, 256, 0.0, 12.0
    obs = tuple(system.pairDistance(i, j) for i, j in system.pairs)
    rng = np.random.default_rng(19)
    pos = system.initialPositions(W, rng, sigma=1.5)
    p0 = (np.repeat(pos[:, :, None], N, axis=2) + rng.normal(0, 0.05, (W, system.DOF, N))).reshape(W, -1)
    ka = _ka(system, W, N, S, beta, 1, 3, True, False, False, "wave", storage="memory")
    ka.histogramObservables, ka.histogramRange, ka.histogramBins = obs, (lo, hi), bins
    k = PIMCKernel(ka)
    k.setPaths(p0)
    k.run()
    centres, dens, counts = k.getObservableHistogram()
    x = k.getPaths().reshape(W, system.nAtoms, 3, N).astype(np.float64) / system.massScale
    r = np.stack([np.sqrt(((x[:, i] - x[:, j]) ** 2).sum(axis=1)) for i, j in system.pairs])
    far, near_v = _recount(r, lo, hi, bins)
    _assert_histogram(counts, far, near_v, bins)
    inside = ((r >= lo) & (r < hi)).mean()
    width = (hi - lo) / bins
    assert abs(dens.sum() * width - inside) <= (len(near_v) + 1) / r.size
    np.testing.assert_allclose(centres, lo + width * (np.arange(bins) + 0.5))


CASES_CORR = [
    # name, W, N, S, beta, lags, double, path storage, correlators
    ("aho", 16, 512, 4, 10.0, 96, False, "registers", ("x1",)),
    ("aho", 16, 512, 4, 10.0, 96, False, "memory", ("x1", "x1 * x1")),
    ("aho", 8, 2048, 4, 20.0, 1024, False, None, ("x1",)),
    ("aho", 8, 4096, 4, 40.0, 512, False, None, ("x1", "x1 * x1")),
    ("ho", 16, 256, 4, 10.0, 128, True, None, ("x1",)),
    ("ho", 32, 32, 3, 4.0, 16, False, None, ("x1",)),
]


@pytest.mark.parametrize("case", CASES_CORR, ids=lambda c: "%s-W%d-N%d-L%d-f64%d-%s-c%d" % (
    c[0], c[1], c[2], c[5], c[6], c[7], len(c[8])))
def test_wave_correlator_matches_oracle(gpu, case):
    """Two-time correlator bins of the wave sampler (FP32 lag sums per measurement added to FP64
    Kahan bins) against the oracle's FP64 lag sums of the same launch, for a whole staged row
    (N <= 1024) and for time chunks with one or two lag chunks (N = 2048, 4096)."""
    name, W, N, S, beta, lags, dbl, storage, corr = case
    _case("wave correlator " + str(case))
    system = _system(name)
    rng = np.random.default_rng(17)
    p0 = _initial_paths(system, W, N, rng)
    opRuns, steps = 3, 2
    ka = _ka(system, W, N, S, beta, opRuns, steps, True, False, dbl, "wave", storage=storage)
    ka.enableCorrelator, ka.correlators, ka.correlatorLags = True, corr, lags
    k = PIMCKernel(ka)
    k.setPaths(p0)
    k.run()
    ft = np.float64 if dbl else np.float32
    lib = O.build(system, N, beta, operators=(system.energyOp,), correlators=corr, double=dbl)
    cfg = O.make_cfg(N, S, system.DOF, W, beta, opRuns, steps, nOps=1, nCorr=len(corr), corLags=lags)
    paths = np.ascontiguousarray(p0.astype(ft))
    ref = O.run_wave(lib, cfg, paths, lane_seeds=O.xorshift_rows(11, 0, W * 64),
                     group_seeds=O.xorshift_rows(11, W * 64, W))
    acc = k._accepts.cpu().numpy().view(np.uint32)
    ok = _compare(k.getPaths(), paths, acc, ref["accepts"], np.arange(W), 2e-4, min_match=0.85)
    raw = k._correlatorValues.cpu().numpy().astype(np.float64).reshape(W, len(corr), lags, 2)
    got = raw[..., 0] - raw[..., 1]
    exp = ref["corBins"].reshape(W, len(corr), lags)
    scale = np.abs(exp[:, :, :1])
    err = (np.abs(got - exp) / scale)[ok]
    print("max error / C(0) bin: %.3g over %d walkers" % (err.max(), ok.sum()))
    assert err.max() < (1e-7 if dbl else 2e-3)
    # C(0) is the measurement sum of c^2, which the oracle forms the same way
    assert np.all(exp[:, :, 0] > 0)


@pytest.mark.parametrize("name,N,S", [("aho", 2048, 4), ("lm2m2", 1024, 6)])
def test_potential_cache_choice_keeps_results(gpu, name, N, S):
    """Memory storage with and without the per-slice potential cache: the cached value is the same FP32
    evaluation of the same coordinates, so paths, operators and streams are bit-identical."""
    _case("potential cache on/off " + name)
    system = _system(name)
    W = 16
    p0 = _initial_paths(system, W, N, np.random.default_rng(23))
    out = []
    for pc in (True, False):
        ka = _ka(system, W, N, S, 10.0, 2, 3, True, False, False, "wave", storage="memory")
        ka.potentialCache = pc
        k = PIMCKernel(ka)
        k.setPaths(p0)
        k.run()
        out.append((k.getPaths(), k.getOperators(), k._seeds.cpu().numpy(), k.getAcceptanceRate()))
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)
