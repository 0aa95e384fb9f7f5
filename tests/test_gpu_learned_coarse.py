"""The whole-path coarse move (LearnedProposal(coarse=True)) against the oracle's coarse_move
(oracle/learned_oracle.py).

Tolerances (floating point): the network runs in BF16 with FP32 accumulation and the GPU sums in
other orders than numpy, so
* the parts of log alpha (log q ratio, Keff change, V change) agree to 2e-2 + 1e-3 relative;
* accept decisions agree wherever |log alpha - log u| > 0.05 (at least 80 % of the paths), and the
  paths of walkers both accept agree to 5e-3;
* with a zero output layer the coarse proposal is the Levy bridge, and log q_old - log q_new must
  equal the change of the coarse kinetic action exactly in the mathematics: 1e-3 relative on the GPU.
Statistics: on the C1 harmonic oscillator the move, alone and trained online, keeps <x^2> at the
exact discretised value within 4 standard errors.  The move's network runs 64 of the 128 hidden units
(the others zero, which training must preserve bit for bit).
"""
import numpy as np
import pytest

from oracle import learned_oracle as L
from oracle import oracle as O
from feynsimul_amd.kernel_args import KernelArgs
from feynsimul_amd.kernel import PIMCKernel
from feynsimul_amd.learned import LearnedProposal
from feynsimul_amd.physical_systems import HarmOsc, AnharmOsc, Lm2m2Cluster
from feynsimul_amd.physical_systems.lm2m2 import lm2m2_kelvin

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case] " + msg, flush=True)


def _kernel(system, W, N, beta, opRuns=1, M=1, seed=5, S=6):
    ka = KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=S, enableOperator=True,
                    enableCorrelator=False, enableBisection=True, enablePathShift=False,
                    enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                    enableGlobalOldPath=False, enableBins=False, enableDouble=False, enableRanlux=False,
                    metroStepsPerOperatorRun=M, operatorRuns=opRuns, nbrOfWalkersPerWorkGroup=1,
                    operators=(system.energyOp,), seed=seed, sampler="wave", pathStorage="memory")
    return PIMCKernel(ka)


def _ho_paths(W, N, beta, rng):
    """Paths from the exact Gaussian of the discretised harmonic action: its matrix is circulant with
    eigenvalues (2 - 2 cos(2 pi k / N)) / eps + eps, so x = F^-1 (F z / sqrt(lambda)) for white z."""
    eps = beta / N
    lam = (2.0 - 2.0 * np.cos(2.0 * np.pi * np.arange(N) / N)) / eps + eps
    z = rng.standard_normal((W, N))
    return np.fft.ifft(np.fft.fft(z, axis=1) / np.sqrt(lam), axis=1).real.astype(np.float32)


def _poly_V(lam):
    return lambda x: 0.5 * x[..., 0] * x[..., 0] + lam * x[..., 0] ** 4


def _lm2m2_V(system):
    vk = np.vectorize(lm2m2_kelvin)

    def V(x):
        x = x.astype(np.float64)
        tot = 0.0
        for i, j in system.pairs:
            r = np.sqrt(((x[..., 3 * i:3 * i + 3] - x[..., 3 * j:3 * j + 3]) ** 2).sum(-1)) / system.massScale
            tot = tot + vk(r)
        return tot
    return V


def _setup(case, rng):
    if case in ("ho128", "ho256", "aho512", "ho8192", "aho2048", "aho4096"):
        N, beta = {"ho128": (128, 2.0), "ho256": (256, 10.0), "aho512": (512, 8.0), "ho8192": (8192, 4.0),
                   "aho2048": (2048, 8.0), "aho4096": (4096, 16.0)}[case]
        system = HarmOsc() if case.startswith("ho") else AnharmOsc(0.1)
        W, D = 64, 1
        paths0 = _ho_paths(W, N, beta, rng)[:, None, :]
        V = _poly_V(0.0 if case.startswith("ho") else 0.1)
        return system, W, N, beta, D, paths0, V, 0.03
    N = {"he3_512": 512, "he2_256": 256}.get(case, 128)
    if case.startswith("he2"):
        # two atoms, DOF 6: at N = 128 a path of 768 floats, which takes the fused move
        system, W, beta, D = Lm2m2Cluster(2), 32, 1.0, 6
        base = system.initialPositions(W, rng)
        paths0 = (base[:, :, None] + 0.02 * rng.standard_normal((W, D, N))).astype(np.float32)
        return system, W, N, beta, D, paths0, _lm2m2_V(system), 0.05
    system, W, beta, D = Lm2m2Cluster(3), 16, 1.0 * N / 128, 9
    base = system.initialPositions(W, rng)
    paths0 = (base[:, :, None] + 0.02 * rng.standard_normal((W, D, N))).astype(np.float32)
    return system, W, N, beta, D, paths0, _lm2m2_V(system), 0.05


def _run_case(case, theta_scale=None, seed=777, sweep=3, width=None, omega2=0.0):
    rng = np.random.default_rng(11)
    system, W, N, beta, D, paths0, V, scale = _setup(case, rng)
    if theta_scale is not None:
        scale = theta_scale
    k = _kernel(system, W, N, beta)
    k.setPaths(paths0.reshape(W, D * N))
    lp = LearnedProposal(k, batch=2048, seed=seed, coarse=True, width=width, bridgeOmega2=omega2)
    theta = L.init_params(D, 5, scale_last=scale, active=lp.width)
    lp.setParameters(theta)
    lp._sweepCount = sweep
    lp.traceSweeps()
    lp.sweep(1)
    got = k.getPaths().reshape(W, D, N)
    tr = lp.getTrace()[:, 0, :]
    ref = paths0.copy()
    _, d = L.coarse_move(ref, L.unflatten(theta, D), V, eps=beta / N, seed=seed, sweep=sweep, atoms3d=(D % 3 == 0),
                         return_detail=True, omega2=omega2)
    return paths0, got, ref, tr, d


# ho8192 (128 intervals) and he3_512 (8 intervals x 9 DOF) exceed 64 values per lane: their accepted bridges are
# drawn a second time from the same counters instead of being held in registers
# The network kernel copies only the active units' weights and biases into LDS and picks its wavefronts per
# block by occupancy, which both depend on the width: the default 64 units, and 32, 96 and 128 below.
# the last five: the harmonic reference bridge (omega2 > 0), in the fused move (ho256, he2), with registers
# holding the new slices (aho2048) and with the bridges drawn twice (ho8192, he3_512)
@pytest.mark.parametrize("case,width,omega2", [("ho128", None, 0.0), ("ho256", None, 0.0), ("aho512", None, 0.0),
                                               ("he3", None, 0.0), ("ho8192", None, 0.0), ("he3_512", None, 0.0),
                                               ("aho512", 32, 0.0), ("he3", 32, 0.0), ("ho256", 96, 0.0),
                                               ("aho512", 128, 0.0), ("he2", None, 0.0),
                                               ("ho256", 32, 1.0), ("he2", 32, 0.3), ("aho2048", 32, 1.5),
                                               ("ho8192", 32, 0.6), ("he3_512", 32, 0.2)])
def test_coarse_move_matches_oracle(gpu, case, width, omega2):
    _case("whole-path move %s width %s omega2 %s" % (case, width, omega2))
    paths0, got, ref, tr, d = _run_case(case, width=width, omega2=omega2)
    W = paths0.shape[0]
    for name, col, tol in (("lq", 0, 1e-3), ("dK", 1, 1e-3), ("dV", 2, 1e-3)):
        err = np.abs(tr[:, col] - d[name])
        lim = 2e-2 + tol * np.abs(d[name])
        print("%s max err %.3e (worst allowed %.3e)" % (name, err.max(), lim[np.argmax(err)]))
        assert (err < lim).all(), (name, tr[:, col][:4], d[name][:4])
    changed = np.array([not np.array_equal(paths0[w], got[w]) for w in range(W)])
    clear = np.abs(d["logA"] - np.log(np.maximum(d["u"], 1e-30))) > 0.05
    print("accepted %d/%d, clear %d" % (d["accept"].sum(), W, clear.sum()))
    assert clear.mean() > 0.8
    assert np.array_equal(changed[clear], d["accept"][clear])
    for w in np.nonzero(changed & d["accept"])[0]:
        assert np.abs(got[w] - ref[w]).max() < 5e-3
    if case in ("ho128", "aho512", "ho8192", "aho2048"):
        assert 0 < d["accept"].sum() < W


def test_zero_output_is_the_levy_bridge(gpu):
    """With the output layer at zero the coarse proposal is the free-particle bridge of the whole
    path, whose density ratio is exactly the change of the coarse kinetic action: log alpha reduces
    to -eps dV, plain bisection of the whole path."""
    _case("whole-path move, zero output layer")
    paths0, got, ref, tr, d = _run_case("ho256", theta_scale=0.0)
    err = np.abs(tr[:, 0] - tr[:, 1])
    print("max |lq - dK| %.3e, max |dK| %.3e" % (err.max(), np.abs(tr[:, 1]).max()))
    assert (err < 1e-3 * np.abs(tr[:, 1]) + 1e-3).all()
    np.testing.assert_allclose(d["lq"], d["dK"], rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("omega2", [0.0, 0.6])
def test_coarse_move_keeps_ho_energy(gpu, omega2):
    """C1 sizes (N = 256, beta = 10, 4096 chains): whole-path moves alone, with the network trained
    online on the coarse nodes, sample the exact discretised <x^2> within 4 standard errors, with the Levy
    bridge and with a harmonic reference bridge of another frequency than the oscillator's."""
    _case("whole-path move, C1 energy, omega2 %s" % omega2)
    system = HarmOsc()
    W, N, beta = 4096, 256, 10.0
    k = _kernel(system, W, N, beta, opRuns=1, M=1500, S=4)
    k.setPaths(np.random.default_rng(0).uniform(-1, 1, (W, N)).astype(np.float32))
    k.run()
    lp = LearnedProposal(k, batch=16384, seed=31, coarse=True, fallbackAcceptance=0.0, bridgeOmega2=omega2)
    first = lp.trainOnline(1)
    last = lp.trainOnline(99)
    ar = lp.sweep(64)
    k.setSeriesCapacity(200)
    lp.run(200, 8)
    s = k.getSeries()[:, :, 0].astype(np.float64)
    per = s.mean(axis=0)
    E, se = per.mean(), per.std() / np.sqrt(W)
    exact = O.ho_discrete_x2(N, beta)
    print("NLL %.4f -> %.4f, whole-path acceptance %.3f, E %.6f +- %.6f exact %.9f" % (first, last, ar, E, se,
                                                                                     exact))
    assert last < first
    assert 0.05 < ar < 1.0
    assert abs(E - exact) < 4 * se
    # training kept the hidden units beyond the active width (64) at exactly zero
    theta = lp.getParameters()
    assert np.array_equal(theta, L.mask_width(theta, 1, lp.width)) and lp.width == 64


def test_coarse_training_rows_match_oracle(gpu):
    """Training draws only coarse nodes (positions that are multiples of 64 of the whole path):
    the loss of one Adam step equals the oracle's on the same rows."""
    _case("whole-path move, training rows")
    rng = np.random.default_rng(4)
    system, W, N, beta, D = HarmOsc(), 128, 512, 8.0, 1
    paths = rng.standard_normal((W, D, N)).astype(np.float32)
    k = _kernel(system, W, N, beta)
    k.setPaths(paths.reshape(W, D * N))
    B = 4096
    lp = LearnedProposal(k, batch=B, seed=78, coarse=True)
    theta = L.init_params(D, 1, scale_last=0.3, active=lp.width)
    lp.setParameters(theta)
    loss = lp.train(1)
    S = 9
    w, s0, p = L.train_rows(78, 1, B, W, N, N, fine=L.FINE_L)
    assert (p % L.FINE_L == 0).all() and p.min() >= L.FINE_L and p.max() <= N - L.FINE_L
    h = p & -p
    n = S - np.log2(h).astype(np.int64)
    a = paths[w, :, (s0 + p - h) % N]
    b = paths[w, :, (s0 + p + h) % N]
    x = paths[w, :, (s0 + p) % N]
    f = L.features(a, b, n, S, False, 1.0)
    sig = np.sqrt(N * (beta / N) / (2.0 * 2.0 ** n)).astype(np.float32)
    loss_o, _ = L.nll_and_grad(L.unflatten(theta, D), f, a, b, x, sig, D)
    print("loss %.6f oracle %.6f" % (loss, loss_o))
    assert abs(loss - loss_o) < 1e-3 * abs(loss_o) + 1e-4


def test_two_stream_halves_match_one_stream(gpu, monkeypatch):
    """With at least 2 x 8192 walkers the move runs the two halves of the walkers on forked streams; every
    walker's moves depend only on its own path and counters, so paths, accept counts and the trace equal
    those of the single-stream launches bit for bit."""
    _case("whole-path move, two streams against one")
    # N DOF > 1024: the two-kernel move (smaller paths take the fused move, which has no second stream)
    system, W, N, beta = AnharmOsc(0.1), 16384, 2048, 8.0
    rng = np.random.default_rng(3)
    paths0 = _ho_paths(W, N, beta, rng)
    theta = L.init_params(1, 5, scale_last=0.03, active=32)
    out = []
    for streams in ("1", "2"):
        monkeypatch.setenv("FS_LPC_STREAMS", streams)
        k = _kernel(system, W, N, beta)
        k.setPaths(paths0)
        lp = LearnedProposal(k, batch=2048, seed=99, coarse=True, width=32)
        lp.setParameters(theta)
        lp.traceSweeps()
        ar = lp.sweep(3)
        out.append((k.getPaths(), lp._acc.cpu().numpy(), lp.getTrace(), ar))
        del lp, k
    print("acceptance %.3f / %.3f" % (out[0][3], out[1][3]))
    assert 0.0 < out[0][3] < 1.0
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("case,width,omega2", [("ho256", 32, 0.0), ("aho512", 64, 0.0), ("ho128", 128, 0.0),
                                               ("he2", 32, 0.0), ("aho512", 32, 1.4), ("he2", 32, 0.3)])
def test_fused_small_path_move_matches_two_kernel_move(gpu, monkeypatch, case, width, omega2):
    """Paths of at most 1024 floats take the fused move (one wavefront per walker, the path in LDS, every move
    of a call in one launch); its random numbers, network rows and sums are the two-kernel move's, so five
    moves in a row give the same paths, accept counts and last-move trace bit for bit."""
    _case("fused whole-path move %s width %d omega2 %s against the two-kernel move" % (case, width, omega2))
    rng = np.random.default_rng(17)
    system, W, N, beta, D, paths0, V, scale = _setup(case, rng)
    theta = L.init_params(D, 5, scale_last=scale, active=width)
    out = []
    for fused in ("0", "1"):
        monkeypatch.setenv("FS_LPC_FUSED", fused)
        k = _kernel(system, W, N, beta)
        k.setPaths(paths0.reshape(W, D * N))
        lp = LearnedProposal(k, batch=2048, seed=321, coarse=True, width=width, bridgeOmega2=omega2)
        lp.setParameters(theta)
        lp._sweepCount = 11
        lp.traceSweeps()
        ar = lp.sweep(5)
        out.append((k.getPaths(), lp._acc.cpu().numpy(), lp.getTrace(), ar))
        del lp, k
    print("acceptance %.3f / %.3f" % (out[0][3], out[1][3]))
    assert 0.0 < out[0][3] < 1.0
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


# two-kernel moves (N DOF > 1024) with 32, 4 x 6, 64 and 128 coarse values per walker
@pytest.mark.parametrize("case", ["aho2048", "he2_256", "aho4096", "ho8192"])
def test_moves_of_one_call_match_single_move_calls(gpu, case):
    """A call of four two-kernel moves reuses the workspace for each move's nodes and counts the moves'
    random numbers by sweep0 + i; it must equal four calls of one move, bit for bit: paths, accept
    counts and the last move's trace."""
    _case("whole-path move %s: one call of four moves against four calls" % case)
    rng = np.random.default_rng(23)
    system, W, N, beta, D, paths0, V, scale = _setup(case, rng)
    assert N * D > 1024  # not the fused move
    theta = L.init_params(D, 5, scale_last=scale, active=32)
    out = []
    for calls in (1, 4):
        k = _kernel(system, W, N, beta)
        k.setPaths(paths0.reshape(W, D * N))
        lp = LearnedProposal(k, batch=2048, seed=57, coarse=True, width=32)
        lp.setParameters(theta)
        lp._sweepCount = 5
        lp.traceSweeps()
        acc = np.zeros(W, np.int64)
        for _ in range(calls):
            lp.sweep(4 // calls)  # each call zeroes and refills the accept counts
            acc += lp._acc.cpu().numpy()
        out.append((k.getPaths(), acc, lp.getTrace()))
        del lp, k
    acc = out[0][1].sum()
    print("accepted %d of %d moves" % (acc, 4 * W))
    assert 0 < acc < 4 * W
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def test_harmonic_reference_of_the_oscillator_leaves_no_fine_action(gpu):
    """With the reference at the oscillator's own frequency (V = x^2 / 2, omega2 = 1, as "auto" finds it) the
    potential part of log alpha, sum (V - x^2 / 2) over the slices, vanishes up to rounding, so the
    decision rests on the coarse action and the network's density alone."""
    _case("whole-path move, harmonic reference of the oscillator")
    rng = np.random.default_rng(5)
    system, W, N, beta, D, paths0, V, scale = _setup("ho256", rng)
    k = _kernel(system, W, N, beta)
    k.setPaths(paths0.reshape(W, D * N))
    lp = LearnedProposal(k, batch=2048, seed=9, coarse=True, width=32, bridgeOmega2="auto")
    assert abs(lp.bridgeOmega2 - 1.0) < 1e-3
    lp.setParameters(L.init_params(D, 5, scale_last=scale, active=32))
    lp.traceSweeps()
    lp.sweep(1)
    tr = lp.getTrace()[:, 0, :]
    print("max |dV| %.3e, max |dK| %.3e" % (np.abs(tr[:, 2]).max(), np.abs(tr[:, 1]).max()))
    assert np.abs(tr[:, 2]).max() < 1e-3 * max(1.0, np.abs(tr[:, 1]).max())


# the two-kernel move with registers holding the new slices (aho2048) and with the bridges drawn twice
# (he3_512), the fused move (ho256), each with the Levy bridge and a harmonic reference
@pytest.mark.parametrize("case,omega2", [("aho2048", 0.0), ("aho2048", 1.5), ("he3_512", 0.0), ("he3_512", 0.2),
                                         ("ho256", 0.0), ("ho256", 1.0)])
def test_measurement_in_the_move_matches_the_measure_only_launch(gpu, case, omega2):
    """LearnedProposal.run with the measurement made by each block's last move (fs_lp_sweep_measure) against
    the measure-only launch of the kernel after each block: the same paths and accept counts bit for bit,
    and the same series, operator means and totals up to the order of the FP32 slice sums."""
    _case("whole-path move %s omega2 %s: measurement in the move" % (case, omega2))
    rng = np.random.default_rng(31)
    system, W, N, beta, D, paths0, V, scale = _setup(case, rng)
    theta = L.init_params(D, 5, scale_last=scale, active=32)
    out = []
    for inMove in (False, True):
        k = _kernel(system, W, N, beta)
        k.setPaths(paths0.reshape(W, D * N))
        k.setSeriesCapacity(8)
        lp = LearnedProposal(k, batch=2048, seed=41, coarse=True, width=32, bridgeOmega2=omega2,
                             fallbackAcceptance=0.0)
        lp.setParameters(theta)
        lp.measureInMove = inMove
        lp.run(3, 2)
        out.append((k.getPaths(), lp._runAcc.cpu().numpy(), k.getSeries(), k.getOperators(),
                    k.getOperatorTotals()))
        del lp, k
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2].shape == out[1][2].shape == (3, W, 1)
    for i in (2, 3, 4):
        np.testing.assert_allclose(out[1][i], out[0][i], rtol=2e-5, atol=1e-6 * np.abs(out[0][i]).max())


def test_measurement_in_the_move_is_refused_where_it_cannot_record(gpu):
    """fs_lp_sweep_measure refuses the segment proposal, kernels that also measure a correlator, and a call
    without a move; LearnedProposal.run then keeps the measure-only launch."""
    from feynsimul_amd import _native
    _case("whole-path move: measurement refusals")
    system, W, N, beta = HarmOsc(), 64, 256, 4.0
    k = _kernel(system, W, N, beta)
    k.setSeriesCapacity(4)
    seg = LearnedProposal(k, batch=2048, seed=1, coarse=False)
    with pytest.raises(_native.FsError, match="coarse = 1"):
        seg._P.sweep_measure(k._paths.data_ptr(), seg._acc.data_ptr(), seg._bufs(), k._buffers(), 0, 1, 0,
                             k._stream.cuda_stream)
    lp = LearnedProposal(k, batch=2048, seed=1, coarse=True, width=32)
    with pytest.raises(_native.FsError, match="at least one move"):
        lp._P.sweep_measure(k._paths.data_ptr(), lp._acc.data_ptr(), lp._bufs(), k._buffers(), 0, 0, 0,
                            k._stream.cuda_stream)
    kc = PIMCKernel(KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=6, enableOperator=True,
                               enableCorrelator=True, enableBisection=True, enablePathShift=False,
                               enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                               enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=1,
                               operatorRuns=1, nbrOfWalkersPerWorkGroup=1, operators=(system.energyOp,),
                               correlators=("x1",), seed=3, sampler="wave"))
    lc = LearnedProposal(kc, batch=2048, seed=1, coarse=True, width=32)
    with pytest.raises(_native.FsError, match="no correlators"):
        lc._P.sweep_measure(kc._paths.data_ptr(), lc._acc.data_ptr(), lc._bufs(), kc._buffers(), 0, 1, 0,
                            kc._stream.cuda_stream)
    lc.run(2, 1)  # the measure-only launch: the correlator is measured
    assert kc._corRuns == 2
