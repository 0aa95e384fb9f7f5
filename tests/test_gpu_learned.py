"""The learned proposal's HIP kernels against the oracle (oracle/learned_oracle.py).

Tolerances (floating point, BF16 MFMA with FP32 accumulation): the GPU and the
oracle round the same values to BF16, but sum in different orders, so a hidden
activation can round to the neighbouring BF16 value now and then.
* network outputs: max |diff| < 1e-2, mean |diff| < 2e-4;
* one sweep: accept decisions agree wherever |log alpha - log u| > 0.05 (at
  least 80 % of the segments), and accepted paths agree to 5e-3;
* a training step: loss to 1e-3 relative, gradient to 3e-2 in relative L2 norm,
  and the Adam update exactly as the oracle's formula applied to the GPU's gradient.
Statistics: a deliberately poor network must keep the exact discretised-HO
energy (4 standard errors).
"""
import numpy as np
import pytest

from oracle import learned_oracle as L
from oracle import oracle as O
from feynsimul_amd.kernel_args import KernelArgs
from feynsimul_amd.kernel import PIMCKernel
from feynsimul_amd.learned import LearnedProposal
from feynsimul_amd.physical_systems import HarmOsc, Lm2m2Cluster
from feynsimul_amd.physical_systems.lm2m2 import lm2m2_kelvin

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case] " + msg, flush=True)


def _kernel(system, W, N, S, beta, opRuns=1, M=1, seed=5):
    ka = KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=S, enableOperator=True,
                    enableCorrelator=False, enableBisection=True, enablePathShift=False,
                    enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                    enableGlobalOldPath=False, enableBins=False, enableDouble=False, enableRanlux=False,
                    metroStepsPerOperatorRun=M, operatorRuns=opRuns, nbrOfWalkersPerWorkGroup=1,
                    operators=(system.energyOp,), seed=seed, sampler="wave", pathStorage="memory")
    return PIMCKernel(ka)


def _ho_V(x):
    return 0.5 * x[..., 0] * x[..., 0]


def _ho_cov(N, beta, k=1.0):
    eps = beta / N
    A = np.zeros((N, N))
    for i in range(N):
        A[i, i] = 2 / eps + eps * k
        A[i, (i + 1) % N] -= 1 / eps
        A[i, (i - 1) % N] -= 1 / eps
    return np.linalg.inv(A)


def _lm2m2_V(system):
    vk = np.vectorize(lm2m2_kelvin)

    def V(x):
        x = x.astype(np.float64)
        tot = 0.0
        for i, j in system.pairs:
            r = np.sqrt(((x[..., 3 * i:3 * i + 3] - x[..., 3 * j:3 * j + 3]) ** 2).sum(-1)) / system.massScale
            tot = tot + vk(r)
        return tot.astype(np.float32)
    return V


@pytest.mark.parametrize("D", [1, 9])
def test_network_matches_oracle(gpu, D):
    _case("learned network D=%d" % D)
    system = HarmOsc() if D == 1 else Lm2m2Cluster(3)
    k = _kernel(system, 64, 128, 4, 10.0)
    lp = LearnedProposal(k, batch=2048, seed=3)
    theta = L.init_params(D, 7, scale_last=0.3)
    lp.setParameters(theta)
    X = np.random.default_rng(1).standard_normal((1000, 2 * D + 1)).astype(np.float32)
    got = lp.evalNetwork(X)
    want, _ = L.mlp_forward(L.unflatten(theta, D), X)
    err = np.abs(got - want)
    print("max %.2e mean %.2e" % (err.max(), err.mean()))
    assert err.max() < 1e-2 and err.mean() < 2e-4


def _segment_decisions(before, after, r, S, N):
    """(W, nseg) bool: the segment's interior changed."""
    W = before.shape[0]
    SEG = 1 << S
    nseg = N // SEG
    out = np.zeros((W, nseg), bool)
    for w in range(W):
        for j in range(nseg):
            sl = (r[w] + j * SEG + np.arange(1, SEG)) % N
            out[w, j] = not np.array_equal(before[w][:, sl], after[w][:, sl])
    return out


@pytest.mark.parametrize("case", ["ho", "he3"])
def test_sweep_matches_oracle(gpu, case):
    _case("learned sweep %s" % case)
    rng = np.random.default_rng(2)
    if case == "ho":
        # paths drawn from the exact Gaussian of the discretised action
        system, W, N, S, beta, D = HarmOsc(), 64, 64, 4, 4.0, 1
        chol = np.linalg.cholesky(_ho_cov(N, beta))
        paths0 = (rng.standard_normal((W, N)) @ chol.T).astype(np.float32)[:, None, :]
        V, scale = _ho_V, 0.03
    else:
        system, W, N, S, beta, D = Lm2m2Cluster(3), 32, 128, 6, 10.0, 9
        base = system.initialPositions(W, rng)
        paths0 = (base[:, :, None] + 0.05 * rng.standard_normal((W, D, N))).astype(np.float32)
        V, scale = _lm2m2_V(system), 0.05
    k = _kernel(system, W, N, S, beta)
    k.setPaths(paths0.reshape(W, D * N))
    lp = LearnedProposal(k, batch=2048, seed=12345)
    theta = L.init_params(D, 5, scale_last=scale)
    lp.setParameters(theta)
    lp._sweepCount = 7
    lp.traceSweeps()
    lp.sweep(1)
    got = k.getPaths().reshape(W, D, N)
    tr = lp.getTrace()
    ref = paths0.copy()
    _, d = L.learned_sweep(ref, L.unflatten(theta, D), V, S=S, eps=beta / N, seed=12345, sweep=7,
                           atoms3d=(D == 9), return_detail=True)
    gpu_acc = _segment_decisions(paths0, got, d["r"], S, N)
    clear = np.abs(d["logA"] - np.log(np.maximum(d["u"], 1e-30))) > 0.05
    print("accepted %d/%d, clear %d" % (d["accept"].sum(), d["accept"].size, clear.sum()))
    SEG = 1 << S
    for w, j in list(zip(*np.nonzero(gpu_acc != d["accept"])))[:8]:
        print("mismatch w %d seg %d r %d: gpu (lq dK dV logA) %s oracle %s" % (
            w, j, d["r"][w], np.array2string(tr[w, j], precision=4),
            np.array2string(np.array([d["lq"][w, j], d["dK"][w, j], d["dV"][w, j], d["logA"][w, j]]), precision=4)))
    assert clear.mean() > 0.8
    assert 0 < d["accept"].sum() < d["accept"].size
    assert np.array_equal(gpu_acc[clear], d["accept"][clear])
    both = gpu_acc & d["accept"]
    for w, j in zip(*np.nonzero(both)):
        sl = (d["r"][w] + j * SEG + np.arange(1, SEG)) % N
        assert np.abs(got[w][:, sl] - ref[w][:, sl]).max() < 5e-3


def test_poor_network_keeps_ho_energy(gpu):
    """C1 sizes: a random network whose proposal means are off by about 0.08 sigma_n and whose
    widths are off by a factor exp(+-0.2) (22 % acceptance); the exact proposal density in the
    acceptance keeps <x^2> at the discretised-HO value."""
    _case("learned HO energy, poor network")
    system = HarmOsc()
    W, N, S, beta = 4096, 256, 4, 10.0
    k = _kernel(system, W, N, S, beta, opRuns=1, M=1500)
    k.setPaths(np.random.default_rng(0).uniform(-1, 1, (W, N)).astype(np.float32))
    k.run()
    lp = LearnedProposal(k, batch=2048, seed=9)
    lp.setParameters(L.init_params(1, 3, scale_last=0.03))
    lp.fallbackAcceptance = 0.0
    lp.sweep(200)
    ar = lp.lastAcceptance
    k.setSeriesCapacity(120)
    lp.run(120, 16)
    s = k.getSeries()[:, :, 0].astype(np.float64)
    per = s.mean(axis=0)
    E, se = per.mean(), per.std() / np.sqrt(W)
    exact = O.ho_discrete_x2(N, beta)
    print("E %.6f +- %.6f exact %.9f, learned AR %.3f" % (E, se, exact, ar))
    assert 0.05 < ar < 0.99
    assert lp.fallbackBlocks == 0
    assert abs(E - exact) < 4 * se


@pytest.mark.parametrize("D", [1, 9])
def test_training_step_matches_oracle(gpu, D):
    _case("learned training step D=%d" % D)
    rng = np.random.default_rng(4)
    if D == 1:
        system, W, N, S, beta = HarmOsc(), 256, 64, 4, 4.0
        paths = rng.standard_normal((W, D, N)).astype(np.float32)
    else:
        system, W, N, S, beta = Lm2m2Cluster(3), 64, 128, 6, 10.0
        base = system.initialPositions(W, rng)
        paths = (base[:, :, None] + 0.1 * rng.standard_normal((W, D, N))).astype(np.float32)
    k = _kernel(system, W, N, S, beta)
    k.setPaths(paths.reshape(W, D * N))
    B = 4096
    lp = LearnedProposal(k, batch=B, seed=77)
    theta = L.init_params(D, 1, scale_last=0.3)
    lp.setParameters(theta)
    loss = lp.train(1)
    g = lp.lastGradient()
    theta1 = lp.getParameters()
    SEG = 1 << S
    w, s0, p = L.train_rows(77, 1, B, W, N, SEG)
    h = p & -p
    n = S - np.log2(h).astype(np.int64)
    a = paths[w, :, (s0 + p - h) % N]
    b = paths[w, :, (s0 + p + h) % N]
    x = paths[w, :, (s0 + p) % N]
    f = L.features(a, b, n, S, D == 9, 1.0)
    sig = np.sqrt(SEG * (beta / N) / (2.0 * 2.0 ** n)).astype(np.float32)
    loss_o, go = L.nll_and_grad(L.unflatten(theta, D), f, a, b, x, sig, D)
    go = np.concatenate([go[k_].reshape(-1) for k_, _ in L.param_shapes(D)])
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    print("loss %.6f oracle %.6f, gradient rel err %.2e" % (loss, loss_o, rel))
    assert abs(loss - loss_o) < 1e-3 * abs(loss_o) + 1e-4
    assert rel < 3e-2
    th = theta.copy()
    L.adam_update(th, g, np.zeros_like(th), np.zeros_like(th), 1)
    np.testing.assert_allclose(theta1, th, rtol=0, atol=2e-6)


@pytest.mark.parametrize("width", [32, 64, 96])
def test_network_and_training_at_active_width(gpu, width):
    """Below the full 128 units the kernels read only the active units' weights and biases into LDS,
    and training computes the hidden layers on the active units only (GEMM sizes of the width, no
    gradient partials of inactive units read): the outputs match the oracle's forward pass of the
    masked parameters, one Adam step's loss and gradient match the oracle's, the gradient of every
    inactive parameter is exactly zero, and those parameters stay exactly zero."""
    _case("learned network and training at width %d" % width)
    D = 9
    rng = np.random.default_rng(6)
    system, W, N, S, beta = Lm2m2Cluster(3), 64, 128, 6, 10.0
    base = system.initialPositions(W, rng)
    paths = (base[:, :, None] + 0.1 * rng.standard_normal((W, D, N))).astype(np.float32)
    k = _kernel(system, W, N, S, beta)
    k.setPaths(paths.reshape(W, D * N))
    B = 4096
    lp = LearnedProposal(k, batch=B, seed=78, width=width)
    theta = L.init_params(D, 2, scale_last=0.3, active=width)
    lp.setParameters(theta)
    X = rng.standard_normal((1000, 2 * D + 1)).astype(np.float32)
    got = lp.evalNetwork(X)
    want, _ = L.mlp_forward(L.unflatten(theta, D), X)
    err = np.abs(got - want)
    loss = lp.train(1)
    g = lp.lastGradient()
    theta1 = lp.getParameters()
    SEG = 1 << S
    w, s0, p = L.train_rows(78, 1, B, W, N, SEG)
    h = p & -p
    n = S - np.log2(h).astype(np.int64)
    a = paths[w, :, (s0 + p - h) % N]
    b = paths[w, :, (s0 + p + h) % N]
    x = paths[w, :, (s0 + p) % N]
    f = L.features(a, b, n, S, True, 1.0)
    sig = np.sqrt(SEG * (beta / N) / (2.0 * 2.0 ** n)).astype(np.float32)
    loss_o, go = L.nll_and_grad(L.unflatten(theta, D), f, a, b, x, sig, D)
    go = np.concatenate([go[k_].reshape(-1) for k_, _ in L.param_shapes(D)])
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    print("outputs max %.2e mean %.2e; loss %.6f oracle %.6f, gradient rel err %.2e"
          % (err.max(), err.mean(), loss, loss_o, rel))
    assert err.max() < 1e-2 and err.mean() < 2e-4
    assert abs(loss - loss_o) < 1e-3 * abs(loss_o) + 1e-4
    assert rel < 3e-2
    assert np.array_equal(L.mask_width(g, D, width), g)
    assert np.array_equal(L.mask_width(theta1, D, width), theta1)
    th = theta.copy()
    L.adam_update(th, g, np.zeros_like(th), np.zeros_like(th), 1)
    np.testing.assert_allclose(theta1, th, rtol=0, atol=2e-6)


def test_training_improves_a_stiff_oscillator(gpu):
    """k = 25 (omega = 5): over a 16-slice segment the Levy bridge misses the oscillator's
    pull (omega T / 2 = 1.56), so online training must lower the NLL and raise the acceptance
    above the plain sampler's on the same paths."""
    _case("learned training, stiff oscillator")
    system = HarmOsc(k=25.0)
    W, N, S, beta = 2048, 256, 4, 10.0
    k = _kernel(system, W, N, S, beta, opRuns=1, M=1000)
    k.setPaths(np.random.default_rng(0).uniform(-0.3, 0.3, (W, N)).astype(np.float32))
    k.run()
    k.run()
    plain = k.getAcceptanceRate()
    lp = LearnedProposal(k, batch=16384, seed=21)
    first = lp.train(1)
    for _ in range(10):
        last = lp.train(20)
    ar = lp.sweep(20)
    print("NLL %.5f -> %.5f, acceptance plain %.3f learned %.3f" % (first, last, plain, ar))
    assert last < first - 0.01
    assert ar > plain


def test_fallback_below_twenty_percent_keeps_ho_energy(gpu):
    """C1 sizes with a network so poor that learned blocks accept under 20 %: the block after each
    such block adds a plain bisection block (the fallback; the first block looks at the acceptance of
    the sweeps before it), and <x^2> stays at the discretised-HO value 0.499950082044 within 4
    standard errors."""
    _case("learned fallback, C1")
    system = HarmOsc()
    W, N, S, beta = 4096, 256, 4, 10.0
    k = _kernel(system, W, N, S, beta, opRuns=1, M=1500)
    k.setPaths(np.random.default_rng(1).uniform(-1, 1, (W, N)).astype(np.float32))
    k.run()
    lp = LearnedProposal(k, batch=2048, seed=10, fallbackAcceptance=0.2)
    lp.setParameters(L.init_params(1, 4, scale_last=0.3))
    ar = lp.sweep(16)
    k.setSeriesCapacity(120)
    lp.run(120, 16)
    s = k.getSeries()[:, :, 0].astype(np.float64)
    per = s.mean(axis=0)
    E, se = per.mean(), per.std() / np.sqrt(W)
    exact = O.ho_discrete_x2(N, beta)
    print("E %.6f +- %.6f exact %.9f, learned AR %.3f, fallback blocks %d" % (E, se, exact, ar, lp.fallbackBlocks))
    assert ar < 0.2
    assert lp.fallbackBlocks > 0
    # the last launch was a fallback block of 16 plain sweeps: its acceptance is a probability
    assert 0.0 < k.getAcceptanceRate() < 1.0
    assert abs(E - exact) < 4 * se


def test_c5_learned_leg_matches_transfer_matrix(gpu):
    """The C5 system (x^2/2 + 0.1 x^4, N = 4096, beta = 40) sampled with the trained learned
    proposal: the virial E0 and the two-time correlator over 96 lags against the FP64 transfer
    matrix fixture (tests/golden/aho_c5.json), with the plain test's bounds (E0 within 4 standard
    errors, correlator max z < 5.5)."""
    import json
    import os
    from feynsimul_amd.physical_systems import AnharmOsc
    _case("learned C5 AHO N=4096")
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "aho_c5.json")))
    system = AnharmOsc(0.1)
    W, N, S, beta, lags = 2048, 4096, 4, 40.0, 96
    ka = KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=S, enableOperator=True,
                    enableCorrelator=True, correlators=("x1",), correlatorLags=lags, enableBisection=True,
                    enablePathShift=False, enableSingleNodeMove=False, enableParallelizePath=True,
                    enableGlobalPath=False, enableGlobalOldPath=False, enableBins=False,
                    metroStepsPerOperatorRun=4000, operatorRuns=1, nbrOfWalkersPerWorkGroup=1,
                    operators=(system.energyOp,), seed=31, sampler="wave", pathStorage="memory")
    k = PIMCKernel(ka)
    k.setPaths(np.random.default_rng(5).uniform(-1, 1, (W, N)).astype(np.float32))
    k.run()
    lp = LearnedProposal(k, batch=65536, seed=32)
    first = lp.train(1)
    last = lp.trainOnline(40)
    k.resetCorrelator()
    k.setSeriesCapacity(48)
    lp.run(48, 8)
    s = k.getSeries()[:, :, 0].astype(np.float64)
    per = s.mean(axis=0)
    E, se = per.mean(), per.std() / np.sqrt(W)
    mean, cse = k.getCorrelator()
    cor = np.array(fix["values"]["correlator"][:lags])
    z = np.abs(mean[0] - cor) / cse[0]
    print("NLL %.4f -> %.4f, learned AR %.3f, fallback %d; E0 %.5f +- %.5f vs %.5f; correlator max z %.2f"
          % (first, last, lp.lastAcceptance, lp.fallbackBlocks, E, se, fix["values"]["E_virial"], z.max()))
    assert lp.fallbackBlocks == 0 and lp.lastAcceptance > 0.5
    assert abs(E - fix["values"]["E_virial"]) < 4 * se
    assert z.max() < 5.5


def test_c4_shape_learned_sweep_and_training_step(gpu):
    """One learned sweep and one Adam step at the learned-proposal workload's shape: 3 He atoms,
    N = 1024, S = 6, 16384 chains, batch 65536.  Properties of the whole sweep (acceptance is a
    probability, paths finite, segment ends and rejected segments unchanged), and a 32-walker slice
    of it against the oracle's sweep of the same walkers (their random numbers are counter-based)."""
    import torch
    _case("learned He3 N=1024 x 16384 chains")
    system = Lm2m2Cluster(3)
    W, N, S, beta, D = 16384, 1024, 6, 10.0, 9
    rng = np.random.default_rng(6)
    base = system.initialPositions(W, rng)
    paths0 = (base[:, :, None] + 0.05 * rng.standard_normal((W, D, N))).astype(np.float32)
    k = _kernel(system, W, N, S, beta)
    k.setPaths(paths0.reshape(W, D * N))
    lp = LearnedProposal(k, batch=65536, seed=4242)
    theta = L.init_params(D, 8, scale_last=0.05)
    lp.setParameters(theta)
    lp._sweepCount = 3
    lp.traceSweeps()
    ar = lp.sweep(1)
    got = k.getPaths().reshape(W, D, N)
    tr = lp.getTrace()
    assert 0.0 < ar < 1.0 and np.isfinite(got).all() and np.isfinite(tr).all()
    # a segment either moved as a whole interior or not at all; log alpha decides (trace column 3)
    changed = (got != paths0).any(axis=1)                        # (W, N)
    SEG = 1 << S
    w0 = 5000
    sl = slice(w0, w0 + 32)
    ref = paths0[sl].copy()
    _, d = L.learned_sweep(ref, L.unflatten(theta, D), _lm2m2_V(system), S=S, eps=beta / N, seed=4242, sweep=3,
                           atoms3d=True, return_detail=True, walker0=w0)
    for i, w in enumerate(range(w0, w0 + 32)):
        ends = (d["r"][i] + SEG * np.arange(N // SEG)) % N
        assert not changed[w, ends].any()
    gpu_acc = _segment_decisions(paths0[sl], got[sl], d["r"], S, N)
    clear = np.abs(d["logA"] - np.log(np.maximum(d["u"], 1e-30))) > 0.05
    assert clear.mean() > 0.8
    assert np.array_equal(gpu_acc[clear], d["accept"][clear])
    np.testing.assert_allclose(tr[sl][..., 3][clear], d["logA"][clear], rtol=2e-2, atol=2e-2)
    both = gpu_acc & d["accept"]
    for i, j in zip(*np.nonzero(both)):
        s_ = (d["r"][i] + j * SEG + np.arange(1, SEG)) % N
        assert np.abs(got[w0 + i][:, s_] - ref[i][:, s_]).max() < 5e-3
    # one Adam step on 65536 rows of the whole ensemble
    loss = lp.train(1)
    th1 = lp.getParameters()
    print("C4 shape: acceptance %.3f, NLL %.4f" % (ar, loss))
    assert np.isfinite(loss) and np.isfinite(th1).all() and not np.array_equal(th1, theta)
    del lp, k
    torch.cuda.empty_cache()
