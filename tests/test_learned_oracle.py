"""Pins of the learned-proposal oracle (oracle/learned_oracle.py), CPU only.

* Philox4x32-10: the known-answer vectors published with Random123
  (Salmon et al., SC'11, kat_vectors).
* BF16 rounding: torch's float32 -> bfloat16 conversion (round to nearest even).
* The NLL gradient: central finite differences in FP64.
* Adam: the closed form of the first step, lr * g / (|g| + eps).
* The MH ratio: with a zero output layer the proposal is the Levy bridge and
  log alpha equals -eps dV; with a deliberately poor random network the chain
  still samples the exact Gaussian of the discretised HO action (brute force
  on a 4-slice ring).
"""
import numpy as np
import pytest

from oracle import learned_oracle as L


def test_philox_known_answers():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        got = L.philox4x32(np.array(ctr, np.uint32), np.array(key, np.uint32))
        assert [int(v) for v in got] == list(want)


def test_bf16_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(100000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 100000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, 1.0 + 3 * 2 ** -8, -0.0, 0.0],
                                 np.float32)]).astype(np.float32)
    ours = L.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    assert L.bf16(np.float32(1.0 + 2 ** -8)) == 1.0          # tie to even
    assert L.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def _batch(D, B, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((B, D)).astype(np.float32)
    b = rng.standard_normal((B, D)).astype(np.float32)
    x = ((a + b) / 2 + 0.3 * rng.standard_normal((B, D))).astype(np.float32)
    lv = rng.integers(1, 5, B)
    f = L.features(a, b, lv, 4, D % 3 == 0, 1.0)
    sig = np.sqrt(16 * 0.05 / (2.0 * 2.0 ** lv)).astype(np.float32)
    return f, a, b, x, sig


@pytest.mark.parametrize("D", [1, 9])
def test_nll_gradient_finite_differences(D):
    theta = L.init_params(D, 3, hidden=16, scale_last=0.3).astype(np.float64)
    f, a, b, x, sig = _batch(D, 64, 1)
    p = L.unflatten(theta, D, hidden=16)
    _, g = L.nll_and_grad(p, f.astype(np.float64), a, b, x, sig, D, emulate_bf16=False, dtype=np.float64)
    gflat = np.concatenate([g[k].reshape(-1) for k, _ in L.param_shapes(D, 16)])
    rng = np.random.default_rng(2)
    idx = rng.choice(theta.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        lp, _ = L.nll_and_grad(L.unflatten(tp, D, 16), f.astype(np.float64), a, b, x, sig, D, False, np.float64)
        lm, _ = L.nll_and_grad(L.unflatten(tm, D, 16), f.astype(np.float64), a, b, x, sig, D, False, np.float64)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gflat[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, gflat[i])


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(5)
    theta = rng.standard_normal(1000).astype(np.float32)
    g = rng.standard_normal(1000).astype(np.float32)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    t0 = theta.copy()
    L.adam_update(theta, g, m, v, 1, lr=1e-3)
    np.testing.assert_allclose(theta, t0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=2e-7)
    # second step with the same gradient: m_hat = g, v_hat = g^2 again
    L.adam_update(theta, g, m, v, 2, lr=1e-3)
    np.testing.assert_allclose(theta, t0 - 2e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=4e-7)


def _ho_V(x):
    return 0.5 * x[..., 0] * x[..., 0]


def test_zero_output_layer_is_levy_bridge():
    """o = 0: the proposal is the Levy bridge, so -dK + log q_old - log q_new cancels and
    log alpha = -eps dV (kernel.c:376-377)."""
    W, N, S, beta = 64, 64, 4, 4.0
    eps = beta / N
    rng = np.random.default_rng(7)
    paths = rng.standard_normal((W, 1, N)).astype(np.float32)
    theta = L.init_params(1, 11, scale_last=0.0)
    p = L.unflatten(theta, 1)
    _, d = L.learned_sweep(paths, p, _ho_V, S=S, eps=eps, seed=99, sweep=3, return_detail=True)
    np.testing.assert_allclose(d["logA"], -eps * d["dV"], atol=2e-3)
    # the proposal is the closed-form bridge: midpoint law N(mean of ends, sigma_1^2)
    SEG = 1 << S
    mid = d["new"][:, :, SEG // 2, 0] - 0.5 * (d["old"][:, :, 0, 0] + d["old"][:, :, SEG, 0])
    sig1 = np.sqrt(SEG * eps / 4.0)
    assert abs(mid.std() / sig1 - 1) < 0.1


def _cov_exact(N, beta):
    eps = beta / N
    A = np.zeros((N, N))
    for i in range(N):
        A[i, i] = 2 / eps + eps
        A[i, (i + 1) % N] -= 1 / eps
        A[i, (i - 1) % N] -= 1 / eps
    return np.linalg.inv(A)


def test_poor_network_keeps_the_exact_law():
    """4-slice HO ring, S = 2, a random network whose proposals are off by about one
    sigma: started in the exact Gaussian of the discretised action, the chain must stay
    there.  Planting a dropped reverse density, a dropped -log sigma or a dropped kinetic
    term in the oracle moves the covariance by 17 %, 6 % and 50 % (and the mean by
    0.30, 0.04 and 0.21 sigma) in the same run."""
    N, S, beta, W, T = 4, 2, 2.0, 16384, 40
    eps = beta / N
    Ce = _cov_exact(N, beta)
    theta = L.init_params(1, 21, hidden=32, scale_last=0.2)
    p = L.unflatten(theta, 1, hidden=32)
    rng = np.random.default_rng(3)
    paths = (rng.standard_normal((W, N)) @ np.linalg.cholesky(Ce).T).astype(np.float32)[:, None, :]
    acc = 0
    samples = []
    for t in range(T):
        acc += L.learned_sweep(paths, p, _ho_V, S=S, eps=eps, seed=17, sweep=t).sum()
        if t >= T - 10:
            samples.append(paths[:, 0, :].astype(np.float64).copy())
    X = np.concatenate(samples)
    ar = acc / (T * W)
    assert 0.05 < ar < 0.95, ar
    assert np.abs(np.cov(X.T) - Ce).max() < 0.035 * Ce[0, 0]
    assert np.abs(X.mean(axis=0)).max() < 0.025 * np.sqrt(Ce[0, 0])
