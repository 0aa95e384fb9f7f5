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
* The whole-path move: with a zero output layer its coarse density ratio equals
  the coarse kinetic change (closed form of the free bridge); the bridge has the
  Brownian-bridge mean and variance; with a poor network whole-path moves keep
  the exact Gaussian of a 256-slice HO ring.
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


# ---- the whole-path coarse move ----
def test_coarse_zero_output_is_levy_bridge():
    """o = 0: the coarse nodes come from the free-particle bridge of the whole path, whose density
    ratio is exactly the change of the coarse kinetic action, so log q_old - log q_new = dKeff and
    log alpha = -eps dV (plain bisection of the whole path)."""
    W, N, beta = 32, 512, 8.0
    eps = beta / N
    paths = np.random.default_rng(8).standard_normal((W, 1, N)).astype(np.float32)
    p = L.unflatten(L.init_params(1, 12, scale_last=0.0), 1)
    _, d = L.coarse_move(paths, p, _ho_V, eps=eps, seed=5, sweep=2, return_detail=True)
    np.testing.assert_allclose(d["lq"], d["dK"], rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(d["logA"], -eps * d["dV"], rtol=1e-4, atol=1e-3)


def test_bridge_law():
    """The Levy bridge between nodes a and b over 64 links of variance eps: mean a + (b - a) i / 64,
    variance eps i (64 - i) / 64 (the Brownian bridge), lane 63 exactly b."""
    W, M, eps = 4096, 2, 0.05
    cn = np.zeros((W, M + 1, 1), np.float32)
    cn[:, 1, 0] = 1.0
    z = L.fine_normals(3, np.arange(W), 1, M, 1)
    x = L.bridge(cn, z, eps)[..., 0].astype(np.float64)       # (W, M, 64)
    i = np.arange(1, 65)
    var = eps * i * (64 - i) / 64.0
    mean0 = i / 64.0                                          # interval 0: 0 -> 1
    assert np.array_equal(x[:, 0, 63], np.ones(W)) and np.array_equal(x[:, 1, 63], np.zeros(W))
    se = np.sqrt(var[:63] / W)
    assert (np.abs(x[:, 0, :63].mean(0) - mean0[:63]) < 5 * se).all()
    assert (np.abs(x[:, 1, :63].mean(0) - (1 - mean0[:63])) < 5 * se).all()
    np.testing.assert_allclose(x[:, 0, :63].var(0), var[:63], rtol=0.1)
    # increments: variance eps each, uncorrelated neighbours
    dx = np.diff(np.concatenate([np.zeros((W, 1)), x[:, 0, :]], axis=1), axis=1)
    assert abs(dx.var(0).mean() / eps - 1) < 0.03
    assert abs(np.mean(dx[:, 1:] * dx[:, :-1]) / eps) < 0.05


def _harmonic_precision(NL, eps, omega2):
    """Precision matrix of the interior slices 1 .. NL - 1 of one interval under the links' kinetic action
    plus eps omega2 x^2 / 2 on the interior (the ends fixed)."""
    T = np.zeros((NL - 1, NL - 1))
    for i in range(NL - 1):
        T[i, i] = 2.0 / eps + eps * omega2
        if i > 0:
            T[i, i - 1] = T[i - 1, i] = -1.0 / eps
    return T


@pytest.mark.parametrize("omega2", [0.0, 0.7, 2.5])
def test_harmonic_bridge_is_the_gaussian_conditional(omega2):
    """The bridge is linear in its normals: fed unit normals it gives the columns of a map A, and A A^T must
    be the inverse of the interior precision matrix (numpy's inverse of the tridiagonal matrix, not the
    closed forms the bridge uses); fed zero normals it must give the conditional mean T^-1 r with
    r = (a e_1 + b e_63) / eps.  A wrong rho, a missing end correction or a wrong increment variance fails."""
    NL, eps = 64, 0.04
    T = _harmonic_precision(NL, eps, omega2)
    C = np.linalg.inv(T)
    z = np.zeros((NL, 1, NL, 1), np.float32)
    z[np.arange(NL), 0, np.arange(NL), 0] = 1.0
    cn = np.zeros((NL, 2, 1), np.float32)
    A = L.bridge(cn, z, eps, omega2)[:, 0, :NL - 1, 0].astype(np.float64).T      # (63 slices, 64 normals)
    np.testing.assert_allclose(A @ A.T, C, rtol=2e-4, atol=2e-5 * C.max())
    a, b = 0.7, -1.3
    cn = np.array([[[a], [b]]], np.float32)
    x = L.bridge(cn, np.zeros((1, 1, NL, 1), np.float32), eps, omega2)[0, 0, :, 0].astype(np.float64)
    r = np.zeros(NL - 1)
    r[0], r[-1] = a / eps, b / eps
    np.testing.assert_allclose(x[:NL - 1], np.linalg.solve(T, r), rtol=1e-5, atol=1e-5)
    assert x[NL - 1] == np.float32(b)


@pytest.mark.parametrize("omega2", [0.0, 0.7, 2.5])
def test_coarse_action_is_the_minimum_of_the_interval_actions(omega2):
    """coarse_action of a cyclic node chain equals the sum over the intervals of the minimum, over the
    interior slices, of the links' kinetic action plus eps omega2 x^2 / 2 on the interior (computed by
    solving the interior's linear system), exactly: the Metropolis-Hastings ratio of the whole-path move
    rests on it."""
    NL, eps, M = 64, 0.04, 4
    T = _harmonic_precision(NL, eps, omega2)
    c = np.random.default_rng(2).standard_normal((3, M + 1, 1))
    c[:, M] = c[:, 0]
    tot = np.zeros(3)
    for w in range(3):
        for j in range(M):
            a, b = c[w, j, 0], c[w, j + 1, 0]
            r = np.zeros(NL - 1)
            r[0], r[-1] = a / eps, b / eps
            y = np.linalg.solve(T, r)
            xs = np.concatenate([[a], y, [b]])
            tot[w] += np.sum(np.diff(xs) ** 2) / (2 * eps) + 0.5 * eps * omega2 * np.sum(y ** 2)
    np.testing.assert_allclose(L.coarse_action(c, eps, omega2), tot, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("omega2", [0.0, 0.6, 1.5])
def test_coarse_log_alpha_is_the_hastings_ratio(omega2):
    """log alpha of the whole-path move, less the network's log q_old - log q_new, equals what the target and
    the bridges' densities give independently: -S(x') + S(x) + sum_j [log p(f_j | c) - log p(f'_j | c')],
    with S the discretised action of the whole path and p(f | a, b) the Gaussian of an interval's interior
    slices under its links' kinetic action and eps omega2 x^2 / 2 inside (numpy's linear algebra, not the
    closed forms of the move).  A node's reference potential counted twice, or a missing coarse term, fails
    it (C1's eps = 0.039, where such a slip biased <x^2> by 8 standard errors)."""
    W, N, beta = 12, 256, 10.0
    eps = beta / N
    V = lambda x: 0.5 * x[..., 0] ** 2 + 0.1 * x[..., 0] ** 4
    rng = np.random.default_rng(6)
    paths = (0.5 * rng.standard_normal((W, 1, N))).astype(np.float32)
    old = paths.copy()
    p = L.unflatten(L.init_params(1, 3, scale_last=0.2), 1)
    _, d = L.coarse_move(paths, p, V, eps=eps, seed=8, sweep=1, return_detail=True, omega2=omega2)
    new = old.copy()
    for w in range(W):
        new[w, 0, d["fine"][w].reshape(-1)] = d["new_fine"][w, :, :, 0].reshape(-1)
    T = _harmonic_precision(64, eps, omega2)
    _, logdet = np.linalg.slogdet(T / (2 * np.pi))

    def action(x):
        x = x[:, 0, :].astype(np.float64)
        return np.sum((np.roll(x, -1, axis=1) - x) ** 2, axis=1) / (2 * eps) + eps * V(x[..., None]).sum(axis=1)

    def log_bridges(x, r):
        x = x[:, 0, :].astype(np.float64)
        out = np.zeros(W)
        for w in range(W):
            for j in range(N // 64):
                s0 = r[w] + 64 * j
                a, b = x[w, s0 % N], x[w, (s0 + 64) % N]
                f = x[w, (s0 + 1 + np.arange(63)) % N]
                rr = np.zeros(63)
                rr[0], rr[-1] = a / eps, b / eps
                e = f - np.linalg.solve(T, rr)
                out[w] += -0.5 * e @ T @ e + 0.5 * logdet
        return out

    want = -action(new) + action(old) + log_bridges(old, d["r"]) - log_bridges(new, d["r"]) + d["lq"]
    print("max |logA - want| %.2e, max |logA| %.2e" % (np.abs(d["logA"] - want).max(), np.abs(want).max()))
    np.testing.assert_allclose(d["logA"], want, rtol=2e-3, atol=2e-3)


def test_decay_scan_is_a_decayed_prefix_sum():
    v = np.random.default_rng(4).standard_normal((2, 2, 64, 1)).astype(np.float32)
    q = 0.97
    ref = np.zeros_like(v, dtype=np.float64)
    acc = np.zeros((2, 2, 1))
    for i in range(64):
        acc = q * acc + v[:, :, i, :]
        ref[:, :, i, :] = acc
    np.testing.assert_allclose(L.decay_scan(v, q), ref, rtol=1e-5, atol=1e-5)


def test_coarse_training_rows_are_coarse_nodes():
    w, s0, p = L.train_rows(5, 3, 8192, 100, 1024, 1024, fine=L.FINE_L)
    assert (p % 64 == 0).all() and p.min() == 64 and p.max() == 1024 - 64
    assert len(np.unique(p)) == 1024 // 64 - 1
    assert w.min() >= 0 and w.max() < 100


@pytest.mark.parametrize("omega2", [0.0, 1.0, 0.4])
def test_coarse_poor_network_keeps_the_exact_law(omega2):
    """256-slice HO ring (four coarse nodes per path, two coarse levels), a random network whose coarse proposals are off
    by about 0.2 sigma: started in the exact Gaussian of the discretised action, whole-path moves keep
    it (1.8 % of the variance at most), with the Levy bridge (omega2 = 0), the harmonic reference bridge
    of the oscillator's own frequency (omega2 = 1: the fine slices drop out of log alpha) and a
    mismatched one (omega2 = 0.4).  Planting a dropped coarse kinetic term, a dropped reverse
    density or a bridge of variance 2 eps in the oracle moves the covariance by 23 %, 22 % and 16 % of
    the variance in the same run (omega2 = 0)."""
    N, beta, W, T = 256, 2.0, 8192, 24
    eps = beta / N
    Ce = _cov_exact(N, beta)
    p = L.unflatten(L.init_params(1, 21, hidden=32, scale_last=0.2), 1, hidden=32)
    rng = np.random.default_rng(3)
    paths = (rng.standard_normal((W, N)) @ np.linalg.cholesky(Ce).T).astype(np.float32)[:, None, :]
    acc = 0
    samples = []
    for t in range(T):
        acc += L.coarse_move(paths, p, _ho_V, eps=eps, seed=17, sweep=t, omega2=omega2).sum()
        if t >= T - 8:
            samples.append(paths[:, 0, ::16].astype(np.float64).copy())
    X = np.concatenate(samples)
    ar = acc / (T * W)
    C = np.cov(X.T)
    Cs = Ce[::16, ::16]
    print("acceptance %.3f, max cov error %.4f of %.4f" % (ar, np.abs(C - Cs).max(), Cs[0, 0]))
    assert 0.05 < ar < 0.95, ar
    assert np.abs(C - Cs).max() < 0.03 * Cs[0, 0]
    assert np.abs(X.mean(axis=0)).max() < 0.03 * np.sqrt(Cs[0, 0])


def test_lane_scan_is_a_prefix_sum():
    v = np.random.default_rng(1).standard_normal((3, 2, 64, 2)).astype(np.float32)
    np.testing.assert_allclose(L.lane_scan(v), np.cumsum(v.astype(np.float64), axis=2), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("active", [32, 64, 96])
def test_inactive_units_get_zero_gradient(active):
    """Units beyond the active width, zero in and out, stay exactly zero under training: their ReLU
    masks stop the gradient of their incoming weights and biases, and their zero activations stop
    that of their outgoing weights; the outputs equal those of the network cut to `active` units."""
    D = 1
    theta = L.init_params(D, 3, scale_last=0.3, active=active)
    p = L.unflatten(theta, D)
    rng = np.random.default_rng(0)
    a, b, x = (rng.standard_normal((256, D)).astype(np.float32) for _ in range(3))
    f = L.features(a, b, 2, 6, False, 1.0)
    _, g = L.nll_and_grad(p, f, a, b, x, np.float32(0.3), D)
    g = np.concatenate([g[k].reshape(-1) for k, _ in L.param_shapes(D)])
    assert np.array_equal(L.mask_width(g, D, active), g)
    cut = {k: (v[:active] if k in ("W1", "b1") else v[:active, :active] if k in ("W2", "W3")
               else v[:active] if k in ("b2", "b3") else v[:, :active] if k == "W4" else v) for k, v in p.items()}
    o_full, _ = L.mlp_forward(p, f)
    o_cut, _ = L.mlp_forward(cut, f)
    assert np.array_equal(o_full, o_cut)
