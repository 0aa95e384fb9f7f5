"""CPU restatement of the learned bisection proposal (numpy).

TEST INFRASTRUCTURE ONLY: tests and the bench's CPU baseline may use this
module as the checker; the product path (feynsimul_amd/) never imports it.

The reference has no learned proposal.  This module states the algorithm the
engine's learned sampler implements, so that the HIP kernels can be checked
against it.  It replaces the Levy midpoint draw of kernel.c:338-360 and
extends the accept test of kernel.c:376-377 with the proposal density ratio:

  level n = 1..S of a segment of 2^S links, node m with ends m - h, m + h,
  h = 2^(S-n), sigma_n^2 = 2^S eps / (2 * 2^n)           (kernel.c:605-619)
  features f = [a', b', n / S] (a, b the ends; a', b' centred on the atoms'
      mean over both ends for translation-invariant clusters, times inScale)
  MLP  h1 = relu(W1 f + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3),
       o = W4 h3 + b4          (BF16 weights, inputs and hidden activations,
                                FP32 accumulation)
  mu_d = (a_d + b_d) / 2 + sigma_n o_d,   s_d = clamp(o_{D+d}, -SMAX, SMAX),
  sigma_d = sigma_n exp(s_d),             new x_d = mu_d + sigma_d z_d
  log q(node) = sum_d (-z_d^2 / 2 - s_d)  (constants shared by old and new dropped)
  reverse density: the same network on the OLD path's ends at the old node
  log alpha = -(K_new - K_old) - eps (V_new - V_old) + log q_old - log q_new
      K = sum over the segment's 2^S links |x_{p+1} - x_p|^2 / (2 eps)
  accept iff !(exp(log alpha) < u)

With o = 0 the proposal is the Levy bridge and log alpha reduces to
-eps dV (the reference's rule).  Training minimises the mean negative log
likelihood sum_d (z_d^2 / 2 + s_d) of midpoints of equilibrium paths given
their ends, with Adam (Kingma & Ba 2015).

Random numbers: Philox4x32-10 (Salmon et al., SC'11), counter-based, keyed by
the 64-bit seed, so every (walker, sweep, slice) draw is independent of the
order in which the GPU processes rows.
"""
import numpy as np

HIDDEN = 128
SMAX = 3.0
M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
TAG_NORMAL, TAG_ACCEPT, TAG_OFFSET, TAG_TRAIN, TAG_FINE = 0x100, 0x200, 0x300, 0x400, 0x500
FINE_L = 64  # slices between the coarse nodes of the whole-path move


# ---------------------------------------------------------------------------
# Philox4x32-10
# ---------------------------------------------------------------------------
def philox4x32(ctr, key, rounds=10):
    """ctr: uint32 array (..., 4); key: uint32 array (..., 2). Returns (..., 4)."""
    c = [np.asarray(ctr[..., i], dtype=np.uint32) for i in range(4)]
    k0 = np.asarray(key[..., 0], dtype=np.uint32).copy()
    k1 = np.asarray(key[..., 1], dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        for r in range(rounds):
            if r > 0:
                k0 = (k0 + W0).astype(np.uint32)
                k1 = (k1 + W1).astype(np.uint32)
            p0 = M0 * c[0].astype(np.uint64)
            p1 = M1 * c[2].astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return np.stack(c, axis=-1)


def seed_key(seed):
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.array([s & 0xFFFFFFFF, s >> 32], dtype=np.uint32)


def _ctr(a, b, c, d):
    a, b, c, d = np.broadcast_arrays(*(np.asarray(v, dtype=np.uint64) & np.uint64(0xFFFFFFFF) for v in (a, b, c, d)))
    return np.stack([a, b, c, d], axis=-1).astype(np.uint32)


def u24(x, half=False):
    """uint32 -> float32 uniform from the top 24 bits: [0,1), or (0,1) with half=True."""
    v = (x >> np.uint32(8)).astype(np.float32)
    if half:
        v = v + np.float32(0.5)
    return (v * np.float32(2.0 ** -24)).astype(np.float32)


def normals(seed, walker, sweep, slices, dof):
    """z[..., d] for each slice: Box-Muller on Philox block (slice, walker, sweep, TAG_NORMAL | d // 4);
    d % 4 = 0, 1 use the pair (x0, x1) (cos, sin); d % 4 = 2, 3 use (x2, x3)."""
    key = seed_key(seed)
    out = np.empty(np.broadcast(walker, slices).shape + (dof,), dtype=np.float32)
    for blk in range((dof + 3) // 4):
        x = philox4x32(_ctr(slices, walker, sweep, TAG_NORMAL | blk), key)
        for pair in range(2):
            u1 = u24(x[..., 2 * pair], half=True)
            u2 = u24(x[..., 2 * pair + 1])
            rho = np.sqrt(np.float32(-2.0) * np.log(u1)).astype(np.float32)
            ang = (np.float32(2.0 * np.pi) * u2).astype(np.float32)
            for t, fn in ((0, np.cos), (1, np.sin)):
                d = 4 * blk + 2 * pair + t
                if d < dof:
                    out[..., d] = rho * fn(ang).astype(np.float32)
    return out


def offsets(seed, walkers, sweep, seg):
    x = philox4x32(_ctr(0, walkers, sweep, TAG_OFFSET), seed_key(seed))
    return (x[..., 0] & np.uint32(seg - 1)).astype(np.int64)


def accept_uniforms(seed, walkers, sweep, segs):
    x = philox4x32(_ctr(segs, walkers, sweep, TAG_ACCEPT), seed_key(seed))
    return u24(x[..., 0])


def train_rows(seed, step, batch, W, N, SEG, fine=1):
    """(walker, segment start slice, node position p in [1, SEG)) of each training row; with
    fine = FINE_L (the whole-path move, SEG = N) only the coarse nodes, p a multiple of FINE_L."""
    b = np.arange(batch, dtype=np.uint64)
    x = philox4x32(_ctr(b, step, np.uint64(int(step) >> 32), TAG_TRAIN), seed_key(seed))
    w = ((x[..., 0].astype(np.uint64) * np.uint64(W)) >> np.uint64(32)).astype(np.int64)
    s0 = (x[..., 1] & np.uint32(N - 1)).astype(np.int64)
    p = fine * (1 + ((x[..., 2].astype(np.uint64) * np.uint64(SEG // fine - 1)) >> np.uint64(32)).astype(np.int64))
    return w, s0, p


# ---------------------------------------------------------------------------
# BF16 and the network
# ---------------------------------------------------------------------------
def bf16(x):
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def param_shapes(dof, hidden=HIDDEN):
    nin, nout = 2 * dof + 1, 2 * dof
    return [("W1", (hidden, nin)), ("b1", (hidden,)), ("W2", (hidden, hidden)), ("b2", (hidden,)),
            ("W3", (hidden, hidden)), ("b3", (hidden,)), ("W4", (nout, hidden)), ("b4", (nout,))]


def param_count(dof, hidden=HIDDEN):
    return sum(int(np.prod(s)) for _, s in param_shapes(dof, hidden))


def unflatten(theta, dof, hidden=HIDDEN):
    out, o = {}, 0
    for name, shape in param_shapes(dof, hidden):
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def features(a, b, level, S, atoms3d, in_scale):
    """a, b: (..., D) float32 ends; returns (..., 2D+1) float32."""
    a = a.astype(np.float32)
    b = b.astype(np.float32)
    if atoms3d:
        A = a.shape[-1] // 3
        c = (a.reshape(a.shape[:-1] + (A, 3)).sum(-2) + b.reshape(b.shape[:-1] + (A, 3)).sum(-2)) \
            * np.float32(1.0 / (2 * A))
        a = a - np.tile(c, A)
        b = b - np.tile(c, A)
    lv = np.broadcast_to(np.float32(level) / np.float32(S), a.shape[:-1])[..., None].astype(np.float32)
    return np.concatenate([a * np.float32(in_scale), b * np.float32(in_scale), lv], axis=-1).astype(np.float32)


def mlp_forward(p, f, emulate_bf16=True, dtype=np.float32):
    """Returns (o, cache).  BF16 emulation rounds inputs, weights and hidden activations."""
    r = bf16 if emulate_bf16 else (lambda v: v)
    cast = lambda v: np.asarray(v, dtype=dtype)
    x0 = cast(r(f) if emulate_bf16 else f)
    Ws = [cast(r(p[k]) if emulate_bf16 else p[k]) for k in ("W1", "W2", "W3", "W4")]
    bs = [cast(p[k]) for k in ("b1", "b2", "b3", "b4")]
    hs = [x0]
    h = x0
    for L in range(3):
        a = h @ Ws[L].T + bs[L]
        h = np.maximum(a, 0).astype(dtype)
        if emulate_bf16:
            h = cast(bf16(h))
        hs.append(h)
    o = (h @ Ws[3].T + bs[3]).astype(dtype)
    return o, (hs, Ws)


def proposal_params(o, a, b, sigma_n, dof):
    mid = (a + b) * np.float32(0.5)
    mu = mid + np.float32(sigma_n) * o[..., :dof]
    s = np.clip(o[..., dof:2 * dof], -SMAX, SMAX).astype(np.float32)
    return mu.astype(np.float32), s, (np.float32(sigma_n) * np.exp(s)).astype(np.float32)


def nll_and_grad(p, f, a, b, x, sigma_n, dof, emulate_bf16=True, dtype=np.float32):
    """Mean over rows of sum_d (z_d^2/2 + s_d) and its gradient wrt every parameter."""
    o, (hs, Ws) = mlp_forward(p, f, emulate_bf16, dtype)
    B = f.shape[0]
    sig_n = np.asarray(sigma_n, dtype=dtype).reshape(-1, 1) * np.ones((B, 1), dtype=dtype)
    mid = (a.astype(dtype) + b.astype(dtype)) * 0.5
    mu = mid + sig_n * o[:, :dof]
    sraw = o[:, dof:2 * dof]
    s = np.clip(sraw, -SMAX, SMAX)
    z = (x.astype(dtype) - mu) / (sig_n * np.exp(s))
    loss = float(np.mean(np.sum(0.5 * z * z + s, axis=1)))
    do = np.zeros_like(o)
    do[:, :dof] = -z * np.exp(-s)
    inside = (sraw > -SMAX) & (sraw < SMAX)
    do[:, dof:2 * dof] = np.where(inside, 1.0 - z * z, 0.0)
    do /= B
    grads = {}
    g = do
    for L in (3, 2, 1, 0):
        # every layer's output gradient is stored in BF16 before it is used
        if emulate_bf16:
            g = bf16(g.astype(np.float32)).astype(dtype)
        grads["W%d" % (L + 1)] = (g.T @ hs[L]).astype(dtype)
        grads["b%d" % (L + 1)] = g.sum(axis=0).astype(dtype)
        if L > 0:
            dh = g @ Ws[L]
            g = (dh * (hs[L] > 0)).astype(dtype)
    return loss, grads


def adam_update(theta, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); arrays are float32 and updated in place."""
    m *= np.float32(b1)
    m += np.float32(1 - b1) * grad
    v *= np.float32(b2)
    v += np.float32(1 - b2) * grad * grad
    mh = m / np.float32(1 - b1 ** step)
    vh = v / np.float32(1 - b2 ** step)
    theta -= np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))


# ---------------------------------------------------------------------------
# One learned sweep over every walker (all segments of a walker share one offset)
# ---------------------------------------------------------------------------
def sigma_table(S, eps):
    SEG = 1 << S
    return [None] + [np.float32(np.sqrt(SEG * eps / (2.0 * (1 << n)))) for n in range(1, S + 1)]


def learned_sweep(paths, p, V, *, S, eps, seed, sweep, atoms3d=False, in_scale=1.0, emulate_bf16=True,
                  return_detail=False, walker0=0):
    """paths: float32 (W, D, N), updated in place.  V: vectorised potential (..., D) -> (...).
    ``walker0``: the walker number of paths[0] in the random-number counters (a slice of a
    larger ensemble draws what those walkers draw in the whole).

    Returns accepted segments per walker (and, with return_detail, the per-segment
    log alpha, uniforms and proposals).
    """
    W, D, N = paths.shape
    SEG = 1 << S
    nseg = N // SEG
    sig = sigma_table(S, eps)
    wi = np.arange(W)
    wg = wi + int(walker0)
    r = offsets(seed, wg, sweep, SEG)                                   # (W,)
    j = np.arange(nseg)
    pos = np.arange(SEG + 1)
    slices = (r[:, None, None] + j[None, :, None] * SEG + pos[None, None, :]) % N   # (W, nseg, SEG+1)
    old = paths[wi[:, None, None, None], np.arange(D)[None, None, None, :], slices[..., None]]
    old = old.astype(np.float32)                                        # (W, nseg, SEG+1, D)
    new = old.copy()
    lq = np.zeros((W, nseg), dtype=np.float32)
    for n in range(1, S + 1):
        h = SEG >> n
        m = np.arange(h, SEG, 2 * h)
        for which in ("new", "old") if n > 1 else ("new",):
            src = new if which == "new" else old
            a, b = src[:, :, m - h, :], src[:, :, m + h, :]
            f = features(a, b, n, S, atoms3d, in_scale)
            o, _ = mlp_forward(p, f.reshape(-1, f.shape[-1]), emulate_bf16)
            o = o.reshape(f.shape[:-1] + (o.shape[-1],))
            mu, s, sd = proposal_params(o, a, b, sig[n], D)
            if which == "new":
                z = normals(seed, wg[:, None, None], sweep, slices[:, :, m], D)
                new[:, :, m, :] = (mu + sd * z).astype(np.float32)
                lq -= np.sum(-0.5 * z * z - s, axis=(2, 3)).astype(np.float32)
                if n == 1:
                    # level 1: the reverse move conditions on the same ends
                    zo = ((old[:, :, m, :] - mu) / sd).astype(np.float32)
                    lq += np.sum(-0.5 * zo * zo - s, axis=(2, 3)).astype(np.float32)
            else:
                zo = ((old[:, :, m, :] - mu) / sd).astype(np.float32)
                lq += np.sum(-0.5 * zo * zo - s, axis=(2, 3)).astype(np.float32)
    inv2e = np.float32(1.0 / (2.0 * eps))
    kin = lambda x: np.sum(np.sum((x[:, :, 1:, :] - x[:, :, :-1, :]) ** 2, axis=3), axis=2) * inv2e
    dK = (kin(new) - kin(old)).astype(np.float32)
    dV = (np.sum(V(new[:, :, 1:SEG, :]), axis=2) - np.sum(V(old[:, :, 1:SEG, :]), axis=2)).astype(np.float32)
    logA = (-dK - np.float32(eps) * dV + lq).astype(np.float32)
    u = accept_uniforms(seed, wg[:, None], sweep, j[None, :])
    with np.errstate(over="ignore"):
        acc = ~(np.exp(logA) < u)
    for w in range(W):
        for k in np.nonzero(acc[w])[0]:
            paths[w][:, slices[w, k, 1:SEG]] = new[w, k, 1:SEG, :].T
    if return_detail:
        return acc.sum(axis=1), dict(logA=logA, u=u, accept=acc, new=new, old=old, slices=slices, lq=lq, dK=dK,
                                     dV=dV, r=r)
    return acc.sum(axis=1)


# ---------------------------------------------------------------------------
# One whole-path coarse move per walker
# ---------------------------------------------------------------------------
def fine_normals(seed, walker, sweep, M, dof):
    """z[w, j, lane, d]: the bridge normals of interval j, lane 0..63; normal k = j dof + d of a lane
    is Box-Muller on Philox block (64 (k // 4) + lane, walker, sweep, TAG_FINE), pair (k // 2) % 2,
    cosine for even k, sine for odd k."""
    walker = np.asarray(walker)
    K = M * dof
    nb = (K + 3) // 4
    lane = np.arange(FINE_L)
    blk = np.arange(nb)
    x = philox4x32(_ctr(blk[None, :, None] * FINE_L + lane[None, None, :], walker[:, None, None], sweep, TAG_FINE),
                   seed_key(seed))                                     # (W, nb, 64, 4)
    k = np.arange(K)
    pair = (k >> 1) & 1
    xa = x[:, k >> 2, :, :][..., 0:4]
    w1 = np.take_along_axis(xa, (2 * pair)[None, :, None, None].repeat(FINE_L, 2), axis=3)[..., 0]
    w2 = np.take_along_axis(xa, (2 * pair + 1)[None, :, None, None].repeat(FINE_L, 2), axis=3)[..., 0]
    rho = np.sqrt(np.float32(-2.0) * np.log(u24(w1, half=True))).astype(np.float32)
    ang = (np.float32(2.0 * np.pi) * u24(w2)).astype(np.float32)
    z = np.where((k & 1)[None, :, None] == 1, rho * np.sin(ang), rho * np.cos(ang)).astype(np.float32)  # (W, K, 64)
    return z.reshape(walker.shape[0], M, dof, FINE_L).transpose(0, 1, 3, 2)    # (W, M, 64, D)


def lane_scan(v):
    """Inclusive prefix sum over axis 2 (64 lanes) in the GPU's order of additions: a log-step scan
    inside each 16-lane row (shifts 1, 2, 4, 8), then row 0's total into row 1 and row 2's into row 3,
    then the first half's total into rows 2 and 3."""
    v = v.astype(np.float32).copy()
    lane = np.arange(FINE_L)
    for o in (1, 2, 4, 8):
        t = np.zeros_like(v)
        t[:, :, o:, :] = v[:, :, :-o, :]
        t[:, :, (lane % 16) < o, :] = 0.0
        v = (v + t).astype(np.float32)
    t = np.zeros_like(v)
    t[:, :, 16:32, :] = v[:, :, 15:16, :]
    t[:, :, 48:64, :] = v[:, :, 47:48, :]
    v = (v + t).astype(np.float32)
    t = np.zeros_like(v)
    t[:, :, 32:64, :] = v[:, :, 31:32, :]
    return (v + t).astype(np.float32)


def harmonic_coefs(eps, omega2):
    """The harmonic reference bridge's constants (float64): rho with rho + 1/rho = 2 + eps^2 omega^2 (the
    AR(1) chain x_{i+1} = rho x_i + sqrt(rho eps) z_i, conditioned on its end, is the Gaussian bridge of
    the links' kinetic action plus eps omega^2 x^2 / 2 on the interior slices), rho^64 and
    V_64 = rho eps (1 - rho^128) / (1 - rho^2), the variance of x_64 given x_0.  omega2 = 0: the free
    bridge, rho = 1, V_64 = 64 eps."""
    if omega2 == 0.0:
        return 1.0, 1.0, eps * FINE_L
    e2w2 = eps * eps * omega2
    rho = 1.0 + 0.5 * e2w2 - np.sqrt(e2w2 * (1.0 + 0.25 * e2w2))
    return rho, rho ** FINE_L, rho * eps * (1.0 - rho ** (2 * FINE_L)) / (1.0 - rho * rho)


def decay_scan(v, q):
    """Inclusive decayed prefix sum over axis 2 (64 lanes), w_i = sum_{m <= i} q^(i - m) v_m, in the
    order of `lane_scan` with each shifted term weighted by q to the power of its distance (the GPU's
    FMA form; q = 1 is `lane_scan`)."""
    v = v.astype(np.float32).copy()
    lane = np.arange(FINE_L)
    for o in (1, 2, 4, 8):
        t = np.zeros_like(v)
        t[:, :, o:, :] = v[:, :, :-o, :]
        t[:, :, (lane % 16) < o, :] = 0.0
        v = (v + np.float32(q ** o) * t).astype(np.float32)
    w1 = np.float32(q) ** ((lane & 15) + 1).astype(np.float32)
    t = np.zeros_like(v)
    t[:, :, 16:32, :] = v[:, :, 15:16, :]
    t[:, :, 48:64, :] = v[:, :, 47:48, :]
    v = (v + w1[None, None, :, None] * t).astype(np.float32)
    w2 = np.float32(q) ** np.maximum(lane - 31, 0).astype(np.float32)
    t = np.zeros_like(v)
    t[:, :, 32:64, :] = v[:, :, 31:32, :]
    return (v + w2[None, None, :, None] * t).astype(np.float32)


def bridge(cn, z, eps, omega2=0.0):
    """Fine slices of every interval: cn (W, M + 1, D) nodes, z (W, M, 64, D) -> (W, M, 64, D), lane 63
    the node j + 1.  omega2 = 0: the Levy bridge x_i = c_j + (c_{j+1} - c_j) i / 64 + W_i - (i / 64) W_64,
    W the prefix sums (`lane_scan`) of the increments sqrt(eps) z.  omega2 > 0: the harmonic reference
    bridge (`harmonic_coefs`) x_i = c_j rho^i + W_i + g_i (c_{j+1} - c_j rho^64 - W_64), W the AR(1)
    chain of `decay_scan` on sqrt(rho eps) z, g_i = rho^(64 - i) (1 - rho^(2i)) / (1 - rho^128)."""
    ca, cb = cn[:, :-1, None, :], cn[:, 1:, None, :]
    if omega2 == 0.0:
        v = lane_scan((np.float32(np.sqrt(np.float32(eps))) * z).astype(np.float32))
        tot = v[:, :, FINE_L - 1:FINE_L, :]
        fi = (np.arange(1, FINE_L + 1, dtype=np.float32) / np.float32(FINE_L))[None, None, :, None]
        x = (ca + (cb - ca) * fi + v - fi * tot).astype(np.float32)
    else:
        rho, rhoL, _ = harmonic_coefs(eps, omega2)
        v = decay_scan((np.float32(np.sqrt(rho * eps)) * z).astype(np.float32), rho)
        tot = v[:, :, FINE_L - 1:FINE_L, :]
        k = np.arange(1, FINE_L + 1, dtype=np.float64)
        pk = (rho ** k).astype(np.float32)[None, None, :, None]
        lam = -np.log(rho)
        g = (np.sinh(lam * k) / np.sinh(lam * FINE_L)).astype(np.float32)    # rho^(64-k) (1 - rho^2k) / (1 - rho^128)
        g = g[None, None, :, None]
        x = (ca * pk + v + g * (cb - ca * np.float32(rhoL) - tot)).astype(np.float32)
    x[:, :, FINE_L - 1, :] = cb[:, :, 0, :]
    return x


def link_action(c, eps, omega2=0.0):
    """sum_j |c_{j+1} - rho^64 c_j|^2 / (2 V_64) over a cyclic node chain c (W, M + 1, D) (node M = node 0),
    float64 (omega2 = 0: sum_j |c_{j+1} - c_j|^2 / (2 eps 64)): the coarse action the whole-path move
    scores beside V - omega2 |x|^2 / 2 on every slice, nodes included."""
    rho, rhoL, VL = harmonic_coefs(eps, omega2)
    c = c.astype(np.float64)
    return np.sum((c[:, 1:, :] - rhoL * c[:, :-1, :]) ** 2, axis=(1, 2)) / (2.0 * VL)


def coarse_action(c, eps, omega2=0.0):
    """The minimum over the fine slices of each interval's links' kinetic action plus eps omega^2 |x|^2 / 2
    on its interior slices, summed over the intervals of a cyclic node chain c, float64:
    link_action - eps omega^2 / 2 sum_j |c_j|^2 (the reference potential of the nodes themselves is not in
    it; the move scores it with the rest of V - omega2 |x|^2 / 2)."""
    k = link_action(c, eps, omega2)
    if omega2 != 0.0:
        k -= 0.5 * eps * omega2 * np.sum(c[:, :-1, :].astype(np.float64) ** 2, axis=(1, 2))
    return k


def coarse_move(paths, p, V, *, eps, seed, sweep, atoms3d=False, in_scale=1.0, emulate_bf16=True,
                return_detail=False, walker0=0, omega2=0.0):
    """The whole-path move: paths float32 (W, D, N), N = 64 M, M = 2^LC, updated in place.

    Anchor r (Philox (0, walker, sweep, TAG_OFFSET) mod N); coarse nodes c_j = x(r + 64 j), node M =
    node 0.  Levels n = 1..LC of the bisection of the whole path (S = log2 N: sigma_n^2 = N eps /
    (2 * 2^n), level feature n / S) draw the new nodes from the network with normals counted by
    (node j, walker, sweep, TAG_NORMAL | block), and score the old ones (the reverse rows); a bridge
    (`bridge`, normals `fine_normals`) fills each interval.  The bridge samples the fine links' kinetic
    action plus the harmonic reference eps omega2 |x|^2 / 2 on its interior exactly given the nodes, so
        log alpha = -eps [(V - omega2 |x|^2 / 2)_new - (...)_old] - (Keff_new - Keff_old) + log q_old - log q_new
    with the potential sums over every slice and Keff = `link_action` (omega2 = 0: the Levy bridge,
    Keff = sum_j |c_{j+1} - c_j|^2 / (2 eps 64)).
    Returns accepted moves per walker (and, with return_detail, the parts of log alpha)."""
    W, D, N = paths.shape
    S = int(np.log2(N))
    M = N // FINE_L
    LC = S - int(np.log2(FINE_L))
    assert (1 << S) == N and M >= 2
    wi = np.arange(W)
    wg = wi + int(walker0)
    r = offsets(seed, wg, sweep, N)
    nodes = (r[:, None] + FINE_L * np.arange(M + 1)[None, :]) % N                 # (W, M + 1)
    cold = paths[wi[:, None, None], np.arange(D)[None, None, :], nodes[..., None]].astype(np.float32)
    cnew = cold.copy()
    lq = np.zeros(W, dtype=np.float32)
    for n in range(1, LC + 1):
        h = M >> n
        m = np.arange(h, M, 2 * h)
        sig = np.float32(np.sqrt(N * eps / (2.0 * (1 << n))))
        for which in ("new", "old") if n > 1 else ("new",):
            src = cnew if which == "new" else cold
            a, b = src[:, m - h, :], src[:, m + h, :]
            f = features(a, b, n, S, atoms3d, in_scale)
            o, _ = mlp_forward(p, f.reshape(-1, f.shape[-1]), emulate_bf16)
            o = o.reshape(f.shape[:-1] + (o.shape[-1],))
            mu, s, sd = proposal_params(o, a, b, sig, D)
            if which == "new":
                z = normals(seed, wg[:, None], sweep, m[None, :], D)
                cnew[:, m, :] = (mu + sd * z).astype(np.float32)
                lq -= np.sum(-0.5 * z * z - s, axis=(1, 2)).astype(np.float32)
                if n == 1:
                    zo = ((cold[:, m, :] - mu) / sd).astype(np.float32)
                    lq += np.sum(-0.5 * zo * zo - s, axis=(1, 2)).astype(np.float32)
            else:
                zo = ((cold[:, m, :] - mu) / sd).astype(np.float32)
                lq += np.sum(-0.5 * zo * zo - s, axis=(1, 2)).astype(np.float32)
        cnew[:, M, :] = cnew[:, 0, :]
    xn = bridge(cnew, fine_normals(seed, wg, sweep, M, D), eps, omega2)           # (W, M, 64, D)
    fine = (r[:, None, None] + FINE_L * np.arange(M)[None, :, None] + 1 + np.arange(FINE_L)[None, None, :]) % N
    xo = paths[wi[:, None, None, None], np.arange(D)[None, None, None, :], fine[..., None]].astype(np.float32)
    Veff = lambda x: V(x) - 0.5 * omega2 * np.sum(x * x, axis=-1)
    dV = (Veff(xn.astype(np.float64)).sum(axis=(1, 2)) - Veff(xo.astype(np.float64)).sum(axis=(1, 2))).astype(np.float32)
    dK = (link_action(cnew, eps, omega2) - link_action(cold, eps, omega2)).astype(np.float32)
    logA = (-dK - np.float32(eps) * dV + lq).astype(np.float32)
    u = accept_uniforms(seed, wg, sweep, 0)
    with np.errstate(over="ignore"):
        acc = ~(np.exp(logA) < u)
    for w in np.nonzero(acc)[0]:
        paths[w][:, fine[w].reshape(-1)] = xn[w].reshape(-1, D).T
    if return_detail:
        return acc.astype(np.int64), dict(logA=logA, u=u, accept=acc, new_nodes=cnew, old_nodes=cold, fine=fine,
                                          new_fine=xn, lq=lq, dK=dK, dV=dV, r=r)
    return acc.astype(np.int64)


def mask_width(theta, dof, active, hidden=HIDDEN):
    """Zero every parameter of the hidden units beyond `active` (their incoming weights and biases and
    their outgoing weights): the network then computes with `active` units exactly."""
    p = {k: v.copy() for k, v in unflatten(np.asarray(theta, np.float32), dof, hidden).items()}
    for L in ("1", "2", "3"):
        p["W" + L][active:, :] = 0.0
        p["b" + L][active:] = 0.0
    for L in ("2", "3", "4"):
        p["W" + L][:, active:] = 0.0
    return np.concatenate([p[k].reshape(-1) for k, _ in param_shapes(dof, hidden)]).astype(np.float32)


def init_params(dof, seed, hidden=HIDDEN, zero_last=True, scale_last=0.0, active=None):
    """He-normal hidden layers, zero biases; the last layer is zero (the Levy bridge) unless
    scale_last is given.  `active`: only the first `active` hidden units are used (the rest zero)."""
    rng = np.random.default_rng(seed)
    p = {}
    fan = {"W1": 2 * dof + 1, "W2": active or hidden, "W3": active or hidden, "W4": active or hidden}
    for name, shape in param_shapes(dof, hidden):
        if name.startswith("b"):
            p[name] = np.zeros(shape, np.float32)
        else:
            p[name] = (rng.standard_normal(shape) * np.sqrt(2.0 / fan[name])).astype(np.float32)
    if zero_last:
        p["W4"] = (rng.standard_normal(p["W4"].shape) * scale_last).astype(np.float32)
    theta = np.concatenate([p[k].reshape(-1) for k, _ in param_shapes(dof, hidden)]).astype(np.float32)
    return mask_width(theta, dof, active, hidden) if active else theta
