"""Pins of the CPU oracle (oracle/pimc_oracle.c) -- CPU only.

What fixes each oracle function, so that a plausible mistake fails a check:

* xorshift128 (kernel.c:214-239): Marsaglia's published xor128 outputs for
  his default seed; the FP32 randFloat end points ([0, 1] inclusive).
* RANLUX SWB core and luxury skipping (pyopencl-ranluxcl.cl:363-557):
  libstdc++'s std::ranlux24_base / std::discard_block_engine loaded with the
  same lagged state produce the same 24-bit stream (small values checked on
  their top 24 bits, the low-bit fill below 2^-12 is RANLUXCL's own).
* ranluxcl_init (pyopencl-ranluxcl.cl:597-725): the expansion equals the
  192-bit concatenation of three xorshift64(13, 7, 17) outputs cut into
  24-bit fields, and the warm-up is 32 x 24 steps of the std engine.
* ranluxcl64: 48-bit composition of two float outputs, checked against the
  float stream.
* bisection move + Metropolis rule (kernel.c:258-399): the reference's own
  printed acceptance rates (doc/tutorial/index.rst:239-256); free particle
  (V = 0) accepts every move; tiny paths against exact Gaussian covariances
  and transfer-matrix quadrature, for both orderings.
* HO and AHO estimators: closed forms and FP64 transfer matrices.
* LM2M2 potential: literature values of the formula and finite-difference
  derivatives of the virial operator; the generated strings of 3-, 4-, 6-
  and 8-atom clusters against the FP64 pair sum.
* histograms (kernel.c:478-506): the observable histogram (pair distances)
  and the DOF-grid bins after one measurement equal an independent numpy
  recount of the oracle's own final paths, bin for bin up to values within
  a few FP32 ulps of a bin edge; a one-bin shift fails the check.
Parity unpinned: none of the functions above.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from feynsimul_amd.physical_systems import HarmOsc, AnharmOsc, Lm2m2Cluster
from feynsimul_amd.physical_systems.lm2m2 import lm2m2_kelvin, mass_scale


class _Free:
    DOF = 1
    potential = "0.0"
    userCode = ""
    energyOp = "x1 * x1"


@pytest.fixture(scope="module")
def ho_lib():
    return O.build(HarmOsc(), 64, 5.0, operators=("1.0 * x1 * x1",))


# ----------------------------------------------------------------------------
# xorshift128
# ----------------------------------------------------------------------------
def test_xorshift_marsaglia_known_answer(ho_lib):
    rows = np.array([[123456789, 362436069, 521288629, 88675123]], dtype=np.uint32)
    ints, flt = O.xs_stream(ho_lib, rows, 5)
    assert list(ints[0]) == [3701687786, 458299110, 2500872618, 3633119408, 516391518]
    np.testing.assert_array_equal(flt[0], (ints[0].astype(np.float32) * np.float32(2.0 ** -32)))


def test_xorshift_float_endpoints(ho_lib):
    # state whose next w is 0 and one whose next w is 2^32 - 1: x = y = z = 0
    rows = np.array([[0, 0, 0, 0], [0, 0, 0, 0xFFFFFFFF]], dtype=np.uint32)
    ints, flt = O.xs_stream(ho_lib, rows, 1)
    assert ints[0, 0] == 0 and flt[0, 0] == 0.0
    assert ints[1, 0] == 0xFFFFFFFF ^ (0xFFFFFFFF >> 19)
    # float32(w) rounds to 2^32 for w >= 2^32 - 128, so randFloat reaches 1.0
    assert np.float32(np.uint32(0xFFFFFF80)) * np.float32(2.0 ** -32) == 1.0
    assert flt[1, 0] == np.float32(np.uint32(ints[1, 0])) * np.float32(2.0 ** -32)


def test_xorshift_double_conversion():
    lib = O.build(HarmOsc(), 64, 5.0, operators=("1.0 * x1 * x1",), double=True)
    rows = O.xorshift_rows(3, 0, 4)
    ints, flt = O.xs_stream(lib, rows, 50)
    np.testing.assert_array_equal(flt, ints.astype(np.float64) * 2.328306437080797e-10)


def test_seed_rows_nonzero_and_distinct():
    rows = O.xorshift_rows(0, 0, 4096)
    assert (rows != 0).any(axis=1).all()
    assert len({tuple(r) for r in rows}) == 4096


# ----------------------------------------------------------------------------
# RANLUX
# ----------------------------------------------------------------------------
def _top24(v):
    return np.floor(v.astype(np.float64) * 16777216.0).astype(np.int64)


@pytest.mark.parametrize("lux", [0, 1, 2, 3, 4])
def test_ranlux_matches_libstdcxx(ho_lib, lux):
    exe = O.std_check_exe()
    seeds = O.ranlux_seeds(2024, 0, 3)
    st = O.ranlux_init(ho_lib, seeds)
    nquad = 250
    for i in range(3):
        std = O.ranlux_std_stream(st[i], O.nskip_of_lux(lux), nquad * 4, exe)
        ours = O.ranlux_stream(ho_lib, st[i:i + 1].copy(), O.nskip_of_lux(lux), nquad)[0]
        np.testing.assert_array_equal(_top24(ours), std)
        small = ours < 2.0 ** -12
        # above 2^-12 the output is exactly the 24-bit state value
        np.testing.assert_array_equal(ours[~small].astype(np.float64) * 16777216.0, std[~small])


def test_ranlux_small_value_fill(ho_lib):
    # a state of zeros except the lag-10 entry produces tiny values that get the fill
    st = np.zeros(1, dtype=O.RANLUX_DTYPE)
    st["s"][0, 9] = 1.0 / 16777216.0  # s10
    st["s"][0, 8] = 0.5               # s09, supplies the low bits
    out = O.ranlux_stream(ho_lib, st, 0, 1)[0]
    assert out[0] == np.float32(2.0 ** -24 + 2.0 ** -24 * 0.5)


def test_ranlux_init_expansion_and_warmup(ho_lib):
    exe = O.std_check_exe()
    for x in (1, 0x123456789ABCDEF, (1 << 64) - 1):
        seeds = np.array([x], dtype=np.uint64)
        ex = O.ranlux_expand(ho_lib, seeds)[0]
        # three xorshift64 (13, 7, 17) outputs, concatenated, cut into 8 fields per output triple
        M = (1 << 64) - 1
        v = x
        fields = []
        for _ in range(3):
            words = []
            for _ in range(3):
                v ^= (v << 13) & M
                v ^= v >> 7
                v ^= (v << 17) & M
                words.append(v)
            big = (words[0] << 128) | (words[1] << 64) | words[2]
            fields += [(big >> (192 - 24 * (k + 1))) & 0xFFFFFF for k in range(8)]
        np.testing.assert_array_equal(np.round(ex["s"] * 16777216.0).astype(np.int64), fields)
        assert float(ex["carry"]) == (2.0 ** -24 if fields[23] == 0 else 0.0)
        warm = O.ranlux_init(ho_lib, seeds)[0]
        # the warm-up is 32 x 24 SWB steps: a std engine that discards 768 values
        # then emits 24 reproduces the warmed state's next 24 values
        std = O.ranlux_std_stream(ex, 0, 24, exe, pre=768)
        ours = O.ranlux_stream(ho_lib, np.array([warm]), 0, 6)[0]
        np.testing.assert_array_equal(_top24(ours), std)


def test_ranlux_double_composition():
    lib64 = O.build(HarmOsc(), 64, 5.0, operators=("1.0 * x1 * x1",), double=True)
    lib32 = O.build(HarmOsc(), 64, 5.0, operators=("1.0 * x1 * x1",))
    st = O.ranlux_init(lib32, O.ranlux_seeds(9, 0, 2))
    f = O.ranlux_stream(lib32, st.copy(), 76, 40)
    d = O.ranlux_stream(lib64, st.copy(), 76, 20, double=True)
    tiny = np.float32(2.0 ** -48)
    f = f.astype(np.float64).reshape(2, -1, 2)
    hi = np.where(f[..., 0] == tiny, 0.0, f[..., 0])
    np.testing.assert_array_equal(d.reshape(2, -1), hi + f[..., 1] / 16777216.0)


# ----------------------------------------------------------------------------
# Sampler laws
# ----------------------------------------------------------------------------
def _ar_ref(N, S, beta, walkers=448, steps=10, runs=64, seed=1):
    sysobj = HarmOsc()
    lib = O.build(sysobj, N, beta, operators=(sysobj.energyOp,))
    cfg = O.make_cfg(N, S, 1, walkers, beta, runs, steps, walkersPerGroup=4, nOps=1)
    rng = np.random.default_rng(seed)
    paths = np.ascontiguousarray(np.repeat(rng.uniform(-0.01, 0.01, (walkers, 1)), N, axis=1).astype(np.float32))
    seeds = O.xorshift_rows(seed, 0, cfg.nThreads + 1)
    O.run_reference(lib, cfg, paths, seeds=seeds)  # burn-in
    out = O.run_reference(lib, cfg, paths, seeds=seeds)
    T = cfg.nThreads // walkers
    e = out["opMeans"].reshape(walkers, T).mean(axis=1)
    return out["accepts"].mean() / (runs * steps), e.mean(), e.std() / math.sqrt(walkers)


@pytest.mark.parametrize("N,S,lo,hi,e_exact", [
    (8, 1, 0.5569 - 0.004, 0.5612 + 0.004, 0.312347556605),
    (16, 2, 0.4741 - 0.004, 0.4792 + 0.004, 0.423999157335),
    (64, 2, 0.8923 - 0.002, 0.8934 + 0.002, 0.494006019083),
])
def test_acceptance_matches_reference_tutorial(N, S, lo, hi, e_exact):
    """doc/tutorial/index.rst:239-256: the reference's modN log at beta = 20."""
    ar, e, se = _ar_ref(N, S, 20.0)
    assert lo <= ar <= hi, "acceptance %.4f outside the reference's [%.4f, %.4f]" % (ar, lo, hi)
    assert abs(O.ho_discrete_x2(N, 20.0) - e_exact) < 1e-11
    assert abs(e - e_exact) < 4 * se


def test_free_particle_always_accepts():
    lib = O.build(_Free(), 32, 4.0, operators=("x1 * x1",))
    cfg = O.make_cfg(32, 3, 1, 16, 4.0, 3, 5, nOps=1)
    paths = np.zeros((16, 32), np.float32)
    out = O.run_reference(lib, cfg, paths, seeds=O.xorshift_rows(5, 0, cfg.nThreads + 1))
    assert (out["accepts"] == 15).all()
    lane = O.xorshift_rows(6, 0, 16 * 64)
    grp = O.xorshift_rows(6, 16 * 64, 16)
    out = O.run_wave(lib, O.make_cfg(32, 3, 1, 16, 4.0, 3, 5, nOps=1), np.zeros((16, 32), np.float32),
                     lane_seeds=lane, group_seeds=grp)
    assert (out["accepts"] == 15 * (32 // 8)).all()


@pytest.mark.parametrize("mode", ["reference", "wave"])
def test_ranlux_offsets_reach_every_slice(mode):
    """RANLUX integers are u * 2^31 with u on the 2^-24 grid (kernel.c:206), so
    the reference's offset draw `int & (2^S - 1)` (kernel.c:715) is always 0 and
    segment ends never move.  The samplers draw the offset from the top bits of
    u instead: with V = 0 every slice of a zero path must leave 0 within a few
    sweeps, and the offsets must cover all 2^S values."""
    N, S, W, nskip = 64, 4, 8, 76
    lib = O.build(_Free(), N, 4.0, operators=("x1 * x1",))
    paths = np.zeros((W, N), np.float32)
    if mode == "reference":
        cfg = O.make_cfg(N, S, 1, W, 4.0, 1, 6, ranlux=True, nskip=nskip, nOps=1)
        rl = O.ranlux_init(lib, O.ranlux_seeds(31, 0, cfg.nThreads))
        grl = O.ranlux_init(lib, O.ranlux_seeds(31, 1 << 32, cfg.nThreads))
        O.run_reference(lib, cfg, paths, rl=rl, grl=grl)
    else:
        cfg = O.make_cfg(N, S, 1, W, 4.0, 1, 6, ranlux=True, nskip=nskip, nOps=1)
        rl = O.ranlux_init(lib, O.ranlux_seeds(31, 0, W * 64))
        grl = O.ranlux_init(lib, O.ranlux_seeds(31, 1 << 32, W))
        O.run_wave(lib, cfg, paths, lane_rl=rl, group_rl=grl)
    assert (paths != 0).all()
    # the offset itself: the top 4 bits of the group stream's uniforms
    g = O.ranlux_init(lib, O.ranlux_seeds(5, 0, 1))
    u = O.ranlux_stream(lib, g, nskip, 64)[0]
    offs = (u * np.float32(1 << S)).astype(np.uint32)
    assert set(offs.tolist()) == set(range(1 << S))
    assert ((u * np.float32(2 ** 31)).astype(np.uint64) & ((1 << S) - 1) == 0).all()


def test_free_bridge_midpoint_law():
    """V = 0, one level: the midpoint of a segment is N(mean of ends, sigma_1^2),
    sigma_1^2 = 2^S eps / 4 (kernel.c:613-618)."""
    N, S, beta, W = 2, 1, 3.0, 20000
    lib = O.build(_Free(), N, beta, operators=("x1 * x1",))
    cfg = O.make_cfg(N, S, 1, W, beta, 1, 1, parallelizePath=False, walkersPerGroup=1, nOps=1)
    paths = np.zeros((W, N), np.float32)
    O.run_reference(lib, cfg, paths, seeds=O.xorshift_rows(8, 0, W + 1))
    # after one move one slice is still 0 and the other is the Levy midpoint
    moved = paths.sum(axis=1).astype(np.float64)
    var = 2 * (beta / N) / 4
    assert abs(moved.mean()) < 4 * math.sqrt(var / W)
    assert abs(moved.var() / var - 1) < 4 * math.sqrt(2.0 / W)
    # chi-square on 20 equiprobable bins of the standard normal
    from scipy import stats
    z = moved / math.sqrt(var)
    edges = stats.norm.ppf(np.linspace(0, 1, 21))
    counts, _ = np.histogram(z, edges)
    chi2 = ((counts - W / 20) ** 2 / (W / 20)).sum()
    assert chi2 < stats.chi2.ppf(0.999, 19)


def _cov_exact(N, beta, k=1.0):
    eps = beta / N
    A = np.zeros((N, N))
    for i in range(N):
        A[i, i] = 2 / eps + eps * k
        A[i, (i + 1) % N] -= 1 / eps
        A[i, (i - 1) % N] -= 1 / eps
    return np.linalg.inv(A)


@pytest.mark.parametrize("mode", ["reference", "wave"])
def test_tiny_path_covariance(mode):
    """N = 4, S = 2: the stationary law is the Gaussian of the discretised action."""
    N, S, beta, W = 4, 2, 2.0, 4096
    sysobj = HarmOsc()
    lib = O.build(sysobj, N, beta, operators=("x1 * x1",))
    cfg = O.make_cfg(N, S, 1, W, beta, 1, 40, nOps=1)
    rng = np.random.default_rng(0)
    paths = np.ascontiguousarray(rng.normal(0, 0.5, (W, N)).astype(np.float32))
    acc = np.zeros((N, N))
    if mode == "reference":
        seeds = O.xorshift_rows(1, 0, cfg.nThreads + 1)
        step = lambda: O.run_reference(lib, cfg, paths, seeds=seeds)
    else:
        lane = O.xorshift_rows(1, 0, W * 64)
        grp = O.xorshift_rows(1, W * 64, W)
        step = lambda: O.run_wave(lib, cfg, paths, lane_seeds=lane, group_seeds=grp)
    step()
    n = 0
    for _ in range(30):
        step()
        p = paths.astype(np.float64)
        acc += p.T @ p / W
        n += 1
    cov = acc / n
    C = _cov_exact(N, beta)
    # 4096 x 30 samples, strongly correlated in time; a 5% band catches wrong widths/actions
    np.testing.assert_allclose(cov, C, rtol=0.05, atol=0.01)


@pytest.mark.parametrize("mode", ["reference", "wave"])
def test_anharmonic_vs_transfer_matrix(mode):
    N, S, beta, W = 16, 2, 4.0, 2048
    sysobj = AnharmOsc(0.5)
    lib = O.build(sysobj, N, beta, operators=(sysobj.positionSquaredOp,))
    cfg = O.make_cfg(N, S, 1, W, beta, 20, 10, nOps=1)
    paths = np.zeros((W, N), np.float32)
    if mode == "reference":
        seeds = O.xorshift_rows(2, 0, cfg.nThreads + 1)
        step = lambda: O.run_reference(lib, cfg, paths, seeds=seeds)
        per = lambda out: out["opMeans"].reshape(W, -1).mean(axis=1)
    else:
        lane = O.xorshift_rows(2, 0, W * 64)
        grp = O.xorshift_rows(2, W * 64, W)
        step = lambda: O.run_wave(lib, cfg, paths, lane_seeds=lane, group_seeds=grp)
        per = lambda out: out["opMeans"]
    step()
    e = per(step())
    x2, _, _ = O.transfer_matrix_1d(sysobj.V, N, beta, L=6.0, M=900)
    se = e.std() / math.sqrt(W)
    assert abs(e.mean() - x2) < 4 * se, (e.mean(), x2, se)


# ----------------------------------------------------------------------------
# Closed forms and systems
# ----------------------------------------------------------------------------
def test_ho_closed_forms():
    assert abs(O.ho_discrete_x2(256, 10.0) - 0.499950082044) < 1e-11
    assert abs(O.ho_discrete_x2(128, 11.0) - 0.499555802927) < 1e-11
    c = O.ho_discrete_correlator(256, 10.0)
    assert abs(c[0] - O.ho_discrete_x2(256, 10.0)) < 1e-14
    # large-beta decay rate = the discretised gap arccosh(1 + eps^2/2)/eps
    eps = 10.0 / 256
    gap = math.acosh(1 + eps * eps / 2) / eps
    assert abs(gap - 0.99993643) < 1e-7


def test_transfer_matrix_reproduces_ho():
    x2, _, cor = O.transfer_matrix_1d(lambda x: 0.5 * x * x, 32, 6.0)
    assert abs(x2 - O.ho_discrete_x2(32, 6.0)) < 1e-9
    np.testing.assert_allclose(cor, O.ho_discrete_correlator(32, 6.0)[:17], atol=1e-9)


def test_transfer_matrix_density_reproduces_ho():
    """The slice density of the HO action is the Gaussian of variance <x^2>_N; its x^4
    moment is 3 <x^2>_N^2."""
    x, p = O.transfer_matrix_density(lambda x: 0.5 * x * x, 64, 8.0, L=7.0, M=900)
    v = O.ho_discrete_x2(64, 8.0)
    assert abs(p.sum() - 1.0) < 1e-12
    assert abs(np.sum(p * x * x) - v) < 1e-9
    assert abs(np.sum(p * x ** 4) - 3 * v * v) < 1e-8


def test_aho_c5_golden_fixture():
    """tests/golden/aho_c5.json (N = 4096, beta = 40): grid-converged (1000 vs 1400 points),
    consistent (C(0) = <x^2>), and within O(eps^2) of the continuum ground state of
    x^2/2 + 0.1 x^4 (E0 = 0.559146, E1 - E0 = 1.210348; beta = 40 suppresses excited
    states by exp(-48))."""
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "aho_c5.json")) as f:
        g = json.load(f)
    a, b = g["values"], g["values_M1000"]
    assert abs(a["E_virial"] - b["E_virial"]) < 1e-10 and abs(a["gap"] - b["gap"]) < 1e-8
    assert abs(a["correlator"][0] - a["x2"]) < 1e-12
    assert abs(a["E_virial"] - 0.559146) < 5e-5
    assert abs(a["gap"] - 1.210348) < 1e-3
    # the correlator decays with the gap between lags 100 and 300
    c = np.array(a["correlator"])
    rate = -np.log(c[300] / c[100]) / (200 * g["eps"])
    assert abs(rate - a["gap"]) < 1e-3 * a["gap"]


def test_lm2m2_literature_values():
    assert abs(lm2m2_kelvin(2.9695) - (-10.97)) < 1e-5
    assert abs(lm2m2_kelvin(3.5) - (-6.284940)) < 2e-5
    assert abs(lm2m2_kelvin(5.0) - (-0.728570)) < 2e-5
    assert abs(lm2m2_kelvin(10.0) - (-0.010414)) < 2e-6
    assert abs(mass_scale(4.0026) ** 2 - 1 / 12.119282) < 2e-6


@pytest.mark.parametrize("n", [3, 4, 6, 8])
def test_lm2m2_system_strings_in_oracle(n):
    """The generated potential and virial strings of an n-atom cluster (DOF 3n: 9, 12, 18, 24),
    compiled by gcc in FP64, against the FP64 pair sum of the formula and its finite-difference
    virial."""
    sysobj = Lm2m2Cluster(n)
    lib = O.build(sysobj, 64, 10.0, operators=(sysobj.energyOp, sysobj.meanPairDistanceOp), double=True)
    rng = np.random.default_rng(4 + n)
    for _ in range(50 if n <= 4 else 20):
        r_atoms = rng.normal(0, 2.5 + 0.3 * n, (n, 3))
        x = np.ascontiguousarray((r_atoms * sysobj.massScale).reshape(3 * n))
        dists = [np.linalg.norm(r_atoms[i] - r_atoms[j]) for i, j in sysobj.pairs]
        V = sum(lm2m2_kelvin(d) for d in dists)
        assert abs(lib.fs_oracle_potential(x.ctypes.data) - V) < 1e-9 * max(1, abs(V))
        acc = np.zeros(3)
        lib.fs_oracle_operators(x.ctypes.data, acc.ctypes.data)
        h = 1e-6
        vir = sum(lm2m2_kelvin(d) + 0.5 * d * (lm2m2_kelvin(d + h) - lm2m2_kelvin(d - h)) / (2 * h) for d in dists)
        assert abs(acc[0] - vir) < 1e-5 * max(1, abs(vir))
        assert abs(acc[1] - np.mean(dists)) < 1e-9
    assert len(sysobj.pairs) == n * (n - 1) // 2


def _recount(values, lo, hi, bins, ulps=8):
    """Independent FP64 recount of a histogram: np.histogram of the values that are not within a few
    FP32 ulps of a bin edge, plus the near-edge values (in bin units) whose bin rounding may decide."""
    v = (np.asarray(values, np.float64).ravel() - lo) * (bins / (hi - lo))
    tol = ulps * np.finfo(np.float32).eps * np.maximum(np.abs(v), 1.0)
    near = np.abs(v - np.round(v)) <= tol
    far = np.histogram(v[~near], bins=bins, range=(0.0, float(bins)))[0]
    return far, v[near]


def _assert_histogram(counts, far, near_v, bins):
    """counts must equal the far-from-edge recount, plus at most one count per near-edge value in one
    of the two bins beside its edge."""
    counts = np.asarray(counts, np.int64)
    diff = counts - far
    assert (diff >= 0).all(), "bins below the recount: %s" % np.nonzero(diff < 0)[0]
    allow = np.zeros(bins, np.int64)
    for e in np.round(near_v).astype(np.int64):
        for b in (e - 1, e):
            if 0 <= b < bins:
                allow[b] += 1
    assert (diff <= allow).all(), "bins above the recount: %s" % np.nonzero(diff > allow)[0]
    assert diff.sum() <= len(near_v)


def test_observable_histogram_matches_numpy_recount():
    """The oracle's histogram of scalar observables (kernel.c:478-506 style accumulation applied to the
    pair distances r_ij of a 3-atom LM2M2 cluster, 256 bins on [0, 12) angstrom) after one launch with
    a single measurement equals an independent numpy recount of the pair distances of the oracle's own
    final paths, bin for bin (values within a few FP32 ulps of a bin edge may fall on either side)."""
    sysobj = Lm2m2Cluster(3)
    W, N, S, beta, bins, lo, hi = 24, 256, 6, 10.0, 256, 0.0, 12.0
    obs = tuple(sysobj.pairDistance(i, j) for i, j in sysobj.pairs)
    lib = O.build(sysobj, N, beta, operators=(sysobj.energyOp,), observables=obs)
    cfg = O.make_cfg(N, S, sysobj.DOF, W, beta, 1, 3, nOps=1, nHistObs=3, histBins=bins, histRange=(lo, hi))
    rng = np.random.default_rng(17)
    pos = sysobj.initialPositions(W, rng, sigma=1.5)
    p0 = np.repeat(pos[:, :, None], N, axis=2) + rng.normal(0, 0.05, (W, sysobj.DOF, N))
    paths = np.ascontiguousarray(p0.reshape(W, sysobj.DOF * N).astype(np.float32))
    ref = O.run_wave(lib, cfg, paths, lane_seeds=O.xorshift_rows(3, 0, W * 64),
                     group_seeds=O.xorshift_rows(3, W * 64, W))
    assert ref["accepts"].sum() > 0
    x = paths.reshape(W, sysobj.nAtoms, 3, N).astype(np.float64) / sysobj.massScale
    r = np.stack([np.sqrt(((x[:, i] - x[:, j]) ** 2).sum(axis=1)) for i, j in sysobj.pairs])
    far, near_v = _recount(r, lo, hi, bins)
    _assert_histogram(ref["obsHist"], far, near_v, bins)
    # some distances fall outside the range: they are not counted
    inside = ((r >= lo) & (r < hi)).sum()
    assert inside < r.size and abs(int(ref["obsHist"].sum()) - int(inside)) <= len(near_v)
    # a planted off-by-one in the bin index fails the same check
    shifted = np.roll(ref["obsHist"], 1)
    with pytest.raises(AssertionError):
        _assert_histogram(shifted, far, near_v, bins)


@pytest.mark.parametrize("dof", [1, 2])
def test_dof_grid_bins_match_numpy_recount(dof):
    """The oracle's DOF-grid histogram (kernel.c:478-506: binIndex = sum_p floor((x_p - xMin) /
    binSize) * res^p over raw coordinates, slices with any coordinate outside [xMin, xMax) dropped)
    after one launch with a single measurement equals numpy's histogramdd of the final paths."""
    class _Osc:
        DOF = dof
        potential = " + ".join("0.5f * x%d * x%d" % (i + 1, i + 1) for i in range(dof))
        userCode = ""
    W, N, S, beta, res, xmin, xmax = 16, 64, 3, 5.0, 12, -1.5, 1.5
    lib = O.build(_Osc(), N, beta, operators=("x1 * x1",))
    cfg = O.make_cfg(N, S, dof, W, beta, 1, 4, nOps=1, bins=True, binRes=res, xMin=xmin, xMax=xmax)
    paths = np.ascontiguousarray(np.random.default_rng(2).uniform(-1.2, 1.2, (W, dof * N)).astype(np.float32))
    ref = O.run_reference(lib, cfg, paths, seeds=O.xorshift_rows(5, 0, cfg.nThreads + 1))
    x = paths.reshape(W, dof, N).transpose(0, 2, 1).reshape(-1, dof).astype(np.float64)
    v = (x - xmin) * (res / (xmax - xmin))
    near = (np.abs(v - np.round(v)) <= 8 * np.finfo(np.float32).eps * np.maximum(np.abs(v), 1)).any(axis=1)
    h = np.histogramdd(v[~near], bins=[res] * dof, range=[(0, res)] * dof)[0]
    # kernel.c:478-506 puts coordinate p at stride res^p: index = sum_p b_p res^p, so axis 0 varies fastest
    flat = h.transpose(tuple(range(dof))[::-1]).reshape(-1) if dof > 1 else h
    got = ref["bins"].astype(np.int64)
    diff = got - flat.astype(np.int64)
    assert (diff >= 0).all() and diff.sum() <= near.sum()
    assert got.sum() > 0.5 * W * N


# pair distances (angstrom) through the core: r = 0, below, at and above the 0.0005 A cut-off of the
# damped dispersion term (lm2m2.h:55-66), where its 1/x^10 factor overflows FP32 first
CORE_DISTANCES = (0.0, 1e-5, 1e-4, 4e-4, 5e-4, 5.2e-4, 1e-3, 1e-2, 0.3, 1.0, 2.0)


def _core_points(sysobj):
    """Atom 2 at distance r from atom 1 on the x axis, atom 3 far away; the exact pair sums."""
    far = np.array([40.0, 40.0, 40.0])
    pts, pot, vir = [], [], []
    h = 1e-7
    for r in CORE_DISTANCES:
        atoms = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0], far])
        pts.append((atoms * sysobj.massScale).reshape(9))
        d = [np.linalg.norm(atoms[i] - atoms[j]) for i, j in sysobj.pairs]
        pot.append(sum(lm2m2_kelvin(x) for x in d))
        vir.append(sum(lm2m2_kelvin(x) + 0.5 * x * (lm2m2_kelvin(x + h) - lm2m2_kelvin(abs(x - h))) / (2 * h)
                       for x in d if x > 0) + (lm2m2_kelvin(0.0) if r == 0.0 else 0.0))
    return np.array(pts), np.array(pot), np.array(vir)


@pytest.mark.parametrize("double", [False, True])
def test_lm2m2_core_matches_reference_branches(double):
    """Inside the core the generated pair energy keeps the reference's branches: the damped
    dispersion term is zero at and below x = 0.0005 / r_m, and r = 0 gives eps/k A (2.08e6 K),
    finite in FP32 as in FP64 (no 0 * inf from the overflowing 1/x^10)."""
    sysobj = Lm2m2Cluster(3)
    assert abs(lm2m2_kelvin(0.0) - 10.97 * 1.89635353e5) < 1e-6
    lib = O.build(sysobj, 64, 10.0, operators=(sysobj.energyOp,), double=double)
    pts, pot, vir = _core_points(sysobj)
    ft = np.float64 if double else np.float32
    tol = 1e-9 if double else 2e-6
    for x, v, e in zip(pts, pot, vir):
        xx = np.ascontiguousarray(x.astype(ft))
        got = lib.fs_oracle_potential(xx.ctypes.data)
        assert np.isfinite(got) and abs(got - v) <= tol * abs(v), (got, v)
        acc = np.zeros(1, ft)
        lib.fs_oracle_operators(xx.ctypes.data, acc.ctypes.data)
        assert np.isfinite(acc[0]) and abs(acc[0] - e) <= 1e-4 * abs(e), (acc[0], e)


@pytest.mark.parametrize("moves", ["shift", "single", "all"])
def test_path_shift_and_single_node_keep_the_exact_law(moves):
    """Path-shift (kernel.c:407-473, 640-663) and single-node (kernel.c:665-703) moves,
    alone and with bisection, one work-item per walker: started in the exact Gaussian of
    the discretised HO action on an 8-slice ring, the chain must stay there."""
    N, S, beta, W = 8, 2, 2.0, 16384
    Ce = _cov_exact(N, beta)
    lib = O.build(HarmOsc(), N, beta, operators=("1.0 * x1 * x1",))
    M = 3 * N
    cfg = O.make_cfg(N, S, 1, W, beta, 1, M, parallelizePath=False, nOps=1, bisection=(moves == "all"),
                     pathShift=moves in ("shift", "all"), singleNode=moves in ("single", "all"),
                     alpha=0.8, psAlpha=0.6)
    rng = np.random.default_rng(11)
    paths = (rng.standard_normal((W, N)) @ np.linalg.cholesky(Ce).T).astype(np.float32)
    seeds = O.xorshift_rows(23, 0, cfg.nThreads + 1)
    acc = 0
    snaps = []
    for t in range(12):
        out = O.run_reference(lib, cfg, paths, seeds=seeds)
        acc += out["accepts"].sum()
        if t >= 4:
            snaps.append(paths.astype(np.float64).copy())
    X = np.concatenate(snaps)
    moves_per_step = {"shift": 1, "single": 1, "all": 3}[moves]
    ar = acc / (12 * W * M * moves_per_step)
    assert 0.1 < ar < 0.95, ar
    assert np.abs(np.cov(X.T) - Ce).max() < 0.03 * Ce[0, 0]
    assert np.abs(X.mean(axis=0)).max() < 0.02 * np.sqrt(Ce[0, 0])



def _eval_system_string(expr, x):
    """Evaluates a generated system string (only fs_sqr, fs_lm2m2_r2 / virial and x1..xD) in FP64."""
    import re
    from feynsimul_amd.physical_systems.lm2m2 import lm2m2_kelvin
    assert set(re.findall(r"[A-Za-z_][A-Za-z_0-9]*", expr)) <= (
        {"fs_sqr", "fs_lm2m2_r2", "fs_lm2m2_virial_r2"} | {"x%d" % (i + 1) for i in range(len(x))})

    def vir(r2):
        r, h = math.sqrt(r2), 1e-6
        return lm2m2_kelvin(r) + 0.5 * r * (lm2m2_kelvin(r + h) - lm2m2_kelvin(r - h)) / (2 * h)
    env = {"fs_sqr": lambda a: a * a, "fs_lm2m2_r2": lambda r2: lm2m2_kelvin(math.sqrt(r2)),
           "fs_lm2m2_virial_r2": vir}
    env.update({"x%d" % (i + 1): float(v) for i, v in enumerate(x)})
    return eval(expr, {"__builtins__": {}}, env)


@pytest.mark.parametrize("n", [2, 3, 4, 6])
def test_jacobi_cluster_matches_cartesian_pair_sum(n):
    """Lm2m2_cluster (lm2m2_cluster.py:7-95, jacobi.py:3-36): unit-mass kinetic metric,
    centre of mass removed, and the generated potential / virial / radius strings equal the
    direct Cartesian LM2M2 pair sums at random configurations."""
    from feynsimul_amd.physical_systems import Lm2m2_cluster
    from feynsimul_amd.physical_systems.lm2m2_cluster import jacobi_matrix
    from feynsimul_amd.physical_systems.lm2m2 import lm2m2_kelvin
    sysobj = Lm2m2_cluster(n, 0.1)
    assert sysobj.DOF == 3 * (n - 1) and abs(sysobj.beta - 10.0) < 1e-12
    assert abs(math.sqrt(sysobj.m) - 0.287251) < 2e-6          # sqrt(m_He) in K^-1/2 / angstrom
    J = jacobi_matrix(n, sysobj.m)
    np.testing.assert_allclose(J @ np.diag(np.full(n, 1.0 / sysobj.m)) @ J.T, np.eye(n), atol=1e-12)
    rng = np.random.default_rng(n)
    for _ in range(5):
        r = rng.normal(0, 3.0, (n, 3)) + rng.normal(0, 20.0, 3)      # angstrom, arbitrary centre
        j = J @ r                                                      # (n, 3) Jacobi vectors
        x = j[: n - 1].reshape(-1)                                     # x_{3i+k+1} = j_i[k]
        pairs = [(a, b) for a in range(n) for b in range(a + 1, n)]
        dist = [np.linalg.norm(r[a] - r[b]) for a, b in pairs]
        V = sum(lm2m2_kelvin(d_) for d_ in dist)
        assert abs(_eval_system_string(sysobj.potential, x) - V) < 1e-9 * max(1.0, abs(V))
        assert abs(_eval_system_string(sysobj.meanSquaredRadiusOp, x) - sum(d_ ** 2 for d_ in dist)) < 1e-8
        # virial V + x . grad V / 2 by central differences in the Jacobi coordinates
        h = 1e-5
        g = np.array([(_eval_system_string(sysobj.potential, x + h * e) -
                       _eval_system_string(sysobj.potential, x - h * e)) / (2 * h) for e in np.eye(len(x))])
        assert abs(_eval_system_string(sysobj.energyOp, x) - (V + 0.5 * x @ g)) < 1e-4 * max(1.0, abs(V))
