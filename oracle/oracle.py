"""CPU oracle for the FeynSimul PIMC hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker; the product
(``feynsimul_amd``) never does.

The oracle compiles ``pimc_oracle.c`` with gcc once per physical system, with
the system's potential / operator / correlator strings pasted into a
generated header in the way the reference pastes them into its OpenCL
template (FeynSimul/kernel.py:253-348): ``x1..xDOF`` arguments, an ``eps``
macro, ``FLOAT_TYPE`` and the system's ``userCode``.  FP32 builds use
``-fsingle-precision-constant`` (the reference builds with
``-cl-single-precision-constant``, kernel.py:376-377).

Parity status of each oracle function (what pins it) is listed in the module
``tests/test_oracle.py`` docstring; nothing here is "parity unpinned" except
where that file says so.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "pimc_oracle.c")
BUILD_DIR = os.environ.get("FS_ORACLE_BUILD", os.path.join(HERE, "_build"))

_lock = threading.Lock()
_libs: dict[str, ctypes.CDLL] = {}

# OpenCL C conveniences the reference's system strings may use
# (physical_systems/__init__.py:6-9 names native_* functions).
_OPENCL_PRELUDE = r"""
#include <stdbool.h>
#include <math.h>
#include <tgmath.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_PI_2
#define M_PI_2 1.57079632679489661923
#endif
#define native_exp exp
#define native_log log
#define native_sqrt sqrt
#define native_sin sin
#define native_cos cos
#define native_powr pow
#define native_recip(a) (1.0 / (a))
#define native_divide(a, b) ((a) / (b))
#define half_exp exp
#define half_log log
#define half_sqrt sqrt
#define rsqrt(a) (1.0 / sqrt(a))
#define mad(a, b, c) ((a) * (b) + (c))
#define pown(a, n) pow((a), (FLOAT_TYPE)(n))
#define fs_sin2pi(a) sin(2.0 * M_PI * (a))
#define fs_cos2pi(a) cos(2.0 * M_PI * (a))
"""


def system_header(system, N: int, beta: float, operators=(), correlators=(), double=False, observables=()) -> str:
    """Generated C header for one (system, N, beta, operators) configuration."""
    dof = int(system.DOF)
    args = ", ".join("FLOAT_TYPE x%d" % (i + 1) for i in range(dof))
    call = ", ".join("x[%d]" % i for i in range(dof))
    lines = []
    lines.append("#define FLOAT_TYPE %s" % ("double" if double else "float"))
    if double:
        lines.append("#define FS_DOUBLE 1")
    lines.append("#define FS_DOF %d" % dof)
    lines.append("#define FS_NOPS %d" % max(1, len(operators)))
    lines.append("#define FS_NCORR %d" % max(1, len(correlators)))
    lines.append("#define FS_NHOBS %d" % max(1, len(observables)))
    lines.append(_OPENCL_PRELUDE)
    lines.append("#define eps %s" % ("%1.17e" % (beta / float(N))))
    lines.append(getattr(system, "userCode", "") or "")
    lines.append("static inline FLOAT_TYPE potential(%s)\n{\n    return %s;\n}" % (args, system.potential))
    body = "".join("    opAccum[%d] += %s;\n" % (i, op) for i, op in enumerate(operators))
    lines.append("static inline void fs_operator(%s, FLOAT_TYPE *opAccum)\n{\n%s}" % (args, body))
    body = "".join("    corProd[%d] *= %s;\n" % (i, cc) for i, cc in enumerate(correlators))
    lines.append("static inline void fs_correlator(%s, FLOAT_TYPE *corProd)\n{\n%s}" % (args, body))
    body = "".join("    obs[%d] = %s;\n" % (i, o) for i, o in enumerate(observables))
    lines.append("static inline void fs_hist_obs(%s, FLOAT_TYPE *obs)\n{\n%s}" % (args, body))
    lines.append("static inline void fs_hist_obs_v(const FLOAT_TYPE *x, FLOAT_TYPE *o) { fs_hist_obs(%s, o); }" % call)
    lines.append("static inline FLOAT_TYPE fs_potential_v(const FLOAT_TYPE *x) { return potential(%s); }" % call)
    lines.append("static inline void fs_operator_v(const FLOAT_TYPE *x, FLOAT_TYPE *a) { fs_operator(%s, a); }" % call)
    lines.append("static inline void fs_correlator_v(const FLOAT_TYPE *x, FLOAT_TYPE *p) { fs_correlator(%s, p); }"
                 % call)
    return "\n".join(lines) + "\n"


def build(system, N: int, beta: float, operators=(), correlators=(), double=False,
          fast_math=False, native=False, observables=()) -> ctypes.CDLL:
    """Compile (or reuse) the oracle for one configuration and load it."""
    hdr = system_header(system, N, beta, operators, correlators, double, observables)
    with open(SRC, "rb") as f:
        src_hash = hashlib.sha1(f.read()).hexdigest()
    flags = ["-O2", "-fPIC", "-shared", "-fopenmp", "-fgnu89-inline", "-std=gnu11",
             "-fvisibility=hidden", "-Wno-unused-function"]
    if not double:
        flags.append("-fsingle-precision-constant")
    if fast_math:
        flags.append("-ffast-math")
    if native:
        flags += ["-O3", "-march=native"]
    key = hashlib.sha1((hdr + src_hash + " ".join(flags)).encode()).hexdigest()[:16]
    with _lock:
        if key in _libs:
            return _libs[key]
        out_dir = os.path.join(BUILD_DIR, key)
        so = os.path.join(out_dir, "oracle.so")
        if not os.path.exists(so):
            os.makedirs(out_dir, exist_ok=True)
            hpath = os.path.join(out_dir, "system.h")
            with open(hpath, "w") as f:
                f.write(hdr)
            tmp = so + ".tmp%d" % os.getpid()
            cmd = ["gcc"] + flags + ["-include", hpath, SRC, "-o", tmp, "-lm"]
            res = subprocess.run(cmd, capture_output=True, text=True)
            if res.returncode != 0:
                raise RuntimeError("oracle build failed:\n%s\n%s" % (" ".join(cmd), res.stderr))
            os.replace(tmp, so)
        lib = ctypes.CDLL(so)
        _declare(lib)
        lib._fs_double = double
        _libs[key] = lib
        return lib


class OracleCfg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "N", "S", "DOF", "nWalkers", "walkersPerGroup", "parallelizePath", "ranlux", "nskip",
        "operatorRuns", "metroSteps", "enableOperator", "nOps", "enableCorrelator", "nCorr",
        "corLags", "enableBins", "binRes", "localSize", "nThreads", "nHistObs", "histBins")] + [
        (n, ctypes.c_double) for n in ("epsilon", "opNorm", "ranluxIntScale", "xMin", "xMax", "invBinSize",
                                       "histLo", "histInv")] + [
        (n, ctypes.c_int32) for n in ("bisection", "pathShift", "singleNode", "pad_")] + [
        (n, ctypes.c_double) for n in ("alpha", "psAlpha", "epsInv2")]


RANLUX_DTYPE = np.dtype([("s", "<f4", (24,)), ("carry", "<f4"), ("dummy", "<f4"),
                         ("in24", "<i4"), ("stepnr", "<i4")])
assert RANLUX_DTYPE.itemsize == 112


def _declare(lib):
    vp = ctypes.c_void_p
    lib.fs_oracle_run_reference.argtypes = [ctypes.POINTER(OracleCfg)] + [vp] * 8 + [ctypes.c_int]
    lib.fs_oracle_run_reference.restype = ctypes.c_int
    lib.fs_oracle_run_wave.argtypes = [ctypes.POINTER(OracleCfg)] + [vp] * 9 + [ctypes.c_int]
    lib.fs_oracle_run_wave.restype = ctypes.c_int
    lib.fs_oracle_xs_stream.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
    lib.fs_oracle_ranlux_init.argtypes = [vp, ctypes.c_int, vp]
    lib.fs_oracle_ranlux_expand.argtypes = [vp, ctypes.c_int, vp]
    lib.fs_oracle_ranlux_stream.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    lib.fs_oracle_ranlux_stream_double.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    lib.fs_oracle_potential.argtypes = [vp]
    lib.fs_oracle_potential.restype = ctypes.c_double
    lib.fs_oracle_operators.argtypes = [vp, vp]
    lib.fs_oracle_float_bytes.restype = ctypes.c_int


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# Seeding: splitmix64 (Steele, Lea, Flood 2014), row r from state seed + r.
# --------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64_pair(state: int):
    out = []
    for _ in range(2):
        state = (state + 0x9E3779B97F4A7C15) & _M64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        out.append(z ^ (z >> 31))
    return out


def xorshift_rows(seed: int, first: int, count: int) -> np.ndarray:
    """xorshift128 state rows first..first+count-1 as (count, 4) uint32."""
    rows = np.empty((count, 4), dtype=np.uint32)
    for i in range(count):
        a, b = splitmix64_pair((seed + first + i) & _M64)
        r = [a & 0xFFFFFFFF, a >> 32, b & 0xFFFFFFFF, b >> 32]
        if not any(r):
            r[0] = 1
        rows[i] = r
    return rows


def ranlux_seeds(seed: int, first: int, count: int) -> np.ndarray:
    """64-bit ranluxcl_init seeds for rows first..first+count-1."""
    return np.array([splitmix64_pair((seed + first + i) & _M64)[0] for i in range(count)], dtype=np.uint64)


# --------------------------------------------------------------------------
# RNG streams
# --------------------------------------------------------------------------
def xs_stream(lib, rows: np.ndarray, ndraw: int):
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    n = rows.shape[0]
    ints = np.empty((n, ndraw), dtype=np.uint32)
    flt = np.empty((n, ndraw), dtype=np.float64 if lib._fs_double else np.float32)
    lib.fs_oracle_xs_stream(_ptr(rows), n, ndraw, _ptr(ints), _ptr(flt))
    return ints, flt


def ranlux_init(lib, seeds: np.ndarray) -> np.ndarray:
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    st = np.zeros(seeds.shape[0], dtype=RANLUX_DTYPE)
    lib.fs_oracle_ranlux_init(_ptr(seeds), seeds.shape[0], _ptr(st))
    return st


def ranlux_expand(lib, seeds: np.ndarray) -> np.ndarray:
    """ranluxcl_init's seed expansion without the warm-up cycles."""
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    st = np.zeros(seeds.shape[0], dtype=RANLUX_DTYPE)
    lib.fs_oracle_ranlux_expand(_ptr(seeds), seeds.shape[0], _ptr(st))
    return st


def ranlux_std_stream(state, nskip: int, count: int, exe: str, pre: int = 0) -> np.ndarray:
    """Integer outputs of libstdc++'s ranlux24_base / discard_block_engine loaded with `state`
    (a RANLUX_DTYPE record at cycle position 0), after discarding `pre` values."""
    assert int(state["stepnr"]) == 0
    lags = [int(round(float(v) * 16777216.0)) for v in state["s"][::-1]]  # x_{i-24} .. x_{i-1}
    carry = 1 if float(state["carry"]) != 0.0 else 0
    txt = "%d %d %d %s %d\n" % (nskip, count, pre, " ".join(map(str, lags)), carry)
    res = subprocess.run([exe], input=txt, capture_output=True, text=True, check=True)
    return np.array([int(v) for v in res.stdout.split()], dtype=np.int64)


def std_check_exe() -> str:
    """Build (once) the libstdc++ RANLUX cross-check program."""
    out = os.path.join(BUILD_DIR, "ranlux_std_check")
    src = os.path.join(HERE, "ranlux_std_check.cpp")
    if not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
        os.makedirs(BUILD_DIR, exist_ok=True)
        tmp = out + ".tmp%d" % os.getpid()
        subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", tmp], check=True)
        os.replace(tmp, out)
    return out


def ranlux_stream(lib, states: np.ndarray, nskip: int, nquad: int, double=False) -> np.ndarray:
    n = states.shape[0]
    if double:
        out = np.empty((n, nquad * 4), dtype=np.float64)
        lib.fs_oracle_ranlux_stream_double(_ptr(states), n, nskip, nquad, _ptr(out))
    else:
        out = np.empty((n, nquad * 4), dtype=np.float32)
        lib.fs_oracle_ranlux_stream(_ptr(states), n, nskip, nquad, _ptr(out))
    return out


def nskip_of_lux(lux: int) -> int:
    """RANLUXCL_NSKIP for a RANLUXCL_LUX value (pyopencl-ranluxcl.cl:305-317)."""
    table = {0: 0, 1: 24, 2: 76, 3: 200, 4: 380}
    if lux in table:
        return table[lux]
    return lux - 24


# --------------------------------------------------------------------------
# Sampler launches
# --------------------------------------------------------------------------
def make_cfg(N, S, DOF, nWalkers, beta, operatorRuns, metroSteps, *, walkersPerGroup=1,
             parallelizePath=True, ranlux=False, nskip=0, ranluxIntMax=2 ** 31 - 1, nOps=0,
             nCorr=0, corLags=None, bins=False, binRes=1, xMin=0.0, xMax=1.0, localSize=None,
             nThreads=None, nHistObs=0, histBins=1, histRange=(0.0, 1.0), bisection=True, pathShift=False,
             singleNode=False, alpha=0.0, psAlpha=0.0) -> OracleCfg:
    c = OracleCfg()
    c.N, c.S, c.DOF = N, S, DOF
    c.nWalkers, c.walkersPerGroup = nWalkers, walkersPerGroup
    c.parallelizePath = int(bool(parallelizePath))
    c.ranlux, c.nskip = int(bool(ranlux)), nskip
    c.operatorRuns, c.metroSteps = operatorRuns, metroSteps
    c.enableOperator, c.nOps = int(nOps > 0), nOps
    c.enableCorrelator, c.nCorr = int(nCorr > 0), nCorr
    c.corLags = corLags if corLags is not None else N // 2
    c.enableBins, c.binRes = int(bool(bins)), binRes
    c.localSize = localSize or walkersPerGroup
    T = (N >> S) if parallelizePath else 1
    c.nThreads = nThreads if nThreads is not None else nWalkers * T
    c.epsilon = beta / float(N)
    c.opNorm = 1.0 / float(operatorRuns * N * nWalkers / float(c.nThreads))
    c.ranluxIntScale = float(np.float32(ranluxIntMax + 1))
    c.xMin, c.xMax = xMin, xMax
    c.invBinSize = float(binRes) / float(xMax - xMin)
    c.nHistObs, c.histBins = nHistObs, histBins
    c.histLo = float(histRange[0])
    c.histInv = float(histBins) / float(histRange[1] - histRange[0])
    c.bisection, c.pathShift, c.singleNode = int(bool(bisection)), int(bool(pathShift)), int(bool(singleNode))
    c.alpha, c.psAlpha = float(alpha), float(psAlpha)
    c.epsInv2 = (float(N) / beta) ** 2
    return c


def run_reference(lib, cfg: OracleCfg, paths, seeds=None, rl=None, grl=None, nthreads=0):
    """One launch in the reference's work-item order; arrays are updated in place."""
    ft = np.float64 if lib._fs_double else np.float32
    assert paths.dtype == ft and paths.flags.c_contiguous
    accepts = np.zeros(cfg.nThreads, dtype=np.uint32)
    opm = np.zeros(cfg.nThreads * max(cfg.nOps, 1), dtype=ft)
    cor = np.zeros(cfg.nWalkers * max(cfg.nCorr, 1) * (cfg.N // 2), dtype=ft)
    bins = np.zeros(cfg.binRes ** cfg.DOF if cfg.enableBins else 1, dtype=np.uint32)
    rc = lib.fs_oracle_run_reference(ctypes.byref(cfg), _ptr(paths), _ptr(accepts), _ptr(seeds), _ptr(rl),
                                     _ptr(grl), _ptr(opm), _ptr(cor), _ptr(bins), nthreads)
    if rc != 0:
        raise RuntimeError("oracle run failed (%d)" % rc)
    return dict(accepts=accepts, opMeans=opm, corMeans=cor, bins=bins)


def run_wave(lib, cfg: OracleCfg, paths, lane_seeds=None, group_seeds=None, lane_rl=None, group_rl=None,
             nthreads=0, cor_bins=None):
    """One launch in the engine's wave-parallel order; arrays are updated in place."""
    ft = np.float64 if lib._fs_double else np.float32
    assert paths.dtype == ft and paths.flags.c_contiguous
    accepts = np.zeros(cfg.nWalkers, dtype=np.uint32)
    opm = np.zeros(cfg.nWalkers * max(cfg.nOps, 1), dtype=np.float64)
    if cor_bins is None:
        cor_bins = np.zeros(cfg.nWalkers * max(cfg.nCorr, 1) * cfg.corLags, dtype=np.float64)
    hist = np.zeros(max(cfg.histBins, 1), dtype=np.uint32)
    rc = lib.fs_oracle_run_wave(ctypes.byref(cfg), _ptr(paths), _ptr(accepts), _ptr(lane_seeds),
                                _ptr(group_seeds), _ptr(lane_rl), _ptr(group_rl), _ptr(opm), _ptr(cor_bins),
                                _ptr(hist), nthreads)
    if rc != 0:
        raise RuntimeError("oracle wave run failed (%d)" % rc)
    return dict(accepts=accepts, opMeans=opm, corBins=cor_bins, obsHist=hist)


# --------------------------------------------------------------------------
# Closed forms (FP64) used to pin the samplers
# --------------------------------------------------------------------------
def ho_discrete_x2(N: int, beta: float, k: float = 1.0) -> float:
    """Exact <x^2> of the primitive periodic HO action sum (x'-x)^2/(2 eps) + eps k x^2/2.

    Modes a_j = (2 - 2 cos(2 pi j / N)) / eps + eps k; <x^2> = (1/N) sum_j 1/a_j.
    For k = 1 this is also the mean of the virial operator k x^2.
    """
    eps = beta / N
    j = np.arange(N)
    a = (2.0 - 2.0 * np.cos(2.0 * np.pi * j / N)) / eps + eps * k
    return float(np.mean(1.0 / a))


def ho_discrete_correlator(N: int, beta: float, k: float = 1.0) -> np.ndarray:
    """Exact <x_t x_{t+tau}> for tau = 0..N-1 of the same action."""
    eps = beta / N
    j = np.arange(N)
    a = (2.0 - 2.0 * np.cos(2.0 * np.pi * j / N)) / eps + eps * k
    tau = np.arange(N)[:, None]
    return np.mean(np.cos(2.0 * np.pi * j[None, :] * tau / N) / a[None, :], axis=1)


def transfer_matrix_gap(V, N: int, beta: float, L: float = 8.0, M: int = 1200) -> float:
    """Gap E1 - E0 of the discretised action: ln(lambda_0 / lambda_1) / eps of its transfer matrix."""
    eps = beta / N
    x = np.linspace(-L, L, M)
    h = x[1] - x[0]
    v = V(x)
    K = np.exp(-(x[:, None] - x[None, :]) ** 2 / (2 * eps) - eps * (v[:, None] + v[None, :]) / 2) * h
    w = np.sort(np.linalg.eigvalsh(K))[::-1]
    return float(np.log(w[0] / w[1]) / eps)


def transfer_matrix_density(V, N: int, beta: float, L: float = 8.0, M: int = 1200):
    """(x, p): grid of [-L, L] and the one-slice density of the discretised action,
    diag(K^N) / tr(K^N) with the transfer matrix K of transfer_matrix_1d, so that the
    mean of any slice function f is sum(p * f(x))."""
    eps = beta / N
    x = np.linspace(-L, L, M)
    h = x[1] - x[0]
    v = V(x)
    K = np.exp(-(x[:, None] - x[None, :]) ** 2 / (2 * eps) - eps * (v[:, None] + v[None, :]) / 2) * h
    w, U = np.linalg.eigh(K)
    lam = (np.clip(w, 0.0, None) / w.max()) ** N
    diag = (U ** 2) @ lam
    return x, diag / lam.sum()


def transfer_matrix_1d(V, N: int, beta: float, L: float = 8.0, M: int = 1200):
    """FP64 transfer-matrix quadrature of the primitive action for a 1D potential.

    K(x, y) = exp(-(x - y)^2 / (2 eps) - eps (V(x) + V(y)) / 2) on an M-point grid of
    [-L, L]; Z = tr K^N.  Returns (<x^2>, <V>, correlator C(tau) for tau = 0..N/2).
    """
    eps = beta / N
    x = np.linspace(-L, L, M)
    h = x[1] - x[0]
    v = V(x)
    K = np.exp(-(x[:, None] - x[None, :]) ** 2 / (2 * eps) - eps * (v[:, None] + v[None, :]) / 2) * h
    w, U = np.linalg.eigh(K)
    w = np.clip(w, 0.0, None)
    lam = (w / w.max()) ** N
    Z = lam.sum()
    # diag of K^N in position basis: sum_k lam_k U_xk^2
    diag = (U ** 2) @ lam
    p = diag / Z
    x2 = float(np.sum(p * x * x))
    ev = float(np.sum(p * v))
    ntau = N // 2 + 1
    cor = np.empty(ntau)
    Ux = U.T @ (x[:, None] * U)  # <k| x |l>
    wn = w / w.max()
    for t in range(ntau):
        # tr(x K^t x K^(N-t)) / Z
        cor[t] = np.sum((Ux ** 2) * np.outer(wn ** t, wn ** (N - t))) / Z
    return x2, ev, cor
