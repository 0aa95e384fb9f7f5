"""PIMCKernel: the drop-in counterpart of FeynSimul/kernel.py for one MI355X.

Same constructor, validation (exception types and messages of
kernel.py:93-198), ``run`` and getters, with the OpenCL program replaced by a
hiprtc-compiled gfx950 module of libfeynsimul.so (ctypes) and the device
buffers held as torch tensors on the current HIP device.  The engine borrows
the buffers; Python owns them.

Differences from the reference, each a fix of a reference defect (the
numbers match the defect list kept with the design notes):
  * RANLUX streams are seeded distinctly per work-item (the reference gives
    every work-item ``seeds[0]``, kernel.c:155);
  * ``getOperators`` groups work-items by their true walker (kernel.py:592-600
    assumes threadId = walker * T + t, wrong for nbrOfWalkersPerWorkGroup > 1);
  * correlator and histogram sums keep accumulating across ``run`` calls as
    in the reference, but the getters normalise by all accumulated runs
    (kernel.py:613-614 and 630-637 normalise by one run);
  * group xorshift rows are read from a snapshot taken before each launch
    (in the reference they alias rows other work-items rewrite during it);
  * seeds are reproducible when ``KernelArgs.seed`` is set.
"""
import functools
import math
import os
import time

import numpy as np

from . import _native
from .system_source import system_section

_SM_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_SM_M1 = np.uint64(0xBF58476D1CE4E5B9)
_SM_M2 = np.uint64(0x94D049BB133111EB)
_GROUP_OFFSET = 1 << 32  # splitmix64 state offset of group streams


def _splitmix64_first(states):
    """First splitmix64 output for each uint64 generator state (vectorised)."""
    with np.errstate(over="ignore"):
        z = states.astype(np.uint64) + _SM_GAMMA
        z = (z ^ (z >> np.uint64(30))) * _SM_M1
        z = (z ^ (z >> np.uint64(27))) * _SM_M2
        return z ^ (z >> np.uint64(31))


def humanReadableSize(size):
    """Size in bytes as a short string (kernel.py:28-42)."""
    if size <= 0 or size >= 10e18:
        return "%.3g" % size + " B"
    p = int(np.log(size) / np.log(1024))
    names = ["", "ki", "Mi", "Gi", "Ti", "Pi", "Ei"]
    return "%.3g" % (size / 1024.0 ** p) + " " + names[p] + "B"


def _on_engine_stream(method):
    """Runs `method` with the engine's stream as torch's current stream: the torch work it does on the
    engine's device buffers (copies, fills, sums, reads back) is then ordered with the engine's launches
    on that stream, whichever stream the caller has made current."""
    @functools.wraps(method)
    def call(self, *args, **kwargs):
        k = getattr(self, "_k", self)
        with k._torch.cuda.stream(k._stream):
            return method(self, *args, **kwargs)
    return call


class _Queue:
    """Stand-in for the reference's ``kernel._queue`` (callers use ``finish()``)."""

    def __init__(self, stream):
        self._stream = stream

    def finish(self):
        self._stream.synchronize()


class PIMCKernel:
    """Compiles the sampler for a KernelArgs configuration and runs it on the GPU."""

    def __init__(self, ka, verbose=False, _batch_member=False):
        self._enableBins = ka.enableBins
        self._enableOperator = ka.enableOperator
        self._enableCorrelator = ka.enableCorrelator
        self._enablePathShift = ka.enablePathShift
        self._enableBisection = ka.enableBisection
        self._enableSingleNodeMove = ka.enableSingleNodeMove
        self._enableParallelizePath = ka.enableParallelizePath
        self._enableGlobalPath = ka.enableGlobalPath
        self._enableGlobalOldPath = ka.enableGlobalOldPath
        self._enableDouble = bool(ka.enableDouble)
        self._nbrOfWalkers = ka.nbrOfWalkers
        self._nbrOfWalkersPerWorkGroup = ka.nbrOfWalkersPerWorkGroup
        self._enableRanlux = ka.enableRanlux
        self._luxuaryFactor = ka.luxuaryFactor
        self._ranluxIntMax = ka.ranluxIntMax
        self._N = ka.N
        self._S = ka.S
        self._beta = ka.beta
        self._operatorRuns = ka.operatorRuns
        self._metroStepsPerOperatorRun = ka.metroStepsPerOperatorRun
        self._system = ka.system
        self._operators = ()
        self._correlators = ()

        # ---- validation, in the reference's order (kernel.py:93-198) ----
        if not isinstance(self._beta, float):
            raise TypeError('beta need to be a float')
        if self._enableOperator:
            self._operators = ka.operators
            if self._operatorRuns is None:
                raise NameError('operatorRuns need to be set if enableOperator is True.')
            if self._operators is None:
                raise NameError('operators need to be set if enableOperator is True.')
            if self._metroStepsPerOperatorRun is None:
                raise NameError('metroStepsPerOperatorRun need to be set if enableOperator is True.')
            if not isinstance(self._operators, tuple):
                raise TypeError('operators need to be a tuple of strings.')
        if self._enableCorrelator:
            self._correlators = ka.correlators
            if self._correlators is None:
                raise NameError('correlators need to be set if enableCorrelator is True.')
            if not isinstance(self._correlators, tuple):
                raise TypeError('correlators need to be a tuple of strings.')
        if self._enableSingleNodeMove:
            self._alpha = ka.alpha
            if self._alpha is None:
                raise NameError('alpha need to be set if enableSingleNodeMove is True.')
        if self._enablePathShift:
            self._PSAlpha = ka.PSAlpha
            if self._PSAlpha is None:
                raise NameError('PSALpha need to be set if enablePathShift is True')
        if self._enableRanlux:
            if self._luxuaryFactor is None:
                raise NameError('luxuaryFactor needs to be an uint >= 0.')
            if not isinstance(self._luxuaryFactor, (int, np.integer)):
                raise NameError('luxuaryFactor needs to be an integer.')
            if self._luxuaryFactor < 0:
                raise NameError('luxuaryFactor needs to be an uint >= 0.')
            try:
                # the reference increments the value it was given and uses the
                # result as RANLUXCL_LUX (kernel.py:137, 331-332)
                self._luxuaryFactor += 1
            except TypeError:
                raise NameError('luxuaryFactor needs to be an integer.')
            if self._ranluxIntMax is None:
                raise NameError('ranluxIntMax needs to be an uint > 0.')
            if not isinstance(self._ranluxIntMax, (int, np.integer)):
                raise NameError('ranluxIntMax needs to be an integer.')
            if self._ranluxIntMax <= 0:
                raise NameError('ranluxIntMax needs to be an uint > 0.')
            try:
                self._ranluxIntMax += 1
            except TypeError:
                raise NameError('ranluxIntMax needs to be an integer.')
            lux = self._luxuaryFactor
            nskip = {0: 0, 1: 24, 2: 76, 3: 200, 4: 380}.get(lux, lux - 24)
            if nskip < 0:
                raise Exception("ranluxcl: nskip is negative (luxury value %d)" % lux)
            if nskip % 4 != 0 or (0 < nskip < 24):
                raise Exception("ranluxcl: nskip must be divisible by 4 and either 0 or >= 24")
        if self._enableBins:
            self._binResolutionPerDOF = ka.binResolutionPerDOF
            self._xMin = ka.xMin
            self._xMax = ka.xMax
            if self._xMin is None:
                raise NameError('xMin need to be set if enableBins is True.')
            if self._xMax is None:
                raise NameError('xMax need to be set if enableBins is True.')
            if self._binResolutionPerDOF is None:
                raise NameError('binResolutionPerDOF need to be set if enableBins is True.')
        if self._nbrOfWalkers % self._nbrOfWalkersPerWorkGroup != 0:
            raise Exception("The number of walkers must be a multiple of the number of walkers per work group.")
        if self._nbrOfWalkers < self._nbrOfWalkersPerWorkGroup:
            raise Exception("The number of walkers per group must be less than or equal to the total number "
                            "of walkers")
        if not (self._N != 0 and ((self._N & (self._N - 1)) == 0)):
            raise Exception("Path node number, N, must be a power of 2.")
        if self._enableBisection:
            if self._S is None:
                raise NameError('S need to be set if enableBisection is True')
            if 2 ** self._S > self._N:
                raise Exception("2^S must be less than or equal to N.")
            if self._enableGlobalPath is None:
                raise NameError('enableGlobalPath need to be set if enableBisection is True.')
            if self._enableGlobalOldPath is None:
                raise NameError('enableGlobalOldPath need to be setif enableBisection is True.')
            if self._enableParallelizePath is None:
                raise NameError('enableParallelizePath need to be setif enableBisection is True')
        if self._enableParallelizePath:
            if self._enableSingleNodeMove:
                raise Exception("Single node move not possible when enableParallelizePath is True.")
            if self._enablePathShift:
                raise Exception("Shift path not possible when enableParallelizePath is True.")
        if self._enableGlobalOldPath and not self._enableBisection:
            raise Exception("Old Path is not used when bisection is not enabled")
        if self._operatorRuns is None or self._metroStepsPerOperatorRun is None:
            raise TypeError("%d format: a real number is required, not NoneType")
        if self._S is None:
            # without bisection S only sizes unused tables; ParallelizePath needs it (kernel.py:200)
            if self._enableParallelizePath:
                raise TypeError("unsupported operand type(s) for ** or pow(): 'int' and 'NoneType'")
            self._S = 1
        if self._enableSingleNodeMove:
            self._alpha = float(self._alpha)
        if self._enablePathShift:
            self._PSAlpha = float(self._PSAlpha)

        # ---- geometry (kernel.py:200-211) ----
        T = self._N // (2 ** self._S)
        if self._enableParallelizePath:
            self._localSize = (T, self._nbrOfWalkersPerWorkGroup)
            self._globalSize = ((self._nbrOfWalkers // self._nbrOfWalkersPerWorkGroup) * T,
                                self._nbrOfWalkersPerWorkGroup)
            self._nbrOfThreads = self._nbrOfWalkers * T
        else:
            self._localSize = None
            self._globalSize = (self._nbrOfWalkers,)
            self._nbrOfThreads = self._nbrOfWalkers

        # ---- sampler choice ----
        sampler = getattr(ka, "sampler", None)
        onlyBisection = self._enableBisection and not (self._enablePathShift or self._enableSingleNodeMove)
        if sampler is None:
            sampler = "wave" if (onlyBisection and self._enableParallelizePath and 2 ** self._S <= 64) else "exact"
        if sampler not in ("wave", "exact"):
            raise ValueError("sampler must be 'wave' or 'exact'")
        if sampler == "wave" and not onlyBisection:
            raise Exception("the wave sampler runs bisection moves only; use sampler='exact'")
        if sampler == "wave" and 2 ** self._S > 64:
            raise Exception("the wave sampler needs 2^S <= 64")
        if sampler == "exact" and self._enableParallelizePath and T * self._nbrOfWalkersPerWorkGroup > 1024:
            raise Exception("work-group of %d items exceeds the device limit of 1024"
                            % (T * self._nbrOfWalkersPerWorkGroup))
        self._sampler = sampler
        self._histObs = tuple(getattr(ka, "histogramObservables", None) or ())
        if self._histObs:
            if sampler != "wave":
                raise Exception("histogramObservables need the wave sampler")
            rng_ = getattr(ka, "histogramRange", None)
            if rng_ is None or len(rng_) != 2 or not rng_[1] > rng_[0]:
                raise Exception("histogramRange must be (lo, hi) with hi > lo")
            self._histRange = (float(rng_[0]), float(rng_[1]))
            self._histBins = int(getattr(ka, "histogramBins", 256))
        lags = getattr(ka, "correlatorLags", None)
        self._corLags = int(lags) if lags else self._N // 2
        if self._enableCorrelator and not (1 <= self._corLags <= self._N // 2):
            raise Exception("correlatorLags must be in [1, N/2]")
        if sampler == "exact":
            self._corLags = self._N // 2

        import torch  # tensor plumbing only
        self._torch = torch
        if not torch.cuda.is_available() or _native.device_count() == 0:
            raise RuntimeError("PIMCKernel needs a GPU: no HIP device is visible")
        self._device = torch.device("cuda", torch.cuda.current_device())
        self._stream = torch.cuda.current_stream(self._device)
        self._queue = _Queue(self._stream)

        # ---- compile (kernel.py:346-382) ----
        cfg = _native.FsConfig()
        cfg.abi_version = _native.FS_ABI_VERSION
        cfg.sampler = _native.FS_SAMPLER_WAVE if sampler == "wave" else _native.FS_SAMPLER_EXACT
        cfg.N, cfg.S, cfg.DOF = self._N, self._S, int(self._system.DOF)
        cfg.walkers_per_group = self._nbrOfWalkersPerWorkGroup
        cfg.n_walkers = self._nbrOfWalkers
        cfg.parallelize_path = int(bool(self._enableParallelizePath))
        cfg.local_size = self._nbrOfWalkersPerWorkGroup
        cfg.double_precision = int(self._enableDouble)
        cfg.ranlux = int(bool(self._enableRanlux))
        cfg.ranlux_lux = int(self._luxuaryFactor) if self._enableRanlux else 0
        cfg.fast_math = int(bool(getattr(ka, "fastMath", True)))
        cfg.ranlux_int_range = int(self._ranluxIntMax) if self._enableRanlux else 1
        cfg.n_ops = len(self._operators) if self._enableOperator else 0
        cfg.n_corr = len(self._correlators) if self._enableCorrelator else 0
        cfg.cor_lags = self._corLags
        cfg.bins = int(bool(self._enableBins))
        cfg.bin_res = int(self._binResolutionPerDOF) if self._enableBins else 1
        cfg.waves_per_block = int(getattr(ka, "wavesPerBlock", None) or 0)
        cfg.max_waves_per_simd = int(getattr(ka, "wavesPerSimd", None) or 0)
        pc = getattr(ka, "potentialCache", None)
        cfg.potential_cache = (_native.FS_POTCACHE_AUTO if pc is None else
                               _native.FS_POTCACHE_ON if pc else _native.FS_POTCACHE_OFF)
        cfg.beta = float(self._beta)
        cfg.x_min = float(self._xMin) if self._enableBins else 0.0
        cfg.x_max = float(self._xMax) if self._enableBins else 1.0
        cfg.op_norm = 1.0 / float(self._operatorRuns * self._N * self._nbrOfWalkers / float(self._nbrOfThreads))
        storage = getattr(ka, "pathStorage", None)
        storages = {None: _native.FS_STORAGE_AUTO, "memory": _native.FS_STORAGE_MEMORY,
                    "registers": _native.FS_STORAGE_REGISTERS}
        if storage not in storages:
            raise ValueError("pathStorage must be 'registers', 'memory' or None")
        cfg.path_storage = storages[storage]
        cfg.no_bisection = int(not self._enableBisection)
        cfg.path_shift = int(bool(self._enablePathShift))
        cfg.single_node = int(bool(self._enableSingleNodeMove))
        cfg.alpha = self._alpha if self._enableSingleNodeMove else 0.0
        cfg.ps_alpha = self._PSAlpha if self._enablePathShift else 0.0
        if self._histObs:
            cfg.n_hist_obs = len(self._histObs)
            cfg.hist_bins = self._histBins
            cfg.hist_lo, cfg.hist_hi = self._histRange
        self._cfg = cfg
        self._systemSource = system_section(self._system, self._operators if cfg.n_ops else (),
                                            self._correlators if cfg.n_corr else (), self._histObs)
        seed = getattr(ka, "seed", None)
        if seed is None:
            seed = int.from_bytes(os.urandom(8), "little")
        self._seed = int(seed) & ((1 << 64) - 1)
        self._ev0 = torch.cuda.Event(enable_timing=True)
        self._ev1 = torch.cuda.Event(enable_timing=True)
        self._ranRuns = 0
        self._corRuns = 0
        self._binRuns = 0
        self._seriesCount = 0
        self._lastRunLength = None
        self._engine = None
        if _batch_member:
            return  # a BatchedPIMC compiles every member in one module, then calls _attach
        self._engine = _native.Engine(cfg, self._systemSource)
        if verbose:
            self._dumpSource(self._engine.source)
        self._attach(self._engine.path_storage)

    @staticmethod
    def _dumpSource(source):
        """The generated module source in cleaned.hip and, run through the C preprocessor with
        comments and line markers dropped, in preprocessed.hip (kernel.py:351-364 writes cleaned.c and
        preprocessed.c)."""
        import subprocess
        with open("cleaned.hip", "w") as f:
            f.write(source)
        dev = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "device")
        try:
            res = subprocess.run(["cpp", "-P", "-I", dev, "-D__HIPCC_RTC__"], input=source, capture_output=True,
                                 text=True, timeout=60)
            text = res.stdout if res.returncode == 0 else "// cpp failed:\n" + res.stderr
        except (OSError, subprocess.SubprocessError) as e:
            text = "// cpp could not run: %s\n" % e
        with open("preprocessed.hip", "w") as f:
            f.write("".join(line + "\n" for line in text.splitlines() if line.strip()))
        print("Saved cleaned and preprocessed code in cleaned.hip and preprocessed.hip")

    def _attach(self, storage):
        """Allocates the device buffers (kernel.py:386-457) for the given path storage."""
        self._storage = storage
        self._allocate()

    # ------------------------------------------------------------------
    def _zeros(self, shape, dtype):
        return self._torch.zeros(shape, dtype=dtype, device=self._device)

    def _allocate(self):
        t = self._torch
        ft = t.float64 if self._enableDouble else t.float32
        W, N, DOF = self._nbrOfWalkers, self._N, int(self._system.DOF)
        nOps, nCorr = self._cfg.n_ops, self._cfg.n_corr
        exact = self._sampler == "exact"
        threads = self._nbrOfThreads
        self._paths = self._zeros((W, N * DOF), ft)
        self._accepts = self._zeros((threads if exact else W,), t.int32)
        self._seeds = self._seedsSnap = self._rl = self._grl = None
        stream = self._stream.cuda_stream
        if self._enableRanlux:
            n_rl = threads if exact else W * 64
            self._rl = self._zeros((n_rl, 28), t.int32)
            self._grl = self._zeros((threads if exact else W, 28), t.int32)
        else:
            rows = threads + 1 if exact else W * 65
            self._seeds = self._zeros((rows, 4), t.int32)
            _native.check(_native.lib().fs_seed_xorshift(self._seeds.data_ptr(), rows, self._seed, 0, stream))
            if exact:
                self._seedsSnap = self._zeros((rows, 4), t.int32)
        self._oldPath = self._zeros((threads, (2 ** self._S - 1) * DOF), ft) if exact else None
        self._potCache = None
        # the potential cache holds V at every slice of the current paths (the last launch left it so and
        # nothing has written the paths since): the next launch need not recompute it
        self._potValid = False
        if not exact and self._storage == _native.FS_STORAGE_MEMORY:
            self._potCache = self._zeros((W, N), ft)
        self._opTotals = None
        self._series = None
        if nOps:
            if exact:
                self._operatorValues = self._zeros((threads * nOps,), ft)
            else:
                self._operatorValues = self._zeros((W, nOps), t.float64)
                self._opTotals = self._zeros((W, nOps, 2), t.float64)
        else:
            self._operatorValues = None
        if nCorr:
            if exact:
                self._correlatorValues = self._zeros((W, nCorr, N // 2), ft)
            else:
                self._correlatorValues = self._zeros((W, nCorr, self._corLags, 2), t.float64)
        else:
            self._correlatorValues = None
        if self._enableBins:
            self._binCounts = self._zeros((self._binResolutionPerDOF ** DOF,), t.int32)
        else:
            self._binCounts = None
        self._obsHist = self._zeros((self._histBins,), t.int32) if self._histObs else None
        self._obsRuns = 0

    def _buffers(self):
        b = _native.FsBuffers()

        def p(x):
            return x.data_ptr() if x is not None else None
        b.paths = p(self._paths)
        b.accepts = p(self._accepts)
        b.seeds = p(self._seeds)
        b.seeds_snap = p(self._seedsSnap)
        b.rl = p(self._rl)
        b.grl = p(self._grl)
        b.old_paths = p(self._oldPath)
        b.op_means = p(self._operatorValues)
        b.op_totals = p(self._opTotals)
        b.series = p(self._series)
        b.cor_means = p(self._correlatorValues)
        b.bins = p(self._binCounts)
        b.pot_cache = p(self._potCache)
        b.obs_hist = p(self._obsHist)
        return b

    def _thread_layout(self):
        """(walker, item) -> threadId of the reference's geometry (kernel.c:538-541)."""
        W, Wg = self._nbrOfWalkers, self._nbrOfWalkersPerWorkGroup
        if not self._enableParallelizePath:
            tid = np.arange(W)[:, None]
            return tid, np.arange(W) // Wg
        T = self._N // 2 ** self._S
        nG = W // Wg
        w = np.arange(W)
        g, l = w // Wg, w % Wg
        tid = (g * T)[:, None] + np.arange(T)[None, :] + (l * nG * T)[:, None]
        return tid, g + l * nG

    # ------------------------------------------------------------------
    def run(self, wait=True):
        """Runs operatorRuns x metroStepsPerOperatorRun steps (kernel.py:459-506)."""
        if self._engine is None:
            raise Exception("this kernel is a member of a BatchedPIMC; run the batch")
        self._launch(int(self._operatorRuns), int(self._metroStepsPerOperatorRun))
        if wait:
            startTime = time.time()
            try:
                self._stream.synchronize()
            except RuntimeError:
                if time.time() - startTime > 5.0:
                    raise Exception("Kernel runtime error. Over 5 seconds had passed when kernel aborted.")
                raise

    def _launch(self, opRuns, M):
        """One engine launch of opRuns measurements, each after M sweeps (M = 0: measure only)."""
        offset = 0
        if self._series is not None:
            if self._seriesCount + opRuns > self._series.shape[0]:
                raise Exception("series buffer full: %d measurements stored, capacity %d"
                                % (self._seriesCount, self._series.shape[0]))
            offset = self._seriesCount
        bufs = self._buffers()
        flags = _native.FS_RUN_POT_CACHE_VALID if self._potValid else 0
        self._ev0.record(self._stream)
        self._engine.run(bufs, opRuns, M, offset, self._stream.cuda_stream, flags)
        self._ev1.record(self._stream)
        self._potValid = self._potCache is not None
        self._account_run(opRuns, M)

    @_on_engine_stream
    def initRanlux(self):
        """Seeds every work-item's RANLUX state (kernel.py:508-530), one distinct seed each."""
        if not self._enableRanlux:
            raise Exception("initRanlux needs enableRanlux = True")
        n = self._rl.shape[0]
        seeds = _splitmix64_first(np.uint64(self._seed) + np.arange(n, dtype=np.uint64))
        self._ranlux_fill(self._rl, seeds)

    @_on_engine_stream
    def groupInitRanlux(self):
        """Seeds the group (segment-offset) RANLUX states (kernel.py:532-554): one stream per walker group."""
        if not self._enableRanlux:
            raise Exception("groupInitRanlux needs enableRanlux = True")
        if self._sampler == "exact":
            tid, grow = self._thread_layout()
            rowOfThread = np.empty(self._nbrOfThreads, dtype=np.uint64)
            rowOfThread[tid.reshape(-1)] = np.repeat(grow, tid.shape[1]).astype(np.uint64)
        else:
            rowOfThread = np.arange(self._nbrOfWalkers, dtype=np.uint64)
        with np.errstate(over="ignore"):
            states = np.uint64(self._seed) + np.uint64(_GROUP_OFFSET) + rowOfThread
        self._ranlux_fill(self._grl, _splitmix64_first(states))

    def _ranlux_fill(self, tab, seeds):
        t = self._torch
        dev = t.from_numpy(seeds.view(np.int64)).to(self._device)
        _native.check(_native.lib().fs_ranlux_init(tab.data_ptr(), dev.data_ptr(), tab.shape[0],
                                                   self._stream.cuda_stream))
        self._stream.synchronize()

    @_on_engine_stream
    def getPaths(self):
        """Paths as (walkers, N * DOF), DOF-major per walker (kernel.py:557-565)."""
        return self._paths.cpu().numpy()

    @_on_engine_stream
    def setPaths(self, paths):
        """Replaces the paths (kernel.py:567-574); cast to the kernel's precision."""
        a = np.asarray(paths).astype(np.float64 if self._enableDouble else np.float32)
        a = a.reshape(self._paths.shape)
        self._paths.copy_(self._torch.from_numpy(np.ascontiguousarray(a)))
        self._potValid = False
        self._stream.synchronize()

    @_on_engine_stream
    def getOperators(self):
        """Per-walker operator means, shape (nOps, walkers) (kernel.py:576-606)."""
        if not self._enableOperator:
            raise Exception("Can only return operator when created withkernelArg.enableOperator = True")
        nOps = len(self._operators)
        raw = self._operatorValues.cpu().numpy().astype(np.float64)
        if self._sampler == "wave":
            return raw.T.copy()
        tid, _ = self._thread_layout()
        per = raw.reshape(self._nbrOfThreads, nOps)
        return per[tid].mean(axis=1).T.copy()

    @_on_engine_stream
    def getOperatorTotals(self):
        """Running FP64 (Kahan) sums of the per-measurement slice sums, (nOps, walkers); wave sampler."""
        if self._opTotals is None:
            raise Exception("operator totals are kept by the wave sampler with operators enabled")
        t = self._opTotals.cpu().numpy()
        return (t[..., 0] - t[..., 1]).T.copy()

    def setWavesPerSimd(self, cap):
        """Occupancy cap of the wave sampler's launches (0 = none); results do not depend on it."""
        if self._engine is None:
            raise Exception("this kernel is a member of a BatchedPIMC; the batch sets its launch shape")
        self._engine.set_waves_per_simd(int(cap or 0))

    @_on_engine_stream
    def setSeriesCapacity(self, measurements):
        """Keeps every measurement's per-walker slice mean (for autocorrelation / ESS); wave sampler."""
        if self._sampler != "wave" or not self._cfg.n_ops:
            raise Exception("measurement series are kept by the wave sampler with operators enabled")
        self._series = self._zeros((int(measurements), self._nbrOfWalkers, self._cfg.n_ops), self._torch.float32)
        self._seriesCount = 0

    @_on_engine_stream
    def getSeries(self):
        """(measurements, walkers, nOps) float32 of the stored measurements."""
        if self._series is None:
            raise Exception("call setSeriesCapacity first")
        return self._series[: self._seriesCount].cpu().numpy()

    def getSeriesTensor(self):
        """The stored measurements as a device tensor (measurements, walkers, nOps), without a copy."""
        if self._series is None:
            raise Exception("call setSeriesCapacity first")
        return self._series[: self._seriesCount]

    @_on_engine_stream
    def getCorrelator(self):
        """(mean, standard error) of the correlators, each (nCorr, lags) (kernel.py:608-617)."""
        if not self._enableCorrelator:
            raise Exception("Can only return correlator when created withkernalArg.enableCorrelator = True")
        runs = max(self._corRuns, 1)
        raw = self._correlatorValues.cpu().numpy()
        if self._sampler == "wave":
            raw = raw[..., 0] - raw[..., 1]
        per = raw.astype(np.float64) / float(runs * self._N)
        mean = per.mean(axis=0)
        se = per.std(axis=0) / np.sqrt(self._nbrOfWalkers)
        return (mean, se)

    @_on_engine_stream
    def resetCorrelator(self):
        if self._correlatorValues is not None:
            self._correlatorValues.zero_()
        self._corRuns = 0

    @_on_engine_stream
    def getBinCounts(self):
        """Probability density on the DOF-dimensional grid (kernel.py:619-637)."""
        if not self._enableBins:
            raise Exception("Can only return bins when kernalArg.enableBins = True")
        DOF = int(self._system.DOF)
        tot = self._nbrOfWalkers * max(self._binRuns, 1) * self._N
        binVol = (float(self._xMax - self._xMin) / float(self._binResolutionPerDOF)) ** DOF
        counts = self._binCounts.cpu().numpy().view(np.uint32).reshape((self._binResolutionPerDOF,) * DOF)
        return counts.astype(np.float64 if self._enableDouble else np.float32) / (binVol * tot)

    @_on_engine_stream
    def resetBins(self):
        if self._binCounts is not None:
            self._binCounts.zero_()
        self._binRuns = 0
        if self._obsHist is not None:
            self._obsHist.zero_()
        self._obsRuns = 0

    @_on_engine_stream
    def getObservableHistogram(self):
        """(bin centres, density, counts) of the histogram of ``histogramObservables``.

        The density integrates to the fraction of observable values inside the
        range: counts / (bin width x walkers x slices x measurements x observables).
        """
        if self._obsHist is None:
            raise Exception("no histogramObservables were configured")
        counts = self._obsHist.cpu().numpy().view(np.uint32).astype(np.float64)
        lo, hi = self._histRange
        width = (hi - lo) / self._histBins
        centres = lo + width * (np.arange(self._histBins) + 0.5)
        tot = self._nbrOfWalkers * self._N * max(self._obsRuns, 1) * len(self._histObs)
        return centres, counts / (width * tot), counts

    def getRunTime(self):
        """Device time of the last run in seconds (kernel.py:638-646)."""
        self._ev1.synchronize()
        return 1e-3 * self._ev0.elapsed_time(self._ev1)

    def getAcceptanceRate(self):
        """Mean acceptance ratio of the bisection moves of the last launch (kernel.py:648-658).

        The denominator is the length of the last launch (operator runs x steps), which is the
        configured length for ``run`` and differs for the launches a learned proposal or a batch
        makes; a launch that only measured (no steps) has no acceptance rate."""
        acc = self._accepts.cpu().numpy().view(np.uint32).astype(np.float64)
        opRuns, M = self._lastRunLength or (self._operatorRuns, self._metroStepsPerOperatorRun)
        moves = float(opRuns * M)
        if moves == 0:
            raise Exception("the last launch made no Metropolis moves (it only measured)")
        if self._sampler == "wave":
            return acc.sum() / (self._nbrOfWalkers * (self._N // 2 ** self._S) * moves)
        return (acc / moves).mean()

    def getGlobalMemory(self):
        """Bytes of device memory held by this kernel (kernel.py:660-684)."""
        tot = 0
        for x in (self._paths, self._accepts, self._seeds, self._seedsSnap, self._rl, self._grl, self._oldPath,
                  self._operatorValues, self._opTotals, self._series, self._correlatorValues, self._binCounts,
                  self._potCache, self._obsHist):
            if x is not None:
                tot += x.numel() * x.element_size()
        return tot

    def getStats(self):
        """Memory and launch-shape summary (kernel.py:686-730)."""
        props = self._torch.cuda.get_device_properties(self._device)
        ret = "Device: %s (%s)\n" % (props.name, getattr(props, "gcnArchName", "gfx"))
        ret += ("Global memory (used/max): " + humanReadableSize(self.getGlobalMemory()) + " / "
                + humanReadableSize(props.total_memory) + "\n")
        ret += "Sampler: %s\n" % self._sampler
        if self._sampler == "wave":
            wpb = self._cfg.waves_per_block or 4
            ret += "Workgroup size: %d (%d wavefronts, one walker each)\n" % (64 * wpb, wpb)
            ret += "Number of workgroups (used): %d" % math.ceil(self._nbrOfWalkers / wpb)
        else:
            ret += "Workgroup dimensions (used/max): " + str(self._localSize) + " / (1024, 1024, 1024)\n"
            ret += "Number of workgroups (used): " + str(self._nbrOfWalkers // self._nbrOfWalkersPerWorkGroup)
        return ret

    def _launchInfo(self):
        if self._engine is None:
            raise Exception("this kernel is a member of a BatchedPIMC; the batch owns its module")
        return self._engine.launch_info()

    def getMaxKernelWGSize(self):
        """Largest work-group the compiled sampler kernel can run (kernel.py:732-734)."""
        return int(self._launchInfo().max_threads_per_block)

    def exceedsLimits(self):
        """True if this configuration exceeds the device's limits (kernel.py:736-775): global memory,
        the kernel's LDS against the device's per-work-group LDS, the work-group size against the
        device's and the kernel's limits, and each work-group dimension against the device's."""
        props = self._torch.cuda.get_device_properties(self._device)
        if self.getGlobalMemory() > props.total_memory:
            return True
        info = self._launchInfo()
        if info.lds_bytes > info.lds_limit:
            return True
        if self._enableParallelizePath and self._sampler == "exact":
            wg = (self._N // 2 ** self._S) * self._nbrOfWalkersPerWorkGroup
            if wg > info.device_max_threads_per_block or wg > info.max_threads_per_block:
                return True
        if info.threads_per_block > min(info.device_max_threads_per_block, info.max_threads_per_block):
            return True
        local = self._localSize if (self._localSize and self._sampler == "exact") else (info.threads_per_block,)
        for i, n in enumerate(local):
            if n > info.device_max_block_dim[i]:
                return True
        return False

    def getBuildLog(self):
        return self._engine.build_log if self._engine is not None else ""

    def _account_run(self, opRuns, M):
        """Bookkeeping of one launch run by a BatchedPIMC on this member's buffers."""
        self._ranRuns += 1
        self._corRuns += opRuns
        self._binRuns += opRuns
        self._obsRuns += opRuns
        if self._series is not None:
            self._seriesCount += opRuns
        self._lastRunLength = (opRuns, M)

    def getMetroStepsPerRun(self):
        """Metropolis moves per run (kernel.py:787-801)."""
        if self._enableParallelizePath:
            T = self._N // (2 ** self._S)
            return (self._nbrOfWalkers * T * self._metroStepsPerOperatorRun * self._operatorRuns
                    * self._enableBisection)
        return (self._nbrOfWalkers * self._metroStepsPerOperatorRun * self._operatorRuns
                * (self._enableBisection + self._enablePathShift + self._enableSingleNodeMove))

    @_on_engine_stream
    def evalSystem(self, x):
        """Potential and operators of the compiled module at points x (n, DOF): (pot[n], ops[n, nOps])."""
        t = self._torch
        ft = t.float64 if self._enableDouble else t.float32
        xd = t.as_tensor(np.ascontiguousarray(x), dtype=ft).reshape(-1, int(self._system.DOF)).to(self._device)
        n = xd.shape[0]
        pot = self._zeros((n,), t.float64)
        ops = self._zeros((n, max(self._cfg.n_ops, 1)), t.float64)
        self._engine.eval_system(xd.data_ptr(), pot.data_ptr(), ops.data_ptr(), n, self._stream.cuda_stream)
        self._stream.synchronize()
        return pot.cpu().numpy(), ops.cpu().numpy()[:, : self._cfg.n_ops]
