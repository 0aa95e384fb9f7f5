"""Learned bisection proposal: a conditional MLP on BF16 MFMA, MH-corrected.

No counterpart in the reference.  It replaces the Levy midpoint draw of
FeynSimul/GPUsrc/kernel.c:338-360 by a Gaussian whose mean and log-width come
from an MLP (2 DOF + 1 -> 128 -> 128 -> 128 -> 2 DOF) fed with a node's two
bisection ends and its level, and extends the accept test of kernel.c:376-377
with the exact ratio of proposal densities, so any weights keep detailed
balance:

  log alpha = -(K_new - K_old) - eps (V_new - V_old) + log q(old) - log q(new).

With the output layer at zero the proposal is the Levy bridge (log alpha =
-eps dV, the reference's rule); the network starts there.  Training is online
maximum likelihood of midpoints of the current paths given their ends
(Adam on FP32 master weights).  ``run`` interleaves learned sweeps with the
kernel's measurements and falls back to plain bisection sweeps for a block
whenever the learned acceptance drops below ``fallbackAcceptance``.

All arithmetic runs in libfeynsimul.so (fs_lp_* entry points); torch only
holds the device buffers.
"""
import numpy as np

from . import _native
from .kernel import _on_engine_stream


def _mask_width(theta, dof, active, hidden=128):
    """Zero the parameters of the hidden units beyond `active` (weights in and out, biases)."""
    theta = np.array(theta, dtype=np.float32, copy=True)
    if active >= hidden:
        return theta
    nin, nout, H = 2 * dof + 1, 2 * dof, hidden
    o = 0
    W1 = theta[o:o + H * nin].reshape(H, nin); o += H * nin
    b1 = theta[o:o + H]; o += H
    W2 = theta[o:o + H * H].reshape(H, H); o += H * H
    b2 = theta[o:o + H]; o += H
    W3 = theta[o:o + H * H].reshape(H, H); o += H * H
    b3 = theta[o:o + H]; o += H
    W4 = theta[o:o + nout * H].reshape(nout, H)
    W1[active:] = 0.0
    b1[active:] = 0.0
    W2[active:] = 0.0
    W2[:, active:] = 0.0
    b2[active:] = 0.0
    W3[active:] = 0.0
    W3[:, active:] = 0.0
    b3[active:] = 0.0
    W4[:, active:] = 0.0
    return theta


def _he_init(dof, seed, hidden=128, active=128):
    """He-normal hidden layers, zero biases, zero output layer (the Levy bridge); units beyond
    `active` zero."""
    rng = np.random.default_rng(seed)
    nin, nout = 2 * dof + 1, 2 * dof
    parts = [rng.standard_normal((hidden, nin)) * np.sqrt(2.0 / nin), np.zeros(hidden),
             rng.standard_normal((hidden, hidden)) * np.sqrt(2.0 / active), np.zeros(hidden),
             rng.standard_normal((hidden, hidden)) * np.sqrt(2.0 / active), np.zeros(hidden),
             np.zeros((nout, hidden)), np.zeros(nout)]
    return _mask_width(np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32), dof, active, hidden)


def harmonic_omega2(kernel, points=65536, h=1e-3, seed=0):
    """omega^2 = <x . grad V(x)> / <|x|^2> over up to `points` slices of the kernel's paths, at least 0:
    the frequency of the harmonic potential that has the same virial as V on these paths (exactly the
    oscillator's own for a harmonic V).  x . grad V is the derivative of V(t x) at t = 1, by central
    differences of the compiled potential (PIMCKernel.evalSystem)."""
    t = kernel._torch
    D, N, W = int(kernel._system.DOF), int(kernel._N), int(kernel._nbrOfWalkers)
    rng = np.random.default_rng(seed)
    n = min(points, W * N)
    iw, isl = t.from_numpy(rng.integers(0, W, n)), t.from_numpy(rng.integers(0, N, n))
    with t.cuda.stream(kernel._stream):
        iw, isl = iw.to(kernel._device), isl.to(kernel._device)
        x = kernel._paths.view(W, D, N)[iw, :, isl].double().cpu().numpy()      # (n, D)
    vp, _ = kernel.evalSystem(x * (1.0 + h))
    vm, _ = kernel.evalSystem(x * (1.0 - h))
    xg = (vp - vm) / (2.0 * h)
    return max(0.0, float(xg.mean() / (x * x).sum(axis=1).mean()))


class LearnedProposal:
    """Learned bisection sweeps on the paths of one PIMCKernel.

    kernel:   a PIMCKernel (FP32, any sampler; its paths are updated in place)
    batch:    training rows (segment midpoints) per Adam step, multiple of 2048
    coarse:   the whole-path move instead of segment sweeps (N = 64 x 2^k): the network proposes
              the nodes every 64 slices of the whole cyclic path (bisection levels 1..k), a Levy
              bridge fills the slices between them, one Metropolis-Hastings decision per path;
              training uses those nodes only
    width:    active hidden units (multiple of 32, at most 128; default 64 for the whole-path move,
              128 otherwise): the parameters keep the 128-wide layout with the other units at zero,
              which training preserves (their ReLU masks stop every gradient)
    lr:       Adam learning rate (beta1 0.9, beta2 0.999, eps 1e-8)
    seed:     Philox key of the proposal's random numbers (default: the kernel's seed + 1)
    atoms3d:  centre the inputs on the atoms' mean (default: system.translationInvariant)
    inScale:  scale of the input coordinates
    bridgeOmega2: the whole-path move's bridges: 0 (default) the free-particle (Levy) bridge; omega^2 > 0
              the exact bridge of the kinetic links plus a harmonic reference omega^2 |x|^2 / 2 per
              slice (the accept test then scores V - omega^2 |x|^2 / 2); "auto" takes omega^2 =
              <x . grad V> / <|x|^2> over the kernel's current paths (at least 0), the harmonic
              potential's own frequency for an oscillator
    """

    def __init__(self, kernel, batch=65536, lr=1e-3, seed=None, atoms3d=None, inScale=1.0, segsPerWave=0,
                 initSeed=0, fallbackAcceptance=0.2, coarse=False, width=None, bridgeOmega2=0.0):
        if kernel._enableDouble:
            raise Exception("the learned proposal runs on FP32 paths")
        if kernel._engine is None:
            raise Exception("the learned proposal needs a stand-alone PIMCKernel")
        self._k = kernel
        t = kernel._torch
        system = kernel._system
        self._dof = int(system.DOF)
        if atoms3d is None:
            atoms3d = bool(getattr(system, "translationInvariant", False)) and self._dof % 3 == 0
        if seed is None:
            seed = (kernel._seed + 1) & ((1 << 64) - 1)
        lp = _native.FsLpConfig()
        lp.abi_version = _native.FS_ABI_VERSION
        lp.atoms3d = int(bool(atoms3d))
        lp.batch = int(batch)
        lp.segs_per_wave = int(segsPerWave)
        lp.in_scale = float(inScale)
        lp.lr, lp.beta1, lp.beta2, lp.adam_eps = float(lr), 0.9, 0.999, 1e-8
        lp.seed = int(seed) & ((1 << 64) - 1)
        lp.coarse = int(bool(coarse))
        self.coarse = bool(coarse)
        self.width = int(width if width is not None else (64 if coarse else 128))
        lp.width = self.width
        if isinstance(bridgeOmega2, str):
            if bridgeOmega2 != "auto":
                raise ValueError("bridgeOmega2 must be a number or 'auto'")
            bridgeOmega2 = harmonic_omega2(kernel)
        self.bridgeOmega2 = float(bridgeOmega2)
        lp.bridge_omega2 = self.bridgeOmega2
        # moves proposed per sweep: every segment, or (whole-path move) every path
        self._movesPerSweep = kernel._nbrOfWalkers * (1 if coarse else kernel._N >> kernel._S)
        self._P = _native.Proposal(kernel._cfg, kernel._systemSource, lp)
        info = self._P.info
        z = lambda n, dt: t.zeros((int(n),), dtype=dt, device=kernel._device)
        self._params = z(info.n_params, t.float32)
        self._m = z(info.n_params, t.float32)
        self._v = z(info.n_params, t.float32)
        self._image = z(info.image_bytes, t.uint8)
        self._tw = z(info.train_bf16, t.int16)
        self._tb = z(info.train_bias, t.float32)
        self._ws = z(info.workspace_bytes, t.uint8)
        self._loss = z(info.loss_parts, t.float32)
        self._acc = z(kernel._nbrOfWalkers, t.int32)
        # run(): accept counters, their running sums in pinned host memory and the events that mark them,
        # for two blocks in flight
        self._runAcc = z(kernel._nbrOfWalkers, t.int32)
        self._accHost = t.zeros((2,), dtype=t.int64, pin_memory=True)
        self._accEv = [t.cuda.Event() for _ in range(2)]
        self._trace = None
        self._bufsCache = None
        self._batch = int(batch)
        self._step = 0
        self._sweepCount = 0
        self.fallbackAcceptance = float(fallbackAcceptance)
        # run(): the whole-path move records each block's measurement itself when it can (False: a
        # measure-only launch of the kernel after every block)
        self.measureInMove = True
        self.lastAcceptance = None
        self.fallbackBlocks = 0
        self.setParameters(_he_init(self._dof, initSeed, active=self.width))

    # ---- buffers and parameters ----
    def _bufs(self):
        # the buffers are allocated once, in the constructor, except the trace (traceSweeps resets the cache)
        if self._bufsCache is not None:
            return self._bufsCache
        b = _native.FsLpBuffers()
        b.params, b.adam_m, b.adam_v = self._params.data_ptr(), self._m.data_ptr(), self._v.data_ptr()
        b.image, b.train_w, b.train_b = self._image.data_ptr(), self._tw.data_ptr(), self._tb.data_ptr()
        b.workspace, b.loss_parts = self._ws.data_ptr(), self._loss.data_ptr()
        b.trace = self._trace.data_ptr() if self._trace is not None else None
        self._bufsCache = b
        return b

    @property
    def info(self):
        return self._P.info

    def parameterCount(self):
        return int(self._P.info.n_params)

    @_on_engine_stream
    def getParameters(self):
        return self._params.cpu().numpy()

    @_on_engine_stream
    def setParameters(self, theta, resetAdam=True):
        """FP32 parameters in logical order W1 [128][2D+1], b1, W2, b2, W3, b3, W4 [2D][128], b4.
        The parameters of hidden units beyond the active width are set to zero."""
        t = self._k._torch
        theta = np.ascontiguousarray(theta, dtype=np.float32).reshape(-1)
        if theta.size != self.parameterCount():
            raise ValueError("expected %d parameters, got %d" % (self.parameterCount(), theta.size))
        theta = _mask_width(theta, self._dof, self.width)
        self._params.copy_(t.from_numpy(theta))
        if resetAdam:
            self._m.zero_()
            self._v.zero_()
            self._step = 0
        self._P.pack(self._bufs(), self._k._stream.cuda_stream)
        self._k._stream.synchronize()

    # ---- sampling ----
    @_on_engine_stream
    def sweep(self, nSweeps=1, wait=True):
        """nSweeps learned sweeps; returns their acceptance rate (segments, or paths for the
        whole-path move, accepted / proposed)."""
        k = self._k
        self._acc.zero_()
        self._P.sweep(k._paths.data_ptr(), self._acc.data_ptr(), self._bufs(), self._sweepCount, int(nSweeps),
                      k._stream.cuda_stream)
        k._potValid = False  # the moves change the paths without the kernel's potential cache
        self._sweepCount += int(nSweeps)
        if not wait:
            return None
        self.lastAcceptance = float(self._acc.sum().item()) / (self._movesPerSweep * int(nSweeps))
        return self.lastAcceptance

    @_on_engine_stream
    def lastAcceptanceOfQueued(self, nSweeps):
        """Acceptance rate of the last sweep(nSweeps, wait=False) call, after it has finished."""
        k = self._k
        k._stream.synchronize()
        self.lastAcceptance = float(self._acc.sum().item()) / (self._movesPerSweep * int(nSweeps))
        return self.lastAcceptance

    @_on_engine_stream
    def run(self, operatorRuns=None, sweepsPerMeasurement=None):
        """operatorRuns blocks of sweepsPerMeasurement learned sweeps, each followed by one
        measurement of the kernel's operators (series, totals, correlator, histograms).  The host
        does not wait for a block: its acceptance is copied back while the next block runs, and a
        block that follows one whose learned acceptance fell below fallbackAcceptance (for the first
        block: the last block of the previous call) adds plain bisection sweeps of the same length
        before its measurement.  The whole-path move of a wave-sampler kernel that measures operators
        only makes the measurement itself, in its last move of the block, instead of a measure-only
        launch of the kernel (the same quantities; slice sums in another order)."""
        k = self._k
        opRuns = int(operatorRuns if operatorRuns is not None else k._operatorRuns)
        M = int(sweepsPerMeasurement if sweepsPerMeasurement is not None else k._metroStepsPerOperatorRun)
        moves = float(self._movesPerSweep * M)
        # the whole-path move records the measurement itself (fs_lp_sweep_measure) when the kernel's
        # measurement is operators only; otherwise a measure-only launch of the kernel follows each block
        c = k._cfg
        selfMeasure = (self.measureInMove and self.coarse and k._sampler == "wave" and M > 0 and c.n_ops > 0
                       and c.n_corr == 0 and not c.bins and c.n_hist_obs == 0)
        kb = k._buffers() if selfMeasure else None
        low = self.lastAcceptance is not None and self.lastAcceptance < self.fallbackAcceptance
        # accept counts accumulate over the call's blocks; a block's acceptance is the difference of the
        # running totals copied back after it and after the block before it
        acc = self._runAcc
        acc.zero_()
        done = 0  # running total after the last block whose copy has been read
        bufs = self._bufs()
        paths, stream = k._paths.data_ptr(), k._stream.cuda_stream
        for b in range(opRuns):
            i = b & 1
            if b > 0:
                # block b - 1 has been copied back once its event has passed; block b is not waited for
                self._accEv[i ^ 1].synchronize()
                total = int(self._accHost[i ^ 1])
                self.lastAcceptance = (total - done) / moves
                done = total
                low = self.lastAcceptance < self.fallbackAcceptance
            if selfMeasure and not low:
                if k._series is not None and k._seriesCount + 1 > k._series.shape[0]:
                    raise Exception("series buffer full: %d measurements stored, capacity %d"
                                    % (k._seriesCount, k._series.shape[0]))
                self._P.sweep_measure(paths, acc.data_ptr(), bufs, kb, self._sweepCount, M,
                                      k._seriesCount if k._series is not None else 0, stream)
                k._account_run(1, 0)
            else:
                self._P.sweep(paths, acc.data_ptr(), bufs, self._sweepCount, M, stream)
            self._sweepCount += M
            k._potValid = False
            self._accHost[i].copy_(acc.sum(), non_blocking=True)
            self._accEv[i].record(k._stream)
            if low:
                self.fallbackBlocks += 1
                k._launch(1, M)
            elif not selfMeasure:
                k._launch(1, 0)
        k._stream.synchronize()
        if opRuns > 0:
            self.lastAcceptance = (int(self._accHost[(opRuns - 1) & 1]) - done) / moves

    # ---- training ----
    @_on_engine_stream
    def train(self, steps=1):
        """Adam steps on rows drawn from the current paths; returns the mean NLL of the last step."""
        k = self._k
        for _ in range(int(steps)):
            self._step += 1
            self._P.train(k._paths.data_ptr(), self._bufs(), self._step, k._stream.cuda_stream)
        k._stream.synchronize()
        return float(self._loss.double().sum().item()) / self._batch

    @_on_engine_stream
    def trainOnline(self, steps, plainSweepsPerStep=2):
        """Online training: each Adam step follows plainSweepsPerStep plain bisection sweeps of
        the kernel, whose paths supply its rows (0: every step draws its rows from the current
        paths, as train does).  Returns the mean NLL of the last step."""
        k = self._k
        if int(plainSweepsPerStep) <= 0:
            return self.train(steps)
        for _ in range(int(steps)):
            k._launch(1, int(plainSweepsPerStep))
            self._step += 1
            self._P.train(k._paths.data_ptr(), self._bufs(), self._step, k._stream.cuda_stream)
        k._stream.synchronize()
        return float(self._loss.double().sum().item()) / self._batch

    def traceSweeps(self, on=True):
        """Keep (log q ratio, dK, dV, log alpha) of every segment of the last sweep (test support)."""
        k = self._k
        per = 1 if self.coarse else k._N >> k._S
        self._bufsCache = None
        self._trace = (k._torch.zeros((k._nbrOfWalkers, per, 4), dtype=k._torch.float32,
                                      device=k._device) if on else None)

    @_on_engine_stream
    def getTrace(self):
        return self._trace.cpu().numpy()

    @_on_engine_stream
    def lastGradient(self):
        """Gradient of the mean NLL used by the last Adam step, logical order (test support)."""
        t = self._k._torch
        g = t.zeros_like(self._params)
        self._P.grad(self._bufs(), g.data_ptr(), self._k._stream.cuda_stream)
        self._k._stream.synchronize()
        return g.cpu().numpy()

    @_on_engine_stream
    def evalNetwork(self, X):
        """Network outputs (rows, 2 DOF) for features X (rows, 2 DOF + 1) (test support)."""
        t = self._k._torch
        X = np.asarray(X, dtype=np.float32)
        rows = X.shape[0]
        Xp = np.zeros((rows, 32), np.float32)
        Xp[:, : X.shape[1]] = X
        Xd = t.from_numpy(Xp).to(self._k._device)
        Od = t.zeros((rows, 32), dtype=t.float32, device=self._k._device)
        self._P.eval(self._bufs(), Xd.data_ptr(), Od.data_ptr(), rows, self._k._stream.cuda_stream)
        self._k._stream.synchronize()
        return Od.cpu().numpy()[:, : 2 * self._dof]
