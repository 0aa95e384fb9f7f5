"""Run configuration of a PIMC kernel (FeynSimul/kernel_args.py:21-217).

``KernelArgs`` is the same mutable attribute bag as the reference, with the
same constructor signature; callers mutate it and build a new ``PIMCKernel``
(doc/tutorial/index.rst:84-98).  Engine-specific settings are extra keyword
arguments at the end, all optional:

``seed``
    seed of every random stream of the kernel (splitmix64-derived rows).  The
    reference draws its seeds from unseeded ``np.random``; ``None`` keeps
    that behaviour (a fresh random seed per kernel).
``sampler``
    ``"wave"`` (one wavefront per walker, slice per lane), ``"exact"`` (the
    reference's work-item order and draw order) or ``None`` for automatic:
    wave when only bisection is enabled and 2^S <= 64, exact otherwise.
``correlatorLags``
    number of correlator lags (default N/2, the reference's range).
``fastMath``
    relaxed floating point like the reference's ``-cl-fast-relaxed-math``
    (default True).
``wavesPerBlock``
    launch shape of the wave sampler (default automatic).
``potentialCache``
    wave sampler with paths in memory: keep the potential of every slice in
    a cache buffer (True), evaluate it again for the old slice of each move
    (False), or decide by the system (None: cache when DOF > 1).
``wavesPerSimd``
    occupancy cap of the wave sampler: at most this many wavefronts per SIMD
    (default None: as registers and LDS allow; ``autotune_kernel`` picks it).
``pathStorage``
    where the wave sampler keeps each walker's path during a launch:
    ``"registers"``, ``"memory"`` or ``None`` (automatic: registers when the
    path fits).
``histogramObservables``, ``histogramRange``, ``histogramBins``
    wave sampler: a tuple of scalar expressions in ``x1..xDOF`` (pair
    distances, say) whose values at every slice are binned, at every
    measurement, into one histogram of ``histogramBins`` bins over
    ``histogramRange = (lo, hi)`` (``getObservableHistogram``).
"""


class KernelArgs:

    def __init__(self,
                 system,
                 nbrOfWalkers,
                 N,
                 beta,
                 enableOperator,
                 enableCorrelator,
                 enableBisection,
                 enablePathShift,
                 enableSingleNodeMove,
                 enableBins,
                 nbrOfWalkersPerWorkGroup,
                 enableDouble=None,
                 enableRanlux=None,
                 luxuaryFactor=None,
                 ranluxIntMax=None,
                 S=None,
                 alpha=None,
                 PSAlpha=None,
                 enableParallelizePath=None,
                 metroStepsPerOperatorRun=None,
                 enableGlobalPath=None,
                 enableGlobalOldPath=None,
                 operators=None,
                 operatorRuns=None,
                 correlators=None,
                 xMin=None,
                 xMax=None,
                 binResolutionPerDOF=None,
                 seed=None,
                 sampler=None,
                 correlatorLags=None,
                 fastMath=True,
                 wavesPerBlock=None,
                 pathStorage=None,
                 histogramObservables=None,
                 histogramRange=None,
                 histogramBins=256,
                 wavesPerSimd=None,
                 potentialCache=None):
        self.system = system
        self.nbrOfWalkers = nbrOfWalkers
        self.N = N
        self.S = S
        self.beta = beta
        self.nbrOfWalkersPerWorkGroup = nbrOfWalkersPerWorkGroup
        self.enablePathShift = enablePathShift
        self.enableDouble = enableDouble
        self.enableSingleNodeMove = enableSingleNodeMove
        self.alpha = alpha
        self.enableOperator = enableOperator
        self.enableRanlux = enableRanlux
        self.luxuaryFactor = luxuaryFactor
        self.ranluxIntMax = ranluxIntMax
        self.operatorRuns = operatorRuns
        self.operators = operators
        self.metroStepsPerOperatorRun = metroStepsPerOperatorRun
        self.enableCorrelator = enableCorrelator
        self.correlators = correlators
        self.enableBins = enableBins
        self.xMin = xMin
        self.xMax = xMax
        self.binResolutionPerDOF = binResolutionPerDOF
        self.enableBisection = enableBisection
        self.enableParallelizePath = enableParallelizePath
        self.enableGlobalPath = enableGlobalPath
        self.enableGlobalOldPath = enableGlobalOldPath
        self.PSAlpha = PSAlpha
        self.seed = seed
        self.sampler = sampler
        self.correlatorLags = correlatorLags
        self.fastMath = fastMath
        self.wavesPerBlock = wavesPerBlock
        self.pathStorage = pathStorage
        self.histogramObservables = histogramObservables
        self.histogramRange = histogramRange
        self.histogramBins = histogramBins
        self.wavesPerSimd = wavesPerSimd
        self.potentialCache = potentialCache
