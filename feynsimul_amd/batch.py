"""Several physical systems in one persistent GPU launch.

The reference runs one compiled program per (system, N, S) and relaunches
them from Python one after another (FeynSimul/kernel.py:346-382,
FeynSimul/pimc_utils.py:230-260).  ``BatchedPIMC`` compiles every member
configuration into one module: a persistent kernel whose wavefronts take
walkers from a device work queue, in a dispatch order that starts with the
most expensive walkers (longest first, so the tail of the launch is short),
and run each with its own type's specialisation of the wave sampler.
Launch loops can be captured once into a HIP graph and replayed.  A graph
keeps its own copy of the dispatch tables (buffer addresses and order) from
the capture; ``replay`` refuses a graph whose buffers or order have changed
since (``setSeriesCapacity``, ``setOrder``, a new RANLUX table ...).

The members are split into register classes (path kept in registers or
streamed, DOF <= 9 or larger) that run in separate persistent kernels,
launched one after the other, so each class runs at its own occupancy
instead of all at the largest type's (``kernels="auto"``; ``"one"`` keeps a
single kernel, or give a class number per member); ``kernelWaves`` asks the
compiler for an occupancy per class ("fit": ``class_waves``).

Members are ordinary ``PIMCKernel`` objects for everything but running:
``batch.kernels[t].getOperators()``, ``setPaths``, ``setSeriesCapacity`` ...
"""
import numpy as np

from . import _native
from .kernel import PIMCKernel


def walker_cost(kernel):
    """Relative cost of one walker's launch: slices times the work per slice-update.

    Work per slice-update is estimated from the coordinates (random draws and
    the bridge, about 26 operations each) and the length of the potential
    expression (a proxy for its evaluation cost).
    """
    system = kernel._system
    return float(kernel._N) * (26.0 * system.DOF + len(system.potential) + len(getattr(system, "userCode", "")) / 8.0)


def register_classes(kernels, mode="auto"):
    """Sweep kernel of each member.  "auto": one class per (path storage, DOF <= 9) pair present --
    members that keep their whole path in registers need far more of them, and the coordinates a
    wavefront holds grow with DOF; "one": a single kernel; a list: the class of each member (numbered
    0, 1, ... without gaps, at most 4)."""
    if isinstance(mode, str):
        if mode == "one":
            return [0] * len(kernels)
        if mode != "auto":
            raise ValueError("kernels must be 'auto', 'one' or a list of class numbers")
        keys = [(_native.storage_of(k._cfg), int(k._system.DOF) > 9) for k in kernels]
        order = sorted(set(keys))
        return [order.index(key) for key in keys]
    classes = [int(c) for c in mode]
    if len(classes) != len(kernels) or sorted(set(classes)) != list(range(max(classes) + 1)) or max(classes) > 3:
        raise ValueError("kernels: one class number per member, 0, 1, ... without gaps, at most 4 classes")
    return classes


def class_waves(kernels, classes):
    """An occupancy target per class for the compiler (wavefronts per SIMD): 5 for streamed paths of
    DOF <= 9, 4 for larger DOF, none for paths kept in registers."""
    out = []
    for c in range(max(classes) + 1):
        members = [k for k, cl in zip(kernels, classes) if cl == c]
        if any(_native.storage_of(k._cfg) == _native.FS_STORAGE_REGISTERS for k in members):
            out.append(0)
        else:
            out.append(5 if max(int(k._system.DOF) for k in members) <= 9 else 4)
    return out


class BatchedPIMC:
    def __init__(self, kernelArgsList, wavesPerBlock=None, order="cost", gridBlocks=0, kernels="auto",
                 kernelWaves=None):
        if not kernelArgsList:
            raise ValueError("at least one KernelArgs is needed")
        self.kernels = [PIMCKernel(ka, _batch_member=True) for ka in kernelArgsList]
        k0 = self.kernels[0]
        for t, k in enumerate(self.kernels):
            if k._sampler != "wave":
                raise Exception("member %d: batches need the wave sampler (enableParallelizePath and 2^S <= 64)" % t)
            if (k._operatorRuns, k._metroStepsPerOperatorRun) != (k0._operatorRuns, k0._metroStepsPerOperatorRun):
                raise Exception("member %d: every member must have the same operatorRuns and "
                                "metroStepsPerOperatorRun" % t)
        self.kernelOfType = register_classes(self.kernels, kernels)
        if kernelWaves == "fit":
            kernelWaves = class_waves(self.kernels, self.kernelOfType)
        self.kernelWaves = kernelWaves
        self._batch = _native.Batch([k._cfg for k in self.kernels], [k._systemSource for k in self.kernels],
                                    wavesPerBlock or 0, kernel_of_type=self.kernelOfType, waves_of_kernel=kernelWaves)
        for t, k in enumerate(self.kernels):
            k._attach(self._batch.path_storage(t))
        self._torch = k0._torch
        self._stream = k0._stream
        self.gridBlocks = int(gridBlocks)
        self.setOrder(order)
        self._ev0 = self._torch.cuda.Event(enable_timing=True)
        self._ev1 = self._torch.cuda.Event(enable_timing=True)
        self._measurements = 0

    def setOrder(self, order):
        """Dispatch order: "cost" (most expensive walkers first), "given", or a permutation of type ids."""
        n = len(self.kernels)
        if isinstance(order, str):
            if order == "cost":
                order = sorted(range(n), key=lambda t: -walker_cost(self.kernels[t]))
            elif order == "given":
                order = list(range(n))
            else:
                raise ValueError("order must be 'cost', 'given' or a permutation")
        order = [int(t) for t in order]
        if sorted(order) != list(range(n)):
            raise ValueError("order must be a permutation of the member indices")
        self.order = order

    def _buffers(self):
        return [k._buffers() for k in self.kernels]

    def _dispatchKey(self):
        """What a captured graph depends on: every member's buffer addresses, series capacity and the order."""
        key = []
        for k, b in zip(self.kernels, self._buffers()):
            cap = k._series.shape[0] if k._series is not None else -1
            key.append(tuple(getattr(b, f) for f, _ in b._fields_) + (cap,))
        return (tuple(key), tuple(self.order))

    def _check_series(self, launches):
        opRuns = int(self.kernels[0]._operatorRuns)
        for t, k in enumerate(self.kernels):
            if k._series is not None and self._measurements + launches * opRuns > k._series.shape[0]:
                raise Exception("member %d: series buffer full" % t)

    def run(self, wait=True, noAlloc=False):
        """One launch: operatorRuns x metroStepsPerOperatorRun steps for every walker of every member.

        With ``noAlloc`` the launch keeps its dispatch tables in the library's
        static per-GPU array instead of the batch's own device memory
        (fs_batch_run_noalloc); launches of that kind must not overlap on one GPU.
        """
        k0 = self.kernels[0]
        opRuns, M = int(k0._operatorRuns), int(k0._metroStepsPerOperatorRun)
        self._check_series(1)
        self._ev0.record(self._stream)
        launch = self._batch.run_noalloc if noAlloc else self._batch.run
        launch(self._buffers(), self.order, opRuns, M, self._measurements, self.gridBlocks, self._stream.cuda_stream)
        self._ev1.record(self._stream)
        self._after(1)
        if wait:
            self._stream.synchronize()

    def _after(self, launches):
        k0 = self.kernels[0]
        opRuns, M = int(k0._operatorRuns), int(k0._metroStepsPerOperatorRun)
        for _ in range(launches):
            for k in self.kernels:
                k._account_run(opRuns, M)
        if any(k._series is not None for k in self.kernels):
            self._measurements += launches * opRuns

    def captureGraph(self, launches, stream=None):
        """Records `launches` consecutive launches into a HIP graph (replay with ``replay``).

        Series slots are fixed at capture: measurement m of replay i lands where a
        plain ``run`` sequence from the current count would put it; read them
        before replaying again.  Capture needs a non-default stream.
        """
        if stream is None:
            stream = self._torch.cuda.Stream()
        self._graphStream = stream
        k0 = self.kernels[0]
        opRuns, M = int(k0._operatorRuns), int(k0._metroStepsPerOperatorRun)
        self._check_series(launches)
        self._torch.cuda.synchronize()
        g = self._batch.capture(self._buffers(), self.order, opRuns, M, self._measurements, int(launches),
                                self.gridBlocks, stream.cuda_stream)
        g.launches = int(launches)
        g.stream = stream
        g.seriesStart = self._measurements
        g.dispatchKey = self._dispatchKey()
        return g

    def replay(self, graph, wait=True):
        """Runs a captured graph on its capture stream (after the current stream's work)."""
        if graph.dispatchKey != self._dispatchKey():
            raise Exception("the batch's buffers or dispatch order changed since this graph was captured; "
                            "capture it again")
        graph.stream.wait_stream(self._stream)
        self._ev0.record(graph.stream)
        graph.launch(graph.stream.cuda_stream)
        self._ev1.record(graph.stream)
        self._stream.wait_stream(graph.stream)
        if any(k._series is not None for k in self.kernels):
            self._measurements = graph.seriesStart
            for k in self.kernels:
                k._seriesCount = graph.seriesStart
        self._after(graph.launches)
        if wait:
            graph.stream.synchronize()

    def resetSeries(self):
        for k in self.kernels:
            k._seriesCount = 0
        self._measurements = 0

    def getRunTime(self):
        """Device time of the last run or replay in seconds."""
        self._ev1.synchronize()
        return 1e-3 * self._ev0.elapsed_time(self._ev1)

    def sliceUpdatesPerLaunch(self):
        k0 = self.kernels[0]
        steps = int(k0._operatorRuns) * int(k0._metroStepsPerOperatorRun)
        tot = 0
        for k in self.kernels:
            seg = 2 ** k._S
            tot += k._nbrOfWalkers * (k._N // seg) * (seg - 1) * steps
        return tot
