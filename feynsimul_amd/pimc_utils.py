"""Run loop around PIMCKernel: N doubling, S adaptation, result files, resume.

Restates FeynSimul/pimc_utils.py:28-355 (``modN`` and ``_output``) for this
engine, with the same arguments, directory layout and file formats:

  results/<experimentName>/<YYYY-MM-DD>/<HH.MM.SS>/
      data.tsv              one row per run: N, time, path changes, AR, S, then
                            every (operator, walker) mean (``_output``)
      pathsN<N>episode<k>   the paths, one walker per tab-separated row
      timestamps            one line per finished N
      systemInfo            empty, as in the reference

Starting from straight paths at ``startXList`` (or a saved path file), each N
runs ``runsPerN(N, S)`` launches, then the paths are doubled by copying every
node to the even slots and linearly interpolating the odd slots (periodic), S is
moved towards 20-50 % acceptance, and N doubles, until ``finalN``.

Differences from the reference: Python 3 number formatting; the S update is
capped at log2(N) (the reference reads a leftover loop variable there,
pimc_utils.py:275-277, which raises NameError when resuming); ``resultsDir``
(default "results") chooses the output root.
"""
import copy
import csv
import os
from datetime import datetime
from time import time

import numpy as np

# the reference's pimc_utils star-imports kernel_args and kernel (pimc_utils.py:25-26), and its
# scripts take KernelArgs from ``from FeynSimul.pimc_utils import *``
from .kernel_args import KernelArgs  # noqa: F401
from .kernel import PIMCKernel, humanReadableSize  # noqa: F401


def doublePaths(oldPaths, N, DOF):
    """Paths with 2N slices from paths with N slices (pimc_utils.py:211-226): old node m goes
    to slot 2m, slot 2m + 1 is the mean of slots 2m and 2m + 2 (periodic)."""
    old = np.asarray(oldPaths, dtype=np.float64).reshape(-1, DOF, N)
    new = np.empty((old.shape[0], DOF, 2 * N))
    new[:, :, 0::2] = old
    new[:, :, 1::2] = 0.5 * (old + np.roll(old, -1, axis=2))
    return new.reshape(old.shape[0], DOF * 2 * N)


def _straightPaths(startXList, W, N, DOF):
    start = np.asarray(startXList, dtype=np.float64).reshape(W, -1)
    return np.repeat(start[:, :DOF, None], N, axis=2).reshape(W, DOF * N)


def _readPaths(cont_path, W, width):
    paths = np.zeros((W, width))
    with open(cont_path) as f:
        for k, row in enumerate(csv.reader(f, delimiter="\t")):
            paths[k] = [float(v) for v in row]
    return paths


def _writePaths(fname, paths):
    with open(fname, "w", newline="") as f:
        w = csv.writer(f, delimiter="\t")
        for aPath in paths:
            w.writerow([repr(float(v)) for v in aPath])


def modN(ka, startXList, savePathsInterval, experimentName, opRunsFormula, mStepsPerOPRun, runsPerN, maxWGSize,
         continueRun=False, finalN=-1, simTime=-1, verbosity=0, cont_S=None, cont_N=None, cont_path=None,
         resultsDir="results"):
    """N-doubling run loop (pimc_utils.py:28-285); see the module docstring."""
    kaMod = copy.copy(ka)
    if kaMod.enableBisection is False:
        raise Exception('This function currently only works with bisection sampling')
    if continueRun:
        if cont_path is None:
            raise NameError('cont_path need to be defined if continueRun is enabled')
        if cont_S is None:
            raise NameError('cont_S need to be defined if continueRun is enabled')
        if cont_N is None:
            raise NameError('cont_N need to be defined if continueRun is enabled')
        kaMod.N = cont_N
        kaMod.S = cont_S
    else:
        kaMod.N = ka.N
        kaMod.S = 1
    startClock = time()
    pathChanges = 0
    timestamp = datetime.now().strftime("%Y-%m-%d/%H.%M.%S")
    runDir = os.path.join(resultsDir, experimentName, timestamp)
    filename = os.path.join(runDir, "data")
    os.makedirs(runDir, exist_ok=True)
    with open(os.path.join(runDir, "systemInfo"), "a"):
        pass

    DOF = int(ka.system.DOF)
    W = ka.nbrOfWalkers
    firstNRun = True
    totRuns = 0
    oldPaths = None
    while finalN == -1 or kaMod.N <= finalN:
        overheadClock = time()
        # keep a work-group of the reference geometry within maxWGSize (pimc_utils.py:167-175)
        minS = int(np.ceil(np.log2(ka.nbrOfWalkersPerWorkGroup * kaMod.N / float(maxWGSize))))
        maxS = int(np.log2(kaMod.N))
        if maxS < minS:
            raise Exception('Too many walkers per workgroup')
        kaMod.S = min(maxS, max(minS, kaMod.S))
        kaMod.operatorRuns = opRunsFormula(kaMod.N, kaMod.S)
        kaMod.metroStepsPerOperatorRun = mStepsPerOPRun(kaMod.N, kaMod.S)

        kernel = PIMCKernel(kaMod, verbose=verbosity >= 2)
        if verbosity > 0:
            print("New kernel compiled, getStats():")
            print(kernel.getStats())
            print("Run results:")
        if ka.enableRanlux:
            kernel.initRanlux()
            kernel._queue.finish()
            kernel.groupInitRanlux()
            kernel._queue.finish()

        if firstNRun:
            if not continueRun:
                kernel.setPaths(_straightPaths(startXList, W, kaMod.N, DOF))
            else:
                kernel.setPaths(_readPaths(cont_path, W, kaMod.N * DOF))
        else:
            kernel.setPaths(doublePaths(oldPaths, kaMod.N // 2, DOF))

        nRuns = 1
        runsThisN = runsPerN(kaMod.N, kaMod.S)
        lastCall = time()
        loopClock = time()
        while nRuns <= runsThisN or kaMod.N == finalN:
            if simTime != -1 and simTime != 0 and time() - startClock > simTime:
                if verbosity > 0:
                    print("Time limit reached!")
                return
            kernel.run()
            if totRuns > 100 and totRuns % 100 == 0:
                print('Letting current queue finish before continuing. This is to prevent memory from choking '
                      'by pile-up.')
                kernel._queue.finish()
            pathChanges += kernel.getMetroStepsPerRun()
            _output(filename, kaMod.N, time() - startClock, pathChanges, kernel.getAcceptanceRate(),
                    kernel.getOperators(), (time() - lastCall), kernel.getRunTime(), ka.beta, kaMod.S,
                    verbose=verbosity > 0)
            lastCall = time()
            if nRuns % savePathsInterval == 0 or nRuns == runsThisN:
                _writePaths(os.path.join(runDir, "pathsN%d" % kaMod.N + "episode%d" % int(nRuns / savePathsInterval)),
                            kernel.getPaths())
                if verbosity > 0:
                    print("Paths saved!")
            if simTime == 0 and nRuns == runsThisN and kaMod.N == finalN:
                if verbosity > 0:
                    print("Runs completed!")
                return
            nRuns += 1
            totRuns += 1
        loopClockEnd = time()
        with open(os.path.join(runDir, "timestamps"), "a") as f2:
            el = time() - startClock
            f2.write('N: ' + str(kaMod.N) + '\tS: ' + str(kaMod.S) + '\ttotRuns: ' + str(totRuns) + '\ttime: '
                     + str(el) + '\tnormtime: ' + str(float(el) / (float(kernel._nbrOfThreads) * float(totRuns)))
                     + '\n')
        if kaMod.N != finalN:
            oldPaths = kernel.getPaths()
            firstNRun = False
            # move the acceptance rate towards 20-50 % (pimc_utils.py:271-278)
            ar = kernel.getAcceptanceRate()
            cap = int(np.log2(kaMod.N))
            if ar > 0.5:
                kaMod.S = min(cap, kaMod.S + 2)
            if 0.2 < ar < 0.5:
                kaMod.S = min(cap, kaMod.S + 1)
            if ar < 0.1:
                kaMod.S = max(1, kaMod.S - 1)
            kaMod.N *= 2
        print('Overhead time: ' + str(time() - overheadClock + loopClock - loopClockEnd))


def _output(filename, N, t, pathChanges, acceptanceRate, operatorMean, hostRunTime, kernelRunTime, beta, S,
            verbose=False):
    """Appends one run to <filename>.tsv (pimc_utils.py:288-355); the first write adds the
    header lines '#Beta: beta' and '#N Time pathChanges AR S Operator i, Walker j ...'."""
    if verbose:
        print("N: " + str(N) + "\tS: " + str(S) + "\tbeta: " + str(beta) + "\tAR: " + str(acceptanceRate)
              + "\tOPs: " + str(np.mean(operatorMean, axis=1)) + "\tTime: [" + str(hostRunTime) + ","
              + str(kernelRunTime) + "]")
    path = filename + ".tsv"
    with open(path, "a") as f_data:
        if os.path.getsize(path) == 0:
            f_data.write("#Beta: " + str(beta) + "\n")
            f_data.write("#N\tTime\tpathChanges\tAR\tS")
            for i in range(len(operatorMean)):
                for j in range(len(operatorMean[0])):
                    f_data.write("\tOperator " + str(i) + ", Walker " + str(j))
            f_data.write("\n")
        f_data.write(str(N) + "\t")
        f_data.write(str(t) + "\t")
        f_data.write(str(pathChanges) + "\t")
        f_data.write(str(acceptanceRate) + "\t")
        f_data.write(str(S) + "\t")
        for walkerOperators in operatorMean:
            for op in walkerOperators:
                f_data.write(str(op) + "\t")
        f_data.write("\n")


def readModN(filename):
    """Parses a data.tsv written by _output: (beta, rows) with rows a float array of
    [N, time, pathChanges, AR, S, op0 walker0, op0 walker1, ...]."""
    beta = None
    rows = []
    with open(filename) as f:
        for line in f:
            if line.startswith("#Beta:"):
                beta = float(line.split(":", 1)[1])
            elif line.startswith("#") or not line.strip():
                continue
            else:
                rows.append([float(v) for v in line.rstrip("\n").rstrip("\t").split("\t")])
    return beta, np.array(rows)
