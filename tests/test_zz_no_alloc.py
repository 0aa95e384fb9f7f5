"""fs_batch_run_noalloc against fs_batch_run through the C ABI.

The allocation-free variant keeps the dispatch tables and the work counter
in the library's static per-GPU array.  Two calls in a row, the second with
fewer types and walkers than the first, must each reproduce what
fs_batch_run computes on identical copies of the inputs, bit for bit: the
stream states, acceptance counts, paths and operator means.
"""
import ctypes

import numpy as np
import pytest

from feynsimul_amd import _native
from feynsimul_amd.kernel_args import KernelArgs
from feynsimul_amd.batch import BatchedPIMC
from feynsimul_amd.physical_systems import HarmOsc, Lm2m2Cluster, AnharmOsc

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case] " + msg, flush=True)


def _ka(system, W, N, S, beta, seed):
    return KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=beta, S=S, enableOperator=True,
                      enableCorrelator=False, enableBisection=True, enablePathShift=False,
                      enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                      enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=2,
                      operatorRuns=2, nbrOfWalkersPerWorkGroup=1, operators=(system.energyOp,),
                      seed=seed, sampler="wave")


def _paths(system, W, N, rng):
    if isinstance(system, Lm2m2Cluster):
        pos = system.initialPositions(W, rng, sigma=1.0)
        return np.repeat(pos[:, :, None], N, axis=2).reshape(W, system.DOF * N)
    return rng.uniform(-1, 1, (W, system.DOF * N))


def _pair(members, seed):
    """Two batches of identical members and initial paths."""
    rng = np.random.default_rng(seed)
    starts = [_paths(s, W, N, rng) for s, W, N, S, b in members]
    kas = [_ka(s, W, N, S, b, seed + i) for i, (s, W, N, S, b) in enumerate(members)]
    out = []
    for _ in range(2):
        batch = BatchedPIMC(kas)
        for k, p in zip(batch.kernels, starts):
            k.setPaths(p)
        out.append(batch)
    return out


def _run_abi(batch, noalloc):
    """One launch through the C ABI: fs_batch_run or fs_batch_run_noalloc."""
    L = _native.lib()
    k0 = batch.kernels[0]
    n = len(batch.kernels)
    bufs = (_native.FsBuffers * n)(*batch._buffers())
    order = (ctypes.c_int32 * n)(*batch.order)
    fn = L.fs_batch_run_noalloc if noalloc else L.fs_batch_run
    rc = fn(batch._batch._h, bufs, order, int(k0._operatorRuns), int(k0._metroStepsPerOperatorRun), 0, 0,
            batch._stream.cuda_stream)
    assert rc == _native.FS_OK, L.fs_last_error().decode()
    batch._stream.synchronize()


def _same(a, b, what):
    for i, (k1, k2) in enumerate(zip(a.kernels, b.kernels)):
        for name in ("_seeds", "_accepts"):
            np.testing.assert_array_equal(getattr(k1, name).cpu().numpy(), getattr(k2, name).cpu().numpy(),
                                          err_msg="%s member %d %s" % (what, i, name))
        np.testing.assert_array_equal(k1.getPaths(), k2.getPaths(), err_msg="%s member %d paths" % (what, i))
        np.testing.assert_array_equal(k1._operatorValues.cpu().numpy(), k2._operatorValues.cpu().numpy(),
                                      err_msg="%s member %d operators" % (what, i))
        assert (k1._accepts.cpu().numpy() > 0).any(), "%s member %d: nothing ran" % (what, i)


def test_noalloc_matches_batch_run_twice(gpu):
    big = [(HarmOsc(), 48, 256, 4, 10.0), (Lm2m2Cluster(3), 24, 512, 6, 10.0), (AnharmOsc(0.1), 16, 1024, 5, 20.0)]
    small = [(Lm2m2Cluster(4), 8, 256, 5, 10.0), (HarmOsc(), 8, 64, 3, 5.0)]
    for what, members, seed in (("first call", big, 40), ("second call, smaller", small, 80)):
        _case("noalloc " + what)
        ref, got = _pair(members, seed)
        _run_abi(ref, noalloc=False)
        _run_abi(got, noalloc=True)
        _same(ref, got, what)


def test_noalloc_refuses_a_compile_only_batch(gpu):
    """A compile-only batch has no module: the entry point returns an error instead of launching."""
    L = _native.lib()
    k = BatchedPIMC([_ka(HarmOsc(), 8, 64, 3, 5.0, 1)]).kernels[0]
    b = _native.Batch([k._cfg], [k._systemSource], compile_only=True)
    bufs = (_native.FsBuffers * 1)(k._buffers())
    rc = L.fs_batch_run_noalloc(b._h, bufs, None, 1, 1, 0, 0, None)
    assert rc == _native.FS_ERR_INVALID
    assert "compile-only" in L.fs_last_error().decode()
