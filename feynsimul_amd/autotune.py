"""Launch-shape autotuner for batched runs.

Measures, on the device, every candidate (wavefronts per workgroup) x
(dispatch order) for a list of KernelArgs and keeps the fastest.  Waves per
workgroup fixes the LDS layout and the launch bounds, so each candidate value
is a separate compile; the dispatch order is a run-time choice.  Results are
cached per configuration and device in a JSON file
($FEYNSIMUL_AMD_CACHE, default ~/.cache/feynsimul_amd/autotune.json).
"""
import hashlib
import json
import os
import time

from .batch import BatchedPIMC


def _cache_path():
    d = os.environ.get("FEYNSIMUL_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "feynsimul_amd"))
    return os.path.join(d, "autotune.json")


def config_key(kernelArgsList, device_name):
    h = hashlib.sha1(device_name.encode())
    for ka in kernelArgsList:
        sysobj = ka.system
        h.update(repr((sysobj.DOF, sysobj.potential, getattr(sysobj, "userCode", ""), ka.nbrOfWalkers, ka.N, ka.S,
                       ka.beta, ka.operators, ka.correlators, ka.enableDouble, ka.enableRanlux,
                       ka.operatorRuns, ka.metroStepsPerOperatorRun, getattr(ka, "pathStorage", None),
                       getattr(ka, "histogramObservables", None))).encode())
    return h.hexdigest()[:20]


def autotune_batch(kernelArgsList, wavesPerBlock=(1, 2, 4), orders=("cost", "given"), launches=3,
                   initialPaths=None, use_cache=True, verbose=False):
    """Returns (best, trials): best = {"wavesPerBlock", "order", "ms_per_launch"}; every trial is timed
    over `launches` launches after one warm-up launch, starting from `initialPaths` (a list with one
    array per member, or None for the members' default paths)."""
    import torch
    dev = torch.cuda.get_device_properties(torch.cuda.current_device()).name
    key = config_key(kernelArgsList, dev)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return cache[key]["best"], cache[key]["trials"]
    trials = []
    for wpb in wavesPerBlock:
        b = BatchedPIMC(kernelArgsList, wavesPerBlock=wpb)
        for order in orders:
            if initialPaths is not None:
                for k, p in zip(b.kernels, initialPaths):
                    k.setPaths(p)
            b.setOrder(order)
            b.run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(launches):
                b.run(wait=False)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / launches * 1e3
            trials.append({"wavesPerBlock": wpb, "order": order, "ms_per_launch": ms})
            if verbose:
                print("autotune: waves/workgroup %d, order %s: %.3f ms" % (wpb, order, ms), flush=True)
        del b
    best = min(trials, key=lambda t: t["ms_per_launch"])
    if use_cache:
        cache[key] = {"best": best, "trials": trials}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path + ".tmp", "w") as f:
                json.dump(cache, f, indent=1)
            os.replace(path + ".tmp", path)
        except OSError:
            pass
    return best, trials
