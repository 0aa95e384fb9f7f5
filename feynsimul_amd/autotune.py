"""Launch-shape autotuner.

autotune_batch: every candidate (wavefronts per workgroup) x (dispatch order)
x (register classes) for a list of KernelArgs in one persistent launch.
autotune_levels: the bisection level S of every member of a ragged batch, from
acceptance rates measured on the device (the reference's run loop adapts S to
keep the acceptance between 20 and 50 %, FeynSimul/pimc_utils.py:270-279).
autotune_kernel: lanes per chain for one kernel (a wavefront per chain, one
work-item per segment, or one per chain) and, for the wave sampler, every
(wavefronts per workgroup) x (wavefronts per SIMD cap).  A batch's members are
wave-sampler specialisations of one module, so their lanes per chain is 64.
Waves per workgroup fixes the LDS layout and the launch bounds, so each
candidate value is a separate compile; the dispatch order and the occupancy
cap are run-time choices.  The fastest is kept; results are cached per
configuration and device in a JSON file ($FEYNSIMUL_AMD_CACHE, default
~/.cache/feynsimul_amd/autotune.json).
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
    h = hashlib.sha1(("v2:" + device_name).encode())  # v2: trials include the register-class split
    for ka in kernelArgsList:
        sysobj = ka.system
        h.update(repr((sysobj.DOF, sysobj.potential, getattr(sysobj, "userCode", ""), ka.nbrOfWalkers, ka.N, ka.S,
                       ka.beta, ka.operators, ka.correlators, ka.enableDouble, ka.enableRanlux,
                       ka.operatorRuns, ka.metroStepsPerOperatorRun, getattr(ka, "pathStorage", None),
                       getattr(ka, "histogramObservables", None))).encode())
    return h.hexdigest()[:20]


def _cache_load(key, use_cache):
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    return cache, cache.get(key) if use_cache else None


def _cache_store(cache, key, best, trials):
    path = _cache_path()
    cache[key] = {"best": best, "trials": trials}
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path + ".tmp", "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(path + ".tmp", path)
    except OSError:
        pass


def autotune_kernel(kernelArgs, initialPaths=None, wavesPerBlock=(2, 4, 8), wavesPerSimd=(0, 8, 6, 5, 4, 3),
                    launches=3, use_cache=True, verbose=False, lanesPerChain=(64,)):
    """Returns (best, trials) for one KernelArgs: best = {"lanesPerChain", "wavesPerBlock", "wavesPerSimd",
    "ms_per_launch"}.  lanesPerChain candidates: 64, the wave sampler (a wavefront per chain, a slice per
    lane), timed for every (waves per workgroup, occupancy cap) pair; "segments", the reference-order
    sampler with one work-item per segment (N / 2^S lanes per chain, the reference's ParallelizePath);
    1, the reference-order sampler with one work-item per chain.  Each trial is one warm-up launch and
    `launches` timed launches from `initialPaths` (or the kernel's default paths); neither the occupancy
    cap nor the choice of sampler changes the sampled distribution."""
    import copy
    import torch
    from .kernel import PIMCKernel
    dev = torch.cuda.get_device_properties(torch.cuda.current_device()).name
    key = "kernel:" + config_key([kernelArgs], dev) + ":" + ",".join(str(x) for x in lanesPerChain)
    cache, hit = _cache_load(key, use_cache)
    if hit is not None:
        return hit["best"], hit["trials"]
    trials = []

    def timed(k):
        if initialPaths is not None:
            k.setPaths(initialPaths)
        k.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(launches):
            k.run(wait=False)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / launches * 1e3

    for lanes in lanesPerChain:
        if lanes != 64:
            ka = copy.copy(kernelArgs)
            ka.sampler = "exact"
            ka.enableParallelizePath = lanes == "segments"
            k = PIMCKernel(ka)
            ms = timed(k)
            trials.append({"lanesPerChain": lanes if lanes == 1 else int(ka.N) >> int(ka.S), "sampler": "exact",
                           "wavesPerBlock": None, "wavesPerSimd": None, "ms_per_launch": ms})
            if verbose:
                print("autotune: reference-order sampler, %s lanes per chain: %.3f ms" % (lanes, ms), flush=True)
            del k
            continue
        for wpb in wavesPerBlock:
            ka = copy.copy(kernelArgs)
            ka.wavesPerBlock, ka.wavesPerSimd, ka.sampler = wpb, None, "wave"
            k = PIMCKernel(ka)
            for cap in wavesPerSimd:
                k.setWavesPerSimd(cap)
                ms = timed(k)
                trials.append({"lanesPerChain": 64, "sampler": "wave", "wavesPerBlock": wpb, "wavesPerSimd": cap,
                               "ms_per_launch": ms})
                if verbose:
                    print("autotune: waves/workgroup %d, waves/SIMD cap %d: %.3f ms" % (wpb, cap, ms), flush=True)
            del k
    best = min(trials, key=lambda t: t["ms_per_launch"])
    if use_cache:
        _cache_store(cache, key, best, trials)
    return best, trials


def autotune_batch(kernelArgsList, wavesPerBlock=(1, 2, 4), orders=("cost", "given"), launches=3,
                   initialPaths=None, use_cache=True, verbose=False,
                   kernels=(("auto", None), ("auto", "fit"), ("one", None))):
    """Returns (best, trials): best = {"wavesPerBlock", "order", "kernels", "kernelWaves", "ms_per_launch"}
    (kernels, kernelWaves: the register-class split of BatchedPIMC and its occupancy targets); every
    trial is timed
    over `launches` launches after one warm-up launch, starting from `initialPaths` (a list with one
    array per member, or None for the members' default paths)."""
    import torch
    dev = torch.cuda.get_device_properties(torch.cuda.current_device()).name
    key = config_key(kernelArgsList, dev)
    cache, hit = _cache_load(key, use_cache)
    if hit is not None:
        return hit["best"], hit["trials"]
    trials = []
    for wpb, (kern, kwaves) in [(w, kk) for kk in kernels for w in wavesPerBlock]:
        b = BatchedPIMC(kernelArgsList, wavesPerBlock=wpb, kernels=kern, kernelWaves=kwaves)
        if kern != "one" and kwaves is None and max(b.kernelOfType) == 0:
            del b
            continue  # the split gives one class: the same module as "one"
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
            trials.append({"wavesPerBlock": wpb, "order": order, "kernels": kern, "kernelWaves": kwaves,
                           "ms_per_launch": ms})
            if verbose:
                print("autotune: waves/workgroup %d, order %s, kernels %s/%s: %.3f ms" % (wpb, order, kern, kwaves,
                                                                                        ms), flush=True)
        del b
    best = min(trials, key=lambda t: t["ms_per_launch"])
    if use_cache:
        _cache_store(cache, key, best, trials)
    return best, trials


def autotune_levels(kernelArgsList, initialPaths=None, levels=(1, 2, 3, 4, 5, 6), warmupSweeps=64, sweeps=32,
                    minAcceptance=0.2, verbose=False):
    """Bisection level per member: the largest S whose measured acceptance is at least `minAcceptance`
    (the lower end of the reference's 20-50 % window), else the S with the highest acceptance.

    Every candidate S is one batch of all members (levels a member cannot take, 2^S > min(N, 64), are
    skipped for it); each starts from the same paths, which `warmupSweeps` sweeps at the members' own S
    first bring away from `initialPaths`.  Returns (levels, table): levels[t] is the chosen S of member t,
    table[t] maps S to the acceptance measured with it.
    """
    import copy
    n = len(kernelArgsList)

    def with_steps(ka, S, steps):
        k = copy.copy(ka)
        k.S, k.metroStepsPerOperatorRun, k.operatorRuns = S, steps, 1
        return k

    warm = BatchedPIMC([with_steps(ka, ka.S, warmupSweeps) for ka in kernelArgsList])
    if initialPaths is not None:
        for k, p in zip(warm.kernels, initialPaths):
            k.setPaths(p)
    warm.run()
    start = [k.getPaths() for k in warm.kernels]
    del warm
    table = [dict() for _ in range(n)]
    for S in levels:
        members = [t for t, ka in enumerate(kernelArgsList) if (1 << S) <= min(int(ka.N), 64)]
        if not members:
            continue
        b = BatchedPIMC([with_steps(kernelArgsList[t], S, sweeps) for t in members])
        for k, t in zip(b.kernels, members):
            k.setPaths(start[t])
        b.run()
        for k, t in zip(b.kernels, members):
            table[t][S] = k.getAcceptanceRate()
        if verbose:
            print("autotune levels: S=%d acceptance %s" % (S, [round(table[t][S], 3) for t in members]), flush=True)
        del b
    chosen = []
    for t in range(n):
        ok = [S for S, a in table[t].items() if a >= minAcceptance]
        chosen.append(max(ok) if ok else max(table[t], key=table[t].get))
    return chosen, table
