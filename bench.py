"""Throughput of the plain bisection sampler on one MI355X.

Workload (one GPU): 3 He atoms in 3D with the LM2M2 pair potential, mass
4.0026 u, N = 1024 slices, beta = 10 1/K, 16384 chains, bisection level 6,
FP32 coordinates, xorshift128 streams, operators virial energy and mean pair
distance, and a 256-bin pair-distance histogram.  One step = one launch of
32 sweeps over every chain followed by one measurement.  value = slice-updates/s (one slice-update = one interior
slice of a segment, all 9 coordinates, proposed by one bisection move).

After the timed steps: the ESS/s of E0 of the plain sampler and, from its
equilibrated paths, of the learned proposal (online training included), on
equal wall-clock budgets.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--storage registers|memory]
                       [--no-ess] [--no-learned] [--learned-train 200] [--learned-measurements 64]
       python bench.py --copy-calibration   (1 GiB device copy, to scale rocprofv3 HBM counters)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, S, W, BETA, DOF = 1024, 6, 16384, 10.0, 9
HIST_MAX_A = 40.0  # pair-distance histogram: 256 bins on [0, 40) angstrom
SWEEPS_PER_STEP = 32
TAU_MEASUREMENTS = 2048  # E0 after every sweep, for the plain sampler's tau_int in sweeps
# standard-error targets of the fixed-sigma comparison: small enough that sampling, not the learned
# proposal's training, takes most of the time (the plain sampler needs about 3 s and 10 s)
HO_SIGMA, AHO_SIGMA = 2e-5, 2e-5
SEG = 1 << S
UPDATES_PER_SWEEP = W * (N // SEG) * (SEG - 1)
# algorithmic HBM bytes per slice-update: a segment reads 2^S + 1 slices and
# writes 2^S - 1 slices of DOF FP32 coordinates, i.e. 8 DOF 2^S / (2^S - 1)
BYTES_PER_UPDATE = 8.0 * DOF * SEG / (SEG - 1)
# VALU operations per slice-update as written (transcendentals and divides
# counted once): 26 per coordinate + 2 potential evaluations of 3 LM2M2 pairs
VALU_PER_UPDATE = 26 * DOF + 2 * (39 * 3 + 1)
HBM_PEAK_GBS = 8000.0
# measured HBM bytes per launch of this workload (rocprofv3 FETCH_SIZE/WRITE_SIZE
# scaled by a 1 GiB copy), kept with the profile summary it comes from
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "fs_wave_sweep_lm2m2_he3_n1024.md")


def measured_traffic(profile=TRAFFIC_PROFILE):
    import re
    try:
        with open(profile) as f:
            m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", f.read(), re.M)
        return int(m.group(1)) if m else None
    except OSError:
        return None


VALU_NOMINAL = 256 * 64 * 2.4e9  # lane-ops/s: one wave64 VALU instruction per 4 cycles per SIMD at 2.4 GHz
VALU_CALIBRATION = os.path.join(ROOT, "profiles", "valu_issue_calibration_xorshift_1m_streams.md")


def _profile_value(profile, key):
    import re
    try:
        with open(profile) as f:
            m = re.findall(r"^%s:\s*([0-9.eE+-]+)" % re.escape(key), f.read(), re.M)
        return float(m[-1]) if m else None
    except OSError:
        return None


def valu_fields(achieved_ops, profile):
    """VALU side of the roofline.  achieved = as-written operations per second of this run; the issue
    fraction comes from the kernel's profile summary (its SQ_INSTS_VALU per SIMD per cycle, the last
    measurement the summary records) divided by the highest issue rate measured on the box
    (fs_k_xs_bench at 1,048,576 streams, profiles/valu_issue_calibration_xorshift_1m_streams.md)."""
    issue = _profile_value(profile, "valu_issue_per_simd_cycle")
    peak_issue = _profile_value(VALU_CALIBRATION, "valu_issue_per_simd_cycle_max")
    out = {"ops_per_update": None, "achieved": achieved_ops, "nominal_peak": VALU_NOMINAL,
           "nominal_frac": achieved_ops / VALU_NOMINAL,
           "issue_per_simd_cycle_from_profile": issue, "issue_peak_per_simd_cycle_measured": peak_issue,
           "issue_frac": issue / peak_issue if issue and peak_issue else None,
           "source": "SQ counters in %s and %s (profile summaries, not this run)"
                     % (os.path.relpath(profile, ROOT), os.path.relpath(VALU_CALIBRATION, ROOT))}
    return out


MFMA_BF16_PEAK = 2.5e15       # dense BF16 MFMA FLOP/s
# learned proposal: one network evaluation = 2 (19*128 + 128*128*2 + 128*18) = 75,008 FLOP; a
# segment evaluates 2^S - 1 proposals and 2^S - 2 reverse rows (level 1 shares its row)
LP_FLOP_PER_ROW = 2 * (19 * 128 + 2 * 128 * 128 + 128 * 18)
LP_ROWS_PER_SEGMENT = 2 * (SEG - 1) - 1
LP_FLOP_PER_UPDATE = LP_FLOP_PER_ROW * LP_ROWS_PER_SEGMENT / (SEG - 1)


ORACLE_LAUNCHES = 23  # 736 sweeps: a sweep count the FP64 reference holds walker by walker


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def reap_children():
    """Waits for (then ends) any child process still running, so the bench leaves nothing behind."""
    try:
        import psutil
    except ImportError:
        return
    kids = psutil.Process().children(recursive=True)
    for c in kids:
        progress("child process still running at exit: pid %d %s" % (c.pid, " ".join(c.cmdline()[:4])))
        try:
            c.terminate()
        except psutil.NoSuchProcess:
            pass
    psutil.wait_procs(kids, timeout=5)


def make_kernel(steps_per_launch, seed, storage=None, operator_runs=1, ranlux=False, waves_per_simd=None):
    from feynsimul_amd.kernel_args import KernelArgs
    from feynsimul_amd.kernel import PIMCKernel
    from feynsimul_amd.physical_systems import Lm2m2Cluster
    system = Lm2m2Cluster(3)
    ka = KernelArgs(system=system, nbrOfWalkers=W, N=N, beta=BETA, S=S, enableOperator=True,
                    enableCorrelator=False, enableBisection=True, enablePathShift=False,
                    enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                    enableGlobalOldPath=False, enableBins=False, enableDouble=False,
                    metroStepsPerOperatorRun=steps_per_launch, operatorRuns=operator_runs,
                    nbrOfWalkersPerWorkGroup=1,
                    operators=(system.energyOp, system.meanPairDistanceOp), seed=seed, sampler="wave",
                    pathStorage=storage,
                    histogramObservables=tuple(system.pairDistance(i, j) for i, j in system.pairs),
                    histogramRange=(0.0, HIST_MAX_A), histogramBins=256,
                    enableRanlux=ranlux, luxuaryFactor=1 if ranlux else None,
                    ranluxIntMax=2 ** 31 - 1 if ranlux else None, wavesPerSimd=waves_per_simd)
    k = PIMCKernel(ka)
    if ranlux:
        k.initRanlux()
        k.groupInitRanlux()
    return system, k


def initial_paths(system, rng, walkers=W, sigma=3.0):
    import numpy as np
    pos = system.initialPositions(walkers, rng, side=3.5, sigma=sigma)
    return np.repeat(pos[:, :, None], N, axis=2).reshape(walkers, DOF * N)


HE3_GOLDEN = os.path.join(ROOT, "tests", "golden", "he3_ensemble_fp64.json")
HE3_GOLDEN_WALKERS = os.path.join(ROOT, "tests", "golden", "he3_ensemble_fp64_walkers.npz")


def he3_oracle(per_walker, sweeps):
    """Error of the ensemble means after `sweeps` sweeps from the benchmark's initial paths against the
    FP64 CPU reference-order run of the same 16384 initial paths (tests/golden), when it holds that
    sweep count: combined standard error, and walker by walker where its per-walker values are kept."""
    import numpy as np
    with open(HE3_GOLDEN) as f:
        g = json.load(f)
    row = next((r for r in g["rows"] if r["sweeps"] == sweeps), None)
    if row is None:
        return {"sweeps": sweeps, "note": "no FP64 reference at this sweep count (warmup + steps = 23 has one)"}
    out = {"kind": "FP64 CPU run of the reference-order sampler from the same initial paths (tests/golden)",
           "sweeps": sweeps}
    for col, key in ((0, "E_K"), (1, "pair_distance_A")):
        v = np.asarray(per_walker[col], dtype=np.float64)
        m, se = v.mean(), v.std(ddof=1) / np.sqrt(len(v))
        ref, ref_se = row[key], row[key.replace("_K", "_se_K").replace("_A", "_se_A")]
        d = {"gpu": m, "gpu_se": se, "fp64": ref, "fp64_se": ref_se, "z": abs(m - ref) / np.hypot(se, ref_se)}
        if sweeps in g.get("paired_sweeps", []) and len(v) == g["walkers"]:
            dw = v - np.load(HE3_GOLDEN_WALKERS)["%s_T%d" % (key, sweeps)]
            d["paired_difference"] = float(dw.mean())
            d["paired_se"] = float(dw.std(ddof=1) / np.sqrt(len(dw)))
            d["paired_z"] = abs(d["paired_difference"]) / d["paired_se"]
        out[key] = d
    return out


def chain_mean_ess(series):
    """ESS = Var_1 / SE^2: single-measurement variance over the variance of the grand mean,
    the latter from the spread of the independent chains' means (the reference's error
    model, examples/lm2m2_plotter.py:100-117)."""
    import numpy as np
    x = np.asarray(series, dtype=np.float64)
    se2 = x.mean(axis=0).var(ddof=1) / x.shape[1]
    return float(x.var() / se2), float(np.sqrt(se2))


def drift(series, sweeps_per_measurement=SWEEPS_PER_STEP):
    """Linear trend of the ensemble mean of a (measurements, chains) series: slope per 1000 sweeps and its
    standard error from the chain-by-chain slopes (chains are independent)."""
    import numpy as np
    x = np.asarray(series, dtype=np.float64)
    t = np.arange(x.shape[0], dtype=np.float64) * sweeps_per_measurement / 1000.0
    tc = t - t.mean()
    slopes = (tc[:, None] * (x - x.mean(axis=0))).sum(axis=0) / (tc * tc).sum()
    return float(slopes.mean()), float(slopes.std(ddof=1) / np.sqrt(len(slopes)))


def ess_phase(paths, storage, burn_launches, meas_launches, runs_per_launch=16):
    """Stationarity of E0 on the He3 workload: the virial energy series of every chain (slice mean, one
    measurement every 32 sweeps) after burn-in.  An ESS is only quoted when Sokal's tau_int is at most
    1/5 of the series; otherwise the series' drift is reported instead."""
    import torch
    from feynsimul_amd.stats import effective_sample_size
    _, k = make_kernel(SWEEPS_PER_STEP, seed=777, storage=storage, operator_runs=runs_per_launch)
    k.setPaths(paths)
    for _ in range(burn_launches):
        k.run(wait=False)
    k.setSeriesCapacity(meas_launches * runs_per_launch)
    ess_phase.start_paths = k.getPaths()  # the learned phase starts from the same state
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(meas_launches):
        k.run(wait=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    series = k.getSeries()[:, :, 0]
    r = effective_sample_size(series)
    slope, slope_se = drift(series)
    M = series.shape[0]
    out = {"metric": "E0 (virial energy) series of the plain bisection sampler", "unit": "K",
           "measurements": M, "sweeps_per_measurement": SWEEPS_PER_STEP,
           "burn_in_sweeps": burn_launches * runs_per_launch * SWEEPS_PER_STEP, "seconds": dt,
           "E0_K": r["mean"], "tau_int_measurements": r["tau_int"],
           "drift_K_per_1000_sweeps": slope, "drift_se": slope_se, "drift_z": abs(slope) / slope_se}
    if r["tau_int"] <= M / 5.0:
        out["ess"], out["ess_per_s"] = r["ess"], r["ess"] / dt
    else:
        out["ess"] = None
        out["note"] = ("not stationary: tau_int exceeds 1/5 of the series and E0 drifts (the trimer is unbound at "
                       "beta = 10/K and slowly dissociates), so no ESS is quoted; the ESS/s comparison at fixed sigma "
                       "runs on the stationary workloads (ess_at_fixed_sigma)")
    return out


def learned_phase(paths, train_steps, measurements):
    """The learned proposal on the He3 workload, started from the paths the plain series starts from:
    learned sweep throughput against the BF16 MFMA peak, online training (Adam, 65,536 rows a step, each
    step after 2 plain sweeps that produce its rows), then learned sweeps with one measurement every 32
    (the E0 series drifts as the plain one does, so no ESS is quoted here)."""
    import torch
    from feynsimul_amd.learned import LearnedProposal
    _, k = make_kernel(SWEEPS_PER_STEP, seed=778, storage="memory")
    k.setPaths(paths)
    lp = LearnedProposal(k, batch=65536, lr=1e-3, seed=779)
    lp.sweep(2)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(k._stream)
    lp.sweep(8, wait=False)
    ev1.record(k._stream)
    torch.cuda.synchronize()
    ms_sweep = ev0.elapsed_time(ev1) / 8
    ar0 = lp.lastAcceptanceOfQueued(8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    first = lp.train(1)
    last = lp.trainOnline(train_steps - 1, plainSweepsPerStep=2)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    k.setSeriesCapacity(measurements)
    lp.run(measurements, SWEEPS_PER_STEP)
    series = k.getSeries()[:, :, 0]
    slope, slope_se = drift(series)
    flop = LP_FLOP_PER_UPDATE * UPDATES_PER_SWEEP
    return {"metric": "learned proposal on He3: throughput, training and acceptance",
            "seconds_train": t_train, "adam_steps": train_steps, "nll_first": first, "nll_last": last,
            "acceptance_untrained": ar0, "acceptance_trained": lp.lastAcceptance, "fallback_blocks": lp.fallbackBlocks,
            "E0_K": float(series.mean()), "drift_K_per_1000_sweeps": slope, "drift_se": slope_se,
            "sweeps": measurements * SWEEPS_PER_STEP, "ms_per_learned_sweep": ms_sweep,
            "slice_updates_per_s": UPDATES_PER_SWEEP / (ms_sweep * 1e-3),
            "mfma": {"bound": "mfma", "kernel": "fs_lp_sweep", "achieved": flop / (ms_sweep * 1e-3),
                     "peak": MFMA_BF16_PEAK, "unit": "FLOP/s", "frac": flop / (ms_sweep * 1e-3) / MFMA_BF16_PEAK,
                     "flop_per_update": LP_FLOP_PER_UPDATE, "updates_per_launch": UPDATES_PER_SWEEP}}


def hbm_ceiling(gib=2, repeats=7):
    """Measured HBM ceilings: fs_hbm_probe streams of `gib` GiB (16-byte vectors, grid-stride) that copy,
    only read or only write, at several grid sizes; median of `repeats` launches, best grid."""
    import torch
    from feynsimul_amd import _native
    L = _native.lib()
    nbytes = gib << 30
    a = torch.zeros(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    out = {}
    for kind, name, moved in ((0, "copy", 2 * nbytes), (1, "read", nbytes), (2, "write", nbytes)):
        best = None
        for blocks in (1024, 2048, 4096, 8192, 16384):
            t = []
            for _ in range(repeats):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                _native.check(L.fs_hbm_probe(kind, a.data_ptr(), b.data_ptr(), nbytes // 16, blocks, None))
                e1.record()
                torch.cuda.synchronize()
                t.append(e0.elapsed_time(e1))
            ms = sorted(t)[repeats // 2]
            gbs = moved / (ms * 1e-3) / 1e9
            if best is None or gbs > best["GBps"]:
                best = {"GBps": gbs, "blocks": blocks, "median_ms": ms}
        out[name] = best
    return {"metric": "HBM ceilings (fs_hbm_probe)", "bytes": nbytes, "ceilings": out, "peak_GBps": HBM_PEAK_GBS}


INFINITY_CACHE_BYTES = 256 << 20  # MI355X last-level (Infinity) cache in front of HBM


def walker_scaling(launches=10, warmup=3):
    """Time per slice-update of the plain sampler against the number of chains, for the He3 workload
    (N = 1024, S = 6, 32 sweeps + 1 measurement per launch) and the anharmonic oscillator (N = 4096,
    S = 4, 32 sweeps + 1 energy measurement, no correlator).  With few chains the paths fit in the
    256 MB Infinity Cache and each chain's wavefront is resident for the whole launch; if HBM
    bandwidth bounded the kernel, those launches would update slices faster than the full workloads."""
    import numpy as np
    from feynsimul_amd.kernel_args import KernelArgs
    from feynsimul_amd.kernel import PIMCKernel
    from feynsimul_amd.physical_systems import Lm2m2Cluster, AnharmOsc
    out = {"metric": "slice-updates/s of the plain sampler against chain count", "cache_bytes": INFINITY_CACHE_BYTES,
           "he3": [], "aho": []}
    he3 = Lm2m2Cluster(3)
    aho = AnharmOsc(AHO_LAM)
    for name, system, NN, beta, S_, counts in (("he3", he3, N, BETA, S, (1024, 2048, 4096, 5120, 8192, 16384)),
                                               ("aho", aho, AHO_N, AHO_BETA, 4, (4096, 8192, 16384, 32768, 65536))):
        for walkers in counts:
            progress("walker scaling: %s, %d chains" % (name, walkers))
            extra = {}
            if name == "he3":
                extra = dict(histogramObservables=tuple(system.pairDistance(i, j) for i, j in system.pairs),
                             histogramRange=(0.0, HIST_MAX_A), histogramBins=256)
            ops = (system.energyOp, system.meanPairDistanceOp) if name == "he3" else (system.energyOp,)
            k = PIMCKernel(KernelArgs(system=system, nbrOfWalkers=walkers, N=NN, beta=beta, S=S_, enableOperator=True,
                                      enableCorrelator=False, enableBisection=True, enablePathShift=False,
                                      enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                                      enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=32,
                                      operatorRuns=1, nbrOfWalkersPerWorkGroup=1, operators=ops, seed=3,
                                      sampler="wave", pathStorage="memory", **extra))
            rng = np.random.default_rng(1)
            if name == "he3":
                pos = system.initialPositions(walkers, rng, side=3.5, sigma=3.0)
                k.setPaths(np.repeat(pos[:, :, None], NN, axis=2).reshape(walkers, system.DOF * NN))
            else:
                k.setPaths(rng.uniform(-1, 1, (walkers, NN)).astype(np.float32))
            ms, _ = timed_launches(k, launches, warmup)
            seg = 1 << S_
            upd = walkers * (NN // seg) * (seg - 1) * 32
            out[name].append({"chains": walkers, "path_bytes": walkers * NN * system.DOF * 4, "ms_per_launch": ms,
                              "slice_updates_per_s": upd / (ms * 1e-3),
                              "paths_fit_in_cache": walkers * NN * system.DOF * 4 <= INFINITY_CACHE_BYTES})
            del k
    return out


def lp_bench(steps, warmup, spw=0):
    """The learned sweep alone on the He3 workload (fs_lp_sweep): `steps` sweeps timed with HIP events on
    the kernel's stream after `warmup`, from paths 512 plain sweeps into the run; a network with a
    non-zero output layer (the trained case's cost)."""
    import numpy as np
    import torch
    from feynsimul_amd.learned import LearnedProposal
    system, k = make_kernel(512, seed=12345, storage="memory")
    k.setPaths(initial_paths(system, np.random.default_rng(0)))
    k.run()
    lp = LearnedProposal(k, batch=65536, lr=1e-3, seed=779, segsPerWave=spw)
    lp.train(4)
    lp.sweep(max(warmup, 1))
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(k._stream)
    lp.sweep(steps, wait=False)
    ev1.record(k._stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    ar = lp.lastAcceptanceOfQueued(steps)
    flop = LP_FLOP_PER_UPDATE * UPDATES_PER_SWEEP
    return {"metric": "learned sweeps: slice-updates/s (fs_lp_sweep, BF16 MFMA MLP)",
            "value": UPDATES_PER_SWEEP / (ms * 1e-3), "unit": "slice-updates/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16/f32", "data": "synthetic: He3 paths after 512 plain sweeps",
            "config": {"workload": "LM2M2 He3, N=1024, 16384 chains, S=6, one learned sweep per step"},
            "acceptance": ar,
            "roofline": {"bound": "mfma", "achieved": flop / (ms * 1e-3) / 1e12, "peak": MFMA_BF16_PEAK / 1e12,
                         "unit": "TFLOP/s", "frac": flop / (ms * 1e-3) / MFMA_BF16_PEAK, "kernel": "fs_lp_sweep",
                         "avg_launch_ms": ms, "flop_per_update": LP_FLOP_PER_UPDATE,
                         "updates_per_launch": UPDATES_PER_SWEEP, "traffic": None}}


def lpc_bench(steps, warmup, width=None, train_steps=None):
    """The learned whole-path move alone on the anharmonic workload (x^2/2 + 0.1 x^4, N = 4096,
    beta = 40, 65536 chains): `steps` moves timed with HIP events on the kernel's stream after
    `warmup`, from paths 2048 plain sweeps (S = 6) into the run, with the network trained as the
    fixed-sigma comparison trains it.  One move = the network kernel on the coarse nodes (fs_lp_sweep)
    and the fine pass (fs_lpc_fine, harmonic reference bridges, decision, write)."""
    import numpy as np
    import torch
    from feynsimul_amd.kernel_args import KernelArgs
    from feynsimul_amd.kernel import PIMCKernel
    from feynsimul_amd.learned import LearnedProposal
    from feynsimul_amd.physical_systems import AnharmOsc
    width = width or LEARNED_WIDTH
    system = AnharmOsc(AHO_LAM)
    k = PIMCKernel(KernelArgs(system=system, nbrOfWalkers=AHO_W, N=AHO_N, beta=AHO_BETA, S=6, enableOperator=True,
                              enableCorrelator=False, enableBisection=True, enablePathShift=False,
                              enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                              enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=2048,
                              operatorRuns=1, nbrOfWalkersPerWorkGroup=1, operators=(system.energyOp,), seed=41,
                              sampler="wave", pathStorage="memory"))
    k.setPaths(np.random.default_rng(40).uniform(-1, 1, (AHO_W, AHO_N)).astype(np.float32))
    k.run()
    lp = LearnedProposal(k, batch=65536, lr=1e-3, seed=44, coarse=True, width=width,
                         bridgeOmega2=FIXED_SIGMA_LEARNED["aho"]["bridge"])
    cfg = FIXED_SIGMA_LEARNED["aho"]
    train_steps = train_steps or cfg["train_steps"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lp.trainOnline(train_steps, plainSweepsPerStep=cfg["plain_sweeps_per_step"])
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    # Adam steps alone (no plain sweeps between them), for the cost of one training step
    t0 = time.perf_counter()
    lp.train(20)
    ms_adam = 1e3 * (time.perf_counter() - t0) / 20
    lp.sweep(max(warmup, 1))
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(k._stream)
    lp.sweep(steps, wait=False)
    ev1.record(k._stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    ar = lp.lastAcceptanceOfQueued(steps)
    slices = AHO_W * AHO_N
    # HBM bytes the move needs: every slice read, the accepted paths written
    algo = 4.0 * slices * (1.0 + ar)
    return {"metric": "learned whole-path moves: slice-updates/s (fs_lp_sweep + fs_lpc_fine)",
            "value": slices / (ms * 1e-3), "unit": "slice-updates/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16/f32", "data": "synthetic: anharmonic-oscillator paths after 2048 plain sweeps",
            "config": {"workload": "anharmonic oscillator x^2/2 + 0.1 x^4, N=4096, beta=40, 65536 chains, "
                                   "one whole-path move per step, %d hidden units, %d Adam steps, harmonic "
                                   "reference bridges (omega^2 %.4g)" % (width, train_steps, lp.bridgeOmega2)},
            "acceptance": ar,
            "training": {"adam_steps": train_steps, "plain_sweeps_per_step": cfg["plain_sweeps_per_step"],
                         "seconds": t_train,
                         "ms_per_adam_step": ms_adam, "batch": 65536},
            "roofline": {"bound": "hbm", "achieved": algo / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "kernel": "fs_lp_sweep + fs_lpc_fine",
                         "avg_launch_ms": ms, "bytes_per_move": algo, "traffic": None}}


# ---- ESS/s of E0 at fixed sigma, learned vs plain, on the stationary workloads ----
TAU_CHAINS = 512  # chains whose autocorrelation estimates tau_int in the stopping rule


def _series_stats(series, exact):
    """Statistics of an E0 series (torch tensor (measurements, chains) on the GPU): the standard error
    of the grand mean from the spread of the chain means (chains are independent), ESS = variance /
    SE^2, and Sokal's tau_int from the first TAU_CHAINS chains."""
    import numpy as np
    from feynsimul_amd.stats import integrated_autocorr_time
    x = series.double()
    M, C = x.shape
    means = x.mean(dim=0)
    se = float(means.std() / np.sqrt(C))
    mean = float(means.mean())
    tau, _ = integrated_autocorr_time(series[:, :TAU_CHAINS].cpu().numpy())
    return {"E0": mean, "sigma": se, "tau_int_measurements": tau, "measurements": int(M),
            "ess": float(x.var() / se ** 2), "z": abs(mean - exact) / se}


def run_to_sigma(block, get_series, sigma, exact, max_seconds, measurements_per_block, capacity):
    """Calls block() (one block of measurements on the GPU) until the chain-mean standard error of E0
    is at most sigma and Sokal's tau_int is at most 1/5 of the series, or max_seconds of sampling have
    passed, or the series buffer (capacity measurements) has no room for another block.  Wall time
    counts the blocks only (each ends synchronised), not the host analysis."""
    import torch
    t = 0.0
    n = 0
    check = measurements_per_block
    while True:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        block()
        torch.cuda.synchronize()
        t += time.perf_counter() - t0
        n += measurements_per_block
        full = n + measurements_per_block > capacity
        # the rule is checked each time the series has doubled (the host analysis stays a small share)
        if n < check and not full and t <= max_seconds:
            continue
        check = 2 * n
        st = _series_stats(get_series(), exact)
        if (st["sigma"] <= sigma and st["tau_int_measurements"] <= st["measurements"] / 5.0) or t > max_seconds \
                or full:
            break
    st["seconds"] = t
    st["sigma_target"] = sigma
    st["reached"] = bool(st["sigma"] <= sigma)
    # time to exactly sigma (the ESS grows linearly with sampling time)
    st["seconds_to_sigma"] = t * (st["sigma"] / sigma) ** 2
    st["measurements_per_block"] = measurements_per_block
    return st


# the learned leg of the fixed-sigma comparison: the whole-path move, its network on 32 hidden units, trained
# with Adam on rows drawn from the equilibrated paths of the plain sampler.  Plain sweeps between the steps buy
# nothing here: on the anharmonic workload 500 steps reach acceptance 0.630 after one sweep each (0.47 s) and
# 0.629 without (0.066 s); 1200 steps 0.77, 2000 steps 0.81 (0.26 s), 3000 steps 0.82.
# bridge: the whole-path move's harmonic reference, omega^2 = <x . grad V> / <|x|^2> of the start paths
FIXED_SIGMA_LEARNED = {"ho": {"train_steps": 300, "batch": 16384, "plain_sweeps_per_step": 0, "bridge": "auto"},
                       "aho": {"train_steps": 2000, "batch": 65536, "plain_sweeps_per_step": 0, "bridge": "auto"}}
LEARNED_WIDTH = 32
# Measurement cadences each leg is run at (sweeps or moves per E0 measurement): a measurement costs about
# as much as reading every path once, so measuring after every sweep wastes time, while measuring far
# less often than tau_int discards information the sweeps paid for.  Each cadence runs for SCAN_SECONDS
# of sampling; the one with the most ESS/s is then run to sigma.
PLAIN_CADENCES = {"ho": (1, 2, 4, 8, 16, 32), "aho": (2, 4, 8, 16, 32)}
LEARNED_CADENCES = (1, 2, 4, 8)
SCAN_SECONDS = 3.0


def _ess_per_s(st, sigma, t_extra=0.0):
    """ESS/s at the target sigma: the ESS grows linearly with sampling time, so a run that stopped at
    another sigma is scaled to the time it would need for sigma; t_extra (training) is added to it."""
    return st["ess"] / (st["seconds_to_sigma"] + t_extra) * (st["sigma"] / sigma) ** 2


def fixed_sigma(workload, sigma, max_seconds=60.0, S=None, copy_GBps=None):
    """Plain bisection and the learned whole-path move from the same equilibrated paths, each at the
    measurement cadence that gives it the most ESS/s (a short run per cadence), then run to the same
    standard error sigma of E0 (virial energy, slice mean).  The learned leg's ESS/s counts its online
    training.  ratio = ESS/s learned / ESS/s plain at that sigma.
    `ceiling`: the plain sampler's tau_int of E0 in sweeps (a measurement after every sweep) and, given
    the measured HBM copy ceiling in GB/s, the largest ESS/s ratio over plain that any move replacing a
    sweep could reach."""
    import numpy as np
    import torch
    from feynsimul_amd.kernel_args import KernelArgs
    from feynsimul_amd.kernel import PIMCKernel
    from feynsimul_amd.learned import LearnedProposal, harmonic_omega2
    from feynsimul_amd.physical_systems import HarmOsc, AnharmOsc
    from feynsimul_amd.stats import integrated_autocorr_time
    if workload == "ho":
        system, W, NN, beta = HarmOsc(), HO_W, HO_N, HO_BETA
        S = S or HO_S
        exact = ho_exact_x2(HO_N, HO_BETA)
        burn = 2000
    else:
        with open(AHO_GOLDEN) as f:
            exact = json.load(f)["values"]["E_virial"]
        system, W, NN, beta = AnharmOsc(AHO_LAM), AHO_W, AHO_N, AHO_BETA
        S = S or AHO_S
        burn = 2048
    runs = 16

    def kernel(M, opRuns, seed):
        return PIMCKernel(KernelArgs(system=system, nbrOfWalkers=W, N=NN, beta=beta, S=S, enableOperator=True,
                                     enableCorrelator=False, enableBisection=True, enablePathShift=False,
                                     enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                                     enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=M,
                                     operatorRuns=opRuns, nbrOfWalkersPerWorkGroup=1, operators=(system.energyOp,),
                                     seed=seed, sampler="wave"))
    kb = kernel(burn, 1, 41)
    kb.setPaths(np.random.default_rng(40).uniform(-1, 1, (W, NN)).astype(np.float32))
    kb.run()
    start = kb.getPaths()
    del kb
    # series buffers of at most 1 GiB
    cap = min(65536, (1 << 30) // (W * 4)) // runs * runs

    # plain: a measurement every k sweeps
    def plain_leg(k, seconds):
        kp = kernel(k, runs, 42)
        kp.setPaths(start)
        kp.setSeriesCapacity(cap)
        st = run_to_sigma(lambda: kp.run(wait=False), lambda: kp.getSeriesTensor()[:, :, 0], sigma, exact,
                          seconds, runs, cap)
        st["sweeps_per_measurement"] = k
        st["acceptance"] = kp.getAcceptanceRate()
        # sweeps and their share of the measurements
        st["ms_per_sweep"] = 1e3 * st["seconds"] / (st["measurements"] * k)
        st["ess_per_s"] = _ess_per_s(st, sigma)
        del kp
        torch.cuda.empty_cache()
        return st
    scan_plain = {k: plain_leg(k, SCAN_SECONDS) for k in PLAIN_CADENCES[workload]}
    k_plain = max(scan_plain, key=lambda k: scan_plain[k]["ess_per_s"])
    plain = plain_leg(k_plain, max_seconds)
    plain["cadence_scan_ess_per_s"] = {str(k): v["ess_per_s"] for k, v in scan_plain.items()}
    plain["cadence_scan_tau_int"] = {str(k): v["tau_int_measurements"] for k, v in scan_plain.items()}
    # the sweep's own cost, from the sparsest cadence (measurements a small share)
    k_sparse = max(scan_plain)
    ms_sweep = scan_plain[k_sparse]["ms_per_sweep"]
    # E0 after every plain sweep: tau_int in sweeps is what a move that replaced the sweep and gave an
    # independent E0 value every sweep (tau_int = 1) could gain per sweep
    kt = kernel(1, runs, 45)
    kt.setPaths(start)
    kt.setSeriesCapacity(TAU_MEASUREMENTS)
    for _ in range(TAU_MEASUREMENTS // runs):
        kt.run(wait=False)
    tau_sweeps, window = integrated_autocorr_time(kt.getSeries()[:, :, 0])
    del kt
    torch.cuda.empty_cache()
    ceiling = {"plain_tau_int_sweeps": tau_sweeps, "window_sweeps": window, "measurements": TAU_MEASUREMENTS,
               "plain_ms_per_sweep": ms_sweep, "plain_ms_per_sweep_from_cadence": k_sparse}
    if copy_GBps:
        # such a move still reads and writes every slice once: at the measured HBM copy ceiling that takes
        # path_ms, so its ESS/s can exceed the plain sampler's (measuring after every sweep at no cost) by
        # at most tau_sweeps * plain_ms / path_ms
        path_ms = 1e3 * 2 * W * NN * 4 / (copy_GBps * 1e9)
        ceiling.update({"path_read_write_ms_at_copy_ceiling": path_ms, "copy_ceiling_GBps": copy_GBps,
                        "max_ess_per_s_ratio_any_sweep_move": tau_sweeps * ms_sweep / path_ms})
    # learned: the whole-path move, trained online from the same paths, then a measurement every k moves
    cfg = FIXED_SIGMA_LEARNED[workload]

    def learned_leg(width, cadences):
        kl = kernel(SWEEPS_PER_STEP, runs, 43)
        kl.setPaths(start)
        # the reference's omega^2 from the start paths (counted with the training time)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        omega2 = harmonic_omega2(kl) if cfg["bridge"] == "auto" else float(cfg["bridge"])
        t_omega = time.perf_counter() - t0
        lp = LearnedProposal(kl, batch=cfg["batch"], lr=1e-3, seed=44, coarse=True, width=width,
                             bridgeOmega2=omega2)
        torch.cuda.synchronize()
        t0 = time.perf_counter() - t_omega
        first = lp.train(1)
        last = lp.trainOnline(cfg["train_steps"] - 1, plainSweepsPerStep=cfg["plain_sweeps_per_step"])
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0

        def leg(k, seconds):
            kl.setSeriesCapacity(cap)
            fb = lp.fallbackBlocks
            r = run_to_sigma(lambda: lp.run(runs, k), lambda: kl.getSeriesTensor()[:, :, 0], sigma, exact,
                             seconds, runs, cap)
            r.update({"moves_per_measurement": k, "acceptance": lp.lastAcceptance,
                      "fallback_blocks": lp.fallbackBlocks - fb,
                      "ms_per_move": 1e3 * r["seconds"] / (r["measurements"] * k)})
            r["ess_per_s"] = _ess_per_s(r, sigma, t_train)
            r["ess_per_s_without_training"] = _ess_per_s(r, sigma)
            return r
        scan = {k: leg(k, SCAN_SECONDS) for k in cadences}
        k_best = max(scan, key=lambda k: scan[k]["ess_per_s"])
        r = leg(k_best, max(max_seconds - t_train, 1.0))
        r.update({"seconds_train": t_train, "adam_steps": cfg["train_steps"], "training_batch": cfg["batch"],
                  "plain_sweeps_per_adam_step": cfg["plain_sweeps_per_step"],
                  "nll_first": first, "nll_last": last, "hidden_units": width, "bridge_omega2": lp.bridgeOmega2,
                  "cadence_scan_ess_per_s": {str(k): v["ess_per_s"] for k, v in scan.items()},
                  "cadence_scan_tau_int": {str(k): v["tau_int_measurements"] for k, v in scan.items()}})
        del kl, lp
        torch.cuda.empty_cache()
        return r

    learned = learned_leg(LEARNED_WIDTH, LEARNED_CADENCES)
    # the full 128-unit network of the segment proposal, for comparison (the anharmonic workload), at the
    # 32-unit network's cadence
    full = learned_leg(128, (learned["moves_per_measurement"],)) if workload == "aho" else None
    # per unit of work: effective samples per learned move over effective samples per plain sweep
    per = [plain["ess"] / (plain["measurements"] * k_plain),
           learned["ess"] / (learned["measurements"] * learned["moves_per_measurement"])]
    out = {"workload": {"ho": "harmonic oscillator N=256, beta=10, 4096 chains",
                         "aho": "anharmonic oscillator x^2/2 + 0.1 x^4, N=4096, beta=40, 65536 chains"}[workload]
                        + ", plain S=%d, E0 every %d sweeps (the best of %s)" % (S, k_plain, list(PLAIN_CADENCES[workload])),
            "learned_form": "whole-path move: network on the nodes every 64 slices (%d hidden units), harmonic "
                            "reference bridges between them (omega^2 %.4g from the start paths), one MH decision "
                            "per path, E0 every %d moves (the best of %s)"
                            % (LEARNED_WIDTH, learned["bridge_omega2"], learned["moves_per_measurement"],
                               list(LEARNED_CADENCES)),
            "sigma": sigma, "E0_exact": exact, "plain": plain, "learned": learned,
            "ratio_learned_over_plain": learned["ess_per_s"] / plain["ess_per_s"],
            "ratio_without_training": learned["ess_per_s_without_training"] / plain["ess_per_s"],
            "ess_per_move_over_ess_per_plain_sweep": per[1] / per[0],
            # against a plain sampler that measured after every sweep at no cost: W chains / (tau_int in sweeps
            # x ms per sweep), the most any measurement cadence could give it
            "ratio_vs_plain_with_free_measurement": learned["ess_per_s"] / (W / (tau_sweeps * ms_sweep * 1e-3)),
            "ceiling": ceiling,
            "bias_within_2_sigma": {"plain": plain["z"] <= 2.0, "learned": learned["z"] <= 2.0}}
    if full is not None:
        out["learned_128_units"] = full
        out["ratio_learned_128_units_over_plain"] = full["ess_per_s"] / plain["ess_per_s"]
        out["bias_within_2_sigma"]["learned_128_units"] = full["z"] <= 2.0
    return out


RAGGED_PROFILE = os.path.join(ROOT, "profiles", "fs_batch_sweep_ragged.md")


def ragged_bench(steps, warmup, cpu=True, levels=None, shape=None, separate=True):
    """Batched ragged LM2M2 systems in one persistent launch (32 sweeps per launch), autotuned."""
    import numpy as np
    import torch
    from feynsimul_amd.kernel_args import KernelArgs
    from feynsimul_amd.physical_systems import Lm2m2Cluster
    from feynsimul_amd.batch import BatchedPIMC
    from feynsimul_amd.autotune import autotune_batch, autotune_levels
    rng = np.random.default_rng(5)
    kas, starts = [], []
    for n in (3, 4, 6, 8):
        system = Lm2m2Cluster(n)
        for NN in (256, 512, 1024, 2048):
            kas.append(KernelArgs(system=system, nbrOfWalkers=2048, N=NN, beta=BETA, S=6, enableOperator=True,
                                  enableCorrelator=False, enableBisection=True, enablePathShift=False,
                                  enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                                  enableGlobalOldPath=False, enableBins=False, metroStepsPerOperatorRun=SWEEPS_PER_STEP,
                                  operatorRuns=1, nbrOfWalkersPerWorkGroup=1, operators=(system.energyOp,),
                                  seed=n * 4096 + NN, sampler="wave"))
            pos = system.initialPositions(2048, rng, side=3.5, sigma=3.0)
            starts.append(np.repeat(pos[:, :, None], NN, axis=2).reshape(2048, system.DOF * NN))
    # bisection level per type from measured acceptance (the reference's 20-50 % window), then launch shape
    if levels is None:
        progress("ragged: bisection level per type from measured acceptance")
        levels, table = autotune_levels(kas, initialPaths=starts, warmupSweeps=512)
    else:
        table = []
    for ka, S in zip(kas, levels):
        ka.S = S
    if shape is None:
        progress("ragged: launch shape (waves per workgroup, order, register classes; cached per device)")
        best, trials = autotune_batch(kas, initialPaths=starts, use_cache=True, verbose=True)
    else:
        best, trials = dict(shape, ms_per_launch=None), []
    batch = BatchedPIMC(kas, wavesPerBlock=best["wavesPerBlock"], order=best["order"], kernels=best["kernels"],
                        kernelWaves=best["kernelWaves"])
    for k, p in zip(batch.kernels, starts):
        k.setPaths(p)
    for _ in range(warmup):
        batch.run()
    graph = batch.captureGraph(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.replay(graph)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_s = batch.getRunTime() / steps   # device time of the replay (HIP events on its stream)
    upd = batch.sliceUpdatesPerLaunch() * steps
    # algorithmic HBM bytes: a sweep reads and writes each slice of a walker once, 8 DOF bytes per slice
    bytes_launch = sum(k._nbrOfWalkers * k._N * 8.0 * int(k._system.DOF) for k in batch.kernels) * SWEEPS_PER_STEP
    acc = [k.getAcceptanceRate() for k in batch.kernels]
    per_type = [k._nbrOfWalkers * (k._N >> k._S) * ((1 << k._S) - 1) * SWEEPS_PER_STEP for k in batch.kernels]
    accepted = sum(u * a for u, a in zip(per_type, acc)) * steps
    out = {"metric": "path-slice updates/s, ragged batch in one persistent launch", "value": upd / wall,
            "unit": "slice-updates/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic: seeded Gaussian blobs (sigma 3 A)",
            "config": {"workload": "LM2M2 clusters n in {3,4,6,8} x N in {256,512,1024,2048}, 2048 chains each "
                                   "(32768 chains), beta=10/K, S per type from measured acceptance, 32 sweeps per "
                                   "step, hipGraph replay",
                       "levels": levels, "autotune": best, "trials": trials,
                       "acceptance_by_level": [{str(S): round(a, 4) for S, a in sorted(t.items())} for t in table]},
            "roofline": {"bound": "hbm", "achieved": bytes_launch / launch_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": bytes_launch / launch_s / 1e9 / HBM_PEAK_GBS, "traffic": measured_traffic(RAGGED_PROFILE),
                         "kernel": "fs_batch_sweep_*", "avg_launch_ms": launch_s * 1e3,
                         "bytes_per_launch": bytes_launch, "updates_per_launch": batch.sliceUpdatesPerLaunch()},
            "register_classes": {"kernel_of_type": batch.kernelOfType, "waves_of_kernel": batch.kernelWaves},
            "accepted_slice_updates_per_s": accepted / wall,
            "acceptance": [round(a, 4) for a in acc]}
    if separate:
        del batch, graph
        torch.cuda.empty_cache()
        out["separate"] = ragged_separate(kas, starts, launch_s)
    if cpu and int(os.environ.get("RANK", "0")) == 0:
        progress("CPU baseline, ragged types")
        out["cpu_baseline"] = cpu_baseline_ragged(kas)
    return out


def ragged_separate(kas, starts, batch_launch_s, fill=8):
    """The same ragged types launched one at a time, as the reference's run loop does (one program per
    system): each type's stand-alone kernel at its 2048 chains, and at `fill` x 2048 chains (enough
    wavefronts to fill the GPU, the type's full-occupancy rate), both at the batch's bisection levels.
    separate_ms is what the sixteen launches take back to back; full_occupancy_ms the sixteen types at
    their full-occupancy rates (the time no scheduler can beat without a faster sampler); the batch's
    efficiency is full_occupancy_ms over its own launch time."""
    import copy
    import numpy as np
    import torch
    from feynsimul_amd.kernel import PIMCKernel
    rows = []
    for ka, start in zip(kas, starts):
        progress("ragged: type n=%d N=%d alone" % (int(ka.system.DOF) // 3, ka.N))
        row = {"atoms": int(ka.system.DOF) // 3, "N": ka.N, "S": ka.S}
        for chains in (ka.nbrOfWalkers, fill * ka.nbrOfWalkers):
            kb = copy.copy(ka)
            kb.nbrOfWalkers = chains
            k = PIMCKernel(kb)
            k.setPaths(np.tile(start, (chains // start.shape[0], 1)))
            ms, _ = timed_launches(k, 3, 1)
            row["ms_%d_chains" % chains] = ms
            del k
            torch.cuda.empty_cache()
        row["ms_at_full_occupancy_rate"] = row["ms_%d_chains" % (fill * ka.nbrOfWalkers)] / fill
        rows.append(row)
    sep = sum(r["ms_%d_chains" % kas[0].nbrOfWalkers] for r in rows)
    full = sum(r["ms_at_full_occupancy_rate"] for r in rows)
    return {"types": rows, "separate_ms": sep, "full_occupancy_ms": full, "batch_ms": batch_launch_s * 1e3,
            "batch_speedup_over_separate": sep / (batch_launch_s * 1e3),
            "scheduler_efficiency": full / (batch_launch_s * 1e3)}


# ---- anharmonic oscillator: the correlator and RNG stress workload ----
AHO_LAM, AHO_N, AHO_BETA, AHO_W, AHO_S, AHO_LAGS = 0.1, 4096, 40.0, 65536, 4, 512
AHO_SEG = 1 << AHO_S
AHO_UPDATES_PER_SWEEP = AHO_W * (AHO_N // AHO_SEG) * (AHO_SEG - 1)
AHO_BYTES_PER_UPDATE = 8.0 * 1 * AHO_SEG / (AHO_SEG - 1)
# 26 per coordinate + 2 potential evaluations (x^2/2 + lam x^4 and the eps-weighted sum: 5 ops)
AHO_VALU_PER_UPDATE = 26 * 1 + 2 * 5
AHO_GOLDEN = os.path.join(ROOT, "tests", "golden", "aho_c5.json")
AHO_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "fs_wave_sweep_aho_n4096.md")


def make_aho_kernel(S, steps, opRuns, corr, ranlux=False, seed=5):
    from feynsimul_amd.kernel_args import KernelArgs
    from feynsimul_amd.kernel import PIMCKernel
    from feynsimul_amd.physical_systems import AnharmOsc
    system = AnharmOsc(AHO_LAM)
    ka = KernelArgs(system=system, nbrOfWalkers=AHO_W, N=AHO_N, beta=AHO_BETA, S=S, enableOperator=True,
                    enableCorrelator=corr, enableBisection=True, enablePathShift=False,
                    enableSingleNodeMove=False, enableParallelizePath=True, enableGlobalPath=False,
                    enableGlobalOldPath=False, enableBins=False, enableDouble=False,
                    enableRanlux=ranlux, luxuaryFactor=1 if ranlux else None,
                    ranluxIntMax=2 ** 31 - 1 if ranlux else None,
                    metroStepsPerOperatorRun=steps, operatorRuns=opRuns, nbrOfWalkersPerWorkGroup=1,
                    operators=(system.energyOp,), correlators=("x1",) if corr else None,
                    seed=seed, sampler="wave", correlatorLags=AHO_LAGS if corr else None)
    k = PIMCKernel(ka)
    if ranlux:
        k.initRanlux()
        k.groupInitRanlux()
    return system, k


def timed_launches(k, steps, warmup):
    """Mean device time (ms) of `steps` launches on the kernel's stream (HIP events) and the wall time (s)."""
    import numpy as np
    import torch
    for _ in range(warmup):
        k.run()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i][0].record(k._stream)
        k.run(wait=False)
        ev[i][1].record(k._stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return float(np.mean([a.elapsed_time(b) for a, b in ev])), wall


def aho_bench(steps, warmup, cpu=True, ess=True):
    """Anharmonic oscillator x^2/2 + 0.1 x^4, N = 4096, beta = 40, 65536 chains, S = 4: one step = 32
    sweeps + one measurement of the virial energy and the two-time correlator <x(0) x(tau)>,
    tau = 0..511 (FP64 Kahan bins); xorshift128 and RANLUX luxury 2; estimators against the FP64
    transfer matrix of the same discretised action; ESS/s of E0, plain and learned."""
    import numpy as np
    import torch
    with open(AHO_GOLDEN) as f:
        gold = json.load(f)["values"]
    rng = np.random.default_rng(0)
    p0 = rng.uniform(-1, 1, (AHO_W, AHO_N)).astype(np.float32)
    _, kb = make_aho_kernel(AHO_S, 512, 1, False, seed=3)
    kb.setPaths(p0)
    for _ in range(4):
        kb.run(wait=False)
    paths = kb.getPaths()
    del kb
    torch.cuda.empty_cache()
    # throughput, xorshift128 (the headline line of this workload)
    _, k = make_aho_kernel(AHO_S, SWEEPS_PER_STEP, 1, True, seed=4)
    k.setPaths(paths)
    ms, wall = timed_launches(k, steps, warmup)
    updates = AHO_UPDATES_PER_SWEEP * SWEEPS_PER_STEP
    mean, se = k.getCorrelator()
    per = k.getOperators()[0]
    E, Ese = float(per.mean()), float(per.std() / np.sqrt(AHO_W))
    cor = np.array(gold["correlator"][:AHO_LAGS])
    z = np.abs(mean[0] - cor) / se[0]
    eps = AHO_BETA / AHO_N
    gap = float(-np.log(mean[0][300] / mean[0][100]) / (200 * eps))
    achieved = AHO_BYTES_PER_UPDATE * updates / (ms * 1e-3) / 1e9
    line = {"metric": "path-slice updates/s (plain bisection sampler)", "value": updates * steps / wall,
            "unit": "slice-updates/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic: paths uniform on [-1, 1], 2048 burn-in sweeps",
            "config": {"workload": "anharmonic oscillator x^2/2 + 0.1 x^4, N=4096, beta=40, 65536 chains, S=4, "
                                   "xorshift128, 32 sweeps + 1 measurement (virial energy, correlator "
                                   "tau=0..511 in FP64 Kahan bins) per step", "sampler": "wave",
                       "chains": AHO_W, "slices": AHO_N, "bisection_level": AHO_S, "lags": AHO_LAGS},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(AHO_TRAFFIC_PROFILE),
                         "kernel": "fs_wave_sweep", "avg_launch_ms": ms, "bytes_per_update": AHO_BYTES_PER_UPDATE,
                         "updates_per_launch": updates},
            "valu": dict(valu_fields(AHO_VALU_PER_UPDATE * updates / (ms * 1e-3), AHO_TRAFFIC_PROFILE),
                         ops_per_update=AHO_VALU_PER_UPDATE),
            "acceptance": k.getAcceptanceRate(),
            "oracle": {"kind": "FP64 transfer matrix of the discretised action (tests/golden/aho_c5.json)",
                       "E_virial": E, "E_virial_se": Ese, "E_virial_exact": gold["E_virial"],
                       "E_virial_z": abs(E - gold["E_virial"]) / Ese,
                       "correlator_max_z": float(z.max()), "correlator_lags": AHO_LAGS,
                       "gap_from_correlator": gap, "gap_exact": gold["gap"]}}
    del k
    torch.cuda.empty_cache()
    # the same with RANLUX luxury 2 streams
    _, kr = make_aho_kernel(AHO_S, SWEEPS_PER_STEP, 1, True, ranlux=True, seed=6)
    kr.setPaths(paths)
    ms_rl, wall_rl = timed_launches(kr, max(2, steps // 2), 1)
    mean_rl, se_rl = kr.getCorrelator()
    line["ranlux"] = {"value": updates * max(2, steps // 2) / wall_rl, "avg_launch_ms": ms_rl,
                      "luxury": 2, "correlator_max_z": float((np.abs(mean_rl[0] - cor) / se_rl[0]).max())}
    del kr
    torch.cuda.empty_cache()
    if not ess:
        if cpu and int(os.environ.get("RANK", "0")) == 0:
            line["cpu_baseline"] = cpu_baseline_aho()
        return line
    # ESS/s of E0 at fixed sigma, plain vs learned, at the configured S = 4 and at S = 6
    progress("ESS/s at fixed sigma: S=4 and S=6")
    copy = hbm_ceiling(1, 3)["ceilings"]["copy"]["GBps"]
    # plain at S = 6, the level that decorrelates E0 fastest (tau_int 17 sweeps against 210 at S = 4)
    line["ess_at_fixed_sigma"] = {"S6": fixed_sigma("aho", AHO_SIGMA, S=6, copy_GBps=copy)}
    if cpu and int(os.environ.get("RANK", "0")) == 0:
        line["cpu_baseline"] = cpu_baseline_aho()
    return line


def cpu_thread_counts():
    """Threads for the CPU baseline: the cores os.sched_getaffinity reports, and OMP_NUM_THREADS when it
    names a different count (a CPU quota smaller than the affinity mask makes more threads slower)."""
    aff = len(os.sched_getaffinity(0))
    counts = [aff]
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and 0 < int(env) != aff:
        counts.append(int(env))
    return aff, counts


def cpu_reference_rate(system, NN, beta, S_, ops, make_paths, updates_per_walker_sweep, describe, run_s=2.5,
                       repeats=5):
    """The oracle's reference-order sampler (kernel.c:514-837 restated in C, xorshift128, FP32) on this
    host: built strict-IEEE and with -ffast-math (the reference builds its kernels with
    -cl-fast-relaxed-math, kernel.py:372-382), each timed as the median of `repeats` runs of a sample
    sized to about run_s seconds, at each thread count of cpu_thread_counts.  value = the best median."""
    import numpy as np
    from oracle import oracle as O
    aff, counts = cpu_thread_counts()
    variants = {}
    best = None
    for name, fast in (("strict", False), ("fast_math", True)):
        lib = O.build(system, NN, beta, operators=ops, native=True, fast_math=fast)
        for threads in counts:
            walkers = max(threads * 4, 16)
            rng = np.random.default_rng(1)
            sweeps = 1
            while True:  # calibration, also the warm-up
                paths = np.ascontiguousarray(make_paths(rng, walkers).astype(np.float32))
                cfg = O.make_cfg(NN, S_, int(system.DOF), walkers, beta, 1, sweeps, nOps=len(ops))
                seeds = O.xorshift_rows(3, 0, cfg.nThreads + 1)
                t0 = time.perf_counter()
                O.run_reference(lib, cfg, paths, seeds=seeds, nthreads=threads)
                dt = time.perf_counter() - t0
                if dt > run_s / 3 or sweeps >= 1 << 18:
                    break
                sweeps = max(sweeps * 2, int(sweeps * run_s / max(dt, 1e-3)))
            times = []
            for _ in range(repeats):
                paths = np.ascontiguousarray(make_paths(rng, walkers).astype(np.float32))
                seeds = O.xorshift_rows(3, 0, cfg.nThreads + 1)
                t0 = time.perf_counter()
                O.run_reference(lib, cfg, paths, seeds=seeds, nthreads=threads)
                times.append(time.perf_counter() - t0)
            med = float(np.median(times))
            rate = walkers * sweeps * updates_per_walker_sweep / med
            key = "%s_%dt" % (name, threads)
            variants[key] = {"value": rate, "threads": threads, "median_s": med, "runs_s": times,
                             "chains": walkers, "sweeps": sweeps}
            if best is None or rate > best[1]["value"]:
                best = (key, variants[key])
    key, b = best
    return {"value": b["value"], "unit": "slice-updates/s", "cores": b["threads"], "kind": "port",
            "affinity_cores": aff, "variant": key, "variants": variants,
            "sample": "%s; median of %d runs of %d chains x %d sweeps; reference work-item order, FP32, "
                      "xorshift128, gcc -O3 -march=native -fopenmp, strict and -ffast-math builds"
                      % (describe, repeats, b["chains"], b["sweeps"])}


def cpu_baseline_aho():
    """Reference-order CPU restatement (oracle) of the anharmonic workload (no correlator)."""
    from feynsimul_amd.physical_systems import AnharmOsc
    system = AnharmOsc(AHO_LAM)
    return cpu_reference_rate(system, AHO_N, AHO_BETA, AHO_S, (system.energyOp,),
                              lambda rng, w: rng.uniform(-1, 1, (w, AHO_N)), (AHO_N // AHO_SEG) * (AHO_SEG - 1),
                              "anharmonic oscillator N=4096, beta=40, S=4, no correlator")


# ---- harmonic oscillator: the small plain-method workload ----
HO_N, HO_BETA, HO_W, HO_S = 256, 10.0, 4096, 4
HO_SEG = 1 << HO_S
HO_UPDATES_PER_SWEEP = HO_W * (HO_N // HO_SEG) * (HO_SEG - 1)
HO_PROFILE = os.path.join(ROOT, "profiles", "fs_wave_sweep_ho_n256.md")


def ho_exact_x2(N, beta):
    """<x^2> (= the virial energy) of the discretised periodic HO action with m = omega = 1: the path's
    normal modes a_k = (2 - 2 cos(2 pi k / N)) / eps + eps give <x^2> = (1/N) sum_k 1/a_k."""
    import numpy as np
    eps = beta / N
    a = (2.0 - 2.0 * np.cos(2.0 * np.pi * np.arange(N) / N)) / eps + eps
    return float(np.mean(1.0 / a))


def ho_bench(steps, warmup, cpu=True, ess=True):
    """1D harmonic oscillator, N = 256, beta = 10, 4096 chains, S = 4, xorshift128, paths uniform on [-1, 1],
    2000 burn-in sweeps, then steps of 32 sweeps + one measurement of the virial energy; ESS/s of E0 over
    256 measurements against the exact discretised value 0.499950082044."""
    import numpy as np
    import torch
    from feynsimul_amd.kernel_args import KernelArgs
    from feynsimul_amd.kernel import PIMCKernel
    from feynsimul_amd.physical_systems import HarmOsc
    from feynsimul_amd.stats import effective_sample_size
    system = HarmOsc()
    exact = ho_exact_x2(HO_N, HO_BETA)

    def kernel(steps_per_run, runs, seed):
        return PIMCKernel(KernelArgs(system=system, nbrOfWalkers=HO_W, N=HO_N, beta=HO_BETA, S=HO_S,
                                     enableOperator=True, enableCorrelator=False, enableBisection=True,
                                     enablePathShift=False, enableSingleNodeMove=False, enableParallelizePath=True,
                                     enableGlobalPath=False, enableGlobalOldPath=False, enableBins=False,
                                     metroStepsPerOperatorRun=steps_per_run, operatorRuns=runs,
                                     nbrOfWalkersPerWorkGroup=1, operators=(system.energyOp,), seed=seed,
                                     sampler="wave"))
    kb = kernel(500, 4, 1)
    kb.setPaths(np.random.default_rng(0).uniform(-1, 1, (HO_W, HO_N)).astype(np.float32))
    kb.run()
    paths = kb.getPaths()
    k = kernel(SWEEPS_PER_STEP, 1, 2)
    k.setPaths(paths)
    ms, wall = timed_launches(k, steps, warmup)
    updates = HO_UPDATES_PER_SWEEP * SWEEPS_PER_STEP
    bytes_upd = 8.0 * HO_SEG / (HO_SEG - 1)
    line = {"metric": "path-slice updates/s (plain bisection sampler)", "value": updates * steps / wall,
            "unit": "slice-updates/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic: paths uniform on [-1, 1], 2000 burn-in sweeps",
            "config": {"workload": "harmonic oscillator, N=256, beta=10, 4096 chains, S=4, xorshift128, 32 sweeps "
                                   "+ 1 measurement (virial energy) per step", "sampler": "wave"},
            "roofline": {"bound": "hbm", "achieved": bytes_upd * updates / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": bytes_upd * updates / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "traffic": None, "kernel": "fs_wave_sweep", "avg_launch_ms": ms,
                         "bytes_per_update": bytes_upd, "updates_per_launch": updates,
                         "note": "the 4 MiB of paths stay in registers for the whole launch; HBM is not a bound"},
            "valu": dict(valu_fields(32 * updates / (ms * 1e-3), HO_PROFILE), ops_per_update=32),
            "acceptance": k.getAcceptanceRate()}
    if ess:
        line["ess_at_fixed_sigma"] = fixed_sigma("ho", HO_SIGMA)
    if cpu and int(os.environ.get("RANK", "0")) == 0:
        line["cpu_baseline"] = cpu_baseline_ho()
    return line


def cpu_baseline_ho():
    """Reference-order CPU restatement (oracle) of the harmonic-oscillator workload."""
    from feynsimul_amd.physical_systems import HarmOsc
    system = HarmOsc()
    return cpu_reference_rate(system, HO_N, HO_BETA, HO_S, (system.energyOp,),
                              lambda rng, w: rng.uniform(-1, 1, (w, HO_N)), (HO_N // HO_SEG) * (HO_SEG - 1),
                              "harmonic oscillator N=256, beta=10, S=4")


# ---- random streams: the reference's RNG benchmark ----
# examples/benchmark_xorshift_vs_ranlux.py:12-26: seconds for 1e7 uniforms per thread on
# 448 * 32 * 4 = 57,344 threads (single precision; RANLUX also double), hardware unstated
REF_RNG_THREADS, REF_RNG_DRAWS = 448 * 32 * 4, 10_000_000
REF_RNG_SECONDS = {("xorshift", None, "f32"): 7.748644,
                   ("ranlux", 0, "f32"): 12.359268, ("ranlux", 0, "f64"): 23.163764,
                   ("ranlux", 1, "f32"): 18.334329, ("ranlux", 1, "f64"): 34.851969,
                   ("ranlux", 2, "f32"): 36.360428, ("ranlux", 2, "f64"): 71.408121,
                   ("ranlux", 3, "f32"): 67.230933, ("ranlux", 3, "f64"): 133.259846,
                   ("ranlux", 4, "f32"): 112.292905, ("ranlux", 4, "f64"): 224.170527}


def rng_bench(streams=REF_RNG_THREADS, draws=None):
    """Uniforms/s of every generator the reference times, through the samplers' wrappers
    (fs_rng_bench), on the reference's 57,344 streams and on a device-filling 1,048,576."""
    import numpy as np
    import torch
    from feynsimul_amd import _native
    L = _native.lib()
    dev = torch.device("cuda", 0)
    rows_out = []

    def run(kind, lux, dbl, n, count):
        if kind == "xorshift":
            st = torch.zeros((n, 4), dtype=torch.int32, device=dev)
            _native.check(L.fs_seed_xorshift(st.data_ptr(), n, 2026, 0, None))
        else:
            seeds = torch.arange(1, n + 1, dtype=torch.int64, device=dev) * 0x9E3779B97F4A7C15 % (1 << 62)
            st = torch.zeros((n, 28), dtype=torch.int32, device=dev)
            _native.check(L.fs_ranlux_init(st.data_ptr(), seeds.data_ptr(), n, None))
        chk = torch.zeros(n, dtype=torch.int32, device=dev)
        k = _native.FS_RNG_XORSHIFT if kind == "xorshift" else _native.FS_RNG_RANLUX
        _native.check(L.fs_rng_bench(k, lux or 0, int(dbl), st.data_ptr(), chk.data_ptr(), n, 1024, None))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _native.check(L.fs_rng_bench(k, lux or 0, int(dbl), st.data_ptr(), chk.data_ptr(), n, count, None))
        e1.record()
        torch.cuda.synchronize()
        return n * count / (e0.elapsed_time(e1) * 1e-3)

    for (kind, lux, prec), ref_s in REF_RNG_SECONDS.items():
        ref = REF_RNG_THREADS * REF_RNG_DRAWS / ref_s
        # draws per stream sized to ~0.2-2 s per launch
        count = draws or (REF_RNG_DRAWS if kind == "xorshift" else REF_RNG_DRAWS // 10)
        v_ref_shape = run(kind, lux, prec == "f64", streams, count)
        v_full = run(kind, lux, prec == "f64", 1 << 20, max(count // 16, 1024))
        rows_out.append({"generator": kind if lux is None else "ranlux luxury %d" % lux, "precision": prec,
                         "uniforms_per_s": v_ref_shape, "uniforms_per_s_1M_streams": v_full,
                         "reference_uniforms_per_s": ref, "vs_reference": v_ref_shape / ref,
                         "vs_reference_1M_streams": v_full / ref})
    head = rows_out[0]
    return {"metric": "uniforms/s, xorshift128 FP32 through the sampler's wrapper (the reference's RNG benchmark)",
            "value": head["uniforms_per_s"], "unit": "uniforms/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": REF_RNG_THREADS * REF_RNG_DRAWS / head["uniforms_per_s"] * 1e3,
            "higher_is_better": True, "scaling": "weak",
            "vs_baseline": head["vs_reference"], "dtype": "u32/f32", "data": "synthetic: splitmix64 / ranluxcl_init seeds",
            "config": {"workload": "57,344 streams x 1e7 uniforms (RANLUX: 1e6), XOR checksum per stream, "
                                   "as examples/benchmark_xorshift_vs_ranlux.py times it",
                       "reference_hardware": "unstated in the reference"},
            "generators": rows_out}


def cpu_baseline_ragged(kas):
    """Reference-order CPU restatement (oracle) of every ragged type (its own S), a shorter sample per
    type; the batch's rate is its total slice-updates over the sum of the types' CPU times."""
    import numpy as np
    per, total_upd, total_s = [], 0.0, 0.0
    for ka in kas:
        system = ka.system
        seg = 1 << ka.S
        rng0 = np.random.default_rng(7)
        r = cpu_reference_rate(system, ka.N, ka.beta, ka.S, (system.energyOp,),
                               lambda rng, w, s_=system, n_=ka.N: np.repeat(
                                   s_.initialPositions(w, rng0, side=3.5, sigma=3.0)[:, :, None], n_,
                                   axis=2).reshape(w, s_.DOF * n_),
                               (ka.N // seg) * (seg - 1), "", run_s=0.3, repeats=3)
        upd = ka.nbrOfWalkers * (ka.N // seg) * (seg - 1) * SWEEPS_PER_STEP
        total_upd += upd
        total_s += upd / r["value"]
        per.append({"atoms": system.nAtoms, "N": ka.N, "S": ka.S, "value": r["value"], "variant": r["variant"],
                    "cores": r["cores"]})
    aff, _ = cpu_thread_counts()
    return {"value": total_upd / total_s, "unit": "slice-updates/s", "cores": max(p_["cores"] for p_ in per),
            "affinity_cores": aff, "kind": "port", "per_type": per,
            "sample": "each of the 16 types: median of 3 runs of about 0.3 s, reference work-item order, FP32, "
                      "xorshift128, gcc -O3 -march=native -fopenmp, strict and -ffast-math builds (the faster "
                      "kept per type); the batch rate is its slice-updates over the summed per-type CPU time"}


def cpu_baseline():
    """Reference-order CPU restatement (oracle) of the He3 workload on this host's cores."""
    from feynsimul_amd.physical_systems import Lm2m2Cluster
    system = Lm2m2Cluster(3)
    return cpu_reference_rate(system, N, BETA, S, (system.energyOp, system.meanPairDistanceOp),
                              lambda rng, w: initial_paths(system, rng, w), (N // SEG) * (SEG - 1),
                              "He3 LM2M2 N=1024, beta=10/K, S=6, energy and mean pair distance")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ess", action="store_true", help="skip the ESS/s phase")
    ap.add_argument("--workload", choices=["he3", "ragged", "aho", "rng", "ho", "lp", "lpc"], default="he3",
                    help="he3: the headline workload; ragged: LM2M2 n in {3,4,6,8} x N in {256..2048}, "
                         "2048 chains each, in one autotuned persistent launch; aho: anharmonic oscillator "
                         "N=4096, beta=40, 65536 chains, 512 correlator lags, xorshift128 and RANLUX; "
                         "rng: stream throughput of every generator, as the reference's RNG benchmark; "
                         "ho: harmonic oscillator N=256, beta=10, 4096 chains; lp: the learned sweep alone on He3; "
                         "lpc: the learned whole-path move alone on the anharmonic oscillator")
    ap.add_argument("--ess-burn", type=int, default=24, help="burn-in launches of 512 sweeps before the E0 series")
    ap.add_argument("--ess-launches", type=int, default=64, help="E0 series launches of 16 x 32 sweeps")
    ap.add_argument("--aho-sigma", type=float, default=AHO_SIGMA, help="standard error target of E0 on the anharmonic "
                                                                   "oscillator in the fixed-sigma comparison")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal phase")
    ap.add_argument("--fixed-sigma", action="store_true",
                    help="only the ESS/s comparison at fixed sigma (harmonic and anharmonic oscillators)")
    ap.add_argument("--learned-train", type=int, default=200, help="online Adam steps before learned sampling")
    ap.add_argument("--learned-measurements", type=int, default=32, help="learned measurements (32 sweeps each)")
    ap.add_argument("--ragged-levels", default=None,
                    help="comma-separated S of the 16 ragged types (skips the level autotuner)")
    ap.add_argument("--ragged-shape", default=None,
                    help="waves per workgroup, order, kernels (e.g. 4,cost,one): skips the launch-shape autotuner")
    ap.add_argument("--no-separate", action="store_true",
                    help="ragged: skip timing each type alone (at 2048 chains and at full occupancy)")
    ap.add_argument("--lp-spw", type=int, default=0, help="segments per wavefront of the learned sweep (0: automatic)")
    ap.add_argument("--hbm-ceiling", action="store_true",
                    help="measure the HBM ceilings (copy, read-only, write-only streams of 2 GiB) and exit")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="run only the xorshift128 stream kernel on 1,048,576 streams (3 launches), the VALU "
                         "issue-rate reference for rocprofv3 SQ counters")
    ap.add_argument("--walker-scaling", action="store_true",
                    help="time the plain sampler of the He3 and anharmonic workloads at several chain counts, "
                         "including counts whose paths fit in the Infinity Cache, and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 1 GiB (5 times), to scale HBM counters")
    ap.add_argument("--waves-per-simd", type=int, default=None,
                    help="occupancy cap of the He3 launches (wavefronts per SIMD; default: the kernel's own)")
    ap.add_argument("--ranlux", action="store_true", help="He3 workload with RANLUX luxury 2 streams instead of xorshift128")
    ap.add_argument("--storage", choices=["registers", "memory"], default=None,
                    help="where the wave sampler keeps each path during a launch (default: automatic)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: chains are independent and this engine runs on a single device; "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    torch.cuda.set_device(0)
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(5)]
        for a, b in ev:
            a.record()
            dst.copy_(src)
            b.record()
        torch.cuda.synchronize()
        ms = sorted(a.elapsed_time(b) for a, b in ev)[2]
        # a copy reads and writes each byte once
        print(json.dumps({"copy_bytes_per_launch": src.numel() * 4, "launches": 5, "median_ms": ms,
                          "hbm_GBps": 2 * src.numel() * 4 / (ms * 1e-3) / 1e9}), flush=True)
        return
    if args.hbm_ceiling:
        print(json.dumps(hbm_ceiling()), flush=True)
        return
    if args.walker_scaling:
        print(json.dumps(walker_scaling()), flush=True)
        return
    if args.fixed_sigma:
        ceil = hbm_ceiling()
        progress("ESS/s at fixed sigma: harmonic oscillator")
        fs = {"ho": fixed_sigma("ho", HO_SIGMA)}
        progress("ESS/s at fixed sigma: anharmonic oscillator, S=6")
        fs["aho_S6"] = fixed_sigma("aho", args.aho_sigma, S=6, copy_GBps=ceil["ceilings"]["copy"]["GBps"])
        print(json.dumps({"ess_at_fixed_sigma": fs}), flush=True)
        return
    if args.valu_calibration:
        from feynsimul_amd import _native
        L = _native.lib()
        n, count = 1 << 20, 1 << 16
        st = torch.zeros((n, 4), dtype=torch.int32, device="cuda")
        chk = torch.zeros(n, dtype=torch.int32, device="cuda")
        _native.check(L.fs_seed_xorshift(st.data_ptr(), n, 2026, 0, None))
        for _ in range(3):
            _native.check(L.fs_rng_bench(_native.FS_RNG_XORSHIFT, 0, 0, st.data_ptr(), chk.data_ptr(), n, count, None))
        torch.cuda.synchronize()
        print(json.dumps({"streams": n, "uniforms_per_stream": count, "launches": 3}), flush=True)
        return
    if args.workload == "ragged":
        lv = [int(v) for v in args.ragged_levels.split(",")] if args.ragged_levels else None
        shape = None
        if args.ragged_shape:
            wpb, order, kern = args.ragged_shape.split(",")
            shape = {"wavesPerBlock": int(wpb), "order": order, "kernels": kern, "kernelWaves": None}
        print(json.dumps(ragged_bench(args.steps, args.warmup, cpu=not args.no_cpu, levels=lv, shape=shape,
                                      separate=not args.no_separate)), flush=True)
        return
    if args.workload == "ho":
        print(json.dumps(ho_bench(args.steps, args.warmup, cpu=not args.no_cpu, ess=not args.no_ess)), flush=True)
        return
    if args.workload == "lp":
        print(json.dumps(lp_bench(args.steps, args.warmup, args.lp_spw)), flush=True)
        return
    if args.workload == "lpc":
        print(json.dumps(lpc_bench(args.steps, args.warmup)), flush=True)
        return
    if args.workload == "rng":
        print(json.dumps(rng_bench()), flush=True)
        return
    if args.workload == "aho":
        print(json.dumps(aho_bench(args.steps, args.warmup, cpu=not args.no_cpu, ess=not args.no_ess)), flush=True)
        return
    # untimed check against the FP64 CPU reference: 23 launches (736 sweeps) from the initial paths the
    # reference run starts from, whatever --warmup and --steps are
    progress("oracle leg: %d launches from the FP64 reference's initial paths" % ORACLE_LAUNCHES)
    system, ko = make_kernel(SWEEPS_PER_STEP, seed=12345, storage=args.storage, ranlux=args.ranlux,
                             waves_per_simd=args.waves_per_simd)
    ko.setPaths(initial_paths(system, np.random.default_rng(0)))
    for _ in range(ORACLE_LAUNCHES):
        ko.run(wait=False)
    torch.cuda.synchronize()
    oracle = he3_oracle(ko.getOperators(), ORACLE_LAUNCHES * SWEEPS_PER_STEP)
    del ko
    torch.cuda.empty_cache()
    progress("timed: %d warmup + %d steps" % (args.warmup, args.steps))
    system, k = make_kernel(SWEEPS_PER_STEP, seed=12345, storage=args.storage, ranlux=args.ranlux,
                            waves_per_simd=args.waves_per_simd)
    rng = np.random.default_rng(0)
    k.setPaths(initial_paths(system, rng))
    for _ in range(args.warmup):
        k.run()
    stream = k._stream
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        k.run(wait=False)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    kern_s = float(np.mean(launch_ms)) * 1e-3
    updates = UPDATES_PER_SWEEP * SWEEPS_PER_STEP
    value = updates * args.steps / wall
    achieved = BYTES_PER_UPDATE * updates / kern_s / 1e9
    per_walker = k.getOperators()
    E = float(per_walker[0].mean())
    line = {
        "metric": "path-slice updates/s (plain bisection sampler)",
        "value": value,
        "unit": "slice-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded Gaussian blobs (sigma 3 A) around a random triangle of side 3.5 A",
        "config": {"workload": "LM2M2 He3, 3D Cartesian, N=1024, beta=10/K, 16384 chains, S=6, %s, "
                               % ("RANLUX luxury 2" if args.ranlux else "xorshift128") +
                               "32 sweeps + 1 measurement (energy, mean pair distance, 256-bin pair-distance "
                               "histogram) per step", "sampler": "wave",
                   "path_storage": args.storage or "auto",
                   "chains": W, "slices": N, "bisection_level": S},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic() if args.storage in (None, "memory") else None,
                     "kernel": "fs_wave_sweep", "avg_launch_ms": kern_s * 1e3,
                     "bytes_per_update": BYTES_PER_UPDATE, "updates_per_launch": updates,
                     "limiter": "not HBM: with the paths in the Infinity Cache (4096-5120 chains) the rate is "
                                "0.92-0.95 of this one and saturates with wavefronts per SIMD; the SIMD's issue of "
                                "the block loop's instruction mix bounds it (VALU active 0.59 of the calibration "
                                "kernel's), and its HBM traffic runs level with the measured copy ceiling",
                     "limiter_source": "profiles/fs_wave_sweep_walker_scaling.md (bench.py --walker-scaling)"},
        "valu": dict(valu_fields(VALU_PER_UPDATE * updates / kern_s, TRAFFIC_PROFILE), ops_per_update=VALU_PER_UPDATE),
        "acceptance": k.getAcceptanceRate(),
        "energy_K": E,
        "oracle": oracle,
    }
    del k
    torch.cuda.empty_cache()
    # the kernel's measured HBM traffic rate against the box's measured stream ceilings
    progress("HBM ceilings")
    ceil = hbm_ceiling()
    line["hbm_ceiling"] = ceil
    if line["roofline"]["traffic"]:
        rate = line["roofline"]["traffic"] / kern_s / 1e9
        line["roofline"]["traffic_GBps"] = rate
        line["roofline"]["frac_of_measured_copy_ceiling"] = rate / ceil["ceilings"]["copy"]["GBps"]
    if not args.no_ess:
        progress("He3 E0 series (stationarity)")
        line["ess"] = ess_phase(initial_paths(system, np.random.default_rng(0)), args.storage, args.ess_burn,
                                args.ess_launches)
        if not args.no_learned:
            progress("He3 learned proposal")
            line["learned"] = learned_phase(ess_phase.start_paths, args.learned_train, args.learned_measurements)
        progress("ESS/s at fixed sigma: harmonic oscillator")
        # the harmonic paths (4 MB) stay in cache, so the HBM copy ceiling bounds only the anharmonic case
        fs = {"ho": fixed_sigma("ho", HO_SIGMA)}
        progress("ESS/s at fixed sigma: anharmonic oscillator, S=6")
        fs["aho_S6"] = fixed_sigma("aho", args.aho_sigma, S=6, copy_GBps=ceil["ceilings"]["copy"]["GBps"])
        line["ess_at_fixed_sigma"] = fs
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        progress("CPU baseline")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)
    reap_children()


if __name__ == "__main__":
    main()
