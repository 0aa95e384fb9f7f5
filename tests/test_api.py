"""Host-side boundary: KernelArgs / PIMCKernel validation and the C ABI (CPU only).

Validation branches mirror FeynSimul/kernel.py:93-198 (exception types and
messages); they are all raised before any device work, so they run here.
The C-ABI tests load libfeynsimul.so, check every declared export and compile
sampler modules for gfx950 with hiprtc (no GPU needed to compile).
"""
import inspect
import os
import re

import pytest

from feynsimul_amd.kernel_args import KernelArgs
from feynsimul_amd.kernel import PIMCKernel, humanReadableSize
from feynsimul_amd.physical_systems import HarmOsc, Lm2m2Cluster, AnharmOsc
from feynsimul_amd.system_source import system_section

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ka(**kw):
    s = HarmOsc()
    base = dict(system=s, nbrOfWalkers=64, N=64, beta=5.0, S=3, enableOperator=True, enableCorrelator=False,
                enableBisection=True, enablePathShift=False, enableSingleNodeMove=False,
                enableParallelizePath=True, enableGlobalPath=False, enableGlobalOldPath=False, enableBins=False,
                metroStepsPerOperatorRun=2, operatorRuns=2, nbrOfWalkersPerWorkGroup=1,
                operators=(s.energyOp,))
    base.update(kw)
    return KernelArgs(**base)


def test_kernel_args_signature_matches_reference():
    params = list(inspect.signature(KernelArgs.__init__).parameters)[1:]
    ref = ["system", "nbrOfWalkers", "N", "beta", "enableOperator", "enableCorrelator", "enableBisection",
           "enablePathShift", "enableSingleNodeMove", "enableBins", "nbrOfWalkersPerWorkGroup", "enableDouble",
           "enableRanlux", "luxuaryFactor", "ranluxIntMax", "S", "alpha", "PSAlpha", "enableParallelizePath",
           "metroStepsPerOperatorRun", "enableGlobalPath", "enableGlobalOldPath", "operators", "operatorRuns",
           "correlators", "xMin", "xMax", "binResolutionPerDOF"]
    assert params[: len(ref)] == ref
    ka = _ka()
    ka.enableBins = True  # mutable bag, as the tutorial uses it
    assert ka.enableBins


@pytest.mark.parametrize("kw,exc,msg", [
    (dict(beta=5), TypeError, "beta need to be a float"),
    (dict(operatorRuns=None), NameError, "operatorRuns need to be set"),
    (dict(operators=None), NameError, "operators need to be set"),
    (dict(metroStepsPerOperatorRun=None), NameError, "metroStepsPerOperatorRun need to be set"),
    (dict(operators=["x1"]), TypeError, "operators need to be a tuple"),
    (dict(enableCorrelator=True), NameError, "correlators need to be set"),
    (dict(enableCorrelator=True, correlators="x1"), TypeError, "correlators need to be a tuple"),
    (dict(enableSingleNodeMove=True, enableParallelizePath=False), NameError, "alpha need to be set"),
    (dict(enablePathShift=True, enableParallelizePath=False), NameError, "PSALpha need to be set"),
    (dict(enableRanlux=True), NameError, "luxuaryFactor needs to be an uint"),
    (dict(enableRanlux=True, luxuaryFactor="2"), NameError, "luxuaryFactor"),
    (dict(enableRanlux=True, luxuaryFactor=2), NameError, "ranluxIntMax needs to be an uint"),
    (dict(enableRanlux=True, luxuaryFactor=4, ranluxIntMax=10), Exception, "nskip is negative"),
    (dict(enableBins=True), NameError, "xMin need to be set"),
    (dict(enableBins=True, xMin=-1.0), NameError, "xMax need to be set"),
    (dict(enableBins=True, xMin=-1.0, xMax=1.0), NameError, "binResolutionPerDOF need to be set"),
    (dict(nbrOfWalkers=63, nbrOfWalkersPerWorkGroup=4), Exception, "multiple"),
    (dict(nbrOfWalkers=2, nbrOfWalkersPerWorkGroup=4), Exception, "multiple"),
    (dict(N=48), Exception, "power of 2"),
    (dict(S=None), NameError, "S need to be set"),
    (dict(S=7), Exception, "2^S must be less than or equal to N"),
    (dict(enableGlobalPath=None), NameError, "enableGlobalPath need to be set"),
    (dict(enableGlobalOldPath=None), NameError, "enableGlobalOldPath need to be set"),
    (dict(enableParallelizePath=None), NameError, "enableParallelizePath need to be set"),
    (dict(enableSingleNodeMove=True, alpha=0.5), Exception, "Single node move not possible"),
    (dict(enablePathShift=True, PSAlpha=0.5), Exception, "Shift path not possible"),
    (dict(enableBisection=False, enableGlobalOldPath=True), Exception, "Old Path is not used"),
])
def test_validation_matches_reference(kw, exc, msg):
    with pytest.raises(exc, match=re.escape(msg)):
        PIMCKernel(_ka(**kw))


def test_no_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="needs a GPU"):
        PIMCKernel(_ka())


def test_human_readable_size():
    assert humanReadableSize(592) == "592 B"
    assert humanReadableSize(1024 ** 3) == "1 GiB"


# ----------------------------------------------------------------------------
# C ABI
# ----------------------------------------------------------------------------
def _declared_functions():
    with open(os.path.join(ROOT, "include", "feynsimul.h")) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fs_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from feynsimul_amd import _native
    L = _native.lib()
    declared = _declared_functions()
    assert set(declared) == set(_native.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    assert L.fs_abi_version() == _native.FS_ABI_VERSION


def _cfg(system, N, S, sampler, nops=1, ncorr=0, **kw):
    from feynsimul_amd import _native
    c = _native.FsConfig()
    c.abi_version, c.sampler, c.N, c.S, c.DOF = _native.FS_ABI_VERSION, sampler, N, S, system.DOF
    c.walkers_per_group, c.n_walkers, c.parallelize_path, c.local_size = 1, 64, 1, 1
    c.fast_math, c.ranlux_int_range, c.n_ops, c.n_corr, c.cor_lags = 1, 2 ** 31, nops, ncorr, N // 2
    c.bin_res, c.beta, c.x_max, c.op_norm = 1, 10.0, 1.0, 1.0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("name", ["ho", "aho", "lm2m2-4"])
def test_module_compiles_for_gfx950(name, sampler):
    from feynsimul_amd import _native
    system = {"ho": HarmOsc(), "aho": AnharmOsc(0.1), "lm2m2-4": Lm2m2Cluster(4)}[name]
    ops = (system.energyOp,)
    corr = ("x1",) if name != "lm2m2-4" else ()
    e = _native.Engine(_cfg(system, 256, 4, sampler, ncorr=len(corr), bins=int(name == "ho"), bin_res=8,
                            x_min=-2.0, x_max=2.0),
                       system_section(system, ops, corr), compile_only=True)
    assert "fs_sampler.h" in e.source
    e.close()


@pytest.mark.parametrize("min_waves", [None, "5"])
@pytest.mark.parametrize("double", [0, 1])
@pytest.mark.parametrize("lux", [0, 2])
def test_ranlux_wave_module_compiles_for_gfx950(lux, double, min_waves, monkeypatch):
    """The wave sampler with RANLUX streams (per-lane streams drawn a block at a time, the group stream's state in
    LDS) builds for both precisions and luxury levels, also with the FS_MIN_WAVES register budget."""
    from feynsimul_amd import _native
    if min_waves:
        monkeypatch.setenv("FS_MIN_WAVES", min_waves)
    else:
        monkeypatch.delenv("FS_MIN_WAVES", raising=False)
    system = Lm2m2Cluster(3)
    e = _native.Engine(_cfg(system, 256, 4, 1, path_storage=1, ranlux=1, ranlux_lux=lux,
                            double_precision=double),
                       system_section(system, (system.energyOp,)), compile_only=True)
    assert "#define FS_RANLUX 1" in e.source
    e.close()


def test_compile_error_reports_build_log():
    from feynsimul_amd import _native

    class Broken:
        DOF = 1
        potential = "x1 +* 2"
        userCode = ""
    with pytest.raises(_native.FsError) as ei:
        _native.Engine(_cfg(Broken(), 64, 3, 1), system_section(Broken(), ("x1",)), compile_only=True)
    assert ei.value.code == _native.FS_ERR_COMPILE
    assert "error" in str(ei.value)


@pytest.mark.parametrize("field,value,msg", [
    ("N", 48, "power of 2"),
    ("S", 7, "2^S"),
    ("abi_version", 99, "ABI version"),
    ("sampler", 5, "unknown sampler"),
    ("cor_lags", 0, "cor_lags"),
])
def test_invalid_config_rejected(field, value, msg):
    from feynsimul_amd import _native
    c = _cfg(HarmOsc(), 64, 3, 1, ncorr=1)
    setattr(c, field, value)
    with pytest.raises(_native.FsError, match=re.escape(msg)) as ei:
        _native.Engine(c, system_section(HarmOsc(), ("x1",), ("x1",)), compile_only=True)
    assert ei.value.code == _native.FS_ERR_INVALID


def test_device_entry_points_need_a_gpu():
    import torch
    from feynsimul_amd import _native
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    assert _native.device_count() == 0
    rc = _native.lib().fs_seed_xorshift(8, 1, 0, 0, None)  # any non-null pointer; rejected before use
    assert rc == _native.FS_ERR_NO_DEVICE


def _lp(**kw):
    from feynsimul_amd import _native
    lp = _native.FsLpConfig()
    lp.abi_version, lp.atoms3d, lp.batch, lp.in_scale = _native.FS_ABI_VERSION, 0, 65536, 1.0
    lp.lr, lp.beta1, lp.beta2, lp.adam_eps, lp.seed = 1e-3, 0.9, 0.999, 1e-8, 1
    for k, v in kw.items():
        setattr(lp, k, v)
    return lp


@pytest.mark.parametrize("name,N,S,spw", [("ho", 256, 4, 16), ("aho", 4096, 4, 32), ("lm2m2-3", 1024, 6, 4)])
def test_learned_module_compiles_for_gfx950(name, N, S, spw):
    """The learned-proposal module (MFMA sweep, training GEMMs, Adam) builds for each system;
    eight wavefronts per block share the network image, and segments per wavefront follow the 160 KiB LDS budget
    (82 KiB network image + eight wavefronts' staging)."""
    from feynsimul_amd import _native
    system = {"ho": HarmOsc(), "aho": AnharmOsc(0.1), "lm2m2-3": Lm2m2Cluster(3)}[name]
    c = _cfg(system, N, S, 1, path_storage=1)
    p = _native.Proposal(c, system_section(system, (system.energyOp,)),
                         _lp(atoms3d=int(name.startswith("lm2m2"))), compile_only=True)
    info = p.info
    assert info.segs_per_wave == spw and info.waves_per_block == 8
    assert info.lds_bytes <= 163840
    assert info.n_params == 128 * (2 * system.DOF + 1) + 128 + 2 * (128 * 128 + 128) + 2 * system.DOF * 129
    if name == "lm2m2-3":
        assert info.n_params == 37906
    p.close()


@pytest.mark.parametrize("field,value,msg", [
    ("batch", 1000, "multiple of 2048"),
    ("atoms3d", 1, "atoms3d needs DOF"),
    ("abi_version", 7, "ABI version"),
    ("bridge_omega2", -1.0, "bridge_omega2 must be finite and >= 0"),
    ("bridge_omega2", float("nan"), "bridge_omega2 must be finite and >= 0"),
    ("bridge_omega2", 1.0, "bridge_omega2 applies to the whole-path move"),
])
def test_learned_config_rejected(field, value, msg):
    from feynsimul_amd import _native
    with pytest.raises(_native.FsError, match=re.escape(msg)):
        _native.Proposal(_cfg(HarmOsc(), 64, 3, 1), system_section(HarmOsc(), ("x1",)), _lp(**{field: value}),
                         compile_only=True)


def test_move_measurement_entry_checks_its_arguments():
    """fs_lp_sweep_measure needs a proposal that was built for the device and the kernel's buffers (what it
    refuses on the device: the GPU tests of the whole-path move)."""
    import ctypes
    from feynsimul_amd import _native
    L = _native.lib()
    ho = HarmOsc()
    p = _native.Proposal(_cfg(ho, 256, 6, 1), system_section(ho, (ho.energyOp,)), _lp(coarse=1), compile_only=True)
    b, kb = _native.FsLpBuffers(), _native.FsBuffers()
    assert L.fs_lp_sweep_measure(p._h, 8, 8, ctypes.byref(b), None, 0, 1, 0, None) != 0
    assert "kernel's buffers are required" in L.fs_last_error().decode()
    assert L.fs_lp_sweep_measure(p._h, 8, 8, ctypes.byref(b), ctypes.byref(kb), 0, 1, 0, None) != 0
    assert "compile-only" in L.fs_last_error().decode()
    p.close()


def test_learned_rejects_double_and_deep_levels():
    from feynsimul_amd import _native
    with pytest.raises(_native.FsError, match="FP32"):
        _native.Proposal(_cfg(HarmOsc(), 64, 3, 1, double_precision=1), system_section(HarmOsc(), ("x1",)), _lp(),
                         compile_only=True)
    with pytest.raises(_native.FsError, match="S <= 6"):
        _native.Proposal(_cfg(HarmOsc(), 256, 7, 1), system_section(HarmOsc(), ("x1",)), _lp(), compile_only=True)


@pytest.mark.parametrize("name", ["ho", "lm2m2-3"])
def test_exact_module_with_shift_and_single_node_compiles(name):
    """Path-shift and single-node moves (kernel.c:407-473, 640-705), with and without bisection."""
    from feynsimul_amd import _native
    system = {"ho": HarmOsc(), "lm2m2-3": Lm2m2Cluster(3)}[name]
    for nob in (0, 1):
        c = _cfg(system, 128, 3, 0, parallelize_path=0, path_shift=1, single_node=1, alpha=0.5, ps_alpha=0.3,
                 no_bisection=nob)
        e = _native.Engine(c, system_section(system, (system.energyOp,)), compile_only=True)
        assert "#define FS_PATH_SHIFT 1" in e.source and "#define FS_BISECTION %d" % (1 - nob) in e.source
        e.close()


@pytest.mark.parametrize("kw,msg", [
    (dict(sampler=1, path_shift=1), "exact sampler"),
    (dict(sampler=0, parallelize_path=1, single_node=1), "exact sampler"),
    (dict(sampler=1, no_bisection=1), "bisection"),
])
def test_move_configs_rejected(kw, msg):
    from feynsimul_amd import _native
    c = _cfg(HarmOsc(), 64, 3, kw.pop("sampler"), **kw)
    with pytest.raises(_native.FsError, match=msg):
        _native.Engine(c, system_section(HarmOsc(), ("x1",)), compile_only=True)


def test_kernel_args_moves_select_the_exact_sampler():
    """kernel.py:193-198: the moves need enableParallelizePath = False; the engine then runs
    one work-item per walker (the exact sampler)."""
    from feynsimul_amd.kernel_args import KernelArgs
    from feynsimul_amd.kernel import PIMCKernel
    base = dict(system=HarmOsc(), nbrOfWalkers=64, N=64, beta=2.0, enableOperator=True, enableCorrelator=False,
                enableBisection=True, enableBins=False, nbrOfWalkersPerWorkGroup=1, S=3, operators=("x1",),
                operatorRuns=1, metroStepsPerOperatorRun=1, enableGlobalPath=False, enableGlobalOldPath=False)
    with pytest.raises(Exception, match="Shift path not possible"):
        PIMCKernel(KernelArgs(enablePathShift=True, enableSingleNodeMove=False, PSAlpha=0.1,
                              enableParallelizePath=True, **base))
    with pytest.raises(NameError, match="alpha need to be set"):
        PIMCKernel(KernelArgs(enablePathShift=False, enableSingleNodeMove=True, enableParallelizePath=False,
                              **base))
    with pytest.raises(Exception, match="wave sampler runs bisection moves only"):
        PIMCKernel(KernelArgs(enablePathShift=True, enableSingleNodeMove=False, PSAlpha=0.1,
                              enableParallelizePath=False, sampler="wave", **base))


def test_occupancy_cap_range():
    """fs_set_waves_per_simd and max_waves_per_simd accept 0..16 (0 = no cap)."""
    from feynsimul_amd import _native
    e = _native.Engine(_cfg(HarmOsc(), 64, 3, 1), system_section(HarmOsc(), ("x1",)), compile_only=True)
    e.set_waves_per_simd(5)
    e.set_waves_per_simd(0)
    for bad in (-1, 17):
        with pytest.raises(_native.FsError, match="waves per SIMD"):
            e.set_waves_per_simd(bad)
    e.close()
    with pytest.raises(_native.FsError, match="max_waves_per_simd"):
        _native.Engine(_cfg(HarmOsc(), 64, 3, 1, max_waves_per_simd=20), system_section(HarmOsc(), ("x1",)),
                       compile_only=True)


@pytest.mark.parametrize("N, D, width, needle", [(96, 1, 32, "power of 2"), (64, 1, 32, "N = 64 x 2^k"),
                                                 (65536, 9, 32, "coarse values per walker"),
                                                 (256, 1, 48, "multiple of 32")])
def test_whole_path_move_rejects_bad_shapes(N, D, width, needle):
    """The whole-path move refuses, before compiling, paths that are not 64 x 2^k slices, coarse node sets
    that would not fit the fine pass's LDS, and an active width that is not a multiple of 32."""
    from feynsimul_amd import _native
    system = HarmOsc() if D == 1 else Lm2m2Cluster(3)
    c = _cfg(system, N, 4, _native.FS_SAMPLER_WAVE)
    lp = _native.FsLpConfig()
    lp.abi_version, lp.atoms3d, lp.batch, lp.in_scale = _native.FS_ABI_VERSION, int(D == 9), 2048, 1.0
    lp.lr, lp.beta1, lp.beta2, lp.adam_eps, lp.seed, lp.coarse, lp.width = 1e-3, 0.9, 0.999, 1e-8, 1, 1, width
    with pytest.raises(Exception, match=needle.replace("^", r"\^")):
        _native.Proposal(c, "", lp, compile_only=True)
