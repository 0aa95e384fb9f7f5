"""feynsimul_amd -- path-integral Monte Carlo (Levy-bisection Metropolis) on one AMD MI355X.

Drop-in for the FeynSimul API: ``kernel_args.KernelArgs``, ``kernel.PIMCKernel``
and ``physical_systems``; the compute path is libfeynsimul.so (hand-written
gfx950 HIP compiled with hiprtc per system), reached through ctypes.
"""
__version__ = "0.1.0"


_ALIASED = ("kernel", "kernel_args", "pimc_utils", "physical_systems", "physical_systems.harm_osc",
            "physical_systems.lm2m2_3part", "physical_systems.lm2m2_4part", "physical_systems.lm2m2_cluster")


def install_feynsimul_alias():
    """Registers this package as ``FeynSimul`` (and its modules under the reference's module
    names), so that scripts written for the reference run unchanged, e.g.
    ``from FeynSimul.pimc_utils import *`` and
    ``from FeynSimul.physical_systems.lm2m2_3part import *``."""
    import importlib
    import sys
    sys.modules["FeynSimul"] = sys.modules[__name__]
    for name in _ALIASED:
        sys.modules["FeynSimul." + name] = importlib.import_module(__name__ + "." + name)
