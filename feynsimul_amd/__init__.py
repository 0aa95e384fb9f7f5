"""feynsimul_amd -- path-integral Monte Carlo (Levy-bisection Metropolis) on one AMD MI355X.

Drop-in for the FeynSimul API: ``kernel_args.KernelArgs``, ``kernel.PIMCKernel``
and ``physical_systems``; the compute path is libfeynsimul.so (hand-written
gfx950 HIP compiled with hiprtc per system), reached through ctypes.
"""
__version__ = "0.1.0"
