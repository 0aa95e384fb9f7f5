"""Physical systems for the PIMC engine.

A system is any object with

* ``potential`` -- a C expression (OpenCL C / HIP C++ common subset) in the
  slice coordinates ``x1 .. xDOF``; it may use ``FLOAT_TYPE``, the ``eps``
  macro (time step beta/N) and helpers defined in ``userCode``;
* ``DOF`` -- the number of coordinates per time slice;
* ``userCode`` -- C function definitions the expressions may call;

plus operator strings in the same form (``energyOp`` ...), which the caller
passes to ``KernelArgs(operators=...)``.  Kinetic units are fixed by the
engine: hbar = m = 1, so coordinates must be mass-scaled.  This is the
contract of FeynSimul/physical_systems/__init__.py:1-18.
"""
from .harm_osc import HarmOsc
from .anharm_osc import AnharmOsc
from .lm2m2 import Lm2m2Cluster
from .lm2m2_cluster import Lm2m2_cluster, JacobiHeCluster
from .lm2m2_3part import Lm2m2_3part
from .lm2m2_4part import Lm2m2_4part

__all__ = ["HarmOsc", "AnharmOsc", "Lm2m2Cluster", "Lm2m2_cluster", "JacobiHeCluster", "Lm2m2_3part", "Lm2m2_4part"]
