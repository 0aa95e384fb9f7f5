"""Three He atoms with the LM2M2 potential in Jacobi coordinates: ``Lm2m2_3part``.

The drop-in for FeynSimul/physical_systems/lm2m2_3part.py:19-153: DOF 6 (two
Jacobi vectors), lengths in xUnit = 8.98888e-11 m and energies in
potentialUnit = 2.07098e-22 J, so that a kernel beta of 1000 is 15 mK, and the
operators energyOp (virial energy), meanSquaredRadiusOp, meanRadiusOp and
meanSquaredAtomDistOp of that file.  The strings come from this package's own
Jacobi generator (``JacobiHeCluster``); the virial estimator uses the exact
derivative of the add-on term (lm2m2_3part.py:146 adds a stray +1.0 there).
"""
from .lm2m2_cluster import JacobiHeCluster


class Lm2m2_3part(JacobiHeCluster):
    def __init__(self):
        JacobiHeCluster.__init__(self, 3, xUnit=8.98888e-11, potentialUnit=2.07098e-22)
        self.groundStateEnergy = -0.126  # K (lm2m2_3part.py:23)
