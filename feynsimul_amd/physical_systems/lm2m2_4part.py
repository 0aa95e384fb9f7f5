"""Four He atoms with the LM2M2 potential in Jacobi coordinates: ``Lm2m2_4part``.

The drop-in for FeynSimul/physical_systems/lm2m2_4part.py:19-158: DOF 9
(three Jacobi vectors), lengths in xUnit = 4.92341e-11 m and energies in
potentialUnit = 6.90325e-22 J (a kernel beta of 1000 is 50 mK), the
literature reference values groundStateEnergy, meanSquaredRadius and
meanSquaredAtomDist of that file, and its operators (``JacobiHeCluster``).
"""
from .lm2m2_cluster import JacobiHeCluster


class Lm2m2_4part(JacobiHeCluster):
    def __init__(self):
        JacobiHeCluster.__init__(self, 4, xUnit=4.92341e-11, potentialUnit=6.90325e-22)
        # literature values quoted by lm2m2_4part.py:26-30 (K, angstrom^2)
        self.groundStateEnergy = -0.4
        self.meanSquaredRadius = 7.15 ** 2.0
        self.meanSquaredAtomDist = 10.7 ** 2.0
