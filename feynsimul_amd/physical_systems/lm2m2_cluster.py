"""n helium atoms with LM2M2 pair interactions in Jacobi coordinates.

The drop-in for FeynSimul/physical_systems/lm2m2_cluster.py:7-95
(``Lm2m2_cluster(n, T)``), generated without sympy.  Jacobi coordinates of
equal masses m (FeynSimul/physical_systems/jacobi.py:3-19):

  j_0     = sqrt(m / 2)               (r_0 - r_1)
  j_i     = sqrt(m (i+1) / (i+2))    (mean(r_0 .. r_i) - r_{i+1}),  0 < i < n-1
  j_{n-1} = sqrt(n m)                 mean(r_0 .. r_{n-1})            (centre of mass)

Each j_i is a 3-vector; slice coordinate x_{3i+k+1} is component k of j_i.
The transformation is orthonormal in the mass metric, so the kinetic term has
unit mass, and the centre of mass drops out of every pair distance: DOF =
3 (n - 1).  A pair difference r_a - r_b is a fixed linear combination of
j_0 .. j_{n-2}; the potential is sum_pairs V(|r_a - r_b|) (partSys.py:3-12),
the virial estimator V + x . grad V / 2 (lm2m2_cluster.py:75-76), which for a
function of one squared distance s is sum_pairs (V + s dV/ds) =
sum_pairs (V + r V'(r) / 2), and the mean squared radius operator
sum_pairs |r_a - r_b|^2.  Units: angstrom and kelvin, hbar = 1, so
m = M_He angstrom^2 k_B / hbar^2 (lm2m2_cluster.py:23-32).
"""
import numpy as np

from .lm2m2 import LM2M2, _user_code, _KB as _KB_SI

_HE_MASS = 6.64647617e-27   # kg, as lm2m2_cluster.py:23
_HBAR = 1.05457148e-34      # m^2 kg / s, lm2m2_cluster.py:24
_KB = 1.3806503e-23         # J / K, lm2m2_cluster.py:25
_GROUND = [0.0872, 0.387, 0.901, 1.605, 2.478, 3.489, 4.641, 5.904]  # K, n = 3 .. 10 (lm2m2_cluster.py:17-18)


def jacobi_matrix(n, m):
    """J with j = J r (per spatial component), equal masses m."""
    J = np.zeros((n, n))
    J[0, 0], J[0, 1] = 1.0, -1.0
    J[0] *= np.sqrt(m / 2.0)
    for i in range(1, n - 1):
        J[i, :i + 1] = 1.0 / (i + 1)
        J[i, i + 1] = -1.0
        J[i] *= np.sqrt(m * (i + 1) / (i + 2))
    J[n - 1, :] = 1.0 / n
    J[n - 1] *= np.sqrt(n * m)
    return J


def pair_coefficients(n, m):
    """{(a, b): c} with r_a - r_b = sum_k c[k] j_k (k < n - 1), per spatial component."""
    Jinv = np.linalg.inv(jacobi_matrix(n, m))
    out = {}
    for a in range(n):
        for b in range(a + 1, n):
            c = Jinv[a] - Jinv[b]
            assert abs(c[n - 1]) < 1e-9 * np.abs(c).max()   # the centre of mass cancels
            out[(a, b)] = c[: n - 1]
    return out


def _linear(c, dim):
    terms = ["%r * x%d" % (float(ck), 3 * k + dim + 1) for k, ck in enumerate(c) if abs(ck) > 1e-14]
    return " + ".join(terms) if terms else "0.0"


class Lm2m2_cluster:
    """Lm2m2_cluster(n, T): n He atoms at temperature T (kelvin) in Jacobi coordinates.

    Attributes as the reference: DOF, beta (= 1/T, kelvin^-1), m, rUnit, potentialUnit,
    groundStateEnergy (n = 3 .. 10), potential, energyOp, meanSquaredRadiusOp, userCode.
    """

    def __init__(self, n, T, useJacobi=True, simplify=True, verbose=False, enableDouble=False):
        d = 3
        if useJacobi:
            self.DOF = (n - 1) * d
        else:
            self.DOF = n * d
            raise Exception("disablement of jacobi isn't implemented yet")
        if n < 2:
            raise Exception("n must be >=2")
        if 3 <= n <= 10:
            self.groundStateEnergy = _GROUND[n - 3]
        self.nAtoms = n
        self.rUnit = 1e-10
        self.potentialUnit = _KB
        self.beta = self.potentialUnit / (_KB * T)
        m = self.m = _HE_MASS * self.rUnit ** 2 / _HBAR ** 2 * self.potentialUnit
        if not simplify:
            raise Exception("disablement of simplify isn't implemented yet")
        if verbose:
            print("Building pair distances in Jacobi coordinates...")
        self.pairs = sorted(pair_coefficients(n, m).items())
        s = ["fs_sqr(%s) + fs_sqr(%s) + fs_sqr(%s)" % tuple(_linear(c, k) for k in range(d)) for _, c in self.pairs]
        self.potential = " + ".join("fs_lm2m2_r2(%s)" % e for e in s)
        self.energyOp = " + ".join("fs_lm2m2_virial_r2(%s)" % e for e in s)
        self.meanSquaredRadiusOp = " + ".join("(%s)" % e for e in s)
        # the pair functions take r^2 in angstrom^2 (qscale = 1 / r_m)
        self.userCode = _user_code(1.0 / LM2M2["rm"])


_EV = 1.602176634e-19   # J


class JacobiHeCluster:
    """n He atoms in unit-mass Jacobi coordinates, lengths in ``xUnit`` metres and energies in
    ``potentialUnit`` joules, with hbar = m_He = 1 in those units (potentialUnit = hbar^2 /
    (m_He xUnit^2)).  This is the unit convention of the reference's fixed-size systems
    (physical_systems/lm2m2_3part.py:22-31, lm2m2_4part.py:22-33); the strings are generated
    here from the Jacobi map of ``jacobi_matrix`` (unit mass), not copied.

    Strings (in x1 .. x{3(n-1)}):
      potential, energyOp      sum_pairs V(r) and the virial sum_pairs V + r V'(r) / 2, in potentialUnit
      meanSquaredRadiusOp      (1/n) sum_i |r_i - R|^2            [xUnit^2]
      meanRadiusOp             (1/n) sum_i |r_i - R|              [xUnit]
      meanSquaredAtomDistOp    mean over pairs of |r_a - r_b|^2   [xUnit^2]
    ``userCode`` also defines ``sqr(a)`` and ``lm2m2(r)`` (pair energy in eV at r angstrom), the
    helper names of the reference's strings.
    """

    def __init__(self, n, xUnit, potentialUnit):
        if n < 2:
            raise Exception("n must be >=2")
        self.nAtoms = n
        self.DOF = 3 * (n - 1)
        self.xUnit = float(xUnit)
        self.potentialUnit = float(potentialUnit)
        xu = self.xUnit * 1e10                        # angstrom per length unit
        scale = _KB_SI / self.potentialUnit           # kelvin -> potentialUnit
        self.pairs = sorted(pair_coefficients(n, 1.0).items())
        r2 = ["fs_sqr(%s) + fs_sqr(%s) + fs_sqr(%s)" % tuple(_linear(c, k) for k in range(3)) for _, c in self.pairs]
        self.potential = "%r * (%s)" % (scale, " + ".join("fs_lm2m2_r2(%s)" % e for e in r2))
        self.energyOp = "%r * (%s)" % (scale, " + ".join("fs_lm2m2_virial_r2(%s)" % e for e in r2))
        sq = " + ".join("fs_sqr(x%d)" % (i + 1) for i in range(self.DOF))
        self.meanSquaredRadiusOp = "%r * (%s)" % (1.0 / n, sq)
        self.meanSquaredAtomDistOp = "%r * (%s)" % (2.0 / (n - 1), sq)
        # r_i - R in Jacobi coordinates: row i of J^-1 minus the mean row; the centre of mass drops out
        Jinv = np.linalg.inv(jacobi_matrix(n, 1.0))
        rel = Jinv - Jinv.mean(axis=0)
        assert np.abs(rel[:, n - 1]).max() < 1e-12
        radii = ["sqrt(fs_sqr(%s) + fs_sqr(%s) + fs_sqr(%s))" % tuple(_linear(rel[i, : n - 1], k) for k in range(3))
                 for i in range(n)]
        self.meanRadiusOp = "%r * (%s)" % (1.0 / n, " + ".join(radii))
        kev = _KB_SI / _EV
        self.userCode = _user_code(xu / LM2M2["rm"]) + (
            "FLOAT_TYPE sqr(FLOAT_TYPE a) { return a * a; }\n"
            "/* LM2M2 pair energy in eV at a distance of r angstrom */\n"
            "FLOAT_TYPE lm2m2(FLOAT_TYPE r) { return %r * fs_lm2m2_pair(fs_sqr(r * %r)); }\n" % (kev, 1.0 / xu))
