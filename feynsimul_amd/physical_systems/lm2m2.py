"""Helium clusters with the LM2M2 pair potential (Aziz and Slaman, J. Chem. Phys. 94, 8047 (1991)).

Energies are in kelvin and ``beta`` in 1/K.  The engine's kinetic term is
(x' - x)^2 / (2 eps) with hbar = m = 1, so atom positions r (angstrom) enter
as mass-scaled Cartesian coordinates x = sqrt(m_eff) r with
m_eff = m k_B / hbar^2 [1/(K A^2)].  The reference reaches the same unit
convention through ``xUnit`` / ``potentialUnit`` (lm2m2_3part.py:22-25) and
``m = HeMass rUnit^2 / hbar^2 potentialUnit`` (lm2m2_cluster.py:21-28); its
n-atom generator works in Jacobi coordinates via sympy, this class writes the
pair sum directly in Cartesian coordinates (DOF = 3 n; the free centre of mass
does not enter the potential or the virial estimator).

The potential and its radial derivative follow the LM2M2 form of
FeynSimul/physical_systems/lm2m2.h:22-147 (HFD-B base, damping
exp(-(D/x - 1)^2) for x < D, zero damping term below x = 0.0005/r_m, sine
add-on on [x_a1, x_a2]).  The add-on derivative here is A_a B cos(...), the
derivative of the add-on term; lm2m2.h:143 adds a stray +1.0 inside it.
"""
import math

# CODATA 2018
_HBAR = 1.054571817e-34   # J s
_KB = 1.380649e-23        # J / K
_AMU = 1.66053906660e-27  # kg

_LOG2E = 1.0 / math.log(2.0)
_SQRT_LOG2E = math.sqrt(_LOG2E)

# LM2M2 parameters (Aziz & Slaman 1991, table V)
LM2M2 = dict(A=1.89635353e5, alpha=10.70203539, c6=1.34687065, c8=0.41308398, c10=0.17060159,
             D=1.4088, beta=-1.90740649, eps_k=10.97, rm=2.9695, Aa=0.0026,
             xa1=1.003535949, xa2=1.454790369)

# ground-state energies (K) of He_n, n = 3..10: the binding energies listed in
# lm2m2_cluster.py:17-19, with the sign of an energy
GROUND_STATES = {3: -0.0872, 4: -0.387, 5: -0.901, 6: -1.605, 7: -2.478, 8: -3.489, 9: -4.641, 10: -5.904}


def mass_scale(mass_u=4.0026):
    """sqrt(m k_B / hbar^2) in 1/(sqrt(K) angstrom)."""
    m_eff = mass_u * _AMU * _KB / _HBAR ** 2 * 1e-20
    return math.sqrt(m_eff)


def lm2m2_kelvin(r):
    """LM2M2 pair energy in K at distance r (angstrom); FP64 reference formula."""
    p = LM2M2
    x = r / p["rm"]
    # the damped dispersion term, zero at and below x = 0.0005 / r_m (lm2m2.h:55-66)
    if x <= 0.0005 / p["rm"]:
        ddisp = 0.0
    else:
        disp = p["c6"] / x ** 6 + p["c8"] / x ** 8 + p["c10"] / x ** 10
        ddisp = math.exp(-(p["D"] / x - 1.0) ** 2) * disp if x < p["D"] else disp
    vb = p["A"] * math.exp(-p["alpha"] * x + p["beta"] * x * x) - ddisp
    va = 0.0
    if p["xa1"] <= x <= p["xa2"]:
        B = 2.0 * math.pi / (p["xa2"] - p["xa1"])
        va = p["Aa"] * (math.sin(B * (x - p["xa1"]) - math.pi / 2) + 1.0)
    return p["eps_k"] * (vb + va)


def _user_code(qscale):
    p = LM2M2
    B = 2.0 * math.pi / (p["xa2"] - p["xa1"])
    return r"""
/* LM2M2 He-He potential in K; r2 is a squared pair distance in mass-scaled
 * coordinates, q = r / r_m is the reduced distance.  Written without
 * branches: the damping exponent max(D/q - 1, 0) gives exactly 1 for q >= D,
 * and the add-on A_a (sin(B (q - x_a1) - pi/2) + 1) = A_a (1 - cos(2 pi w)),
 * w = (q - x_a1) / (x_a2 - x_a1), is selected on [x_a1, x_a2]. */
FLOAT_TYPE fs_sqr(FLOAT_TYPE a) { return a * a; }
/* The pair energy at squared distance r2: one reciprocal square root gives
 * both 1/q and q, the exponentials are base 2 with log2(e) folded into their
 * coefficients, eps/k is folded into every term (into the repulsive
 * exponent as log2(eps/k A)), and the add-on window w in [0, 1] is one
 * comparison |w - 1/2| <= 1/2 (same function).  At and below q = 0.0005 / r_m
 * the damped dispersion term is zero, as in the reference: there the damping
 * factor exp2(-t^2) is exactly 0 (t > 7000), and the dispersion sum, whose
 * 1/q^10 overflows first, is held finite by the clamp so that the product
 * stays 0.  The same clamp on 1/r gives q = 0 at r2 = 0, where the pair
 * energy is eps/k A; above the cut-off neither clamp changes a value. */
FLOAT_TYPE fs_lm2m2_pair(FLOAT_TYPE r2)
{
    const FLOAT_TYPE s = 1.0 / sqrt(r2);
    const FLOAT_TYPE q = r2 * fmin(s, 3.0e38) * %(qscale)r;
    const FLOAT_TYPE iq = s * %(iqscale)r;
    const FLOAT_TYPE iq2 = iq * iq;
    const FLOAT_TYPE disp = fmin(iq2 * iq2 * iq2 * (%(ec6)r + iq2 * (%(ec8)r + iq2 * %(ec10)r)), 3.0e38);
    const FLOAT_TYPE t = fmax(%(Dl)r * iq - %(sl)r, 0.0);
    const FLOAT_TYPE w = (q - %(xa1)r) * %(iwidth)r;
    const FLOAT_TYPE va = fabs(w - 0.5) <= 0.5 ? %(eAa)r * (1.0 - fs_cos2pi(w)) : 0.0;
    return exp2(q * (%(betal)r * q - %(alphal)r) + %(log2eA)r) - exp2(-t * t) * disp + va;
}
/* V(q) + r dV/dr / 2 in K, the pair term of the virial energy estimator */
FLOAT_TYPE fs_lm2m2_virial_q(FLOAT_TYPE q)
{
    const FLOAT_TYPE iq = 1.0 / q;
    const FLOAT_TYPE iq2 = iq * iq;
    const FLOAT_TYPE iq6 = iq2 * iq2 * iq2;
    const FLOAT_TYPE disp = iq6 * (%(c6)r + iq2 * (%(c8)r + iq2 * %(c10)r));
    const FLOAT_TYPE qdisp = -iq6 * (6.0 * %(c6)r + iq2 * (8.0 * %(c8)r + iq2 * 10.0 * %(c10)r));
    const FLOAT_TYPE t = fmax(%(D)r * iq - 1.0, 0.0);
    const FLOAT_TYPE damp = exp(-t * t);
    const FLOAT_TYPE qdamp = damp * 2.0 * t * (%(D)r * iq);
    const FLOAT_TYPE rep = %(A)r * exp(q * (%(beta)r * q - %(alpha)r));
    const FLOAT_TYPE w = (q - %(xa1)r) * %(iwidth)r;
    const bool win = w >= 0.0 && w <= 1.0;
    const FLOAT_TYPE va = win ? %(Aa)r * (1.0 - fs_cos2pi(w)) : 0.0;
    const FLOAT_TYPE qva = win ? %(Aa)r * %(B)r * q * fs_sin2pi(w) : 0.0;
    /* the damped dispersion term and its q-derivative, zero below q = 0.0005 / r_m */
    const bool core = q > %(qmin)r;
    const FLOAT_TYPE ddisp = core ? damp * disp : 0.0;
    const FLOAT_TYPE qddisp = core ? qdamp * disp + damp * qdisp : 0.0;
    const FLOAT_TYPE v = rep - ddisp + va;
    const FLOAT_TYPE qdv = rep * q * (2.0 * %(beta)r * q - %(alpha)r) - qddisp + qva;
    return %(eps_k)r * (v + 0.5 * qdv);
}
FLOAT_TYPE fs_lm2m2_r2(FLOAT_TYPE r2) { return fs_lm2m2_pair(r2); }
FLOAT_TYPE fs_lm2m2_virial_r2(FLOAT_TYPE r2) { return fs_lm2m2_virial_q(sqrt(r2) * %(qscale)r); }
""" % dict(p, B=B, iwidth=1.0 / (p["xa2"] - p["xa1"]), qmin=0.0005 / p["rm"], qscale=qscale,
                  iqscale=1.0 / qscale, ec6=p["eps_k"] * p["c6"], ec8=p["eps_k"] * p["c8"],
                  ec10=p["eps_k"] * p["c10"], Dl=p["D"] * _SQRT_LOG2E, sl=_SQRT_LOG2E, eAa=p["eps_k"] * p["Aa"],
                  log2eA=math.log2(p["eps_k"] * p["A"]), betal=p["beta"] * _LOG2E, alphal=p["alpha"] * _LOG2E)


class Lm2m2Cluster:
    """n helium atoms in 3D with pairwise LM2M2 interactions (DOF = 3 n).

    Operators (strings in x1..x{3n}):
      energyOp            virial energy  sum_pairs V(r) + r V'(r) / 2   [K]
      meanPairDistanceOp  mean pair distance                            [angstrom]
      meanSquaredPairDistanceOp                                         [angstrom^2]
      potentialOp         potential energy                              [K]
    ``pairDistance(i, j)`` is the expression of one pair distance in angstrom.
    """

    def __init__(self, nAtoms=3, mass=4.0026):
        if nAtoms < 2:
            raise Exception("n must be >=2")
        self.nAtoms = n = int(nAtoms)
        self.DOF = 3 * n
        self.translationInvariant = True   # V depends on relative positions only
        self.mass = mass
        self.massScale = mass_scale(mass)             # x = massScale * r
        self.rUnit = 1e-10                             # coordinates before scaling are angstrom
        self.potentialUnit = _KB                       # energies in K
        self.groundStateEnergy = GROUND_STATES.get(n)
        qscale = 1.0 / (self.massScale * LM2M2["rm"])
        self.userCode = _user_code(qscale)
        self.pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
        r2 = [self._r2(i, j) for i, j in self.pairs]
        self.potential = " + ".join("fs_lm2m2_r2(%s)" % e for e in r2)
        self.potentialOp = self.potential
        self.energyOp = " + ".join("fs_lm2m2_virial_r2(%s)" % e for e in r2)
        inv = 1.0 / self.massScale
        self.meanPairDistanceOp = "%r * (%s)" % (inv / len(self.pairs), " + ".join("sqrt(%s)" % e for e in r2))
        self.meanSquaredPairDistanceOp = "%r * (%s)" % (inv * inv / len(self.pairs), " + ".join(r2))

    @staticmethod
    def _r2(i, j):
        a, b = 3 * i + 1, 3 * j + 1
        return "fs_sqr(x%d - x%d) + fs_sqr(x%d - x%d) + fs_sqr(x%d - x%d)" % (a, b, a + 1, b + 1, a + 2, b + 2)

    def pairDistance(self, i, j):
        return "%r * sqrt(%s)" % (1.0 / self.massScale, self._r2(i, j))

    def initialPositions(self, nWalkers, rng, side=3.5, sigma=3.0):
        """Gaussian blobs (sigma, angstrom) around a random equilateral triangle / simplex-like
        arrangement of the given side, in mass-scaled coordinates; shape (nWalkers, DOF)."""
        import numpy as np
        n = self.nAtoms
        base = np.zeros((n, 3))
        if n >= 2:
            base[1] = [side, 0, 0]
        if n >= 3:
            base[2] = [side / 2, side * math.sqrt(3) / 2, 0]
        for k in range(3, n):
            base[k] = rng.normal(0, side, 3)
        base -= base.mean(0)
        out = np.empty((nWalkers, n, 3))
        for w in range(nWalkers):
            # random orientation of the reference arrangement
            qm, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            out[w] = base @ qm.T + rng.normal(0, sigma, (n, 3))
        return (out * self.massScale).reshape(nWalkers, 3 * n)
