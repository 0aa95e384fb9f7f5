"""One-dimensional harmonic oscillator (FeynSimul/physical_systems/harm_osc.py:34-53)."""


class HarmOsc:
    """V(x) = k x^2 / 2 with the virial energy operator k x^2.

    ``lb=True`` selects the higher-order ("Li-Broughton"-corrected) potential
    and energy operator of the reference, written with the ``eps`` macro.
    """

    def __init__(self, k=1.0, lb=False):
        self.k = k
        self.DOF = 1
        self.userCode = "FLOAT_TYPE sqr(FLOAT_TYPE x){return x*x;}"
        if lb:
            self.potential = "0.5 * (1.0 + sqr(eps)/12.0)*sqr(x1)"
            self.energyOp = "2.0 * 0.5*(1.0 + sqr(eps)/4.0)*sqr(x1)"
        else:
            self.potential = "0.5 * " + str(k) + " * x1 * x1"
            self.energyOp = str(k) + " * x1 * x1"
        self.positionSquaredOp = "x1 * x1"
