"""One-dimensional anharmonic oscillator V(x) = x^2/2 + lam x^4 (m = omega = hbar = 1)."""


class AnharmOsc:
    """Quartic oscillator in the harmonic-oscillator style of the reference systems.

    The virial energy operator is V + x V'/2 = x^2 + 3 lam x^4.
    """

    def __init__(self, lam=0.1):
        self.lam = lam
        self.DOF = 1
        self.userCode = "FLOAT_TYPE sqr(FLOAT_TYPE x){return x*x;}"
        self.potential = "0.5 * sqr(x1) + %r * sqr(sqr(x1))" % float(lam)
        self.energyOp = "sqr(x1) + %r * sqr(sqr(x1))" % (3.0 * float(lam))
        self.positionSquaredOp = "x1 * x1"

    def V(self, x):
        return 0.5 * x * x + self.lam * x ** 4
