"""The reference's fixed-size He systems and its He3 harness through the ``FeynSimul`` alias (CPU).

* ``Lm2m2_3part`` / ``Lm2m2_4part`` (physical_systems/lm2m2_3part.py:19-153,
  lm2m2_4part.py:19-158): DOF 3(n-1) unit-mass Jacobi coordinates in xUnit,
  energies in potentialUnit = hbar^2 / (m_He xUnit^2).  Their strings,
  compiled by gcc in FP64, against the FP64 LM2M2 pair sum and the geometric
  operators computed from Cartesian atom positions.
* examples/lm2m2_3particles.py:28-77: the script's imports and calls, written
  out here (not copied) through ``install_feynsimul_alias``; without a GPU the
  call sequence must reach PIMCKernel and stop there with its no-GPU error.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
import feynsimul_amd
from feynsimul_amd.physical_systems import Lm2m2_3part, Lm2m2_4part
from feynsimul_amd.physical_systems.lm2m2 import lm2m2_kelvin
from feynsimul_amd.physical_systems.lm2m2_cluster import jacobi_matrix

HBAR, KB, M_HE = 1.054571817e-34, 1.380649e-23, 4.002602 * 1.66053906660e-27


@pytest.mark.parametrize("cls,n", [(Lm2m2_3part, 3), (Lm2m2_4part, 4)])
def test_fixed_size_systems_match_cartesian_pair_sums(cls, n):
    sysobj = cls()
    assert sysobj.DOF == 3 * (n - 1)
    # the unit convention: unit kinetic mass in xUnit and potentialUnit
    assert abs(HBAR ** 2 / (M_HE * sysobj.xUnit ** 2) / sysobj.potentialUnit - 1.0) < 2e-4
    ops = (sysobj.energyOp, sysobj.meanSquaredRadiusOp, sysobj.meanRadiusOp, sysobj.meanSquaredAtomDistOp)
    lib = O.build(sysobj, 64, 1000.0, operators=ops, double=True)
    xu = sysobj.xUnit * 1e10
    J = jacobi_matrix(n, 1.0)
    scale = KB / sysobj.potentialUnit
    rng = np.random.default_rng(n)
    for _ in range(40):
        r = rng.normal(0, 3.0, (n, 3))                          # angstrom
        j = J @ (r / xu)                                        # Jacobi vectors, row k = j_k
        x = np.ascontiguousarray(j[: n - 1].reshape(-1))        # x_{3k+d+1} = component d of j_k
        dists = [np.linalg.norm(r[a] - r[b]) for a in range(n) for b in range(a + 1, n)]
        V = sum(lm2m2_kelvin(d) for d in dists) * scale
        h = 1e-6
        vir = sum(lm2m2_kelvin(d) + 0.5 * d * (lm2m2_kelvin(d + h) - lm2m2_kelvin(d - h)) / (2 * h)
                  for d in dists) * scale
        rel = np.linalg.norm(r - r.mean(0), axis=1) / xu
        acc = np.zeros(5)
        lib.fs_oracle_operators(x.ctypes.data, acc.ctypes.data)
        assert abs(lib.fs_oracle_potential(x.ctypes.data) - V) < 1e-9 * max(1.0, abs(V))
        assert abs(acc[0] - vir) < 1e-5 * max(1.0, abs(vir))
        assert abs(acc[1] - np.mean(rel ** 2)) < 1e-9 * np.mean(rel ** 2)
        assert abs(acc[2] - np.mean(rel)) < 1e-9 * np.mean(rel)
        assert abs(acc[3] - np.mean(np.square(dists)) / xu ** 2) < 1e-9 * np.mean(np.square(dists)) / xu ** 2
    # kernel beta 1000 is 15 mK (3 atoms) and 50 mK (4 atoms), as the reference's files state
    assert abs(sysobj.potentialUnit / (KB * 1000.0) - {3: 0.015, 4: 0.050}[n]) < 0.001


def test_lm2m2_helper_in_ev():
    """userCode's lm2m2(r) is the pair energy in eV, as the helper of the reference's strings."""
    sysobj = Lm2m2_3part()

    class _Probe:
        DOF = 1
        userCode = sysobj.userCode
        potential = "lm2m2(x1)"
    lib = O.build(_Probe(), 64, 1.0, double=True)
    for r in (2.5, 2.9695, 3.5, 5.0, 10.0):
        x = np.array([r])
        ev = lm2m2_kelvin(r) * KB / 1.602176634e-19
        assert abs(lib.fs_oracle_potential(x.ctypes.data) - ev) < 1e-9 * max(abs(ev), 1e-6)


HARNESS = """
import numpy as np
from FeynSimul.pimc_utils import *
from FeynSimul.physical_systems.lm2m2_3part import *

ka = KernelArgs(nbrOfWalkers=64, N=8, enablePathShift=False, enableBisection=True,
                enableSingleNodeMove=False, enableGlobalPath=True, enableGlobalOldPath=True,
                enableParallelizePath=True, enableBins=False, enableDouble=False, enableRanlux=False,
                luxuaryFactor=2, ranluxIntMax=2**32-1, beta=1000.0, nbrOfWalkersPerWorkGroup=4,
                operatorRuns=100, enableOperator=True, enableCorrelator=False,
                metroStepsPerOperatorRun=40, system=Lm2m2_3part())
ka.operators = (ka.system.energyOp, ka.system.meanSquaredRadiusOp, ka.system.meanRadiusOp)

def opRunsFormula(N, S):
    return max(2 ** 10 / 2 ** S, 1)

def mStepsPerOPRun(N, S):
    return 10

def runsPerN(N, S):
    return max(N / 8, 10)

startXList = np.random.uniform(size=(ka.nbrOfWalkers, ka.system.DOF), low=-1.0, high=1.0) * 0.01
modN(ka, startXList, SAVE, "lm2m2_3part", opRunsFormula, mStepsPerOPRun, runsPerN, 512,
     simTime=SIMTIME, finalN=FINALN, verbosity=VERBOSITY)
"""


def run_harness(finalN, simTime, verbosity=2, save=3000):
    """The call sequence of examples/lm2m2_3particles.py:28-77 (finalN and simTime as given)."""
    feynsimul_amd.install_feynsimul_alias()
    src = HARNESS.replace("SAVE", str(save)).replace("SIMTIME", repr(simTime)).replace("FINALN", str(finalN))
    src = src.replace("VERBOSITY", str(verbosity))
    ns = {"__name__": "lm2m2_3particles"}
    exec(compile(src, "lm2m2_3particles_sequence", "exec"), ns)
    return ns


def test_he3_harness_resolves_through_the_alias(tmp_path, monkeypatch):
    import torch
    if torch.cuda.is_available():
        pytest.skip("this is the CPU check; the GPU run is in test_gpu_modn.py")
    monkeypatch.chdir(tmp_path)
    with pytest.raises(RuntimeError, match="needs a GPU"):
        run_harness(finalN=32, simTime=0)
    # modN made its result directory before compiling the first kernel
    assert os.path.isdir(tmp_path / "results" / "lm2m2_3part")
