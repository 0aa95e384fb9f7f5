"""modN (FeynSimul/pimc_utils.py:28-285) end to end on the GPU: N doubling, result files, resume."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from feynsimul_amd.kernel_args import KernelArgs
from feynsimul_amd.physical_systems import HarmOsc
from feynsimul_amd.pimc_utils import modN, readModN

pytestmark = pytest.mark.gpu


def _ka(W):
    s = HarmOsc()
    return KernelArgs(system=s, nbrOfWalkers=W, N=16, beta=4.0, S=1, enableOperator=True, enableCorrelator=False,
                      enableBisection=True, enablePathShift=False, enableSingleNodeMove=False,
                      enableParallelizePath=True, enableGlobalPath=False, enableGlobalOldPath=False,
                      enableBins=False, nbrOfWalkersPerWorkGroup=1, operators=(s.energyOp,),
                      metroStepsPerOperatorRun=1, operatorRuns=1, seed=42)


def test_modn_doubles_n_writes_files_and_resumes(gpu, tmp_path):
    print("[case] modN HO 16 -> 64", flush=True)
    W = 2048
    runs = lambda N, S: 6
    rng = np.random.default_rng(0)
    modN(_ka(W), rng.uniform(-1, 1, (W, 1)), 3, "ho", lambda N, S: 40, lambda N, S: 8, runs, 1024, finalN=64,
         simTime=0, resultsDir=str(tmp_path))
    runDirs = glob.glob(str(tmp_path / "ho" / "*" / "*"))
    assert len(runDirs) == 1
    d = runDirs[0]
    beta, rows = readModN(os.path.join(d, "data.tsv"))
    assert beta == 4.0
    assert list(rows[:, 0]) == [16] * 6 + [32] * 6 + [64] * 6
    assert np.all(np.diff(rows[:, 2]) > 0)              # path changes accumulate
    assert rows.shape[1] == 5 + W
    for N in (16, 32, 64):
        assert os.path.exists(os.path.join(d, "pathsN%depisode1" % N))
        assert os.path.exists(os.path.join(d, "pathsN%depisode2" % N))
    assert len(open(os.path.join(d, "timestamps")).read().strip().split("\n")) == 2
    last = rows[-1]
    per = last[5:]
    E, se = per.mean(), per.std() / np.sqrt(W)
    exact = O.ho_discrete_x2(64, 4.0)
    print("N=64 E %.5f +- %.5f exact %.6f, S %d" % (E, se, exact, int(last[4])))
    assert abs(E - exact) < 5 * se
    # resume from the last checkpoint at N = 64
    modN(_ka(W), None, 3, "ho_resume", lambda N, S: 40, lambda N, S: 8, runs, 1024, continueRun=True,
         finalN=64, simTime=0, cont_S=int(last[4]), cont_N=64, cont_path=os.path.join(d, "pathsN64episode2"),
         resultsDir=str(tmp_path))
    d2 = glob.glob(str(tmp_path / "ho_resume" / "*" / "*"))[0]
    _, rows2 = readModN(os.path.join(d2, "data.tsv"))
    assert list(rows2[:, 0]) == [64] * 6 and int(rows2[0, 4]) == int(last[4])
    per2 = rows2[-1, 5:]
    assert abs(per2.mean() - exact) < 5 * per2.std() / np.sqrt(W)


def test_reference_he3_harness_runs_through_the_alias(gpu, tmp_path, monkeypatch):
    """examples/lm2m2_3particles.py:28-77 through ``install_feynsimul_alias``: Lm2m2_3part (Jacobi,
    DOF 6, kernel beta 1000 = 15 mK), three operators, modN from N = 8 (straight paths in a 0.01-unit
    cube) to N = 32 with the script's run formulas; verbosity 2 dumps each generated module source."""
    from test_reference_harness import run_harness
    print("[case] lm2m2_3particles harness N 8 -> 32", flush=True)
    monkeypatch.chdir(tmp_path)
    run_harness(finalN=32, simTime=0, verbosity=2, save=5)
    d = glob.glob(str(tmp_path / "results" / "lm2m2_3part" / "*" / "*"))[0]
    beta, rows = readModN(os.path.join(d, "data.tsv"))
    assert beta == 1000.0
    assert list(rows[:, 0]) == [8] * 10 + [16] * 10 + [32] * 10
    assert rows.shape[1] == 5 + 3 * 64
    assert np.isfinite(rows).all()
    ar, S = rows[:, 3], rows[:, 4]
    assert ((ar > 0) & (ar <= 1)).all()
    assert ((S >= 1) & (S <= np.log2(rows[:, 0]))).all()
    ops = rows[:, 5:].reshape(len(rows), 3, 64)
    # the atoms start within 0.01 length units of each other (mean squared radius <= 2e-4), deep in
    # the repulsive core; the sampler pulls them apart
    msr = ops[:, 1].mean(axis=1)
    assert msr[-1] > 100 * 2e-4 and (ops[:, 1] > 0).all() and (ops[:, 2] > 0).all()
    assert os.path.exists("cleaned.hip")
    for N in (8, 16, 32):
        assert os.path.exists(os.path.join(d, "pathsN%depisode2" % N))
