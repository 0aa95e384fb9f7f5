"""modN helpers (FeynSimul/pimc_utils.py) that need no GPU: path doubling and the TSV format."""
import os

import numpy as np

from feynsimul_amd.pimc_utils import doublePaths, _output, readModN, _writePaths, _readPaths


def test_double_paths_matches_reference_loop():
    """pimc_utils.py:211-226 written as the reference's per-walker loop."""
    rng = np.random.default_rng(0)
    W, N, DOF = 5, 8, 3
    old = rng.standard_normal((W, N * DOF))
    new = doublePaths(old, N, DOF)
    Nn = 2 * N
    ref = np.zeros((W, Nn * DOF))
    for walker in range(W):
        for d in range(DOF):
            second = d * Nn + np.arange(0, Nn, 2)
            ref[walker, second] = old[walker, second // 2]
            right = second + 2
            right[-1] -= Nn
            ref[walker, second + 1] = (ref[walker, second] + ref[walker, right]) / 2.0
    np.testing.assert_array_equal(new, ref)


def test_output_tsv_format(tmp_path):
    fn = str(tmp_path / "data")
    ops = np.array([[1.5, 2.5, 3.5], [0.25, 0.5, 0.75]])
    _output(fn, 16, 1.25, 4096, 0.5, ops, 0.01, 0.002, 10.0, 2)
    _output(fn, 32, 2.5, 8192, 0.375, ops * 2, 0.01, 0.002, 10.0, 3)
    lines = open(fn + ".tsv").read().split("\n")
    assert lines[0] == "#Beta: 10.0"
    assert lines[1] == ("#N\tTime\tpathChanges\tAR\tS\tOperator 0, Walker 0\tOperator 0, Walker 1\tOperator 0, "
                        "Walker 2\tOperator 1, Walker 0\tOperator 1, Walker 1\tOperator 1, Walker 2")
    assert lines[2] == "16\t1.25\t4096\t0.5\t2\t1.5\t2.5\t3.5\t0.25\t0.5\t0.75\t"
    beta, rows = readModN(fn + ".tsv")
    assert beta == 10.0 and rows.shape == (2, 11)
    np.testing.assert_array_equal(rows[1, 5:], (ops * 2).reshape(-1))


def test_path_files_round_trip(tmp_path):
    p = np.random.default_rng(1).standard_normal((4, 24)).astype(np.float32)
    fn = str(tmp_path / "pathsN8episode1")
    _writePaths(fn, p)
    back = _readPaths(fn, 4, 24)
    np.testing.assert_array_equal(back.astype(np.float32), p)
