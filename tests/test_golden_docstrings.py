"""The reference's docstring examples (tests/golden/docstring_examples.json)
against the oracle (CPU) and, marked gpu, against the device classes."""
import json
import os

import numpy as np
import pytest

from oracle import dense

HERE = os.path.dirname(os.path.abspath(__file__))
EX = json.load(open(os.path.join(HERE, "golden", "docstring_examples.json")))


def test_oracle_permutation_diag_hadamard_lfsr():
    p = EX["permutation"]
    np.testing.assert_array_equal(dense.permutation(np.array(p["sigma"])), p["matrix"])
    d = EX["diag"]
    np.testing.assert_array_equal(dense.diag(np.array(d["d"])), d["matrix"])
    H1 = np.array(EX["hadamard"]["H1"])
    np.testing.assert_array_equal(dense.hadamard(1), H1)
    Hn = np.array([[1]])
    for n in range(1, 6):
        Hn = np.kron(H1, Hn)
        np.testing.assert_array_equal(dense.hadamard(n), Hn)
    lf = EX["lfsr_circulant"]
    np.testing.assert_array_equal(dense.lfsr_sequence(lf["polynomial"], lf["start"]),
                                  lf["sequence"])


@pytest.mark.gpu
def test_device_classes_reproduce_docstring_examples():
    import fastmat_amd as fm
    from fastmat_amd import _lib
    _lib.load()
    p = EX["permutation"]
    P = fm.Permutation(np.array(p["sigma"]))
    np.testing.assert_array_equal(P.forward(np.eye(4, dtype=np.int64)), p["matrix"])
    d = EX["diag"]
    D = fm.Diag(np.array(d["d"]))
    np.testing.assert_array_equal(D.forward(np.eye(4, dtype=np.int64)), d["matrix"])
    H1 = np.array(EX["hadamard"]["H1"])
    Hn = np.array([[1]])
    for n in range(1, 6):
        Hn = np.kron(H1, Hn)
        np.testing.assert_array_equal(fm.Hadamard(n).forward(np.eye(2 ** n, dtype=np.int32)), Hn)
    lf = EX["lfsr_circulant"]
    L = fm.LFSRCirculant(lf["polynomial"], lf["start"])
    e0 = np.zeros(L.numCols, dtype=np.int32)
    e0[0] = 1
    np.testing.assert_array_equal(L.forward(e0), lf["sequence"])
