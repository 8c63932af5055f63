"""Writes docstring_examples.json: the worked examples printed in the
reference's class docstrings, transcribed as data (inputs and the matrices /
sequences the docstrings state).  The reference ships no other golden
vectors or fixtures.

  Permutation.pyx:36-50   sigma = [3, 1, 2, 0] and its matrix J
  Diag.pyx:38-60          d = (1, 0, 3, 6) and diag(d)
  Hadamard.pyx:68-78      H_1 = [[1, 1], [1, -1]], H_n = H_1 kron H_{n-1}
  LFSRCirculant.pyx:95-122  polynomial 0b11001, start 0b1010 and the
                          sequence s (the first column)

Run: python tests/golden/make_docstring_examples.py
"""
import json
import os

EXAMPLES = {
    "permutation": {
        "sigma": [3, 1, 2, 0],
        "matrix": [[0, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0], [1, 0, 0, 0]],
    },
    "diag": {
        "d": [1, 0, 3, 6],
        "matrix": [[1, 0, 0, 0], [0, 0, 0, 0], [0, 0, 3, 0], [0, 0, 0, 6]],
    },
    "hadamard": {
        "H1": [[1, 1], [1, -1]],
        "rule": "H_n = kron(H_1, H_{n-1}), H_0 = [[1]]",
    },
    "lfsr_circulant": {
        "polynomial": 25,
        "start": 10,
        "sequence": [1, -1, 1, -1, -1, 1, 1, -1, 1, 1, 1, -1, -1, -1, -1],
    },
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "docstring_examples.json")
    with open(out, "w") as f:
        json.dump(EXAMPLES, f, indent=1)
    print("wrote", out)
