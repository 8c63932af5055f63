"""CPU oracle for fastmat's structured-matvec hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``fastmat_amd``) imports,
calls or links anything in this package.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only to check results (or, for ``cpu_baseline``, as the timed CPU
restatement of the reference's own CPU path).

Every function here is pinned by ``tests/test_oracle_pins.py`` against
something that does not restate its own formula (a library routine, a closed
form or invariant, brute force on tiny inputs, or the reference's own test
statements); none is left unpinned.

Modules
-------
dense        closed-form dense matrices, restating each class's ``_reference()``
             (e.g. Fourier.pyx:241-246, Circulant.pyx:349-425,
             Toeplitz.pyx:637-734, Hadamard.pyx:242-248, Kron.pyx:344-353).
fast64       FP64 / exact-integer evaluators of the same operators at large N
             (np.fft in complex128, exact int64 FWHT, mode-wise tensordot).
heuristics   restatement of the reference's FFT-size heuristics
             (core/cmath.pyx:35-214) and of the Fourier radix/Bluestein
             decision (Fourier.pyx:109-122).
cpu_path     restatement of the reference's CPU forward/backward (same np.fft
             calls, same Product(F.H, Diag, F) structure, same Bluestein
             set-up), used as the timed CPU baseline.
fwht_c       ctypes binding of ``fwht.c``: the reference's typed in-place
             radix-2 Hadamard butterfly loop (Hadamard.pyx:36-61, 164-230),
             restated in C with the same wrap-around integer arithmetic.
accept       the reference's acceptance rule (inspect/test.py:104-202).

Parity pins
-----------
Every function here is checked in ``tests/test_oracle_*.py`` against a check
that would catch a plausible mistake (a dropped term, a wrong sign or index,
a transposed operand): library routines (scipy.linalg.hadamard / toeplitz /
circulant, numpy.kron), closed forms and invariants (Parseval, F^H F = N I,
H H = N I, <Ax, y> = <x, A^H y>), brute force on tiny inputs, and the
reference's own test-file statements (e.g. "41 is the first size bluestein is
faster", Circulant.pyx:436).  The reference ships no golden vectors.
"""
