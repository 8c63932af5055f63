"""Restatement of the reference's FFT-size heuristics (test infrastructure).

The reference picks FFT lengths and the radix/Bluestein path with three
functions in core/cmath.pyx and one decision in Fourier.pyx.  They decide only
*which* algorithm and padding fastmat's CPU path uses (not the matrix), so the
oracle needs them to restate that CPU path faithfully for timing.

C ``float`` arithmetic (single precision) in the reference is reproduced with
numpy.float32 where it can change a comparison.
"""
import math

import numpy as np


def find_fft_factors(target, max_factor, state, best_state):
    """Greedy recursion of ``_findFFTFactors`` (core/cmath.pyx:35-85).

    ``state`` packs (complexity << 16) + length; the lower packed value wins.
    """
    for ff in range(max_factor, 0, -1):
        length = (state & 0xFFFF) * ff
        complexity = (state >> 16) + ff + 1
        new_state = (complexity << 16) + length
        if new_state <= best_state and length < target:
            best_state = find_fft_factors(target, ff, new_state, best_state)
        else:
            if new_state < best_state:
                best_state = new_state
    return best_state


def find_optimal_fft_size(order, max_stage=4):
    """``_findOptimalFFTSize`` (core/cmath.pyx:88-153): radix-4 stages while
    more than 64 remains (float32 ``remaining /= 4``, :131-138), then the best
    'last mile' product of factors <= max_stage."""
    padded = 1
    remaining = np.float32(order)
    while remaining > 64:
        padded *= 4
        remaining = np.float32(remaining / np.float32(4))
    x = int(math.ceil(float(remaining)))
    if x != 1:
        length = 64
        complexity = 3 * (4 + 1)
        factor = find_fft_factors(x, max_stage, 1,
                                  (complexity << 16) + length) & 0xFFFF
        padded *= factor
    return padded


def fft_complexity(n):
    """``_getFFTComplexity`` (core/cmath.pyx:156-214), returned as float32.

    Radix-4 stages cost 5, one leftover factor 2 costs 3, each odd factor p
    costs p + 1 (trial division stops when ii*ii >= nn, so e.g. 9 counts as
    one factor of cost 10), plus one stage; times N.
    """
    complexity = np.float32(0)
    float_n = np.float32(n)
    nn = int(n)
    while nn % 4 == 0:
        complexity += np.float32(5)
        nn //= 4
    if nn > 1 and (nn & 1) == 0:
        complexity += np.float32(3)
        nn //= 2
    ii = 3
    while nn > 1 and ii * ii < nn:
        if nn % ii == 0:
            complexity += np.float32(ii + 1)
            nn //= ii
        else:
            ii += 2
    if nn > 1:
        complexity = np.float32(complexity + np.float32(nn + 1))
    return np.float32(float_n * np.float32(complexity + np.float32(1)))


def fourier_bluestein_length(order, optimize=True, max_stage=4):
    """Return L > 0 when ``Fourier(order)`` uses Bluestein, else 0.

    Fourier.pyx:109-122:
    ``numL = 0 if cplx(N) < 2 cplx(L) + 2 L + 2 N else L`` with
    ``L = _findOptimalFFTSize(2N - 1, maxStage)``.  The first two terms of the
    right side are C float arithmetic; ``2 * self.order`` is a Python int, so
    the last addition and the comparison are in double.
    """
    if not optimize:
        return 0
    padded = find_optimal_fft_size(2 * order - 1, max_stage)
    assert padded >= 2 * order - 1
    lhs = float(fft_complexity(order))
    rhs32 = np.float32(np.float32(2) * fft_complexity(padded)) + \
        np.float32(2 * padded)
    rhs = float(np.float32(rhs32)) + 2 * order
    return 0 if lhs < rhs else padded


def circulant_embedding_length(n, optimize=True, max_stage=4):
    """FFT length used by a 1-D ``Circulant`` of size n (Circulant.pyx:104-137):
    ``_findOptimalFFTSize(2n - 1)`` if that is cheaper than n, else n."""
    if not optimize:
        return n
    padded = find_optimal_fft_size(2 * n - 1, max_stage)
    assert padded >= 2 * n - 1
    return padded if fft_complexity(n) > fft_complexity(padded) else n


def toeplitz_embedding_length(m, n, optimize=True, max_stage=4):
    """FFT length used by a 1-D ``Toeplitz`` of size m x n
    (Toeplitz.pyx:211-227): the defining vector has m + n - 1 entries and is
    padded to ``_findOptimalFFTSize(m + n - 1)`` only when that is cheaper."""
    d = m + n - 1
    if not optimize:
        return d
    opt = find_optimal_fft_size(d, max_stage)
    return opt if fft_complexity(opt) < fft_complexity(d) else d
