"""fastmat_amd: MI355X-native structured matrix-vector products.

A GPU-backed drop-in for the hot path of fastmat (EMS-TU-Ilmenau/fastmat
0.2.2): the ``Matrix`` API (forward, backward, .H/.T/.conj, composition)
over hand-written HIP kernels for gfx950 reached through the C ABI in
include/fastmat_amd.h.  Classes keep fastmat's names and argument meaning.
"""
from .matrix import Matrix, Hermitian, Conjugate, Transpose, flags
from .product import Product
from .sum import Sum
from .partial import Partial
from .diag import Diag
from .eye import Eye
from .permutation import Permutation
from .fourier import Fourier
from .circulant import Circulant, Toeplitz
from .hadamard import Hadamard
from .kron import Kron
from .blocks import BlockDiag, Blocks
from .lfsr import LFSRCirculant

__version__ = "0.1.0"

__all__ = [
    "Matrix", "Hermitian", "Conjugate", "Transpose", "flags", "Product", "Sum",
    "Partial", "Diag", "Eye", "Permutation", "Fourier", "Circulant", "Toeplitz",
    "Hadamard", "Kron", "BlockDiag", "Blocks", "LFSRCirculant",
]
