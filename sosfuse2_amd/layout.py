"""SH and source-field layout (`src/sosfuse2/backends/sh_layout.py:11-18,120-123`).

The source field of one solve chunk is f32[cells, rows, W] with rows = 10 + K (one aerosol SH plane) or 10 + 2K (a
sulfate plane after it), K = (L+1)(L+2)/2 cosine-only coefficients with k = l(l+1)/2 + m:
rows 0-3 the Rayleigh tensor source (xx, yy, zz, xz), rows 4-9 the first six SH moments times the Rayleigh
eigenvalues, rows 10.. the aerosol planes (`src/sosfuse2/solve/source_iteration.py:176-208`).
"""

import numpy as np

SUPPORTED_L = (8, 16, 24, 32)


def count(L):
    if type(L) is not int or L not in SUPPORTED_L:
        raise ValueError('E_SH_L: expected 8, 16, 24 or 32')
    return (L + 1) * (L + 2) // 2


def source_rows(L, strato):
    return 10 + count(L) * (2 if strato else 1)


def degrees(L):
    """l of every coefficient k."""
    return np.array([l for l in range(L + 1) for m in range(l + 1)])


def orders(L):
    """m of every coefficient k."""
    return np.array([m for l in range(L + 1) for m in range(l + 1)])


def rayleigh_factors():
    """Eigenvalues of rows 4-9: 1 for l = 0, 0 for l = 1, 0.1 for l = 2 (`src/sosfuse2/solve/engine_core.py:93-94`)."""
    lv = degrees(2)
    return (lv == 0) + 0.1 * (lv == 2)
