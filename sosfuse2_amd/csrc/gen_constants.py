"""Write sh_constants.h: the normalised associated-Legendre recurrence constants as FP32 literals, per L.

The recurrence is the reference's (`src/sosfuse2/backends/arithmetic.py:69-109`): diagonal seed
1/sqrt(4 pi) * prod dd[m] st, first step sqrt(2m+3) mu, then z_l = aa[k] mu z_{l-1} - bb[k] z_{l-2}.
"""

import math
import sys
from pathlib import Path


def constants(L):
    K = (L + 1) * (L + 2) // 2
    aa = [0.0] * K
    bb = [0.0] * K
    dd = [0.0] * (L + 1)
    for m in range(1, L + 1):
        dd[m] = -math.sqrt((2.0 * m + 1) / (2 * m))
    for l in range(1, L + 1):
        for m in range(l):
            k = l * (l + 1) // 2 + m
            aa[k] = math.sqrt((4.0 * l * l - 1) / (l * l - m * m))
            if l > m + 1:
                bb[k] = math.sqrt((2.0 * l + 1) * ((l - 1.0) ** 2 - m * m) / ((2.0 * l - 3) * (l * l - m * m)))
    sq = [math.sqrt(2.0 * m + 3) for m in range(L + 1)]
    return aa, bb, dd, sq


def fmt(v):
    return f'{float(v):.9e}f'


def main(out):
    lines = ['// Generated by gen_constants.py: SH recurrence constants (arithmetic.py:69-109 of the reference).',
             '#pragma once', '']
    for L in (8, 16, 24, 32):
        aa, bb, dd, sq = constants(L)
        for name, arr in (('AA', aa), ('BB', bb), ('DD', dd), ('SQ', sq)):
            body = ', '.join(fmt(v) for v in arr)
            lines.append(f'__constant__ float SH{name}_{L}[{len(arr)}] = {{{body}}};')
        lines.append('')
    lines += ['template <int L> struct ShConst;']
    for L in (8, 16, 24, 32):
        lines.append(
            f'template <> struct ShConst<{L}> {{ __device__ static const float* aa() {{ return SHAA_{L}; }} '
            f'__device__ static const float* bb() {{ return SHBB_{L}; }} '
            f'__device__ static const float* dd() {{ return SHDD_{L}; }} '
            f'__device__ static const float* sq() {{ return SHSQ_{L}; }} }};')
    Path(out).write_text('\n'.join(lines) + '\n')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else Path(__file__).with_name('sh_constants.h'))
