"""Oracle arithmetic: FP32 flush-to-zero mode, double-single (two-word FP32) operations, the SH basis recurrence,
cos(m phi), the 32-lane tree sum and Kahan steps.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product package.

`Arith(np.float32)` restates `src/sosfuse2/backends/arithmetic.py:15-66` (DS add/sub/mul/root/divide with an exactly
rounded product error) and runs under the x86-64 MXCSR FTZ+DAZ mode of `numpy_reference.py:19-78`.
`Arith(np.float64)` is the FP64 twin: the same call sequence with one-word FP64 values (the low word is 0), used as the
accuracy oracle.
"""

import contextlib
import ctypes
import ctypes.util
import math

import numpy as np

TRACE_MIN32 = np.float32(np.finfo(np.float32).tiny)


class _FenvX86(ctypes.Structure):
    _fields_ = [('x87', ctypes.c_uint8 * 28), ('mxcsr', ctypes.c_uint32)]


_LIBM = None


@contextlib.contextmanager
def flush_to_zero(active=True):
    """FP32 FTZ+DAZ for the block (numpy_reference.py:54-78). Stops with E_FTZ_INACTIVE if subnormals survive."""
    global _LIBM
    if not active:
        yield
        return
    if _LIBM is None:
        _LIBM = ctypes.CDLL(ctypes.util.find_library('m') or ctypes.util.find_library('c'))
    old, new = _FenvX86(), _FenvX86()
    if _LIBM.fegetenv(ctypes.byref(old)) != 0:
        raise RuntimeError('E_FTZ_FENV: fegetenv failed')
    ctypes.memmove(ctypes.byref(new), ctypes.byref(old), ctypes.sizeof(_FenvX86))
    new.mxcsr |= 0x8040
    if _LIBM.fesetenv(ctypes.byref(new)) != 0:
        raise RuntimeError('E_FTZ_FENV: fesetenv failed')
    try:
        tiny = TRACE_MIN32
        if tiny * np.float32(0.5) != 0 or (np.full(64, tiny, 'f4') * np.float32(0.5))[0] != 0:
            raise RuntimeError('E_FTZ_INACTIVE: subnormal FP32 values survived')
        yield
    finally:
        _LIBM.fesetenv(ctypes.byref(old))


class Arith:
    """Two-word arithmetic in dtype F. For FP64 the low word is identically zero and every operation is one FP64
    operation, so the same transport code yields the FP64 twin."""

    def __init__(self, dtype):
        self.F = np.dtype(dtype).type
        self.ds = self.F is np.float32
        self.R = self.F(6371)

    def split(self, x):
        x = np.asarray(x, dtype=np.float64)
        hi = x.astype(self.F)
        if not self.ds:
            return hi, np.zeros_like(hi)
        return hi, (x - hi.astype('f8')).astype('f4')

    def add(self, a, b):
        if not self.ds:
            return a[0] + b[0], np.zeros_like(a[0] + b[0])
        s = a[0] + b[0]
        v = s - a[0]
        e = ((a[0] - (s - v)) + (b[0] - v)) + (a[1] + b[1])
        h = s + e
        return h, e - (h - s)

    def neg(self, a):
        return -a[0], -a[1]

    def sub(self, a, b):
        return self.add(a, self.neg(b))

    def mul(self, a, b):
        if not self.ds:
            p = a[0] * b[0]
            return p, np.zeros_like(p)
        p = a[0] * b[0]
        # the exactly rounded error of the FP32 product (an FP32 fma), as arithmetic.py:37-43
        e = (a[0].astype('f8') * b[0].astype('f8') - p.astype('f8')).astype('f4')
        e = (e + a[0] * b[1]) + a[1] * b[0]
        h = p + e
        return h, e - (h - p)

    def root(self, a):
        F = self.F
        h = np.sqrt(np.maximum(a[0], F(0)))
        if not self.ds:
            return h, np.zeros_like(h)
        z = np.zeros_like(h)
        e = self.sub(a, self.mul((h, z), (h, z)))
        lo = (e[0] + e[1]) / np.maximum(F(2) * h, F(1e-30))
        return self.add((h, z), (lo, np.zeros_like(lo)))

    def divide(self, a, b):
        h = a[0] / b[0]
        if not self.ds:
            return h, np.zeros_like(h)
        z = np.zeros_like(h)
        e = self.sub(a, self.mul((h, z), b))
        lo = (e[0] + e[1]) / b[0]
        return self.add((h, z), (lo, np.zeros_like(lo)))

    def rounded(self, a):
        return a[0] + a[1]

    def quotient(self, a, b):
        return self.rounded(self.divide(a, b))


def count(L):
    if L not in (8, 16, 24, 32):
        raise ValueError('E_SH_L: expected 8, 16, 24 or 32')
    return (L + 1) * (L + 2) // 2


def norm_constants(L):
    """Recurrence constants of the normalised associated Legendre functions (arithmetic.py:69-83)."""
    K = count(L)
    aa = np.zeros(K)
    bb = np.zeros(K)
    dd = np.zeros(L + 1)
    for m in range(1, L + 1):
        dd[m] = -math.sqrt((2.0 * m + 1) / (2 * m))
    for l in range(1, L + 1):
        for m in range(l):
            k = l * (l + 1) // 2 + m
            aa[k] = math.sqrt((4.0 * l * l - 1) / (l * l - m * m))
            if l > m + 1:
                bb[k] = math.sqrt((2.0 * l + 1) * ((l - 1.0) ** 2 - m * m) / ((2.0 * l - 3) * (l * l - m * m)))
    return aa, bb, dd


def basis(mu, L, dtype=np.float32):
    """Real SH (cosine half) at azimuth 0, k = l(l+1)/2 + m, computed in FP64 and rounded (arithmetic.py:86-109)."""
    mu = np.clip(np.asarray(mu, 'f8'), -1, 1)
    st = np.sqrt(np.maximum(0.0, 1 - mu * mu))
    aa, bb, dd = norm_constants(L)
    out = np.empty((len(mu), count(L)), dtype)
    diag = np.full_like(mu, 1 / math.sqrt(4 * math.pi))
    for m in range(L + 1):
        if m:
            diag = (diag * dd[m]) * st
        mult = math.sqrt(2.0) if m else 1.0
        z0 = diag
        out[:, m * (m + 1) // 2 + m] = mult * z0
        if m < L:
            z1 = (math.sqrt(2.0 * m + 3) * mu) * z0
            out[:, (m + 1) * (m + 2) // 2 + m] = mult * z1
            for l in range(m + 2, L + 1):
                k = l * (l + 1) // 2 + m
                z2 = (aa[k] * mu) * z1 - bb[k] * z0
                out[:, k] = mult * z2
                z0, z1 = z1, z2
    return out


def cosine(cp, M, F=np.float32):
    """cos(m phi) for m = 0..M-1 by the Chebyshev recurrence in F (arithmetic.py:112-119)."""
    cp = np.clip(np.asarray(cp, F), F(-1), F(1)).astype(F)
    out = np.empty((len(cp), M), F)
    out[:, 0] = F(1)
    if M > 1:
        out[:, 1] = cp
    for m in range(2, M):
        out[:, m] = (F(2) * cp) * out[:, m - 1] - out[:, m - 2]
    return out


def tree_sum(x):
    """Sum over axis 1 of x[:, M, ...] in the 32-lane shuffle-tree order (arithmetic.py:122-130): lanes >= 32 fold
    into lane 0 first, then strides 16, 8, 4, 2, 1. For M <= 32 the missing lanes are 0."""
    x = x.copy()
    M = x.shape[1]
    if M < 32:
        pad = np.zeros((x.shape[0], 32 - M) + x.shape[2:], x.dtype)
        x = np.concatenate([x, pad], axis=1)
    if x.shape[1] > 32:
        x[:, 0] = x[:, 0] + x[:, 32]
        x = x[:, :32]
    for stride in (16, 8, 4, 2, 1):
        old = x.copy()
        x[:, : 32 - stride] = old[:, : 32 - stride] + old[:, stride:]
    return x[:, 0]


def kahan(total, correction, value):
    """One Kahan step (arithmetic.py:133-136)."""
    y = value - correction
    t = total + y
    return t, (t - total) - y
