"""Oracle transport and projection: a NumPy restatement of the reference's operation-matched device mirror.

TEST INFRASTRUCTURE ONLY (the checker of the HIP path and the timed CPU baseline); the product package never imports
it.

What it follows, function by function:
- `transport` = `src/sosfuse2/backends/numpy_reference.py:461-643` (`_transport`): per ray, segments of the compacted
  edge list, Gauss-Legendre nodes with quadratic grading (`sample` :294-387), the Simpson optical-depth prefix
  (`SimpsonPrefix` :236-291), attenuation ds exp(-tau), four bilinear corners with the log-space trace blend
  (:89-103, :610-628), SH reconstruction per (ray, corner cell) (:579-600), the Rayleigh tensor rotation (:327-351),
  first light with sun columns, ground pools and direct beams (:515-568), Kahan accumulation, and the ground boundary
  term (:631-642).
- `project` = `numpy_reference.py:683-711`: weighted FP32 products, ray-order FP64 sums.
- `sun_columns` = `numpy_reference.py:431-458`.

Deliberate differences from the reference, each shared with the HIP kernels (so parity is unaffected):
- profiles are the synthetic analytic shapes of `sosfuse2_amd.atmosphere` (exp, exp, Gaussian, Gaussian, air^2, and a
  Gaussian sulfate layer) on the whole domain, without the 80 km support switch of the reference's extended domain;
- the aerosol asymmetry, albedo and depolarisation are per lane (several states per launch), and the sulfate phase is
  the closed Henyey-Greenstein form (the reference's 33-term Legendre sum of an HG layer differs by < g^33);
- phase 2 may return six tensor components (xx, yy, zz, xz, xy, yz) so that Stokes Q and U exist; the reference
  forms the first four only (`numpy_reference.py:14`).

What pins each part (tests/test_oracle_pins.py, tests/test_warm_pins.py): `Simpson` with `sample`'s path coordinates
and weights by fourth-order convergence to an FP64 integral and the attenuation closure; `sun_columns` by FP64 path
integrals toward the sun; `hg`, the log-space blend, the SH basis and cos(m phi) by closed forms; `project` by the
projection of isotropic light. The first-light assembly of `_transport` is pinned by the textbook single-scattering
integral of a pure Rayleigh atmosphere, and with aerosol, sulfate and a Lambertian ground by the FP64 single-
scattering source at cell centres (direct sun, and the sunlit ground integrated over every direction that sees it).
The warm assembly is pinned by FP64 ray integrals of closed-form sources: an isotropic source (tensor and SH (0, 0)
parts), a zz-only cell tensor (the rotation into the ray frame) and an (l = 1, m = 1) plane coefficient (cos(m phi) at
the node); the boundary term by E exp(-tau_end) on ground-hitting rays.
"""

import math

import numpy as np

from .arith import Arith, TRACE_MIN32, basis, cosine, count, flush_to_zero, kahan, tree_sum

PAIRS4 = ((0, 0), (1, 1), (2, 2), (0, 2))
PAIRS6 = PAIRS4 + ((0, 1), (1, 2))


# Profile kinds of a state's shapes (the product's atmosphere.ProfileSpec codes)
_EXP, _GAUSS, _TRIANGLE, _TABLE, _AIR_SQUARED, _ZERO = 1, 2, 3, 4, 5, 6


def _profile(spec, k, h, air, F):
    """Profile k at FP heights h (>= 0) from a shape spec (kinds, a, b, table): exponential exp(-h / a), Gaussian
    exp(-((h - b) / a)^2 / 2), triangle max(0, 1 - |h - b| / a) (numpy_reference.py:205), table interpolated as
    interpolate32 (x = h / step, i = min(int x, n - 2), t[i] + (x - i)(t[i + 1] - t[i]); density.py:677-690), the
    square of the air profile over its surface value (O4, numpy_reference.py:219-220), zero."""
    kind, a, b = spec.kinds[k], F(spec.a[k]), F(spec.b[k])
    if kind == _EXP:
        return np.exp(-h / a)
    if kind == _GAUSS:
        t = (h - b) / a
        return np.exp(F(-0.5) * (t * t))
    if kind == _TRIANGLE:
        return np.maximum(F(0), F(1) - np.abs(h - b) / a)
    if kind == _TABLE:
        row = np.asarray(spec.table[k], F)
        x = np.maximum(h, F(0)) * (F(1) / a)
        i = np.minimum(x.astype('i8'), len(row) - 2)
        f = x - i.astype(F)
        return row[i] + f * (row[i + 1] - row[i])
    if kind == _AIR_SQUARED:
        return (air * air) / (a * a)
    if kind == _ZERO:
        return np.zeros_like(h)
    raise ValueError(f'E_PROFILE_KIND: {kind}')


def _profiles(h, c, F):
    """Five component profiles and the sulfate profile at FP heights h (profile_at, numpy_reference.py:187-223)."""
    spec = c['profile']['spec']
    air = _profile(spec, 0, h, None, F)
    prof = np.stack([air] + [_profile(spec, k, h, air, F) for k in (1, 2, 3, 4)], -1)
    st = _profile(spec, 5, h, air, F) if c['strato'] else np.zeros_like(h)
    return prof, st


def _extinction(prof, st, c, F):
    """Extinction [n, W] from the five components plus the sulfate layer (numpy_reference.py:226-233)."""
    out = np.zeros((len(prof), c['W']), F)
    for k in range(5):
        out += prof[:, k, None] * c['ext'][k]
    if c['strato']:
        out += st[:, None] * c['strato_ext']
    return out


def hg(u, g, F):
    """Henyey-Greenstein per steradian, u[..., 1] against per-lane g[W] (numpy_reference.py:390-392)."""
    den = np.maximum(F(1e-20), (F(1) + g * g) - (F(2) * g) * u)
    return (F(1) - g * g) / (F(4 * math.pi) * (den * np.sqrt(den)))


class _Ops:
    def __init__(self, F):
        self.F = F
        self.a = Arith(F)

    def path_point(self, pp, dd, t):
        a = self.a
        q = a.add(pp, a.mul((t[0][:, None], t[1][:, None]), dd))
        q2 = a.mul(q, q)
        rr = a.root(a.add(a.add((q2[0][:, 0], q2[1][:, 0]), (q2[0][:, 1], q2[1][:, 1])), (q2[0][:, 2], q2[1][:, 2])))
        return q, q2, rr

    def path_height(self, rr, top):
        a = self.a
        F = self.F
        return np.clip(a.rounded(a.sub(rr, (a.R, F(0)))), F(0), F(top))


class Simpson:
    """Optical depth from the ray start to every node: Simpson panels between consecutive nodes with the extinction
    evaluated anew at segment ends and panel midpoints, Kahan into the ray total at each segment end
    (numpy_reference.py:236-291)."""

    def __init__(self, g, c, ops):
        self.g, self.c, self.o = g, c, ops
        a = ops.a
        self.pp, self.dd = a.split(g['p']), a.split(g['d'])
        self.tau = np.zeros((len(g['d']), c['W']), ops.F)
        self.tc = np.zeros_like(self.tau)

    def beta(self, t):
        h = self.o.path_height(self.o.path_point(self.pp, self.dd, t)[2], self.c['top'])
        prof, st = _profiles(h, self.c, self.o.F)
        return _extinction(prof, st, self.c, self.o.F)

    def piece(self, t1, b1):
        a = self.o.a
        F = self.o.F
        t0, b0 = self.t, self.b
        tm = a.add(t0, t1)
        tm = (tm[0] * F(0.5), tm[1] * F(0.5))
        w = a.rounded(a.sub(t1, t0)) / F(6)
        self.acc = self.acc + np.where(self.ok[:, None], w[:, None] * ((b0 + b1) + F(4) * self.beta(tm)), F(0))
        self.t, self.b = t1, b1

    def begin(self, seg):
        e = self.g['edges']
        a = self.o.a
        self.ok = e[:, seg + 1] - e[:, seg] > 1e-8
        self.end = a.split(e[:, seg + 1])
        self.t = a.split(e[:, seg])
        self.b = self.beta(self.t)
        self.acc = np.zeros_like(self.tau)

    def node(self, q):
        self.piece(q['t'], _extinction(q['prof'], q['st'], self.c, self.o.F))
        return (self.tau - self.tc) + self.acc

    def finish(self):
        if not self.ok.any():
            return
        self.piece(self.end, self.beta(self.end))
        tau, tc = kahan(self.tau, self.tc, self.acc)
        self.tau = np.where(self.ok[:, None], tau, self.tau)
        self.tc = np.where(self.ok[:, None], tc, self.tc)


def sample(g, seg, step, c, ops, st, pairs):
    """Node `step` of segment `seg` for every ray (numpy_reference.py:294-387)."""
    a = ops.a
    F = ops.F
    A_ = a.split(g['edges'][:, seg])
    B_ = a.split(g['edges'][:, seg + 1])
    length = a.sub(B_, A_)
    u = a.split(c['u'][step])
    wu = a.split(c['wu'][step])
    lower = g['lower'][:, seg]
    v1 = a.sub((F(1), F(0)), u)
    v = (np.where(lower, u[0], v1[0]), np.where(lower, u[1], v1[1]))
    tl = a.add(A_, a.mul(length, a.mul(u, u)))
    tr = a.sub(B_, a.mul(length, a.mul(v, v)))
    tt = tuple(np.where(lower, x, y) for x, y in zip(tl, tr))
    ds = a.rounded(a.mul(a.mul(length, v), wu)) * F(2)
    valid = g['edges'][:, seg + 1] - g['edges'][:, seg] > 1e-8
    ds = np.where(valid, ds, F(0))
    pp, dd = a.split(g['p']), a.split(g['d'])
    q, q2, rr = ops.path_point(pp, dd, tt)
    h = ops.path_height(rr, c['top'])
    xy = a.root(a.add((q2[0][:, 0], q2[1][:, 0]), (q2[0][:, 1], q2[1][:, 1])))
    rho = a.rounded(xy)
    xyd = (np.maximum(xy[0], F(1e-30)), xy[1])
    cx = np.where(rho > F(1e-12), a.quotient((q[0][:, 0], q[1][:, 0]), xyd), F(1))
    sx = np.where(rho > F(1e-12), a.quotient((q[0][:, 1], q[1][:, 1]), xyd), F(0))
    angle = np.arctan2(rho, a.rounded((q[0][:, 2], q[1][:, 2])))
    hg_, chi = c['h'], c['chi']
    ir = np.clip(np.searchsorted(hg_, h, side='right') - 1, 0, len(hg_) - 2)
    j = np.clip(np.searchsorted(chi, angle, side='right') - 1, 0, len(chi) - 2)
    ah = np.clip((h - hg_[ir]) / (hg_[ir + 1] - hg_[ir]), F(0), F(1))
    ac = np.clip((angle - chi[j]) / (chi[j + 1] - chi[j]), F(0), F(1))
    ang = chi[j] + ac * (chi[j + 1] - chi[j])
    sn, cc = np.sin(ang), np.cos(ang)
    Q = np.stack([-cc * cx, sx, sn * cx, -cc * sx, -cx, sn * sx, sn, np.zeros_like(sn), cc], -1).reshape(-1, 3, 3)
    B = g['target'].astype(F)
    dloc = g['dloc'].astype(F)
    d = g['d'].astype(F)
    A = np.zeros((len(h), 3, 3), F)
    dot = np.zeros((len(h), 3), F)
    for k in range(3):
        A += B[:, k, :, None] * Q[:, k, None, :]
        dot += d[:, k, None] * Q[:, k]
    Kx = A - dloc[:, :, None] * dot[:, None, :]
    T = np.stack(
        [
            np.stack(
                [Kx[:, i, 0] * Kx[:, k, 0], Kx[:, i, 1] * Kx[:, k, 1], Kx[:, i, 2] * Kx[:, k, 2],
                 Kx[:, i, 0] * Kx[:, k, 2] + Kx[:, i, 2] * Kx[:, k, 0]],
                -1,
            )
            for i, k in pairs
        ],
        1,
    )
    cp = np.where(st > F(1e-12), (d[:, 0] * cx + d[:, 1] * sx) / np.maximum(st, F(1e-30)), F(1))
    cm = cosine(cp, c['L'] + 1, F)
    prof, stp = _profiles(h, c, F)
    nc = len(chi)
    cells = np.stack([ir * nc + j, ir * nc + j + 1, (ir + 1) * nc + j, (ir + 1) * nc + j + 1], -1)
    sw = np.stack([(F(1) - ah) * (F(1) - ac), (F(1) - ah) * ac, ah * (F(1) - ac), ah * ac], -1)
    return dict(h=h, ds=ds, cells=cells, sw=sw, ir=ir, T=T, cm=cm, prof=prof, st=stp, Q=Q, d=d, valid=valid,
                r=rr, qz=(q[0][:, 2], q[1][:, 2]), t=tt)


def ftz_value(a, F):
    """A two-word value as an FTZ build sees it: components and their sum below FLT_MIN are 0 (numpy_reference.py:81-86)."""
    h, l = (np.where(np.abs(x) < TRACE_MIN32, F(0), x) for x in a)
    v = h + l
    return np.where(np.abs(v) < TRACE_MIN32, F(0), v)


def sun_columns(q, c, ops):
    """Sun slant columns (5 components + sulfate) and the lit flag at the nodes (numpy_reference.py:431-458)."""
    a = ops.a
    F = ops.F
    R = a.R
    hh = np.clip(q['h'], F(0), F(c['top']))
    height = a.sub(q['r'], (R, F(0)))
    height = tuple(np.where(ftz_value(height, F) > F(0), x, F(0)) for x in height)
    rad = a.add((R, F(0)), height)
    mg = a.neg(a.root(a.divide(a.mul(height, a.add((F(2) * R, F(0)), height)), a.mul(rad, rad))))
    delta = a.sub(a.divide(q['qz'], q['r']), mg)
    zero = ftz_value(delta, F) == F(0)
    delta = tuple(np.where(zero, F(0), x) for x in delta)
    lit = a.rounded(delta) >= F(0)
    delta = tuple(np.where(lit, x, F(0)) for x in delta)
    x = np.clip(a.rounded(a.root(a.divide(delta, a.sub((F(1), F(0)), mg)))), F(0), F(1))
    sh, scol = c['sun_h'], c['sun_col']
    i = np.clip(np.searchsorted(sh, hh, side='right') - 1, 0, len(sh) - 2)
    al = (hh - sh[i]) / (sh[i + 1] - sh[i])
    f = x * F(scol.shape[1] - 1)
    j = np.clip(f.astype('i4'), 0, scol.shape[1] - 2)
    b = f - j.astype(F)
    col = ((F(1) - al) * (F(1) - b))[:, None] * scol[i, j]
    col += (al * (F(1) - b))[:, None] * scol[i + 1, j]
    col += ((F(1) - al) * b)[:, None] * scol[i, j + 1]
    col += (al * b)[:, None] * scol[i + 1, j + 1]
    return col, lit


def trace_accumulate(tr, weight, logs, wg, am, F):
    good = tr >= TRACE_MIN32
    t = np.where(good, tr, F(1))
    w = np.where(good, weight, F(0))
    logs += w * np.log(t)
    wg += w
    am += w * t


def trace_scale(val, logs, wg, am, F):
    ok = (wg > F(0)) & (am >= TRACE_MIN32)
    mean = np.where(ok, am, F(1)) / np.where(ok, wg, F(1))
    scale = np.exp(np.where(ok, logs - wg * np.log(mean), F(0)))
    return val * scale[:, :, None]


def combined_plane(field, c, F):
    """SH plane with both aerosol layers folded in by their densities at each cell's height:
    field[:, 10:10+K] ha[ih] + field[:, 10+K:] sd[ih] (`src/sosfuse2/backends/sh_layout.py:135-152`)."""
    K = c['K']
    nc = len(c['chi'])
    ha = np.repeat(c['ha'], nc)[:, None, None]
    out = field[:, 10 : 10 + K] * ha
    if c['strato']:
        out = out + field[:, 10 + K :] * np.repeat(c['strato_density'], nc)[:, None, None]
    return out.astype(F)


def _transport(g, c, field, first, boundary, pairs, stats):
    F = c['F']
    ops = _Ops(F)
    n, W, L = len(g['d']), c['W'], c['L']
    K = count(L)
    P = len(pairs)
    Y = basis(g['d'][:, 2], L, F)
    d = g['d'].astype(F)
    dloc = g['dloc'].astype(F)
    pu = np.stack([(F(1) if i == j else F(0)) - dloc[:, i] * dloc[:, j] for i, j in pairs], -1)
    u = g['target'].astype(F)[:, 2]
    dotu = (dloc[:, 0] * u[:, 0] + dloc[:, 1] * u[:, 1]) + dloc[:, 2] * u[:, 2]
    e = u - dloc * dotu[:, None]
    direct = np.stack([(pu[:, k] - e[:, i] * e[:, j]) / F(2) for k, (i, j) in enumerate(pairs)], -1)
    st = np.sqrt(np.maximum(0.0, 1 - g['d'][:, 2] ** 2)).astype(F)
    total = np.zeros((n, W, P), F)
    corr = np.zeros_like(total)
    prefix = Simpson(g, c, ops)
    cache = np.zeros((n, 4, L + 1, W), F)
    tags = np.full((n, 4), -1, 'i8')
    gl_ = c['g']
    mix = c['mix']
    rows = np.arange(n)
    phase = hg(d[:, 2, None], gl_, F)
    strat = c['strato']
    if strat:
        sphase = hg(d[:, 2, None], c['strato_g'], F)
    plane = None
    if not first:
        plane = combined_plane(field, c, F) if c['combine'] else field[:, 10 : 10 + K].astype(F)
        if strat and not c['combine']:
            splane = field[:, 10 + K : 10 + 2 * K].astype(F)
            scache = np.zeros_like(cache)
    src = None if first else field.astype(F)
    for seg in range(g['edges'].shape[1] - 1):
        prefix.begin(seg)
        for step in range(len(c['u'])):
            q = sample(g, seg, step, c, ops, st, pairs)
            if not q['valid'].any():
                continue
            stats['samples'] += int(q['valid'].sum())
            mid = prefix.node(q)
            att = q['ds'][:, None] * np.exp(-mid)
            val = np.zeros((n, W, P), F)
            logs = np.zeros((n, W), F)
            wg = np.zeros_like(logs)
            am = np.zeros_like(logs)
            if first:
                sc, lit = sun_columns(q, c, ops)
                optical = np.zeros((n, W), F)
                for k in range(5):
                    optical += sc[:, k, None] * c['ext'][k]
                if strat:
                    optical += sc[:, 5, None] * c['strato_ext']
                sun = np.exp(-optical) * lit[:, None]
                ground = np.zeros((n, W), F)
                sground = np.zeros((n, W), F)
                mu = np.zeros(n, F)
                exd = np.zeros(n, F)
                eyd = np.zeros(n, F)
                for k in range(3):
                    mu += q['d'][:, k] * q['Q'][:, k, 2]
                    exd += q['d'][:, k] * q['Q'][:, k, 0]
                    eyd += q['d'][:, k] * q['Q'][:, k, 1]
                for corner in range(4):
                    cells = q['cells'][:, corner]
                    gd = c['gd'][cells]
                    m0 = mu[:, None] * gd[:, :, 2] + exd[:, None] * gd[:, :, 0]
                    m1 = abs(eyd)[:, None] * gd[:, :, 1]
                    up = np.clip(m0 + m1, F(-1), F(1))[..., None]
                    dn = np.clip(m0 - m1, F(-1), F(1))[..., None]
                    wph = q['sw'][:, corner, None, None] * ((hg(up, gl_, F) + hg(dn, gl_, F)) / F(2))
                    if strat:
                        sph = q['sw'][:, corner, None, None] * ((hg(up, c['strato_g'], F) + hg(dn, c['strato_g'], F)) / F(2))
                    gv = c['gv'][cells]
                    for gi in range(gd.shape[1]):
                        ground += wph[:, gi] * gv[:, gi]
                        if strat:
                            sground += sph[:, gi] * gv[:, gi]
                beam = (q['prof'][:, 1, None] * c['asca']) * (sun * phase + ground)
                ray = (sun * q['prof'][:, 0, None]) * c['ext'][0]
                for k in range(P):
                    val[:, :, k] = ray * (mix * F(3 / (8 * math.pi)) * direct[:, k, None]
                                          + (F(1) - mix) * F(1 / (8 * math.pi)) * pu[:, k, None]) + beam * pu[:, k, None] / F(2)
                    if strat:
                        sbeam = (q['st'][:, None] * c['strato_sca']) * (sun * sphase + sground)
                        val[:, :, k] += (sbeam * pu[:, k, None]) / F(2)
            for corner in range(4):
                cell = q['cells'][:, corner]
                ir = q['ir'] + corner // 2
                if first:
                    rm = c['gm'][cell].transpose(0, 2, 1)
                    aero = np.zeros((n, W), F)
                    density = q['prof'][:, 0][:, None]
                else:
                    change = (tags[:, corner] != cell) & q['valid']
                    if change.any():
                        rr = rows[change]
                        cc = cell[change]
                        cache[rr, corner] = F(0)
                        if strat and not c['combine']:
                            scache[rr, corner] = F(0)
                        for l in range(L + 1):
                            for m in range(l + 1):
                                k = l * (l + 1) // 2 + m
                                cache[rr, corner, m] += Y[rr, k, None] * plane[cc, k]
                                if strat and not c['combine']:
                                    scache[rr, corner, m] += Y[rr, k, None] * splane[cc, k]
                        tags[rr, corner] = cc
                        stats['angular_pairs'] += len(rr)
                    aero = tree_sum(q['cm'][:, :, None] * cache[:, corner])
                    if not c['combine']:
                        aero = aero * c['ha'][ir, None]
                        if strat:
                            aero = aero + tree_sum(q['cm'][:, :, None] * scache[:, corner]) * c['strato_density'][ir, None]
                    rm = src[cell, :4].transpose(0, 2, 1)
                    density = np.broadcast_to(c['hr'][ir][:, None], (n, 1))
                trace = (rm[:, :, 0] + rm[:, :, 1]) + rm[:, :, 2]
                tensor = np.zeros((n, W, P), F)
                for k in range(P):
                    rotated = np.zeros((n, W), F)
                    for uu in range(4):
                        rotated += q['T'][:, k, uu, None] * rm[:, :, uu]
                    tensor[:, :, k] = density * (mix * rotated + ((F(1) - mix) / F(3) * trace) * pu[:, k, None])
                    tensor[:, :, k] += (aero * pu[:, k, None]) / F(2)
                    val[:, :, k] += q['sw'][:, corner, None] * tensor[:, :, k]
                if not first:
                    tr = (tensor[:, :, 0] + tensor[:, :, 1]) + tensor[:, :, 2]
                    trace_accumulate(tr, q['sw'][:, corner, None], logs, wg, am, F)
            if not first:
                val = trace_scale(val, logs, wg, am, F)
            total, corr = kahan(total, corr, val * att[:, :, None])
        prefix.finish()
    if boundary is not None:
        ef = g['end_fraction'].astype(F)[:, None]
        bnd = boundary.astype(F)
        E = (F(1) - ef) * bnd[g['end_index']] + ef * bnd[g['end_index'] + 1]
        endcol = g['end_columns'].astype(F)
        optical = np.zeros((n, W), F)
        for k in range(5):
            optical += endcol[:, k, None] * c['ext'][k]
        if strat:
            optical += endcol[:, 5, None] * c['strato_ext']
        E = E * (np.exp(-optical) * g['hits_ground'][:, None].astype(F))
        for k in range(P):
            total[:, :, k] += (E * pu[:, k, None]) / F(2)
    return total


def transport(g, c, field, first=False, boundary=None, pairs=PAIRS4, stats=None):
    """Per-ray radiance tensor [n, W, len(pairs)] in the target frame. FP32 runs under flush-to-zero."""
    stats = {} if stats is None else stats
    stats.setdefault('samples', 0)
    stats.setdefault('angular_pairs', 0)
    with flush_to_zero(c['F'] is np.float32):
        return _transport(g, c, field, first, boundary, pairs, stats)


def project(g, radiance, c, ground_only=False):
    """Moments [4 + K, W] and ground flux [W] of one cell's rays, FP64 sums in ray order (numpy_reference.py:683-711)."""
    F = c['F']
    with flush_to_zero(F is np.float32):
        L = c['L']
        K = count(L)
        Y = basis(g['d'][:, 2], L, F)
        st = np.sqrt(np.maximum(0, 1 - g['d'][:, 2] ** 2))
        cp = np.where(st > 1e-12, g['d'][:, 0] / np.maximum(st, 1e-30), 1.0).astype(F)
        cm = cosine(cp, L + 1, F)
        weight = g['weight'].astype(F)
        trace = (radiance[:, :, 0] + radiance[:, :, 1]) + radiance[:, :, 2]
        weighted = trace * weight[:, None]
        out = np.zeros((4 + K, c['W']), 'f8')
        if not ground_only:
            for k in range(4):
                out[k] = np.sum(radiance[:, :, k] * weight[:, None], axis=0, dtype='f8')
            for l in range(L + 1):
                for m in range(l + 1):
                    k = l * (l + 1) // 2 + m
                    out[4 + k] = np.sum(weighted * (Y[:, k] * cm[:, m])[:, None], axis=0, dtype='f8')
        ground = np.sum((trace * (weight * np.maximum(F(0), g['dloc'][:, 2].astype(F)))[:, None]).astype('f8'), axis=0)
        ground = ground * np.asarray(c['albedo'], 'f8') / math.pi
    return out, ground
