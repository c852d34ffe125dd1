"""Pins of the oracle's warm-transport assembly, its boundary term, its first-light ground term and its composition
factors, each against an independent FP64 evaluation of what the mathematics fixes.

The oracle (`oracle/transport.py`, `oracle/solve.py`) restates the reference's NumPy mirror
(`src/sosfuse2/backends/numpy_reference.py:467-643`) and its composition (`src/sosfuse2/solve/source_iteration.py:
176-212`). The GPU parity tests compare the HIP path with it, so a slip shared by both would go unseen there; these
checks compute the answers from the physics instead:

- an isotropic warm source (cell tensors S / hr I / 3 and an SH plane with only its (0, 0) coefficient) gives every
  ray (S / 3 + Y00 c00 / 2) (1 - d d^T) int e^-tau ds, whatever the corner blend and the frames;
- a zz-only cell tensor rotates into Z (P r)(P r)^T along the ray (r the local vertical, P the projector normal to
  the ray), and an (l = 1, m = 1) coefficient follows Y11(mu) cos(phi - phi_node);
- the ground boundary term is E exp(-tau_end) (1 - d d^T) / 2 on ground-hitting rays;
- first light at a cell centre is the single-scattering source of the direct sun and of the sunlit Lambertian ground,
  integrated over every direction that sees the ground;
- composition followed by reconstruction obeys the Funk-Hecke theorem: incident radiance P_l(mu) scattered by a
  Henyey-Greenstein layer gives g^l P_l(mu) asca, by Rayleigh scattering 0.1 mix P_2(mu) ext_R; the Rayleigh rows'
  factors are the Legendre moments of the Rayleigh phase function.

The optical-depth integrals here are Gauss-Legendre on many pieces of every segment in FP64, far finer than the
oracle's node quadrature, and use only the state's profile shapes and extinction coefficients.
"""

import math
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import transport as ot
from oracle.arith import count
from oracle.solve import composition_factors, compose
from sosfuse2_amd import api, atmosphere, geometry

R = 6371.0
PAIRS6 = ot.PAIRS6
NQ = 16  # oracle nodes per segment in these pins (the oracle's own quadrature error stays far below the bounds)


# ------------------------------------------------------------------------------------------------ FP64 physics
def _gauss(h, c, s):
    t = (h - c) / s
    return np.exp(-0.5 * t * t)


def beta64(c, h):
    """Extinction [..., W] at heights h (km) from the state's profile shapes (exponential air and aerosol, Gaussian
    ozone, NO2 and sulfate, O4 as air squared) and its per-lane coefficients, in FP64."""
    p = c['profile']
    h = np.clip(np.asarray(h, 'f8'), 0.0, c['top'])[..., None]
    air = np.exp(-h / p['air_H'])
    out = (air * c['ext'][0] + np.exp(-h / p['aer_H']) * c['ext'][1] + _gauss(h, p['o3_c'], p['o3_s']) * c['ext'][2]
           + _gauss(h, p['no2_c'], p['no2_s']) * c['ext'][3] + air * air * c['ext'][4])
    if c['strato']:
        out = out + _gauss(h, p['st_c'], p['st_s']) * c['strato_ext']
    return out


def height(q):
    return np.linalg.norm(q, axis=-1) - R


def tau64(p, d, s0, s1, c, pieces=24, n=16):
    """Optical depth [rays, W] along p + s d from s0 to s1 (arrays over rays): Gauss-Legendre n on `pieces` equal
    pieces."""
    x, w = np.polynomial.legendre.leggauss(n)
    x, w = (x + 1) / 2, w / 2
    L = (s1 - s0)[:, None, None]
    a = s0[:, None, None] + L * np.arange(pieces)[None, :, None] / pieces
    s = a + L / pieces * x[None, None, :]
    q = p[:, None, None, :] + s[..., None] * d[:, None, None, :]
    b = beta64(c, height(q))
    return np.einsum('rpnw,n->rw', b, w) * (L[:, 0] / pieces)


def ray_integral(g, c, f, pieces=16, n=12):
    """int_0^end exp(-tau(0, s)) f(s) ds per ray [rays, W, P], FP64, over the pack's segments (the edges split the
    ray at the shells, the tangent point and the shadow boundary, where the integrand is smooth on each side).
    f(s, q) takes path coordinates [rays, m] and points [rays, m, 3] and returns [rays, m, W, P]."""
    x, w = np.polynomial.legendre.leggauss(n)
    x, w = (x + 1) / 2, w / 2
    E = g['edges']
    p, d = g['p'], g['d']
    rays = len(E)
    tau0 = np.zeros((rays, c['W']))
    total = 0.0
    for seg in range(E.shape[1] - 1):
        A, B = E[:, seg], E[:, seg + 1]
        for k in range(pieces):
            a = A + (B - A) * k / pieces
            L = (B - A) / pieces
            s = a[:, None] + L[:, None] * x[None, :]  # outer nodes [rays, n]
            # optical depth from the piece start to every outer node: n inner nodes on [a, s]
            si = a[:, None, None] + (s - a[:, None])[:, :, None] * x[None, None, :]
            qi = p[:, None, None, :] + si[..., None] * d[:, None, None, :]
            part = np.einsum('ronw,n->row', beta64(c, height(qi)), w) * (s - a[:, None])[..., None]
            q = p[:, None, :] + s[..., None] * d[:, None, :]
            val = f(s, q) * np.exp(-(tau0[:, None, :] + part))[..., None]
            total = total + np.einsum('rmwp,m->rwp', val, w) * L[:, None, None]
            qb = p[:, None, :] + s[..., None] * d[:, None, :]
            tau0 = tau0 + np.einsum('rmw,m->rw', beta64(c, height(qb)), w) * L[:, None]
    return total


def hg64(u, g):
    den = 1 + g * g - 2 * g * u
    return (1 - g * g) / (4 * math.pi * den * np.sqrt(den))


def y11(mu):
    """Cosine-only real harmonic (l = 1, m = 1) without its cos(phi): sqrt(3 / (4 pi)) P_1^1(mu), with the Condon-
    Shortley phase of scipy's lpmv (the basis convention pinned in test_oracle_pins)."""
    from scipy.special import lpmv

    return math.sqrt(3 / (4 * math.pi)) * lpmv(1, 1, mu)


# --------------------------------------------------------------------------------------------------- contexts
def _context(state, L=16, nr=5, nmu=2, nphi=4):
    cfg = dict(nr=nr, chi_grid='tiny', nmu=nmu, nphi=nphi, path_n=4, L=L, top_km=100.0, observer_km=0.1,
               query_n=6, anderson_depth=4, sun_scale=1, workers=1)
    scope = api.Scope(cfg, state, SimpleNamespace(L=L))
    return scope.context(atmosphere.lane_optics(state, np.array([450.0, 550.0, 650.0]))), scope


def _pool_context(state, qscale, L=8):
    """A context whose ground pools use qscale times the incident quadrature (32 qscale azimuths, 3 qscale nodes
    per mu interval)."""
    from sosfuse2_amd import engine

    grid = api.make_grid(dict(nr=5, chi_grid='tiny', nmu=2, nphi=4, path_n=4, L=L, top_km=100.0))
    sun = geometry.sun_table(state, grid.domain, 1)
    pools = geometry.ground_pools(state, grid, sun, qscale=qscale)
    optics = atmosphere.lane_optics(state, np.array([450.0, 550.0, 650.0]))
    return engine.Context(grid, optics, SimpleNamespace(L=L), sun, pools)


def make_warm_case():
    """The warm pins' state, context, scope and a concatenated pack of rays sampled from four cells."""
    st = atmosphere.State(aerosol_height_km=4.0, sulfate_aod550=0.02, no2_column=5e15, o4=True, aod550=0.15)
    ctx, scope = _context(st)
    nc = ctx.geo.nc
    picks = []
    for cell in (1 * nc + 1, 2 * nc + 3, 3 * nc + 6, 4 * nc + 8):
        g = scope.packs[cell]
        picks.append({k: (v[::max(1, len(g['d']) // 9)] if isinstance(v, np.ndarray) and v.shape[:1] == (len(g['d']),)
                          else v) for k, v in g.items()})
    g = geometry.concat_packs(picks)
    g['end_columns'] = np.concatenate([p['end_columns'] for p in picks])
    return ctx, scope, g


@pytest.fixture(scope='module')
def warm_case():
    return make_warm_case()


def _pu(g):
    dl = g['dloc']
    return np.stack([(1.0 if i == j else 0.0) - dl[:, i] * dl[:, j] for i, j in PAIRS6], -1)


def _compare(value, truth, bound):
    err = float(np.max(np.abs(value - truth)) / np.max(np.abs(truth)))
    assert err < bound, err
    return err


# --------------------------------------------------------------------------------------------------- (i) isotropic
@pytest.mark.parametrize('dtype, bound', [(np.float64, 2e-6), (np.float32, 2e-5)])
def test_isotropic_warm_source(warm_case, dtype, bound):
    """Cell tensors S / hr(h) I / 3 and an SH plane holding only c00 / ha(h): every corner carries the same isotropic
    source, the rotation of an isotropic tensor is the projector (1 - d d^T) whatever the frames, the log blend of
    equal traces is the identity, so every ray gives (S / 3 + Y00 c00 / 2) (1 - d d^T)_k int_0^end e^-tau ds.
    The tensor part and the SH part are checked separately: a wrong SH factor (1/3 for 1/2) or a permuted rotation
    index moves them by O(1)."""
    ctx, scope, g = warm_case
    c = ctx.constants(NQ, dtype)
    c64 = ctx.constants(NQ, np.float64)
    W, K, ns = c['W'], count(c['L']), ctx.geo.ns
    ih = np.arange(ns) // ctx.geo.nc
    S = np.array([2e-3, 5e-3, 1e-3])
    c00 = np.array([4e-3, 1e-3, 3e-3])
    atten = ray_integral(g, c64, lambda s, q: np.ones(s.shape + (c['W'], 1)))[..., 0]
    pu = _pu(g)
    y00 = 1 / math.sqrt(4 * math.pi)
    for part in ('tensor', 'sh', 'both'):
        field = np.zeros((ns, ctx.rows, W), dtype)
        s_, a_ = (S if part != 'sh' else 0 * S), (c00 if part != 'tensor' else 0 * c00)
        field[:, 0:3] = (s_[None, :] / 3 / c64['hr'][ih][:, None])[:, None, :]
        field[:, 10] = a_[None, :] / c64['ha'][ih][:, None]
        v = ot.transport(g, c, field, False, None, PAIRS6)
        truth = (s_ / 3 + y00 * a_ / 2)[None, :, None] * pu[:, None, :] * atten[:, :, None]
        _compare(v.astype('f8'), truth, bound)


# ---------------------------------------------------------------------------------------------------- (ii) rotation
@pytest.mark.parametrize('dtype, bound', [(np.float64, 2e-6), (np.float32, 3e-5)])
def test_rotation_and_order_one_harmonic(warm_case, dtype, bound):
    """A zz-only cell tensor Z / hr(h) e_z e_z^T (the local vertical of each cell) seen along a ray rotates into
    Z [mix (P r)(P r)^T + (1 - mix) / 3 P] in the ray's target frame, r the unit vertical at the node and P = 1 - d d^T;
    an SH plane holding c00 and the (l = 1, m = 1) coefficient c11 reconstructs to Y00 c00 + c11 Y11(mu_d)
    cos(phi_d - phi_node) at every node, added as (...) P / 2. Both integrated with e^-tau in FP64."""
    ctx, scope, g = warm_case
    c = ctx.constants(NQ, dtype)
    c64 = ctx.constants(NQ, np.float64)
    W, ns = c['W'], ctx.geo.ns
    ih = np.arange(ns) // ctx.geo.nc
    Z = np.array([3e-3, 1e-3, 2e-3])
    c00 = np.array([4e-3, 2e-3, 3e-3])
    c11 = np.array([1e-3, -5e-4, 8e-4])
    mix = np.asarray(c64['mix'], 'f8')
    B = g['target']
    d = g['d']
    P = np.eye(3)[None] - d[:, :, None] * d[:, None, :]
    pu = _pu(g)
    y00 = 1 / math.sqrt(4 * math.pi)
    phi_d = np.arctan2(d[:, 1], d[:, 0])
    st_d = np.sqrt(np.maximum(0, 1 - d[:, 2] ** 2))

    def tensor_part(s, q):
        r = q / np.linalg.norm(q, axis=-1, keepdims=True)
        w = np.einsum('rki,rkl,rml->rmi', B, P, r)  # B^T P r in the target frame
        ww = np.stack([w[..., i] * w[..., j] for i, j in PAIRS6], -1)  # [rays, m, 6]
        return Z[None, None, :, None] * (mix[None, None, :, None] * ww[:, :, None, :]
                                         + ((1 - mix) / 3)[None, None, :, None] * pu[:, None, None, :])

    def sh_part(s, q):
        phi_q = np.arctan2(q[..., 1], q[..., 0])
        rho = np.hypot(q[..., 0], q[..., 1])
        cosrel = np.where((st_d[:, None] > 1e-12) & (rho > 1e-9), np.cos(phi_d[:, None] - phi_q), 1.0)
        aero = y00 * c00[None, None, :] + c11[None, None, :] * (y11(d[:, 2])[:, None] * cosrel)[..., None]
        return aero[..., None] * pu[:, None, None, :] / 2

    for part in ('tensor', 'sh'):
        field = np.zeros((ns, ctx.rows, W), dtype)
        if part == 'tensor':
            field[:, 2] = Z[None, :] / c64['hr'][ih][:, None]
            truth = ray_integral(g, c64, tensor_part)
        else:
            field[:, 10] = c00[None, :] / c64['ha'][ih][:, None]
            field[:, 12] = c11[None, :] / c64['ha'][ih][:, None]  # k = l (l + 1) / 2 + m = 2
            truth = ray_integral(g, c64, sh_part)
        v = ot.transport(g, c, field, False, None, PAIRS6)
        _compare(v.astype('f8'), truth, bound)


# ---------------------------------------------------------------------------------------------------- (iii) boundary
@pytest.mark.parametrize('dtype, bound', [(np.float64, 1e-5), (np.float32, 2e-5)])
def test_ground_boundary_term(warm_case, dtype, bound):
    """A constant upwelling boundary E (Lambertian, unpolarised) seen by ground-hitting rays arrives attenuated by the
    whole path: E exp(-tau(0, end)) (1 - d d^T) / 2; rays that leave through the top see nothing (a zero source)."""
    ctx, scope, g = warm_case
    c = ctx.constants(NQ, dtype)
    c64 = ctx.constants(NQ, np.float64)
    W, ns = c['W'], ctx.geo.ns
    hits = g['hits_ground'].astype(bool)
    assert hits.sum() >= 5 and (~hits).sum() >= 5
    E = np.array([0.02, 0.05, 0.01])
    boundary = np.broadcast_to(E, (ctx.geo.nc, W)).astype(dtype)
    field = np.zeros((ns, ctx.rows, W), dtype)
    v = ot.transport(g, c, field, False, boundary, PAIRS6)
    tau = tau64(g['p'], g['d'], np.zeros(len(g['d'])), g['edges'][:, -1], c64)
    truth = (E[None, :, None] * np.exp(-tau)[:, :, None] * _pu(g)[:, None, :] / 2) * hits[:, None, None]
    _compare(v.astype('f8'), truth, bound)


# ----------------------------------------------------------------------------------------- (iv) first light
def _tiny_pack(ctx, cell, dirs, eps):
    """Rays of length eps from the centre of `cell` along dirs: their radiance over eps is eps times the source there."""
    ih, ic = divmod(cell, ctx.geo.nc)
    p, up, ex = geometry.frame(np.array(R + ctx.geo.h[ih]), ctx.geo.chi[ic])
    n = len(dirs)
    P = np.broadcast_to(p, (n, 3)).copy()
    edges = np.column_stack([np.zeros(n), np.full(n, eps)])
    B = geometry.basis_at(P)
    return dict(p=P, d=dirs, edges=edges, lower=geometry.lower_flags(P, dirs, edges), target=B,
                dloc=np.einsum('nij,ni->nj', B, dirs), weight=np.ones(n), end_index=np.zeros(n, 'i4'),
                end_fraction=np.zeros(n), hits_ground=np.zeros(n, bool), end_columns=np.zeros((n, 6)), nq=NQ)


def _directions(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _sun_tau(q, c, n=64):
    """Optical depth [m, W] from points q [m, 3] toward the sun (+z) to the top, nodes graded toward the start."""
    x, w = np.polynomial.legendre.leggauss(n)
    x, w = (x + 1) / 2, w / 2
    r = np.linalg.norm(q, axis=1)
    mu = q[:, 2] / r
    s_top = -r * mu + np.sqrt(r * r * mu * mu - r * r + (R + c['top']) ** 2)
    s = s_top[:, None] * x[None, :] ** 2
    pts = q[:, None, :] + s[..., None] * np.array([0.0, 0.0, 1.0])
    b = beta64(c, height(pts))
    return np.einsum('mnw,n->mw', b, 2 * x * w) * s_top[:, None]


def _ground_field(ctx, c, cell):
    """Directions [n, 3], weights [n] and the radiance L [n, W] of the sunlit Lambertian ground seen from the centre
    of `cell`: (albedo / pi) max(0, cos chi_g) exp(-tau_sun(g)) exp(-tau(p, g)) for every direction hitting the
    ground at g. mu (about the local vertical) on [-1, mu_horizon] in pieces shrinking geometrically toward the
    horizon, Gauss-Legendre 8 in each; 96 azimuths."""
    ih, ic = divmod(cell, ctx.geo.nc)
    r = R + ctx.geo.h[ih]
    p, up, ex = geometry.frame(np.array(r), ctx.geo.chi[ic])
    ey = np.cross(up, ex)
    mg = -math.sqrt(1 - (R / r) ** 2)
    x, w = np.polynomial.legendre.leggauss(8)
    edges = [-1.0] + [mg - (1 + mg) * 0.5**k for k in range(1, 28)] + [mg]
    mus, wm = [], []
    for a, b in zip(edges, edges[1:]):
        mus.append((a + b) / 2 + (b - a) / 2 * x)
        wm.append((b - a) / 2 * w)
    mu, wm = np.concatenate(mus), np.concatenate(wm)
    nphi = 96
    phi = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    M, PHI = np.meshgrid(mu, phi, indexing='ij')
    st = np.sqrt(1 - M * M)
    loc = np.stack([st * np.cos(PHI), st * np.sin(PHI), M], -1).reshape(-1, 3)
    wt = (wm[:, None] * np.full(nphi, 2 * math.pi / nphi)[None, :]).ravel()
    dirs = loc[:, 0, None] * ex + loc[:, 1, None] * ey + loc[:, 2, None] * up
    pp = np.broadcast_to(p, dirs.shape)
    b = pp @ np.zeros(3) + np.einsum('ni,ni->n', pp, dirs)
    t_hit = -b - np.sqrt(np.maximum(b * b - r * r + R * R, 0.0))
    gpt = pp + t_hit[:, None] * dirs
    cos_g = gpt[:, 2] / np.linalg.norm(gpt, axis=1)
    lit = cos_g > 0
    L = np.zeros((len(dirs), c['W']))
    if lit.any():
        tp = tau64(pp[lit], dirs[lit], np.zeros(lit.sum()), t_hit[lit], c, pieces=16, n=12)
        ts = _sun_tau(gpt[lit], c)
        L[lit] = c['albedo'][None, :] / math.pi * cos_g[lit, None] * np.exp(-tp - ts)
    return dirs, wt, L


@pytest.mark.parametrize('cell_ih, cell_ic', [(1, 1), (2, 0), (3, 2)])
def test_first_light_source_at_cell_centres(cell_ih, cell_ic):
    """First light with aerosol (HG), a sulfate layer (HG), Rayleigh with depolarisation and a Lambertian ground
    (albedo 0.3): the oracle's radiance over a 0.1 m ray from a cell centre, divided by its length, is the
    single-scattering source there. Sun part: exp(-tau_sun) [ext_R hr p_R(d_z) + asca ha HG(d_z, g) + ssca sd
    HG(d_z, g_s)]. Ground part (albedo 0.3 minus albedo 0): the sunlit ground's radiance L(w) over every direction
    that sees the ground, times the same phase functions of w . d. The oracle carries it by the cell's ground pools;
    with the production pool quadrature (32 azimuths, three nodes per mu interval) the HG forward peak of downward
    rays is under-resolved (up to 32 % off here), and with four times that quadrature (pools of qscale 4) the
    oracle reaches the FP64 integral to 3e-3, while a doubled or dropped pool sum misses by 100 %."""
    base = atmosphere.State(aod550=0.2, sulfate_aod550=0.03, no2_column=5e15, o4=True, albedo=0.3,
                            aerosol_height_km=2.0)
    dark = atmosphere.with_albedo(base, 0.0)
    ctx = _pool_context(base, 4)
    ctxd = _pool_context(dark, 4)
    cell = cell_ih * ctx.geo.nc + cell_ic
    eps = 1e-4
    dirs = _directions(24, cell)
    g = _tiny_pack(ctx, cell, dirs, eps)
    ih, ic = divmod(cell, ctx.geo.nc)
    r = R + ctx.geo.h[ih]
    # drop rays that would reach the ground within eps (none at these heights) and keep the check on the source
    assert np.all(r + eps >= R)
    c = ctx.constants(NQ, np.float64)
    cd = ctxd.constants(NQ, np.float64)
    zero = np.zeros((ctx.geo.ns, ctx.rows, c['W']))
    tr = lambda v: (v[..., 0] + v[..., 1]) + v[..., 2]
    J = tr(ot.transport(g, c, zero, True, None, ot.PAIRS4)) / eps
    Jd = tr(ot.transport(g, cd, zero, True, None, ot.PAIRS4)) / eps
    p = g['p'][0]
    h = r - R
    pr = c['profile']
    hr = math.exp(-h / pr['air_H'])
    ha = math.exp(-h / pr['aer_H'])
    sd = math.exp(-0.5 * ((h - pr['st_c']) / pr['st_s']) ** 2)
    mix = np.asarray(c['mix'])
    extR, asca, ssca = np.asarray(c['ext'][0]), np.asarray(c['asca']), np.asarray(c['strato_sca'])
    gg, gs = np.asarray(c['g']), np.asarray(c['strato_g'])

    def phase_sum(u):
        """Scattered intensity per unit incident radiance for scattering cosine u [n] -> [n, W]."""
        u = u[:, None]
        ray = extR * hr * (mix * 3 * (1 + u * u) / (16 * math.pi) + (1 - mix) / (4 * math.pi))
        return ray + asca * ha * hg64(u, gg) + ssca * sd * hg64(u, gs)

    sun = np.exp(-_sun_tau(p[None], c))[0]
    J_sun = sun[None, :] * phase_sum(dirs[:, 2])
    _compare(Jd, J_sun, 2e-3)
    gdirs, gw, Lg = _ground_field(ctx, c, cell)
    J_ground = np.stack([(gw[:, None] * Lg * phase_sum(gdirs @ d)).sum(0) for d in dirs])
    err = _compare(J - Jd, J_ground, 3e-3)
    print(f'cell ({ih}, {ic}): ground term within {err:.3%} of the FP64 hemisphere integral')


# ------------------------------------------------------------------------------------------ (v) composition
def test_rayleigh_row_factors_are_the_phase_function_moments():
    """Rows 4-9 carry the first six SH moments (degrees 0, 1, 1, 2, 2, 2) times the Rayleigh phase function's
    normalised Legendre coefficients beta_l = 2 pi int p_R(x) P_l(x) dx (1, 0, 0.1 for p_R = 3 (1 + x^2) / (16 pi)),
    times ext_R (`src/sosfuse2/solve/engine_core.py:93-94`)."""
    from scipy.special import eval_legendre

    x, w = np.polynomial.legendre.leggauss(16)
    p_r = 3 * (1 + x * x) / (16 * math.pi)
    beta = [2 * math.pi * np.sum(w * p_r * eval_legendre(l, x)) for l in (0, 1, 1, 2, 2, 2)]
    c = dict(L=8, strato=False, ext=np.array([[0.02], [0], [0], [0], [0]]), asca=np.array([0.01]),
             g=np.array([0.7]))
    fac = composition_factors(c)
    np.testing.assert_allclose(fac[4:10, 0], np.array(beta) * np.float32(0.02), rtol=1e-6)
    np.testing.assert_allclose(fac[:4, 0], 3 / (8 * math.pi) * 0.02, rtol=1e-6)


def _moments_of_legendre(ctx, l, W):
    """Incident moments [cells, 4 + K, W] of the radiance field L(w) = P_l(w_z) (z the sun axis): the SH moments are
    int P_l Y_l0 = sqrt(4 pi / (2l + 1)) on (l, 0); the tensor moments int L (1 - w w^T) / 2 dw in each cell's frame
    (ex, ey, up): 4 pi / 3 I for l = 0, and diag(2 pi / 15, 2 pi / 15, -4 pi / 15) in the sun frame for l = 2."""
    K = count(ctx.L)
    ns = ctx.geo.ns
    m = np.zeros((ns, 4 + K, W))
    m[:, 4 + l * (l + 1) // 2] = math.sqrt(4 * math.pi / (2 * l + 1))
    if l == 0:
        M = np.eye(3) * 4 * math.pi / 3
    elif l == 2:
        M = np.diag([2 * math.pi / 15, 2 * math.pi / 15, -4 * math.pi / 15])
    else:
        M = np.zeros((3, 3))  # odd l: int P_l w w^T = 0
    for cell in range(ns):
        ih, ic = divmod(cell, ctx.geo.nc)
        _, up, ex = geometry.frame(np.array(1.0), ctx.geo.chi[ic])
        F = np.stack([ex, np.cross(up, ex), up], -1)
        Mc = F.T @ M @ F
        m[cell, 0:4] = np.array([Mc[0, 0], Mc[1, 1], Mc[2, 2], Mc[0, 2]])[:, None]
    return m


@pytest.mark.parametrize('l', [0, 1, 2, 3])
def test_composition_obeys_funk_hecke(l):
    """Incident radiance P_l(mu) composed into a source (oracle.solve.compose) and reconstructed by the warm transport
    at a cell centre gives, by the Funk-Hecke theorem, asca ha g^l P_l(d_z) for the Henyey-Greenstein aerosol
    (Legendre coefficients g^l) plus ssca sd g_s^l P_l(d_z) for the sulfate layer plus ext_R hr beta_l^R P_l(d_z) for
    Rayleigh scattering with depolarisation (beta_0 = 1, beta_2 = mix / 10, others 0). A wrong eigenvalue or a
    shifted row fails."""
    from scipy.special import eval_legendre

    st = atmosphere.State(sulfate_aod550=0.03, aerosol_height_km=2.0)
    ctx, scope = _context(st, L=8)
    c = ctx.constants(NQ, np.float64)
    W = c['W']
    field = compose(_moments_of_legendre(ctx, l, W), c, np.float64)
    nc = ctx.geo.nc
    eps = 1e-4
    for ih, ic in ((1, 1), (2, 4), (3, 7)):
        cell = ih * nc + ic
        dirs = _directions(16, 100 + cell)
        g = _tiny_pack(ctx, cell, dirs, eps)
        v = ot.transport(g, c, field, False, None, ot.PAIRS4)
        J = ((v[..., 0] + v[..., 1]) + v[..., 2]) / eps
        h = ctx.geo.h[ih]
        pr = c['profile']
        hr, ha = math.exp(-h / pr['air_H']), math.exp(-h / pr['aer_H'])
        sd = math.exp(-0.5 * ((h - pr['st_c']) / pr['st_s']) ** 2)
        beta_r = {0: 1.0, 2: 0.1}.get(l, 0.0) * (np.asarray(c['mix']) if l == 2 else 1.0)
        Pl = eval_legendre(l, dirs[:, 2])[:, None]
        truth = Pl * (np.asarray(c['asca']) * ha * np.asarray(c['g']) ** l
                      + np.asarray(c['strato_sca']) * sd * np.asarray(c['strato_g']) ** l
                      + np.asarray(c['ext'][0]) * hr * beta_r)
        err = float(np.max(np.abs(J - truth)) / np.max(np.abs(truth)))
        assert err < 2e-3, (l, ih, ic, err)
