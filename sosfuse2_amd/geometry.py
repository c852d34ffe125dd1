"""Host FP64 geometry of the spherical shell: cell frames, incident-direction quadrature, ray endpoints, column
integrals, the sun-column table and the ground pools.

This is the input preparation below the transport seam. Each routine restates the reference's algorithm (cited per
function) in this project's own code; the device receives the packed FP32/FP64 arrays of `Bundle`.

Conventions (the reference's, `src/sosfuse2/rays/shell_segments.py:37-66`): the sun is +z; a cell at height h and
solar zenith angle chi sits at r (sin chi, 0, cos chi); its local frame is (ex, ey, up) with ex pointing toward the
sun; a ray from point p along the unit direction d collects the radiance arriving at p from direction d.
"""

import math
from dataclasses import dataclass

import numpy as np

from . import atmosphere

R = 6371.0
NORMAL_TOP_KM = 80.0


# ----------------------------------------------------------------------------------------------------------- domain
@dataclass(frozen=True)
class Domain:
    """Calculation top and the fixed interior shells (`src/sosfuse2/foundation/top.py:13-48`)."""

    top_km: float = 100.0

    @property
    def extra_shells(self):
        return (10.0, 25.0, 40.0) + ((NORMAL_TOP_KM,) if self.top_km > NORMAL_TOP_KM else ())

    @property
    def ring_heights(self):
        return (0.0, 1.2, 8.0, 25.0, 40.0) + (
            (NORMAL_TOP_KM, self.top_km) if self.top_km > NORMAL_TOP_KM else (self.top_km,)
        )

    def monitors(self):
        """Solar elevations (deg) of the observer monitor (`src/sosfuse2/foundation/top.py:55-57`): the stop set
        decides convergence and feeds the Anderson residual; a top above 80 km adds -15 to it and records -18 for
        diagnostics only."""
        extended = self.top_km > NORMAL_TOP_KM
        return (2.0, -3.0, -6.0, -9.0, -12.0) + ((-15.0,) if extended else ()), ((-18.0,) if extended else ())


def gl(n):
    """Gauss-Legendre nodes and weights on [0, 1] (`shell_segments.py:21-34`)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1) / 2, w / 2


def frame(r, chi):
    """Cell position and local frame (`shell_segments.py:37-40`)."""
    chi = np.asarray(chi, 'f8')
    up = np.stack([np.sin(chi), np.zeros_like(chi), np.cos(chi)], -1)
    ex = np.stack([-np.cos(chi), np.zeros_like(chi), np.sin(chi)], -1)
    return np.asarray(r, 'f8')[..., None] * up, up, ex


def basis_at(q):
    """Target frame (ex, ey, up) as columns at points q (`src/sosfuse2/solve/engine_core.py:8-20`)."""
    r = np.linalg.norm(q, axis=-1)
    up = q / r[..., None]
    cos = up[..., 2]
    ex = np.array([0.0, 0.0, 1.0]) - cos[..., None] * up
    ln = np.linalg.norm(ex, axis=-1)
    ex = np.divide(ex, ln[..., None], out=np.zeros_like(ex), where=ln[..., None] > 1e-12)
    ex = np.where(
        (ln < 1e-12)[..., None],
        np.stack([-np.where(cos >= 0, 1.0, -1.0), np.zeros_like(cos), np.zeros_like(cos)], -1),
        ex,
    )
    return np.stack([ex, np.cross(up, ex), up], -1)


# ------------------------------------------------------------------------------------------------------- ray edges
def boundaries(p, d, domain):
    """Path coordinates of the shell exit or ground hit, the tangent point, the two crossings of the back half of the
    Earth's shadow cylinder, and the interior shells, sorted (`shell_segments.py:43-66`, `solve/ray_geometry.py:28-45`)."""
    b = np.sum(p * d, -1)
    pp = np.sum(p * p, -1)
    end = -b + np.sqrt(np.maximum(b * b - pp + (R + domain.top_km) ** 2, 0))
    disc = b * b - pp + R * R
    hit = -b - np.sqrt(np.maximum(disc, 0))
    ground = (disc >= 0) & (b < 0) & (hit >= -1e-7)
    end = np.where(ground, np.maximum(hit, 0), end)
    end = np.maximum(end, 0)
    tangent = np.where((-b > 0) & (-b < end), -b, end)
    aa = d[:, 0] ** 2 + d[:, 1] ** 2
    bb = p[:, 0] * d[:, 0] + p[:, 1] * d[:, 1]
    cc = p[:, 0] ** 2 + p[:, 1] ** 2 - R * R
    ds = bb * bb - aa * cc
    cols = [np.zeros_like(end), tangent]
    for sign in (-1, 1):
        root = (-bb + sign * np.sqrt(np.maximum(ds, 0))) / np.maximum(aa, 1e-30)
        valid = (aa > 1e-20) & (ds >= 0) & (root > 0) & (root < end) & (p[:, 2] + root * d[:, 2] < 0)
        cols.append(np.where(valid, root, end))
    cols.append(end)
    for height in domain.extra_shells:
        dd = b * b - pp + (R + height) ** 2
        for sign in (-1, 1):
            rr = -b + sign * np.sqrt(np.maximum(dd, 0))
            cols.append(np.where((dd > 0) & (rr > 0) & (rr < end), rr, end))
    return np.sort(np.stack(cols, -1), axis=-1)


def lower_flags(p, d, edges):
    """Per segment: True when the start is the lower end (radius grows along the segment), which places the
    quadratic node grading at the start (`solve/ray_geometry.py:43-45`)."""
    qa = p[:, None, :] + edges[:, :-1, None] * d[:, None, :]
    qb = p[:, None, :] + edges[:, 1:, None] * d[:, None, :]
    return np.sum(qa * qa, axis=-1) <= np.sum(qb * qb, axis=-1)


# -------------------------------------------------------------------------------------------------- quadratures
def ring_edges(r, chi, phi, heights, prepared):
    """mu of the intersections of the view plane at azimuth phi with the shadow ring of each shell
    (`src/sosfuse2/rays/incoming_rings.py:11-60`)."""
    p, up, ex, ey, levels = prepared
    horizontal = np.cos(phi) * ex + np.sin(phi) * ey
    normal = np.array(
        [
            up[1] * horizontal[2] - up[2] * horizontal[1],
            up[2] * horizontal[0] - up[0] * horizontal[2],
            up[0] * horizontal[1] - up[1] * horizontal[0],
        ]
    )
    amp = R * math.hypot(normal[0], normal[1])
    if amp < 1e-10:
        return []
    angle = math.atan2(normal[1], normal[0])
    out = []
    for z in levels:
        target = -normal[2] * z / amp
        if abs(target) > 1:
            continue
        delta = math.acos(target)
        for az in (angle - delta, angle + delta):
            q = np.array([R * math.cos(az), R * math.sin(az), z])
            dv = q - p
            length = math.sqrt(dv @ dv)
            if length < 1e-8:
                continue
            dv /= length
            if dv @ horizontal <= 1e-10:
                continue
            b = p @ dv
            disc = b * b - r * r + R * R
            hit = -b - math.sqrt(max(0.0, disc))
            if disc >= 0 and b < 0 and hit >= -1e-7 and hit < length - 1e-5:
                continue
            out.append(float(dv @ up))
    return out


def _mu_edges(r, chi, phi, mg, tangent_pair):
    sinchi, coschi = math.sin(chi), math.cos(chi)
    edges = [-1.0, 1.0, mg, 0.0, mg - 0.05234, mg + 0.05234, mg - 0.0061, mg + 0.0061]
    rp = r * sinchi
    if rp > R + 1e-8 and abs(sinchi) > 1e-8:
        kk = R / math.sqrt(rp * rp - R * R)
        for sign in (-1, 1):
            t = math.cos(phi) * coschi / sinchi + sign * math.sin(phi) / (kk * sinchi)
            mu = t / math.sqrt(1 + t * t)
            st = math.sqrt(1 - mu * mu)
            dx = mu * sinchi - st * math.cos(phi) * coschi
            dy = -st * math.sin(phi)
            dz = mu * coschi + st * math.cos(phi) * sinchi
            tmin = -rp * dx / max(dx * dx + dy * dy, 1e-25)
            if abs(dx) + abs(dy) < 1e-10 or (tmin > 0 and r * coschi + tmin * dz < 0):
                edges.append(mu)
    if phi < math.pi / 6:
        mu = coschi / math.sqrt(coschi**2 + math.cos(phi) ** 2 * sinchi**2 + tangent_pair)
        edges.extend([mu, mu - 0.012, mu + 0.012])
    return edges


def _sinh_nodes(edges, mg, u, wu, a=4.0):
    sa = math.sinh(a)
    edges = np.unique(np.clip(edges, -1, 1))
    delta = edges - mg
    tt = np.sign(delta) * np.arcsinh(np.abs(delta) / np.where(delta < 0, 1 + mg, 1 - mg) * sa) / a
    keep = tt[1:] - tt[:-1] >= 1e-10
    left = tt[:-1][keep, None]
    right = tt[1:][keep, None]
    t = left + (right - left) * u
    den = np.where(t < 0, 1 + mg, 1 - mg)
    mus = (mg + np.sign(t) * den * np.sinh(a * np.abs(t)) / sa).ravel()
    w = ((right - left) * wu * den * a * np.cosh(a * np.abs(t)) / sa).ravel()
    return mus, w


def incoming(r, chi, nmu, nphi, domain, rings=True):
    """Incident directions (local frame) and solid-angle weights of one cell: sinh-graded Gauss-Legendre nodes in mu
    between the horizon, shadow and ring edges, trapezoid in phi over [0, pi] with the y-mirror folded into the weights
    (`src/sosfuse2/rays/incoming_rings.py:63-126`). The weights sum to 4 pi."""
    mg = -math.sqrt(max(0.0, 1 - (R / r) ** 2))
    u, wu = gl(nmu)
    prepared = None
    if rings:
        p, up, ex = frame(np.array(r), np.array(chi))
        ey = np.cross(up, ex)
        levels = [-math.sqrt(max(0.0, (R + h) ** 2 - R**2)) for h in domain.ring_heights]
        prepared = (p, up, ex, ey, levels)
    mus, phis, weights = [], [], []
    for j, phi in enumerate(np.linspace(0, math.pi, nphi + 1)):
        edges = _mu_edges(r, chi, phi, mg, 0.0)
        if rings:
            edges += ring_edges(r, chi, phi, domain.ring_heights, prepared)
        m, w = _sinh_nodes(edges, mg, u, wu)
        mus.append(m)
        weights.append(w * (2 * math.pi / nphi) * (0.5 if j in (0, nphi) else 1.0))
        phis.append(np.full(m.size, phi))
    mu = np.concatenate(mus)
    phi = np.concatenate(phis)
    st = np.sqrt(np.maximum(0, 1 - mu * mu))
    return np.column_stack([st * np.cos(phi), st * np.sin(phi), mu]), np.concatenate(weights)


def ground_incoming(r, chi, qscale=1):
    """Incident quadrature of the ground pools: 32 qscale azimuths, 3 qscale nodes per mu interval, no rings
    (`src/sosfuse2/rays/incoming_quadrature.py:6-52`)."""
    mg = -math.sqrt(max(0.0, 1 - (R / r) ** 2))
    nphi = 32 * qscale
    u, wu = gl(3 * qscale)
    mus, phis, weights = [], [], []
    for j, phi in enumerate(np.linspace(0, math.pi, nphi + 1)):
        edges = _mu_edges(r, chi, phi, mg, 1e-30)
        m, w = _sinh_nodes(edges, mg, u, wu)
        mus.append(m)
        weights.append(w * 2 * math.pi / nphi * (0.5 if j in (0, nphi) else 1.0))
        phis.append(np.full(m.size, phi))
    mu = np.concatenate(mus)
    phi = np.concatenate(phis)
    st = np.sqrt(np.maximum(0, 1 - mu * mu))
    return np.column_stack([st * np.cos(phi), st * np.sin(phi), mu]), np.concatenate(weights)


# ------------------------------------------------------------------------------------------------ column integrals
def columns(state, p, d, edges, n, domain, strato=False):
    """Slant columns of the five component profiles (and the sulfate profile) from p along d to the last edge:
    Gauss-Legendre with quadratic node grading on every segment (`shell_segments.py:115-147,223-227`). FP64."""
    u, wu = gl(n)
    m = len(d)
    out = np.zeros((m, 6))
    for k in range(edges.shape[1] - 1):
        active = np.flatnonzero(edges[:, k + 1] - edges[:, k] > 1e-8)
        if not len(active):
            continue
        a, b = edges[active, k, None], edges[active, k + 1, None]
        pa, da = p[active], d[active]
        ra = np.sum((pa + a * da) ** 2, -1)
        rb = np.sum((pa + b * da) ** 2, -1)
        lower = (ra <= rb)[:, None]
        t = np.where(lower, a + (b - a) * u * u, b - (b - a) * (1 - u) ** 2)
        ds = np.where(lower, 2 * (b - a) * u * wu, 2 * (b - a) * (1 - u) * wu)
        q = pa[:, None, :] + t[..., None] * da[:, None, :]
        h = np.sqrt(np.sum(q * q, -1)) - R
        prof = atmosphere.profiles(state, h)
        out[active, :5] += np.sum(prof * ds[..., None], axis=1)
        if strato:
            out[active, 5] += np.sum(atmosphere.strato_profile(state, h) * ds, axis=1)
    return out


# ---------------------------------------------------------------------------------------------------- sun table
@dataclass
class SunTable:
    """Slant columns toward the sun on (height, x) with mu = mg + (1 - mg) x^2 (`shell_segments.py:109-112,267-307`)."""

    h: np.ndarray
    x: np.ndarray
    col: np.ndarray  # [nh, nx, 6]: five components and the sulfate layer


def _sun_rows(args):
    """Sun-table rows of the heights h (one job of sun_table)."""
    state, domain, h, x, scale = args
    out = np.empty((len(h), len(x), 6))
    for i, hh in enumerate(h):
        r = R + hh
        mg = -math.sqrt(max(0.0, 1 - (R / r) ** 2))
        mu = mg + (1 - mg) * x**2
        mu[0] = mg + 1e-9
        p = np.column_stack([np.zeros(len(mu)), np.zeros(len(mu)), np.full(len(mu), r)])
        d = np.column_stack([np.sqrt(np.maximum(1 - mu * mu, 0)), np.zeros(len(mu)), mu])
        e = boundaries(p, d, domain)
        out[i] = columns(state, p, d, e, 32 * scale, domain, strato=True)
    return out


def sun_table(state, domain, scale=1, workers=1):
    """Sun-column table over heights (square-root spaced) and the horizon-relative cosine x; rows split over a
    process pool (a thread pool once the GPU runtime has started) when workers > 1."""
    nh = 80 * scale + 1
    h = np.linspace(0, math.sqrt(domain.top_km), nh) ** 2
    x = np.linspace(0, 1, 400 * scale + 1)
    if workers <= 1 or nh < 2 * workers:
        return SunTable(h, x, _sun_rows((state, domain, h, x, scale)))
    from concurrent.futures import ProcessPoolExecutor, ThreadPoolExecutor
    import multiprocessing as mp

    parts = np.array_split(np.arange(nh), 4 * workers)
    jobs = [(state, domain, h[idx], x, scale) for idx in parts if len(idx)]
    pool = (ThreadPoolExecutor(workers) if gpu_runtime_started()
            else ProcessPoolExecutor(workers, mp_context=mp.get_context('spawn')))
    with pool:
        return SunTable(h, x, np.concatenate(list(pool.map(_sun_rows, jobs))))


def lit_flag(q, r):
    """Whether points q at radius r see the sun: their cosine to the sun axis reaches the horizon's (sun_lookup's lit
    flag without the table)."""
    mg = -np.sqrt(np.maximum(0, 1 - (R / np.maximum(r, R)) ** 2))
    return q[..., 2] / r >= mg


def sun_lookup(table, q, r, top_km):
    """Bilinear sun columns and the lit flag at points q (`shell_segments.py:309-325`)."""
    h = np.clip(r - R, 0, top_km)
    mu = q[..., 2] / r
    mg = -np.sqrt(np.maximum(0, 1 - (R / np.maximum(r, R)) ** 2))
    x = np.sqrt(np.clip((mu - mg) / (1 - mg), 0, 1))
    i = np.clip(np.searchsorted(table.h, h, side='right') - 1, 0, len(table.h) - 2)
    a = (h - table.h[i]) / (table.h[i + 1] - table.h[i])
    f = x * (len(table.x) - 1)
    j = np.clip(f.astype(int), 0, len(table.x) - 2)
    b = f - j
    col = (
        ((1 - a) * (1 - b))[..., None] * table.col[i, j]
        + (a * (1 - b))[..., None] * table.col[i + 1, j]
        + ((1 - a) * b)[..., None] * table.col[i, j + 1]
        + (a * b)[..., None] * table.col[i + 1, j + 1]
    )
    return col, mu >= mg


# ------------------------------------------------------------------------------------------------------- grids
def height_grid(n, top, first=None):
    """n heights from 0 to top, geometrically graded (finer near the ground)."""
    if n < 2:
        raise ValueError('E_GRID: at least two heights')
    k = 4.5 if first is None else first
    s = np.linspace(0, 1, n)
    return top * np.expm1(k * s) / math.expm1(k)


def chi_grid(kind='twilight'):
    """Solar zenith nodes (rad). 'twilight' is the reference's base grid with 0.5 degree steps over 85-110 degrees
    (`shell_segments.py:88-102`); 'coarse' is a 1 degree twilight band."""
    if kind == 'twilight':
        deg = np.unique(np.r_[np.arange(0, 81, 10), np.arange(80, 85, 1), np.arange(85, 110, 0.5),
                              np.arange(110, 120, 2), np.arange(120, 181, 10)])
    elif kind == 'coarse':
        deg = np.unique(np.r_[np.arange(0, 81, 20), np.arange(84, 111, 2), np.arange(115, 181, 15)])
    elif kind == 'tiny':
        deg = np.array([0.0, 60.0, 85.0, 90.0, 93.0, 96.0, 105.0, 110.0, 140.0, 180.0])
    else:
        raise ValueError(f'E_GRID: unknown chi grid {kind}')
    return np.deg2rad(deg)


@dataclass
class Grid:
    h: np.ndarray
    chi: np.ndarray
    nmu: int
    nphi: int
    path_n: int
    domain: Domain
    query_n: int = 24

    @property
    def nh(self):
        return len(self.h)

    @property
    def nc(self):
        return len(self.chi)

    @property
    def ns(self):
        return self.nh * self.nc


# ---------------------------------------------------------------------------------------------------------- packs
def pack(p, d, n, chi_grid_, domain, weights=None):
    """Endpoint pack of rays from p along d (`solve/ray_geometry.py:28-82`): edges, lower flags, target frame,
    direction in the target frame, weights, and the ray end's chi interpolation (index, fraction) and ground hit."""
    p = np.ascontiguousarray(np.broadcast_to(p, d.shape), 'f8')
    d = np.ascontiguousarray(d, 'f8')
    edges = boundaries(p, d, domain)
    end = edges[:, -1]
    lower = lower_flags(p, d, edges)
    B = basis_at(p)
    qe = p + end[:, None] * d
    re = np.linalg.norm(qe, axis=-1)
    ch = np.arccos(np.clip(qe[:, 2] / re, -1, 1))
    j = np.clip(np.searchsorted(chi_grid_, ch, side='right') - 1, 0, len(chi_grid_) - 2)
    f = np.clip((ch - chi_grid_[j]) / (chi_grid_[j + 1] - chi_grid_[j]), 0, 1)
    return dict(
        p=p,
        d=d,
        edges=edges,
        lower=lower,
        target=B,
        dloc=np.einsum('nij,ni->nj', B, d),
        weight=np.ones(len(d)) if weights is None else np.asarray(weights, 'f8'),
        end_index=j.astype('i4'),
        end_fraction=f,
        hits_ground=re < R + 1e-5,
        nq=n,
    )


def cell_pack(grid, cell):
    """The incident rays of one cell (`solve/ray_geometry.py:406-423`)."""
    ih, ic = divmod(cell, grid.nc)
    r = R + grid.h[ih]
    chi = grid.chi[ic]
    loc, w = incoming(r, chi, grid.nmu, grid.nphi, grid.domain)
    p, up, ex = frame(np.array(r), chi)
    ey = np.cross(up, ex)
    d = loc[:, 0, None] * ex + loc[:, 1, None] * ey + loc[:, 2, None] * up
    return pack(p, d, grid.path_n, grid.chi, grid.domain, w)


def observer_pack(grid, solar_elevation_deg, el_deg, az_deg, observer_km=0.1, n=None):
    """Line-of-sight rays of an observer at observer_km and solar zenith 90 - h_s (`solve/engine_core.py:208-226`)."""
    ee, aa = np.broadcast_arrays(np.asarray(el_deg, 'f8'), np.asarray(az_deg, 'f8'))
    e, z = np.deg2rad(ee.ravel()), np.deg2rad(aa.ravel())
    p, up, ex = frame(np.array(R + observer_km), np.deg2rad(90 - solar_elevation_deg))
    ey = np.cross(up, ex)
    d = np.sin(e)[:, None] * up + np.cos(e)[:, None] * (np.cos(z)[:, None] * ex + np.sin(z)[:, None] * ey)
    g = pack(p, d, grid.query_n if n is None else n, grid.chi, grid.domain)
    g['view'] = np.stack([e, z], -1)
    return g


def concat_packs(packs):
    """One pack of many (all with the same edge count and node count)."""
    keys = ('p', 'd', 'edges', 'lower', 'target', 'dloc', 'weight', 'end_index', 'end_fraction', 'hits_ground')
    out = {k: np.concatenate([g[k] for g in packs]) for k in keys}
    out['nq'] = packs[0]['nq']
    if any(g['nq'] != out['nq'] for g in packs):
        raise ValueError('E_QUADRATURE: mixed path orders')
    if 'view' in packs[0]:
        out['view'] = np.concatenate([g['view'] for g in packs])
    return out


def with_end_columns(state, g, domain, n=24):
    """Slant columns from the ray start to its end for rays that hit the ground (`solve/ray_geometry.py:53-65`)."""
    col = np.zeros((len(g['d']), 6))
    hit = np.flatnonzero(g['hits_ground'])
    if len(hit):
        col[hit] = columns(state, g['p'][hit], g['d'][hit], g['edges'][hit], n, domain, strato=True)
    g['end_columns'] = col
    return g


# ------------------------------------------------------------------------------------------------- ground pools
@dataclass
class GroundPools:
    """Per cell: incident directions that see the sunlit ground, as directions in the cell frame, their
    albedo-free geometric factor cos(zenith at the ground point)/pi * weight, and the slant column vector (path to the
    ground plus the ground point's sun column), from which each lane's values follow
    (`solve/engine_core.py:103-147`)."""

    off: np.ndarray  # [C + 1]
    loc: np.ndarray  # [P, 3]
    factor: np.ndarray  # [P]
    column: np.ndarray  # [P, 6]
    path: tuple = None  # (starts, dirs, ground points) [P, 3] while the columns are left to the device

    def values(self, optics):
        """gv[P, W] = albedo * factor * exp(-column . ext)."""
        tau = self.column[:, :5] @ optics.ext + self.column[:, 5:6] * optics.strato_ext[None, :]
        return optics.albedo[None, :] * self.factor[:, None] * np.exp(-tau)

    def tensor(self, optics, cells, gv=None):
        """gm[C, 4, W]: the Rayleigh tensor moments of the ground-reflected light at each cell, times 3/(8 pi) ext_R
        (`engine_core.py:135-144`). gv: the pool values when already at hand. The pools of a cell are contiguous, so
        the per-cell sums are one sparse product with the cells' CSR pattern (off)."""
        import scipy.sparse

        gv = self.values(optics) if gv is None else gv
        loc = self.loc
        factor = np.stack(
            [(1 - loc[:, 0] ** 2) / 2, (1 - loc[:, 1] ** 2) / 2, (1 - loc[:, 2] ** 2) / 2, -loc[:, 0] * loc[:, 2] / 2],
            -1,
        )
        P = len(loc)
        out = np.zeros((cells, 4, optics.W))
        if P:
            for j in range(4):
                seg = scipy.sparse.csr_matrix((factor[:, j], np.arange(P), self.off[: cells + 1]), shape=(cells, P))
                out[:, j] = seg @ gv
        return out * (3 / (8 * math.pi)) * optics.ext[0][None, None, :]

    def padded(self, optics, cells, gv=None):
        """gd[C, ng, 3], gv[C, ng, W] zero-padded per cell (`source_solver.py:1281-1287`)."""
        ng = int(max(np.diff(self.off), default=0))
        gd = np.zeros((cells, max(ng, 1), 3))
        gvp = np.zeros((cells, max(ng, 1), optics.W))
        gv = self.values(optics) if gv is None else gv
        for c in range(cells):
            a, b = self.off[c : c + 2]
            gd[c, : b - a] = loc = self.loc[a:b]
            gvp[c, : b - a] = gv[a:b]
        return gd, gvp


def _pools_of_cells(state, grid, sun, cells, qscale, columns_too=True):
    """Pools of some cells; columns_too=False leaves out the path columns (cell to ground point) and returns the rays
    they follow instead, for device_geometry.pool_columns."""
    domain = grid.domain
    locs, facs, cols, counts, paths = [], [], [], [], []
    for cell in cells:
        ih, ic = divmod(cell, grid.nc)
        r = R + grid.h[ih]
        loc, w = ground_incoming(r, grid.chi[ic], qscale)
        p, up, ex = frame(np.array(r), grid.chi[ic])
        ey = np.cross(up, ex)
        d = loc[:, 0, None] * ex + loc[:, 1, None] * ey + loc[:, 2, None] * up
        pp = np.broadcast_to(p, d.shape).copy()
        e = boundaries(pp, d, domain)
        end = e[:, -1]
        q = pp + end[:, None] * d
        rq = np.linalg.norm(q, axis=-1)
        hit = rq < R + 1e-5
        if columns_too:
            sc, lit = sun_lookup(sun, q, np.maximum(R, rq), domain.top_km)
        else:
            lit = lit_flag(q, np.maximum(R, rq))
        ch = np.arccos(np.clip(q[:, 2] / rq, -1, 1))
        factor = np.maximum(0, np.cos(ch)) / math.pi * lit * hit * w
        take = np.flatnonzero(factor > 0)
        col = np.zeros((len(take), 6))
        if len(take):
            if columns_too:
                col = columns(state, pp[take], d[take], e[take], 24, domain, strato=True) + sc[take]
            else:
                paths.append((pp[take], d[take], q[take]))
        locs.append(loc[take])
        facs.append(factor[take])
        cols.append(col)
        counts.append(len(take))
    return locs, facs, cols, counts, paths


def _cells_job(args):
    kind, state, grid, sun, cells, qscale, columns_too = args
    if kind == 'pools':
        return _pools_of_cells(state, grid, sun, cells, qscale, columns_too)
    if columns_too:
        return [with_end_columns(state, cell_pack(grid, c), grid.domain) for c in cells]
    return [cell_pack(grid, c) for c in cells]


def gpu_runtime_started():
    """Whether this process has brought up the GPU runtime (the HIP library of this package, or torch's CUDA state).
    A process pool started after that forks a process holding the GPU runtime and replaces it with a new interpreter,
    which the GPU host may refuse, leaving the pool waiting for workers that never start."""
    import sys

    hip_mod = sys.modules.get(__package__ + '.hip')
    if hip_mod is not None and hip_mod.started():
        return True
    torch = sys.modules.get('torch')
    return bool(torch is not None and torch.cuda.is_initialized())


def _parallel(kind, state, grid, sun, workers, qscale=1, columns_too=True):
    """Run a per-cell preparation over a spawn process pool (numpy only in the workers), in cell order; a thread
    pool once the process has started the GPU runtime (gpu_runtime_started)."""
    cells = list(range(grid.ns))
    if workers <= 1 or grid.ns < 64:
        return [_cells_job((kind, state, grid, sun, cells, qscale, columns_too))]
    from concurrent.futures import ProcessPoolExecutor, ThreadPoolExecutor
    import multiprocessing as mp

    chunk = max(1, (grid.ns + 4 * workers - 1) // (4 * workers))
    jobs = [(kind, state, grid, sun, cells[i : i + chunk], qscale, columns_too) for i in range(0, grid.ns, chunk)]
    if gpu_runtime_started():
        pool = ThreadPoolExecutor(workers)
    else:
        pool = ProcessPoolExecutor(workers, mp_context=mp.get_context('spawn'))
    with pool:
        return list(pool.map(_cells_job, jobs))


def default_workers():
    import os

    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def ground_pools(state, grid, sun, qscale=1, workers=1, columns_too=True):
    """Build the ground pools of every cell (`engine_core.py:103-147`). columns_too=False: the columns (path to the
    ground and the sun table at the ground point) are left for device_geometry.pool_columns, and `sun` may be None."""
    parts = _parallel('pools', state, grid, sun, workers, qscale, columns_too)
    locs, facs, cols, counts, paths = [], [], [], [], []
    for a, b, c, d, e in parts:
        locs += a
        facs += b
        cols += c
        counts += d
        paths += e
    path = None
    if not columns_too:
        path = tuple(np.concatenate([x[i] for x in paths]) if paths else np.zeros((0, 3)) for i in range(3))
    return GroundPools(
        off=np.r_[0, np.cumsum(counts)].astype('i8'),
        loc=np.concatenate(locs) if locs else np.zeros((0, 3)),
        factor=np.concatenate(facs) if facs else np.zeros(0),
        column=np.concatenate(cols) if cols else np.zeros((0, 6)),
        path=path,
    )


def cell_packs(state, grid, workers=1, end_columns=True):
    """Incident-ray packs of every cell, in cell order, with their end columns (end_columns=False: left for the
    device, device_geometry.end_columns)."""
    return [p for part in _parallel('packs', state, grid, None, workers, 1, end_columns) for p in part]


# ------------------------------------------------------------------------------------------------------- bundles
def compact_edges(edges, lower):
    """Keep every strictly positive interval: values are the left ends of kept intervals plus the right end of the
    last one; a ray with none keeps a single 0. Flag bit i is the lower flag of the i-th kept interval
    (`solve/ray_geometry.py:148-180`)."""
    n, m1 = edges.shape
    m = m1 - 1
    if m > 16:
        raise OverflowError('E_LOWER_BITS: more than 16 intervals per ray')
    mask = np.diff(edges, axis=1) > 0
    has = mask.any(1)
    last = m - 1 - np.argmax(mask[:, ::-1], axis=1)
    select = np.zeros((n, m1), bool)
    select[:, :m] = mask
    rows = np.flatnonzero(has)
    select[rows, last[rows] + 1] = True
    empty = np.flatnonzero(~has)
    select[empty, 0] = True
    source = edges.astype('f8', copy=True)
    source[empty, 0] = 0.0
    counts = select.sum(1).astype('i8')
    rank = np.cumsum(mask, axis=1) - 1
    bits = np.where(mask, np.asarray(lower[:, :m], 'u8') << np.maximum(rank, 0).astype('u8'), np.uint64(0))
    return counts, source[select], bits.sum(1, dtype='u8').astype('u2')


class Bundle:
    """CSR of the rays of many destination cells (or query groups), the device input of the transport
    (`solve/ray_geometry.py:290-392`). Group g owns rays [group_off[g], group_off[g+1]); every group has one start."""

    def __init__(self, packs, end_columns=True):
        self.groups = len(packs)
        self.nq = packs[0]['nq']
        if any(g['nq'] != self.nq for g in packs):
            raise ValueError('E_QUADRATURE: mixed path orders')
        counts = [len(g['d']) for g in packs]
        self.group_off = np.r_[0, np.cumsum(counts)].astype('i8')
        self.rays = int(self.group_off[-1])
        starts = []
        for g in packs:
            if len(g['d']) and not np.all(g['p'] == g['p'][0]):
                raise ValueError('E_START: one start per group')
            starts.append(g['p'][0] if len(g['d']) else np.zeros(3))
        self.starts = np.array(starts, 'f8')
        self.targets = np.array([g['target'][0] if len(g['d']) else np.eye(3) for g in packs], 'f8')
        edges = np.concatenate([g['edges'] for g in packs])
        lower = np.concatenate([g['lower'] for g in packs])
        ec, ev, fl = compact_edges(edges, lower)
        self.edge_off = np.r_[0, np.cumsum(ec)].astype('i8')
        self.edges = ev.astype('f8')
        self.lower = fl
        self.dirs = np.concatenate([g['d'] for g in packs])
        self.dloc = np.concatenate([g['dloc'] for g in packs])
        self.weights = np.concatenate([g['weight'] for g in packs]).astype('f4')
        self.end_j = np.concatenate([g['end_index'] for g in packs]).astype('i4')
        self.end_f = np.concatenate([g['end_fraction'] for g in packs]).astype('f4')
        self.hits = np.concatenate([g['hits_ground'] for g in packs]).astype('u1')
        if end_columns and 'end_columns' in packs[0]:
            self.end_col = np.concatenate([g['end_columns'] for g in packs]).astype('f4')
        else:
            self.end_col = np.zeros((self.rays, 6), 'f4')
        self.ray_group = np.repeat(np.arange(self.groups, dtype='i4'), counts)
        self.samples = int(np.sum(np.diff(edges, axis=1) > 1e-8)) * self.nq

    def nbytes(self):
        return sum(getattr(self, k).nbytes for k in ('group_off', 'starts', 'targets', 'edge_off', 'edges', 'lower',
                                                      'dirs', 'dloc', 'weights', 'end_j', 'end_f', 'end_col', 'hits'))
