"""Oracle source iteration, observer queries and table: a NumPy restatement of the reference's solver loop.

TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline); the product package never imports it.

Follows:
- `evaluate` = the NumPy branch of `src/sosfuse2/solve/ray_geometry.py:443-451` (per cell transport + projection);
- `initialize` / `apply` = `src/sosfuse2/solve/source_iteration.py:123-216` (first light once; ground pass, boundary,
  warm pass, composition with the eigenvalues, FP32 flush of |source| < FLT_MIN);
- `solve` = `source_iteration.py:309-663`, list-history branch :618-635 (per-wavelength Anderson, FP32 Gram, FP64
  regularised solve `_anderson_weights` :25-33, stop on max(observer, source) relative change from evaluation 2);
- `query` / `boxes` = `src/sosfuse2/solve/engine_core.py:208-238`; `make_table` = `src/sosfuse2/outputs/radiance_table.py:167-219`.

Inputs (grid, optics, packs, constants) come from the product's host preparation, so the oracle checks the transport
and the iteration, not the input geometry. The composition factors are restated here from the optical constants
(`composition_factors`), not taken from the product; tests/test_warm_pins.py pins them by the Funk-Hecke theorem
(composition then reconstruction of P_l(mu) radiance gives g^l P_l) and the Rayleigh rows by the Legendre moments of
the Rayleigh phase function; `monitor_elevations` restates the monitor rule of the reference's domain.
"""

import math
import time

import numpy as np

from .arith import TRACE_MIN32, count
from .transport import PAIRS4, PAIRS6, project, transport


def monitor_elevations(top_km):
    """(stop, record) solar elevations of the observer monitor (`src/sosfuse2/foundation/top.py:55-57`): five stop
    elevations, plus -15 when the top is above 80 km, which also records -18 for diagnostics only."""
    extended = float(top_km) > 80.0
    return [2.0, -3.0, -6.0, -9.0, -12.0] + ([-15.0] if extended else []), ([-18.0] if extended else [])


def anderson_weights(products):
    """Regularised Gram solve, normalised weights, guard |w| < 8 (source_iteration.py:25-33)."""
    gram = np.array(products, dtype=np.float64, copy=True)
    gram[np.diag_indices_from(gram)] += max(float(np.trace(gram)) * 1e-10, 1e-30)
    try:
        weights = np.linalg.solve(gram, np.ones(len(gram)))
    except np.linalg.LinAlgError:
        return None
    weights /= weights.sum()
    if np.isfinite(weights).all() and np.max(np.abs(weights)) < 8:
        return weights.astype(np.float32)
    return None


def stokes(comp, view, F=np.float64):
    """I, Q, U [..., 3] from the six tensor components (xx, yy, zz, xz, xy, yz) in the observer frame.
    e_h = (-sin az, cos az, 0) and e_v = (-sin el cos az, -sin el sin az, cos el) span the plane normal to the view;
    I = tr P, Q = e_v P e_v - e_h P e_h, U = 2 e_v P e_h."""
    el, az = view[:, 0], view[:, 1]
    eh = np.stack([-np.sin(az), np.cos(az), np.zeros_like(az)], -1)
    ev = np.stack([-np.sin(el) * np.cos(az), -np.sin(el) * np.sin(az), np.cos(el)], -1)
    xx, yy, zz, xz, xy, yz = (comp[..., k].astype('f8') for k in range(6))

    def quad(a, b):
        a = a[:, None, :]
        b = b[:, None, :]
        return (a[..., 0] * b[..., 0] * xx + a[..., 1] * b[..., 1] * yy + a[..., 2] * b[..., 2] * zz
                + (a[..., 0] * b[..., 2] + a[..., 2] * b[..., 0]) * xz
                + (a[..., 0] * b[..., 1] + a[..., 1] * b[..., 0]) * xy
                + (a[..., 1] * b[..., 2] + a[..., 2] * b[..., 1]) * yz)

    I = (xx + yy) + zz
    return np.stack([I, quad(ev, ev) - quad(eh, eh), 2 * quad(ev, eh)], -1)


def composition_factors(c):
    """Per-row factors [rows, W] turning incident moments [4 + K] into source rows (source_iteration.py:176-208):
    the Rayleigh tensor by F(3 / (8 pi)) ext_R; rows 4-9 (the first six SH moments, degrees 0, 1, 1, 2, 2, 2) by the
    Rayleigh eigenvalues rf = 1 (l = 0), 0 (l = 1), 0.1 (l = 2) (engine_core.py:93-94) times ext_R; the aerosol plane
    by g^l asca (engine_core.py:95); a sulfate plane by g_s^l sca_s. Each factor is the product of FP32-rounded
    parts taken in FP64, as the reference multiplies FP64 moments by FP32 arrays."""
    f4 = lambda a: np.asarray(a, 'f8').astype('f4').astype('f8')
    L = c['L']
    K = count(L)
    deg = np.array([l for l in range(L + 1) for m in range(l + 1)], 'f8')
    rf = (deg[:6] == 0) + 0.1 * (deg[:6] == 2)
    ext_r = f4(c['ext'][0])
    rows = 10 + K * (2 if c['strato'] else 1)
    out = np.zeros((rows, len(ext_r)))
    out[:4] = f4(np.float32(3 / (8 * math.pi)) * np.asarray(c['ext'][0], 'f4'))[None, :]
    out[4:10] = f4(rf)[:, None] * ext_r[None, :]
    out[10 : 10 + K] = f4(np.asarray(c['g'], 'f8')[None, :] ** deg[:, None]) * f4(c['asca'])[None, :]
    if c['strato']:
        out[10 + K :] = f4(np.asarray(c['strato_g'], 'f8')[None, :] ** deg[:, None]) * f4(c['strato_sca'])[None, :]
    return out


def compose(moment, c, F):
    """Source field [cells, rows, W] from incident moments [cells, 4 + K, W]; FP32 flushes |source| < FLT_MIN to 0
    (source_iteration.py:209-212)."""
    fac = composition_factors(c)
    K = count(c['L'])
    moment = np.asarray(moment, 'f8')
    out = np.empty((moment.shape[0], fac.shape[0], moment.shape[2]), F)
    out[:, :4] = (moment[:, :4] * fac[None, :4]).astype(F)
    out[:, 4:10] = (moment[:, 4:10] * fac[None, 4:10]).astype(F)
    out[:, 10 : 10 + K] = (moment[:, 4:] * fac[None, 10 : 10 + K]).astype(F)
    if c['strato']:
        out[:, 10 + K :] = (moment[:, 4:] * fac[None, 10 + K :]).astype(F)
    if F is np.float32:
        out[np.abs(out) < TRACE_MIN32] = 0
    return out


_WORKER = {}


def _worker_init(c, cq, packs):
    """Process-pool worker state: the constants of both node counts and the cell packs, sent once."""
    _WORKER.update(c=c, cq=cq, packs=packs)


def _cells_job(args):
    """Transport and projection of a run of cells in a pool worker (the per-cell loop of evaluate)."""
    from sosfuse2_amd import geometry

    cells, field, first, phase, boundary = args
    c = _WORKER['c']
    sel = [_WORKER['packs'][i] for i in cells]
    cat = geometry.concat_packs(sel)
    cat['end_columns'] = np.concatenate([p['end_columns'] for p in sel])
    off = np.r_[0, np.cumsum([len(p['d']) for p in sel])]
    stats = dict(samples=0, angular_pairs=0)
    value = transport(cat, c, field, first, boundary, PAIRS4, stats)
    K = count(c['L'])
    moments = np.zeros((len(cells), 4 + K, c['W']), 'f8')
    flux = np.zeros((len(cells), c['W']), 'f8')
    for i in range(len(cells)):
        sub = {k: cat[k][off[i] : off[i + 1]] for k in ('d', 'dloc', 'weight')}
        moments[i], flux[i] = project(sub, value[off[i] : off[i + 1]], c, ground_only=phase == 1)
    return moments, flux, stats


def _query_job(args):
    """Observer-ray transport of one pack in a pool worker."""
    pack, field, first, pairs = args
    stats = dict(samples=0, angular_pairs=0)
    return transport(pack, _WORKER['cq'], field, first, None, pairs, stats), stats


def make_pool(ctx, packs, dtype=np.float32, workers=1):
    """A spawn process pool whose workers hold the context's constants and packs (for OracleSolver(pool=...)). The
    pool must be created before the process starts a GPU runtime: its workers are new interpreters."""
    import multiprocessing as mp

    c = ctx.constants(ctx.geo.path_n, dtype)
    cq = ctx.constants(ctx.geo.query_n, dtype)
    return mp.get_context('spawn').Pool(workers, initializer=_worker_init, initargs=(c, cq, packs))


class OracleSolver:
    """CPU solve of one chunk. dtype float32 = the FP32-FTZ restatement; float64 = the FP64 twin. pool: a process
    pool from make_pool over the same context and packs, which runs the per-cell transports and the observer
    queries in parallel (the same arithmetic, cell by cell)."""

    def __init__(self, ctx, dtype=np.float32, depth=4, packs=None, pool=None, workers=1):
        from sosfuse2_amd import geometry  # input preparation only

        self.ctx = ctx
        self.F = np.dtype(dtype).type
        self.c = ctx.constants(ctx.geo.path_n, dtype)
        self.cq = ctx.constants(ctx.geo.query_n, dtype)
        self.depth = depth
        self.g = geometry
        if packs is None:
            packs = [geometry.with_end_columns(ctx.state, geometry.cell_pack(ctx.geo, i), ctx.geo.domain)
                     for i in range(ctx.geo.ns)]
        self.packs = packs
        self.first_moments = None
        self.first_flux = None
        self.pool = pool
        self.workers = workers
        self.stats = dict(samples=0, angular_pairs=0)
        self.timings = dict(apply_seconds=[])

    def _cat(self, count_):
        key = f'_cat{count_}'
        if not hasattr(self, key):
            sel = self.packs[:count_]
            cat = self.g.concat_packs(sel)
            cat['end_columns'] = np.concatenate([p['end_columns'] for p in sel])
            off = np.r_[0, np.cumsum([len(p['d']) for p in sel])]
            setattr(self, key, (cat, off))
        return getattr(self, key)

    def _merge(self, stats):
        for k, v in stats.items():
            self.stats[k] += v

    def evaluate(self, field, first=False, phase=0, ground_only=False, boundary=None):
        count_ = self.ctx.geo.nc if ground_only else self.ctx.geo.ns
        if self.pool is not None:
            step = max(1, -(-count_ // (4 * self.workers)))
            jobs = [(list(range(a, min(a + step, count_))), field, first, phase, boundary)
                    for a in range(0, count_, step)]
            parts = self.pool.map(_cells_job, jobs)
            for _, _, st in parts:
                self._merge(st)
            return np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts])
        cat, off = self._cat(count_)
        value = transport(cat, self.c, field, first, boundary, PAIRS4, self.stats)
        K = count(self.c['L'])
        moments = np.zeros((count_, 4 + K, self.c['W']), 'f8')
        flux = np.zeros((count_, self.c['W']), 'f8')
        for i in range(count_):
            sub = {k: cat[k][off[i] : off[i + 1]] for k in ('d', 'dloc', 'weight')}
            moments[i], flux[i] = project(sub, value[off[i] : off[i + 1]], self.c, ground_only=phase == 1)
        return moments, flux

    def initialize(self):
        if self.first_moments is None:
            zero = np.zeros((self.ctx.geo.ns, self.ctx.rows, self.c['W']), self.F)
            self.first_moments, self.first_flux = self.evaluate(zero, first=True, phase=0)

    def compose(self, moment):
        """Source rows from incident moments (source_iteration.py:176-212)."""
        return compose(moment, self.c, self.F)

    def apply(self, field):
        self.initialize()
        t = time.perf_counter()
        _, flux = self.evaluate(field, phase=1, ground_only=True)
        boundary = flux + self.first_flux[: self.ctx.geo.nc]
        moment, _ = self.evaluate(field, phase=0, boundary=boundary)
        moment += self.first_moments
        out = self.compose(moment)
        self.timings['apply_seconds'].append(time.perf_counter() - t)
        return out

    # ----------------------------------------------------------------------------------------------- queries
    def query_raw(self, h_s, el, az, field, first, pairs=PAIRS4):
        packs = self.ctx.query_packs(h_s, el, az)
        cat = self.g.concat_packs(packs)
        if self.pool is not None:
            parts = self.pool.map(_query_job, [(p, field, first, pairs) for p in packs])
            for _, st in parts:
                self._merge(st)
            return np.concatenate([p[0] for p in parts]), cat
        return transport(cat, self.cq, field, first, None, pairs, self.stats), cat

    def query(self, h_s, el, az, field, first=False):
        """Radiance I [n, W] times the solar irradiance (engine_core.py:208-226)."""
        v, _ = self.query_raw(h_s, el, az, field, first)
        return ((v[..., 0] + v[..., 1]) + v[..., 2]).astype('f8') * self.ctx.solar[None, :]

    def boxes(self, field, record=False):
        """Observer monitor [n_solar, 7, W] of the stop elevations, or of the record ones (engine_core.py:228-238)."""
        el, az, ww = self.ctx.box_directions()
        out = []
        stop, rec = monitor_elevations(self.ctx.geo.domain.top_km)
        for h in rec if record else stop:
            v = self.query(h, el, az, field, False).reshape(len(self.ctx.directions), 16, self.c['W'])
            out.append(np.einsum('dnw,n->dw', v, ww))
        return np.stack(out)

    def solve(self, max_eval=40, tolerance=1e-5, log=None):
        W = self.c['W']
        x = np.zeros((self.ctx.geo.ns, self.ctx.rows, W), self.F)
        gs = [[] for _ in range(W)]
        fs = [[] for _ in range(W)]
        done = np.zeros(W, bool)
        history = []
        t0 = time.perf_counter()
        for iteration in range(1, max_eval + 1):
            old = self.boxes(x)
            g = self.apply(x)
            bad = int(np.count_nonzero(~np.isfinite(g)))
            if bad:
                raise RuntimeError(f'E_NONFINITE_SOURCE: {bad} values after evaluation {iteration}')
            if done.any():
                g = np.where(done[None, None, :], x, g)
            new = self.boxes(g)
            rr = np.max(abs(new - old) / np.maximum(abs(new), 1e-30), axis=(0, 1))
            ss = np.max(abs(g - x), axis=(0, 1)) / np.maximum(np.max(abs(g), axis=(0, 1)), 1e-30)
            if iteration >= 2:
                done |= np.maximum(rr, ss) <= tolerance
            history.append(dict(evaluation=iteration, observer_relative=rr.tolist(), source_relative=ss.tolist(),
                                converged=done.tolist()))
            if log:
                log(f'ORACLE EVAL {iteration} obs={max(rr):.6g} source={max(ss):.6g} done={done.sum()}/{W}')
            if done.all():
                x = g
                break
            nxt = g.copy()
            for w in range(W):
                if done[w]:
                    gs[w], fs[w] = [], []
                    continue
                gw = g[:, :, w]
                norm = np.linalg.norm(gw, axis=1)
                scale = np.maximum(norm, norm.max() * 1e-8)
                feature = ((gw - x[:, :, w]) / scale[:, None]).astype('f4').ravel()
                obs = np.asarray(
                    ((new[:, :, w] - old[:, :, w]) / np.maximum(abs(new[:, :, w]), 1e-20)).ravel()
                    * math.sqrt(len(feature) / (new.shape[0] * new.shape[1])), dtype=np.float32)
                gs[w].append(gw.copy())
                fs[w].append(np.concatenate([feature, obs]))
                if len(fs[w]) > self.depth:
                    gs[w].pop(0)
                    fs[w].pop(0)
                if len(fs[w]) > 1:
                    f = np.stack(fs[w], 1)
                    ah = anderson_weights(f.T @ f)
                    if ah is not None:
                        nxt[:, :, w] = sum(cw * q for cw, q in zip(ah, gs[w]))
            x = nxt
        if not done.all():
            raise RuntimeError(f'E_NOT_CONVERGED: {done.sum()}/{W} within {tolerance}')
        return x, dict(evaluations=iteration, history=history, seconds=time.perf_counter() - t0)

    def table(self, source, hs, el, az, stokes_out=True):
        """Table [len(hs), len(el), len(az), W] of I, or [..., 3] of I, Q, U (radiance_table.py:167-219)."""
        ee, aa = np.meshgrid(el, az, indexing='ij')
        pairs = PAIRS6 if stokes_out else PAIRS4
        out = []
        for h in hs:
            a, cat = self.query_raw(h, ee, aa, source, True, pairs)
            b, _ = self.query_raw(h, ee, aa, source, False, pairs)
            tot = a.astype('f8') + b.astype('f8')
            if stokes_out:
                v = stokes(tot, cat['view'])
            else:
                v = ((tot[..., 0] + tot[..., 1]) + tot[..., 2])[..., None]
            v = v * self.ctx.solar[None, :, None]
            out.append(v.reshape(len(el), len(az), self.c['W'], -1))
        t = np.stack(out)
        return t if stokes_out else t[..., 0]
