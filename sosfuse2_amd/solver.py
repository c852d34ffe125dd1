"""Source iteration on the device: first light, apply (ground pass, boundary, warm pass, composition), the observer
monitor and Anderson mixing (`src/sosfuse2/solve/source_iteration.py:100-663`, `src/sosfuse2/solve/ray_geometry.py:395-451`).

Everything of size cells x rows stays in HBM: the iterate, its Anderson history rings (sources and residuals), the
Gram matrices (FP64) and their solve. Per evaluation the host receives the per-lane metrics and the observer boxes
(a few hundred floats) to decide convergence, as the reference does.
"""

import math
import time

import numpy as np
import torch

from . import geometry, hip
from .kernel import lane_count


class Geometry:
    """Cell packs of one grid and their device bundles (the reference's TransientGeometry). bundles: the (warm,
    ground) device bundles when they were built on the device (device_geometry.cell_bundles); no host packs then."""

    def __init__(self, s, packs=None, kernel=None, bundles=None):
        self.s = s
        k = kernel or s.kernel
        t = time.perf_counter()
        if bundles is not None:
            self.packs = None
            self.bundle, self.ground_bundle = bundles
            self.stats = dict(cells=s.geo.ns, rays=self.bundle.rays, samples=self.bundle.samples,
                              seconds=time.perf_counter() - t, device_bytes=self.bundle.nbytes())
            return
        if packs is None:
            packs = [geometry.with_end_columns(s.state, geometry.cell_pack(s.geo, i), s.geo.domain)
                     for i in range(s.geo.ns)]
        self.packs = packs
        self.host_bundle = geometry.Bundle(packs)
        self.host_ground = geometry.Bundle(packs[: s.geo.nc])
        self.bundle = k.bundle(self.host_bundle)
        self.ground_bundle = k.bundle(self.host_ground)
        self.stats = dict(cells=s.geo.ns, rays=self.host_bundle.rays, samples=self.host_bundle.samples,
                          seconds=time.perf_counter() - t, device_bytes=self.bundle.nbytes())

    def evaluate(self, s, field, first=False, phase=0, ground_only=False, boundary=None, plane=None):
        b = self.ground_bundle if ground_only else self.bundle
        moments, flux, _ = s.kernel.launch(b, s, field, first, phase, boundary, plane)
        return moments, flux


def observer_bundle(s, hs, el, az, kernel=None):
    """One device bundle of the observer rays of several solar elevations (packs of 64 rays, as engine_core.query).
    The bundle depends on the grid and the view axes only, so it is kept per kernel and reused by every state."""
    k = kernel or s.kernel
    el, az = np.asarray(el, 'f8'), np.asarray(az, 'f8')
    g = s.geo
    key = (np.asarray(g.h).tobytes(), np.asarray(g.chi).tobytes(), g.query_n, float(g.domain.top_km),
           float(s.observer_km), tuple(float(h) for h in hs), el.shape, el.tobytes(), az.tobytes())
    cache = k.__dict__.setdefault('_observer_bundles', {})
    if key in cache:
        return cache[key]
    if len(cache) >= 16:
        cache.pop(next(iter(cache)))
    cache[key] = _observer_bundle(s, hs, el, az, k)
    return cache[key]


def drop_observer_bundles(kernel):
    """Release the observer bundles (and their plans) kept on a kernel; they are rebuilt on next use."""
    kernel.__dict__.pop('_observer_bundles', None)


def _observer_bundle(s, hs, el, az, k):
    packs = []
    sizes = []
    for h in hs:
        p = s.query_packs(h, el, az)
        packs += p
        sizes.append(sum(len(x['d']) for x in p))
    host = geometry.Bundle(packs, end_columns=False)
    host.sizes = sizes
    host.view = np.concatenate([x['view'] for x in packs])
    return host, k.bundle(host)


class Solver:
    """Fixed-point iteration of the source field of one chunk."""

    def __init__(self, s, geometry_, depth=4, log=print, graph=False):
        self.s = s
        self.geometry = geometry_
        self.kernel = s.kernel
        self.depth = depth
        self.log = log
        self.graph = graph  # replay the evaluation A(x) as one captured HIP graph
        self._graph = None
        self.first_moments = None
        self.first_flux = None
        self.G = lane_count(s.W)
        self.timings = dict(apply_seconds=[], setup_seconds=0.0)
        el, az, ww = s.box_directions()
        self._box = observer_bundle(s, tuple(s.monitors) + tuple(getattr(s, 'record_monitors', ())), el, az)
        self._ww = ww

    @property
    def device(self):
        return self.kernel.device

    def zeros(self):
        return torch.zeros((self.s.geo.ns, self.s.rows, self.G), dtype=torch.float32, device=self.device)

    def initialize(self):
        if self.first_moments is None:
            t = time.perf_counter()
            self.first_moments, self.first_flux = self.geometry.evaluate(self.s, None, first=True, phase=0)
            self.kernel.sync()
            self.timings['setup_seconds'] = time.perf_counter() - t

    def apply(self, field):
        """g = A(x) (`source_iteration.py:161-216`). With graph=True the launches of one evaluation (plane, SH
        records, ground and warm transports, projections, composition) are captured once into a HIP graph and
        replayed on a static input buffer; the result is a fresh tensor either way."""
        if self.graph:
            return self._apply_graph(field)
        return self._apply(field)

    def _apply_graph(self, field):
        if self._graph is None:
            self.initialize()  # first light before the capture
            bundle = self.geometry.bundle
            if len(self.kernel.parts(bundle, self.kernel.constants(self.s, bundle.nq))) > 1:
                # a bundle run in parts builds each part's plan at its launch (host synchronisation): no capture
                self.graph = False
                return self._apply(field)
            self._gx = torch.empty_like(field)
            self._gx.copy_(field)
            timed, self.kernel.timed = self.kernel.timed, None
            main = torch.cuda.current_stream(self.device)
            side = torch.cuda.Stream(self.device)
            side.wait_stream(main)
            with torch.cuda.stream(side):
                # the first evaluation runs eagerly on the capture stream: it builds the plans and constants the
                # capture needs, settles the stream's allocator state, and is this call's result
                first = self._apply(self._gx)
            main.wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._gy = self._apply(self._gx)
            self.kernel.timed = timed
            self._graph = g
            return first
        self._gx.copy_(field)
        self._graph.replay()
        return self._gy.clone()

    def _apply(self, field):
        self.initialize()
        t = time.perf_counter()
        s, k = self.s, self.kernel
        dc = k.constants(s, self.geometry.bundle.nq)
        plane = self._plane_of(field, take=True)
        _, flux = self.geometry.evaluate(s, field, phase=1, ground_only=True, plane=plane)
        boundary = flux + self.first_flux[: s.geo.nc]
        moment, _ = self.geometry.evaluate(s, field, phase=0, boundary=boundary, plane=plane)
        out = torch.empty_like(field)
        hip.call('sos_compose', dc.struct, s.rows, hip.ptr(moment), hip.ptr(self.first_moments), hip.ptr(dc.factor),
                 hip.ptr(out), k.stream())
        self.timings['apply_seconds'].append(time.perf_counter() - t)
        return out

    def _plane_of(self, field, take=False):
        """The SH plane of a field. The observer boxes of x and the evaluation of x read the same plane, so the
        boxes keep it for the evaluation that follows (matched by tensor identity, used once)."""
        last = getattr(self, '_last_plane', None)
        if last is not None and last[0] is field:
            plane = last[1]
        else:
            plane = self.kernel.plane(self.s, field, self.kernel.constants(self.s, self.geometry.bundle.nq))
        self._last_plane = None if take else (field, plane)
        return plane

    def boxes(self, field):
        """Observer monitor [stop + record solar elevations, 7, G] (`engine_core.py:228-238`)."""
        host, db = self._box
        rad = self.kernel.radiance(db, self.s, field, False, None, 4, self._plane_of(field))
        I = ((rad[..., 0] + rad[..., 1]) + rad[..., 2]).double().cpu().numpy()
        solar = np.r_[self.s.solar, np.full(self.G - self.s.W, self.s.solar[-1])] if self.G > self.s.W else self.s.solar
        I = I * solar[None, :]
        nd = len(self.s.directions)
        return np.einsum('hdnw,n->hdw', I.reshape(-1, nd, 16, self.G), self._ww)

    def metrics(self, g, x, done):
        out = torch.empty((3, self.G), dtype=torch.float32, device=self.device)
        frozen = int(sum(1 << i for i in np.flatnonzero(done)))
        hip.call('sos_source_metrics', g.numel() // self.G, self.G, hip.ptr(g), hip.ptr(x), frozen, hip.ptr(out),
                 self.kernel.stream())
        m = out.cpu().numpy()
        return int(m[0].sum()), m[2] / np.maximum(m[1], np.float32(1e-30))

    def solve(self, max_eval=40, tolerance=1e-5, x0=None, fallback=True):
        """Iterate to max(observer, source) relative change <= tolerance on every lane (from evaluation 2).
        x0: a starting source (a warm start); it only seeds the iteration. With fallback, a seeded iteration that does
        not converge within max_eval evaluations is run again from the cold start, so a start can cost evaluations
        but never the answer; info['evaluations'] then counts both runs and info['fallback'] is set."""
        x, info, ok = self._iterate(max_eval, tolerance, x0)
        if not ok and x0 is not None and fallback:
            if self.log:
                self.log(f'SOSFUSE seeded iteration not converged after {info["evaluations"]} evaluations; '
                         f'solving from the cold start')
            spent = info['evaluations']
            x, info, ok = self._iterate(max_eval, tolerance, None)
            info = dict(info, evaluations=info['evaluations'] + spent, evaluations_seeded=spent, fallback=True)
        if not ok:
            W = self.s.W
            raise RuntimeError(f'E_NOT_CONVERGED: {sum(info["converged"][:W])}/{W} within {tolerance}')
        return x, info

    def _iterate(self, max_eval, tolerance, x0):
        """One run of the iteration: (x, info, converged on every lane)."""
        s, k = self.s, self.kernel
        G, D = self.G, self.depth
        W = s.W
        C, rows = s.geo.ns, s.rows
        x = self.zeros() if x0 is None else x0.clone()
        done = np.zeros(G, bool)
        history = []
        nobs = len(s.monitors) * len(s.directions)
        n = C * rows + nobs
        gring = torch.empty((D, C, rows, G), dtype=torch.float32, device=self.device)
        fring = torch.empty((D, G, n), dtype=torch.float32, device=self.device)
        gram = torch.zeros((G, D, D), dtype=torch.float64, device=self.device)
        work = torch.empty((C * G + G,), dtype=torch.float32, device=self.device)
        gwork = torch.empty((128 * G * D,), dtype=torch.float64, device=self.device)
        wts = torch.empty((G, D), dtype=torch.float32, device=self.device)
        order_dev = torch.zeros((D,), dtype=torch.int32, device=self.device)
        count = 0
        started = time.perf_counter()
        evaluations = 0
        for iteration in range(1, max_eval + 1):
            evaluations = iteration
            t0 = time.perf_counter()
            old = self.boxes(x)
            g = self.apply(x)
            nonfinite, ss = self.metrics(g, x, done)
            if nonfinite:
                raise RuntimeError(f'E_NONFINITE_SOURCE: {nonfinite} values after evaluation {iteration}')
            if done.any():
                mask = torch.from_numpy(done).to(self.device)
                g = torch.where(mask[None, None, :], x, g)
            new = self.boxes(g)
            # the stop monitors decide convergence and enter the residual; record monitors are logged only
            # (`source_iteration.py:369-391`)
            nstop = len(s.monitors)
            rec_old, rec_new = old[nstop:], new[nstop:]
            old, new = old[:nstop], new[:nstop]
            rr = np.max(abs(new - old) / np.maximum(abs(new), 1e-30), axis=(0, 1))
            if iteration >= 2:
                done |= np.maximum(rr, ss) <= tolerance
            history.append(dict(evaluation=iteration, observer_relative=rr[:W].tolist(),
                                source_relative=ss[:W].tolist(), converged=done[:W].tolist(),
                                seconds=time.perf_counter() - t0))
            if len(rec_new):
                history[-1]['record_observer_relative'] = np.max(
                    abs(rec_new - rec_old) / np.maximum(abs(rec_new), 1e-30), axis=1)[:, :W].tolist()
            if self.log:
                self.log(f'SOSFUSE EVAL {iteration} obs={max(rr[:W]):.6g} source={max(ss[:W]):.6g} '
                         f'done={done[:W].sum()}/{W}')
            if done[:W].all():
                x = g
                break
            slot = count % D
            count += 1
            m = min(count, D)
            order = [(count - m + j) % D for j in range(m)]
            gring[slot].copy_(g)
            obs = ((new - old) / np.maximum(abs(new), 1e-20)).reshape(nobs, G).T * math.sqrt(C * rows / nobs)
            obs_dev = torch.from_numpy(np.ascontiguousarray(obs, dtype=np.float32)).to(self.device)
            hip.call('sos_anderson_residual', C, rows, G, hip.ptr(g), hip.ptr(x), hip.ptr(obs_dev), nobs,
                     hip.ptr(work), hip.ptr(fring[slot]), k.stream())
            filled = sum(1 << o for o in order)
            hip.call('sos_anderson_gram', n, G, D, slot, filled, hip.ptr(fring), hip.ptr(gwork), hip.ptr(gram),
                     k.stream())
            x = g.clone()
            if m > 1:
                order_dev.copy_(torch.tensor(order + [0] * (D - m), dtype=torch.int32))
                frozen = int(sum(1 << i for i in np.flatnonzero(done)))
                hip.call('sos_anderson_solve', G, D, m, hip.ptr(order_dev), hip.ptr(gram), frozen, hip.ptr(wts),
                         k.stream())
                hip.call('sos_anderson_mix', C, rows, G, D, m, hip.ptr(gring), hip.ptr(order_dev), hip.ptr(wts),
                         hip.ptr(x), k.stream())
        k.sync()
        self._last_plane = None
        return x, dict(evaluations=evaluations, tolerance=tolerance, converged=done[:W].tolist(), history=history,
                       seconds=time.perf_counter() - started), bool(done[:W].all())
