"""The reference's public entry points for this path, over the HIP backend:

- `create_configured(output, cfg, wavelengths, backend, state, ...)` -> (s, scope)
  (`src/sosfuse2/solve/build.py:14-118`);
- `Solver(s, geometry).solve(max_eval, tolerance)` (`src/sosfuse2/solve/source_iteration.py:309-663`), in
  `sosfuse2_amd.solver`;
- `make_table(s, source, hs, el, az, group, batch)` -> (table, seconds) (`src/sosfuse2/outputs/radiance_table.py:167-219`),
  here with Stokes I, Q, U.

Settings (`cfg`) name the grid and quadrature the way the reference's settings do
(`src/sosfuse2/solve/configuration.py:154-269`): nr, nchi, nmu, nphi, path_n, L, anderson_tol, observer_km, plus
top_km, the chi grid kind, query_n and the Anderson depth of this engine.
"""

import os
import time
from pathlib import Path

import numpy as np
import torch

from . import atmosphere, engine, geometry
from .kernel import Kernel, lane_count
from .solver import Geometry, Solver, observer_bundle

CONFIG0 = dict(nr=60, nchi=10, chi_grid='tiny', nmu=1, nphi=2, path_n=4, L=16, top_km=100.0, anderson_tol=1e-5,
               anderson_depth=5, observer_km=0.1, query_n=6, wavelengths_nm=[550.0], solar_elevation_deg=[-3.0, -6.0],
               view_shape=[4, 6])
CONFIG1 = dict(nr=150, nchi=None, chi_grid='twilight', nmu=4, nphi=8, path_n=6, L=32, top_km=120.0, anderson_tol=1e-5,
               anderson_depth=5, observer_km=0.1, query_n=12,
               wavelengths_nm=atmosphere.log_spaced_wavelengths(30).tolist(),
               solar_elevation_deg=np.arange(6.0, -19.0, -1.0).tolist(), view_shape=[32, 36])


def axes(cfg):
    """(wavelengths, solar elevations, view elevations, relative azimuths) (`radiance_table.py:34-40`)."""
    nr, nc = cfg['view_shape']
    el = 90 - (np.arange(nr) + 0.5) * 90 / nr
    az = (np.arange(nc) + 0.5) * 180 / nc
    return np.asarray(cfg['wavelengths_nm'], float), np.asarray(cfg['solar_elevation_deg'], float), el, az


def make_grid(cfg):
    top = float(cfg.get('top_km', 100.0))
    h = np.asarray(cfg['h'], float) if 'h' in cfg else geometry.height_grid(int(cfg['nr']), top)
    chi = geometry.chi_grid(cfg.get('chi_grid', 'twilight'))
    if cfg.get('nchi') not in (None, len(chi)):
        raise ValueError('E_SETTINGS_GRID_COUNT: nchi disagrees with the chi grid')
    return geometry.Grid(h=h, chi=chi, nmu=int(cfg['nmu']), nphi=int(cfg['nphi']), path_n=int(cfg['path_n']),
                         domain=geometry.Domain(top), query_n=int(cfg.get('query_n', 24)))


class Scope:
    """State-independent preparation shared by every chunk and state on one grid: the grid, the sun-column table,
    the ground pools, the cell packs and their device bundles."""

    def __init__(self, cfg, profile_state, kernel, device_columns=False, device_geometry=False):
        """device_columns: leave the column integrals (sun table, ground-pool and end columns) to the device;
        device_geometry: also build the incident quadrature, ray packs and bundles there (finish_on_device, after the
        kernel exists). The ground pools' directions are chosen here either way; host packs are built on first use
        of `packs` when the bundles come from the device."""
        t = time.perf_counter()
        self.cfg = cfg
        self.grid = make_grid(cfg)
        self.kernel = kernel
        self.state = profile_state
        self.device_columns = device_columns or device_geometry
        self.device_geometry = device_geometry
        self.workers = int(cfg.get('workers', geometry.default_workers()))
        self.bundles = None
        self._packs = None
        if self.device_columns:
            self.sun = None
            self.pools = geometry.ground_pools(profile_state, self.grid, None, workers=self.workers, columns_too=False)
            if not device_geometry:
                self._packs = geometry.cell_packs(profile_state, self.grid, workers=self.workers, end_columns=False)
        else:
            self.sun = geometry.sun_table(profile_state, self.grid.domain, int(cfg.get('sun_scale', 1)), self.workers)
            self.pools = geometry.ground_pools(profile_state, self.grid, self.sun, workers=self.workers)
            self._packs = geometry.cell_packs(profile_state, self.grid, workers=self.workers)
        self.geometry = None
        self.prepare_seconds = time.perf_counter() - t

    @property
    def packs(self):
        """Host ray packs of every cell with their end columns (built on first use when the device built the
        bundles)."""
        if self._packs is None:
            self._packs = geometry.cell_packs(self.state, self.grid, workers=self.workers)
        return self._packs

    def finish_on_device(self):
        """The column integrals on the device (device_geometry): the sun-column table, the ground pools' path columns
        and the packs' end columns, or with device_geometry the whole cell bundles."""
        from . import device_geometry as dg

        if self.sun is None:
            t = time.perf_counter()
            st, dom = self.state, self.grid.domain
            self.sun = dg.sun_table(st, dom, int(self.cfg.get('sun_scale', 1)))
            self.pools = dg.pool_columns(st, dom, self.pools, self.sun)
            if self.device_geometry:
                self.bundles = dg.cell_bundles(st, self.grid)
            else:
                dg.end_columns(st, dom, self._packs)
            self.prepare_seconds += time.perf_counter() - t
        return self

    def context(self, optics):
        return engine.Context(self.grid, optics, self.kernel, self.sun, self.pools, self.cfg.get('observer_km', 0.1),
                              self.state)

    def solver(self, s, log=print, graph=False):
        if self.geometry is None:
            if self.bundles is not None:
                self.geometry = Geometry(s, kernel=self.kernel, bundles=self.bundles)
            else:
                self.geometry = Geometry(s, self.packs, self.kernel)
        return Solver(s, self.geometry, depth=int(self.cfg.get('anderson_depth', 4)), log=log, graph=graph)


def create_configured(output, cfg, wavelengths, backend, state, block_size=256, geometry_cache=None, *,
                      share_sun_geometry=True, sun_scale=1, scope=None, mixed=False, records=True, optics=None):
    """Solver context `s` over `wavelengths` (nm) for one state, and its scope (`solve/build.py:14-118`).
    `scope` reuses another state's preparation on the same grid and profile shapes. `optics` (an atmosphere.Optics
    over the same wavelengths, e.g. atmosphere.optics_from_coefficients) replaces the state's own lane optics."""
    if output is not None:
        Path(output).mkdir(parents=True, exist_ok=True)
    if scope is None:
        cfg = dict(cfg, sun_scale=sun_scale)
        if backend != 'hip':
            raise ValueError(f"E_BACKEND: backend {backend!r}; this engine runs 'hip' only")
        # the host preparation (process pools) runs before the kernel brings up the GPU runtime; the column
        # integrals follow on the device
        scope = Scope(cfg, state, None, device_columns=bool(cfg.get('device_columns', True)),
                      device_geometry=bool(cfg.get('device_geometry', True)))
        scope.kernel = Kernel(backend=backend, L=int(cfg['L']), block_size=block_size, mixed=mixed, records=records)
        scope.finish_on_device()
    else:
        atmosphere.check_profile_shapes(state, scope.state)
    wl = np.asarray(wavelengths, float)
    if wl.ndim != 1 or len(wl) < 1:
        raise ValueError('E_WAVELENGTHS: a one-dimensional wavelength list is required')
    if optics is None:
        optics = atmosphere.lane_optics(state, wl)
    elif not np.array_equal(np.asarray(optics.wavelengths, float), wl):
        raise ValueError('E_WAVELENGTHS: the optics are over other wavelengths')
    lane_count(optics.W)
    s = scope.context(optics)
    return s, scope


def stokes_from_tensor(comp, view):
    """I, Q, U [rays, W, 3] from the six radiance tensor components [rays, W, 6] in the observer frame, on the
    components' device in their dtype (torch tensors; view [rays, 2] = elevation, azimuth in radians). With
    e_h = (-sin az, cos az, 0) and e_v = (-sin el cos az, -sin el sin az, cos el) spanning the plane normal to the view
    and P = [[c0, c4, c3], [c4, c1, c5], [c3, c5, c2]]: I = trace, Q = e_v.P.e_v - e_h.P.e_h, U = 2 e_v.P.e_h, the
    quadratic forms written out over the six components (e_h has no z component)."""
    el, az = view[:, 0, None], view[:, 1, None]
    se, ce, sa, ca = torch.sin(el), torch.cos(el), torch.sin(az), torch.cos(az)
    h0, h1 = -sa, ca
    v0, v1, v2 = -se * ca, -se * sa, ce
    c0, c1, c2, c3, c4, c5 = comp.unbind(-1)
    I = (c0 + c1) + c2
    vpv = (v0 * v0) * c0 + (v1 * v1) * c1 + (v2 * v2) * c2 + 2 * ((v0 * v2) * c3 + (v0 * v1) * c4 + (v1 * v2) * c5)
    hph = (h0 * h0) * c0 + (h1 * h1) * c1 + 2 * (h0 * h1) * c4
    vph = (v0 * h0) * c0 + (v1 * h1) * c1 + (v2 * h0) * c3 + (v0 * h1 + v1 * h0) * c4 + (v2 * h1) * c5
    return torch.stack([I, vpv - hph, 2 * vph], -1)


def make_table(s, source, hs, el, az, group=64, batch=25, log=print, stokes=True):
    """Table [len(hs), len(el), len(az), W, 3] of I, Q, U (or [..., W] of I) and per-batch seconds.
    first + multiple orders times the solar irradiance, as `radiance_table.make_table`."""
    if group < 1:
        raise ValueError('E_TABLE_GROUP: group must be positive')
    if batch < 1:
        raise ValueError('E_TABLE_BATCH: batch must be positive')
    k = s.kernel
    G = lane_count(s.W)
    if not torch.is_tensor(source):
        src = np.asarray(source, 'f4')
        if src.shape != (s.geo.ns, s.rows, s.W):
            raise ValueError(f'E_SOURCE_SHAPE: {src.shape}')
        if G > s.W:
            src = np.concatenate([src, np.repeat(src[..., -1:], G - s.W, -1)], -1)
        source = torch.from_numpy(np.ascontiguousarray(src)).to(k.device)
    elif tuple(source.shape) != (s.geo.ns, s.rows, G):
        raise ValueError(f'E_SOURCE_SHAPE: {tuple(source.shape)}')
    ee, aa = np.meshgrid(el, az, indexing='ij')
    pairs = 6 if stokes else 4
    table = np.empty((len(hs), len(el), len(az), s.W) + ((3,) if stokes else ()), 'f4')
    seconds = []
    dc = k.constants(s, s.geo.query_n)
    plane = k.plane(s, source, dc)
    solar = s.solar
    for b0 in range(0, len(hs), batch):
        t = time.monotonic()
        part = list(range(b0, min(b0 + batch, len(hs))))
        host, db = observer_bundle(s, [hs[j] for j in part], ee, aa)
        one = k.radiance(db, s, None, True, None, pairs)
        many = k.radiance(db, s, source, False, None, pairs, plane)
        # I (Q, U) and the solar scaling in FP64 on the device; the table leaves it in FP32
        tot = (one.double() + many.double())[:, : s.W]
        if stokes:
            view = getattr(host, 'view_dev', None)
            if view is None or view.device != tot.device:
                view = host.view_dev = torch.from_numpy(np.ascontiguousarray(host.view, 'f8')).to(tot.device)
            v = stokes_from_tensor(tot, view)
        else:
            v = (tot[..., 0] + tot[..., 1]) + tot[..., 2]
        sol = torch.from_numpy(np.asarray(solar, 'f8')).to(tot.device)
        v = (v * (sol[None, :, None] if stokes else sol[None, :])).float().cpu().numpy()
        off = 0
        for j, n in zip(part, host.sizes):
            spec = v[off : off + n].reshape(len(el), len(az), *v.shape[1:])
            off += n
            I = spec[..., 0] if stokes else spec
            if not np.isfinite(spec).all() or (I < 0).any():
                raise RuntimeError('E_RADIANCE_INVALID: nonfinite/negative')
            table[j] = spec
        seconds.append(time.monotonic() - t)
        if log:
            log(f'TABLE sun={hs[part[0]]:+g}..{hs[part[-1]]:+g} rays={db.rays} seconds={seconds[-1]:.3f}')
    return table, seconds


UNITS = 'W m^-2 sr^-1 nm^-1'
TABLE_AXES = ('solar_elevation_deg', 'view_elevation_deg', 'relative_azimuth_deg', 'wavelength_nm')


def save_table(path, table, wl, hs, el, az):
    """Write a table as the reference's table.npz (`src/sosfuse2/outputs/radiance_table.py:319-340`): `radiance`
    f32 [solar elevation, view elevation, relative azimuth, wavelength] of total radiance in W m^-2 sr^-1 nm^-1, the
    four axes as f32 (`wavelength_nm`, `solar_elevation_deg`, `view_elevation_deg`, `relative_azimuth_deg`) and
    `units`. A table with Stokes components ([..., W, 3]) also stores `stokes_q` and `stokes_u` of the same shape,
    which readers of the reference's format ignore. Written to a temporary name and renamed into place."""
    table = np.asarray(table, 'f4')
    wl, hs, el, az = (np.asarray(a, 'f8') for a in (wl, hs, el, az))
    shape = (len(hs), len(el), len(az), len(wl))
    extra = {}
    if table.shape == shape + (3,):
        extra = dict(stokes_q=np.ascontiguousarray(table[..., 1]), stokes_u=np.ascontiguousarray(table[..., 2]))
        table = table[..., 0]
    if table.shape != shape:
        raise ValueError(f'E_TABLE_SHAPE: {table.shape} for axes {shape}')
    path = Path(path)
    tmp = path.with_name(path.name + '.tmp')
    with tmp.open('wb') as f:
        np.savez_compressed(f, radiance=np.ascontiguousarray(table), wavelength_nm=wl.astype('f4'),
                            solar_elevation_deg=hs.astype('f4'), view_elevation_deg=el.astype('f4'),
                            relative_azimuth_deg=az.astype('f4'), units=np.array(UNITS), **extra)
    os.replace(tmp, path)
    return path


def read_table(path):
    """((solar elevation, view elevation, relative azimuth, wavelength) axes, radiance) of a table.npz, with the
    reference's checks (`radiance_table.py:291-300`): finite one-dimensional axes, radiance f32 of their shape,
    finite, and the units."""
    with np.load(path, allow_pickle=False) as archive:
        axes = tuple(archive[k] for k in TABLE_AXES)
        radiance = archive['radiance']
        if (any(a.ndim != 1 or not a.size or a.dtype.kind not in 'fiu' or not np.isfinite(a).all() for a in axes)
                or radiance.shape != tuple(len(a) for a in axes) or radiance.dtype != np.dtype('f4')
                or not np.isfinite(radiance).all() or archive['units'].item() != UNITS):
            raise ValueError('E_TABLE_ARRAY: invalid axes, radiance or units')
        return axes, radiance
