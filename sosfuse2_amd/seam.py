"""The reference's backend seam for backend 'hip': a `Kernel` with the constructor, members and call contracts of
`src/sosfuse2/solve/source_solver.py:1109-2484` (`setup` :1243-1339, `launch` :2428-2456, `transport` :2458-2484)
over this package's HIP kernels.

It consumes what the reference hands its kernel:
- a `ray_geometry.Bundle` (`src/sosfuse2/solve/ray_geometry.py:290-392`): `.cells`, `.nq`, `.rays` and the host CSR
  arrays `.arrays` = cell_off i8[C+1], starts f8[C,3], targets f8[C,3,3], dirs f8[R,3], weights f4[R], edge_off
  u4[R+1], edges f8[E], lower u2[R], end_j i4[R], end_f f4[R], end_col f4[R,5], hits u1[R] (the bundle built with
  xp=numpy, before the reference's own FP32 device packing);
- the solver context `s` with the attributes `setup` reads (`source_solver.py:1243-1339`): s.W, s.geo.h, s.geo.chi,
  s.geo.ns, s.geo.sun.h, s.geo.sun.col [nh, nx, 5], s.ext [5, W], s.asca [W], s.hr, s.ha, s.delta,
  s.fuse_state['parameters']['albedo'], the ground pools s.goff, s.gdir, s.gval [P, W] and s.gm [C, 4, W], and
  s.density_optics (the table the reference's attach_density returns: table [n, 3] of air, ozone and NO2 on the
  profile grid, mask, air_surface) when the profiles are tabulated;
- the kernel's own height (aerosol scale height), g (aerosol asymmetry), L and domain (top, profile spacing).

Profiles follow the reference's profile_at (`src/sosfuse2/backends/numpy_reference.py:187-223`): exp(-h / 8) air
(the table's air column when mask bit 1), exp(-h / height) aerosol, a triangle ozone peaking at 25 km with 15 km half
width (the table's ozone column when mask bit 2), the table's NO2 column, O4 = (air / air_surface)^2. Not supported
here, each raising an error: the sulfate layer's tabulated Mie phase (STRATO), the tropospheric Mie and two-layer
aerosol options, and tops above 80 km (the reference's extended-domain support switch).

Array module: with array_module='numpy' (the default) the kernel's `xp` is NumPy and launch takes and returns host
arrays, as the reference's numpy backend does, so the reference's own Solver (its NumPy branch of
`src/sosfuse2/solve/source_iteration.py:309-663`), TransientGeometry and make_table run over the HIP transport
unchanged, the moments crossing PCIe once per launch. With array_module='torch' the arrays stay on the device as
torch tensors (the role CuPy arrays play for the reference's cupy backend). launch gives moments f32[cells, 4+K, W]
and flux f32[cells, W] (phase 0, 1) or the query f32[rays, W, 4] (phase 2); transport gives host f8[n, W, 2, 4]
with lane 0 filled, as the reference's adapter does.
"""

import math
from dataclasses import dataclass
from types import SimpleNamespace

import numpy as np
import torch

from . import atmosphere, geometry, layout
from .kernel import DeviceBundle, Kernel as _Core, lane_count, _pad_lanes

BUNDLE_FIELDS = ('cell_off', 'starts', 'targets', 'dirs', 'weights', 'edge_off', 'edges', 'lower', 'end_j', 'end_f',
                 'end_col', 'hits')


@dataclass(frozen=True)
class Domain(geometry.Domain):
    """The reference's top.Domain as this seam needs it (`src/sosfuse2/foundation/top.py:13-57`): the top, the
    profile spacing, and the monitor sets in the reference's dict form."""

    step_km: float = 0.5

    def monitors(self):
        stop, record = geometry.Domain.monitors(self)
        return dict(stop=list(stop), acceleration=list(stop), record=list(record))


def sh_basis(mu, L):
    """Normalised cosine-only SH basis without its azimuth factor, FP64 [n, K]: sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!)
    P_l^m(mu) (Condon-Shortley phase), times sqrt(2) for m > 0, k = l (l + 1) / 2 + m (`arithmetic.py:86-109`)."""
    mu = np.asarray(mu, 'f8')
    K = layout.count(L)
    out = np.zeros((len(mu), K))
    st = np.sqrt(np.maximum(0.0, 1 - mu * mu))
    pmm = np.ones_like(mu)
    for m in range(L + 1):
        if m > 0:
            pmm = -pmm * (2 * m - 1) * st
        p_prev, p = None, pmm
        for l in range(m, L + 1):
            if l == m + 1:
                p_prev, p = p, mu * (2 * m + 1) * pmm
            elif l > m + 1:
                p_prev, p = p, ((2 * l - 1) * mu * p - (l + m - 1) * p_prev) / (l - m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.exp(math.lgamma(l - m + 1) - math.lgamma(l + m + 1)))
            out[:, l * (l + 1) // 2 + m] = norm * p * (math.sqrt(2.0) if m else 1.0)
    return out


class _Context:
    """This package's view of the reference's solver context `s` (what its kernels and device constants read)."""

    def __init__(self, s, kernel):
        if getattr(s, 'strato', None) is not None:
            raise ValueError('E_SEAM_STRATO: the sulfate layer with a tabulated Mie phase is not supported by the hip '
                             'seam')
        for name in ('tropo', 'aer2'):
            if getattr(s, name, None) is not None:
                raise ValueError(f'E_SEAM_OPTION: {name} is not supported by the hip seam')
        dom = kernel.domain
        if dom.top_km > geometry.NORMAL_TOP_KM:
            raise ValueError('E_SEAM_TOP: tops above 80 km (the extended-domain support switch) are not supported')
        self.s = s
        self.kernel = kernel.core
        W = int(s.W)
        self.W = W
        h = np.asarray(s.geo.h, 'f8')
        chi = np.asarray(s.geo.chi, 'f8')
        self.geo = SimpleNamespace(h=h, chi=chi, nh=len(h), nc=len(chi), ns=len(h) * len(chi),
                                   domain=geometry.Domain(dom.top_km))
        self.strato = False
        mix = (1 - float(s.delta)) / (1 + float(s.delta) / 2)
        albedo = float(s.fuse_state['parameters']['albedo'])
        zeros = np.zeros(W)
        self.optics = SimpleNamespace(
            ext=np.asarray(s.ext, 'f8').reshape(5, W), asca=np.asarray(s.asca, 'f8').reshape(W),
            g=np.full(W, float(kernel.g)), albedo=np.full(W, albedo), mix=np.full(W, mix), strato_ext=zeros,
            strato_sca=zeros, strato_g=zeros, W=W, strato=False)
        self.given_pools = (np.asarray(s.goff, 'i8'), np.asarray(s.gdir, 'f8').reshape(-1, 3),
                            np.asarray(s.gval, 'f8').reshape(-1, W), np.asarray(s.gm, 'f8').reshape(-1, 4, W))
        self.spec = self._profiles(s, kernel)
        sun_col = np.asarray(s.geo.sun.col, 'f8')
        self.sun_h = np.asarray(s.geo.sun.h, 'f8')
        self.sun_col = np.concatenate([sun_col[..., :5], np.zeros(sun_col.shape[:2] + (1,))], -1)
        self.hr = np.asarray(s.hr, 'f8')
        self.ha = np.asarray(s.ha, 'f8')
        self._constants = {}

    @staticmethod
    def _profiles(s, kernel):
        """The reference's profile_at shapes as a ProfileSpec (numpy_reference.py:187-223)."""
        kinds = [atmosphere.EXP, atmosphere.EXP, atmosphere.TRIANGLE, atmosphere.ZERO, atmosphere.AIR_SQUARED,
                 atmosphere.ZERO]
        a = [8.0, float(kernel.height), 15.0, 1.0, 1.0, 1.0]
        b = [0.0, 0.0, 25.0, 0.0, 0.0, 0.0]
        density = getattr(s, 'density_optics', None)
        table = None
        if density is not None:
            tab = np.asarray(density['table'], 'f8')
            mask = int(density.get('mask', 0))
            rows = np.zeros((6, tab.shape[0]))
            step = float(kernel.domain.step_km)
            if mask & 1:
                kinds[0], a[0], rows[0] = atmosphere.TABLE, step, tab[:, 0]
            if mask & 2:
                kinds[2], a[2], b[2], rows[2] = atmosphere.TABLE, step, 0.0, tab[:, 1]
            if tab.shape[1] > 2:
                kinds[3], a[3], rows[3] = atmosphere.TABLE, step, tab[:, 2]
            a[4] = float(density.get('air_surface', 1.0))
            table = tuple(tuple(r) for r in rows)
        return atmosphere.ProfileSpec(kinds=tuple(kinds), a=tuple(a), b=tuple(b), table=table)

    @property
    def L(self):
        return self.kernel.L

    @property
    def K(self):
        return layout.count(self.kernel.L)

    @property
    def rows(self):
        return layout.source_rows(self.kernel.L, False)

    def constants(self, nq, dtype=np.float32, padded=False):
        c = self._constants.get(nq)
        if c is None:
            u, wu = geometry.gl(nq)
            F = np.dtype(dtype).type
            c = dict(u=u, wu=wu, h=self.geo.h.astype(F), chi=self.geo.chi.astype(F), hr=self.hr.astype(F),
                     ha=self.ha.astype(F), strato_density=np.zeros(self.geo.nh, F), sun_h=self.sun_h.astype(F),
                     sun_col=self.sun_col.astype(F), strato=False, top=float(self.geo.domain.top_km),
                     profile=dict(air_H=1.0, aer_H=1.0, o3_c=0.0, o3_s=1.0, no2_c=0.0, no2_s=1.0, st_c=0.0, st_s=1.0,
                                  spec=self.spec))
            self._constants[nq] = c
        return c

    def eigen(self):
        """Composition factors [rows, W] (`source_iteration.py:176-208`): the device constants carry them, though a
        launch does not compose."""
        o = self.optics
        lv = layout.degrees(self.L)
        out = np.zeros((self.rows, self.W))
        out[:4] = (3 / (8 * math.pi)) * o.ext[0][None, :]
        out[4:10] = layout.rayleigh_factors()[:, None] * o.ext[0][None, :]
        out[10:] = o.g[None, :] ** lv[:, None] * o.asca[None, :]
        return out


class _HostBundle:
    """A reference Bundle's host CSR arrays in the fields DeviceBundle uploads."""

    def __init__(self, bundle):
        a = {k: np.asarray(bundle.arrays[k]) for k in BUNDLE_FIELDS}
        if a['starts'].shape[1:] != (3,) or a['starts'].dtype != np.float64:
            raise ValueError('E_SEAM_BUNDLE: the hip seam takes the host bundle (xp=numpy: FP64 starts [cells, 3])')
        self.groups = int(bundle.cells)
        self.nq = int(bundle.nq)
        self.rays = int(bundle.rays)
        self.group_off = a['cell_off'].astype('i8')
        counts = np.diff(self.group_off)
        self.ray_group = np.repeat(np.arange(self.groups, dtype='i4'), counts)
        self.starts = a['starts'].astype('f8')
        self.targets = a['targets'].astype('f8').reshape(-1, 3, 3)
        self.dirs = a['dirs'].astype('f8')
        # the direction in each ray's target frame (the frame's vectors are the columns of targets)
        self.dloc = np.einsum('nij,ni->nj', self.targets[self.ray_group], self.dirs)
        self.weights = a['weights'].astype('f4')
        self.edge_off = a['edge_off'].astype('i8')
        self.edges = a['edges'].astype('f8')
        self.lower = a['lower'].astype('u2')
        self.end_j = a['end_j'].astype('i4')
        self.end_f = a['end_f'].astype('f4')
        end = a['end_col'].astype('f4').reshape(self.rays, -1)
        self.end_col = np.concatenate([end[:, :5], np.zeros((self.rays, 1), 'f4')], 1)
        self.hits = a['hits'].astype('u1')
        spans = np.diff(self.edges)
        last = self.edge_off[1:] - 1
        inside = np.ones(len(spans), bool)
        inside[last[last < len(spans)]] = False
        self.samples = int(np.sum((spans > 1e-8) & inside)) * self.nq


class _RayPack:
    """A one-group Bundle of one raw query pack (the reference's `transport` adapter builds Bundle([geometry]))."""

    def __init__(self, g):
        self.cells, self.nq, self.rays = 1, int(g['nq']), len(g['d'])
        ec, ev, fl = geometry.compact_edges(np.asarray(g['edges'], 'f8'), np.asarray(g['lower'], bool))
        end = np.asarray(g.get('end_columns', np.zeros((self.rays, 5))), 'f4')
        self.arrays = dict(cell_off=np.array([0, self.rays], 'i8'), starts=np.asarray(g['p'][:1], 'f8'),
                           targets=np.asarray(g['target'][:1], 'f8'), dirs=np.asarray(g['d'], 'f8'),
                           weights=np.asarray(g.get('weight', np.ones(self.rays)), 'f4'),
                           edge_off=np.r_[0, np.cumsum(ec)].astype('u4'), edges=ev, lower=fl,
                           end_j=np.asarray(g['end_index'], 'i4'), end_f=np.asarray(g['end_fraction'], 'f4'),
                           end_col=end[:, :5], hits=np.asarray(g['hits_ground'], 'u1'))


class Kernel:
    """Backend 'hip' behind the reference's Kernel contract (`source_solver.py:1109-1242`): the constructor's
    arguments, `setup`, `launch`, `transport`, and the members the reference's Solver, TransientGeometry and
    radiance_table read (xp, L, K, height, g, domain, deterministic, metrics, constants, device_constants, sync,
    host, memory, basis, constants_device, ksplit, ksplit_memory, ksplit_preflight)."""

    def __init__(self, threads=4, height=1.2, g=0.8, backend='hip', block_size=128, flags=None, L=32, *,
                 profile_step_km=0.5, domain=None, device='cuda', array_module='numpy'):
        if backend != 'hip':
            raise ValueError(f"E_BACKEND: this kernel runs backend 'hip' (got {backend!r})")
        if array_module not in ('numpy', 'torch'):
            raise ValueError(f"E_ARRAY_MODULE: numpy or torch (got {array_module!r})")
        if block_size not in (32, 64, 128, 256):
            raise ValueError('E_BLOCK_SIZE: select 32,64,128,256')
        self.core = _Core(backend='hip', L=L, device=device)
        self.backend = backend
        self.threads = threads
        self.block_size = block_size
        self.L, self.K = L, layout.count(L)
        self.height, self.g = float(height), float(g)
        self.domain = domain or Domain(top_km=geometry.NORMAL_TOP_KM, step_km=float(profile_step_km))
        self.profile_step_km = float(getattr(self.domain, 'step_km', profile_step_km))
        self.flags = tuple(flags or ())
        self.deterministic = True  # fixed-order sums, no floating-point atomics: repeat runs are bitwise equal
        self.ksplit = False
        self.ksplit_mode = 0
        self.xp = np if array_module == 'numpy' else torch
        self.constants = {}  # setup() results, keyed (id(s), nq)
        self.device_constants = {}
        self.metrics = dict(samples=0, angular_pairs=0, launches=0, query_launches=0)
        self.attributes = {}
        self.cuda_events = []
        self.pool_peaks = dict(used_bytes_sampled=0, reserved_bytes_sampled=0)
        self._contexts = {}

    # ------------------------------------------------------------------------------------------------- members
    @property
    def device(self):
        return self.core.device

    def sync(self):
        self.core.sync()

    def host(self, array):
        if torch.is_tensor(array):
            return array.detach().cpu().numpy()
        return np.asarray(array)

    def memory(self):
        self.pool_peaks['used_bytes_sampled'] = max(self.pool_peaks['used_bytes_sampled'],
                                                    int(torch.cuda.memory_allocated(self.device)))
        self.pool_peaks['reserved_bytes_sampled'] = max(self.pool_peaks['reserved_bytes_sampled'],
                                                        int(torch.cuda.memory_reserved(self.device)))

    def basis(self, dirs, L):
        """SH basis rows times cos(m phi) of each direction, FP64 [n, K] (`source_solver.py` Kernel.basis)."""
        layout.count(L)
        if L != self.L:
            raise ValueError('E_SH_L_MISMATCH: basis and kernel degree differ')
        dirs = np.asarray(dirs, 'f8')
        y = sh_basis(dirs[:, 2], L)
        st = np.sqrt(np.maximum(0.0, 1 - dirs[:, 2] ** 2))
        cp = np.clip(np.where(st > 1e-12, dirs[:, 0] / np.maximum(st, 1e-30), 1.0), -1.0, 1.0)
        m = layout.orders(L)
        return y * np.cos(m[None, :] * np.arccos(cp)[:, None])

    def ksplit_memory(self):
        """Free device bytes (the reference's estimate of device memory left during a warm pass)."""
        free, _ = torch.cuda.mem_get_info(self.device)
        return int(free)

    def ksplit_preflight(self, s, bundles):
        return dict(requested=0, effective=0, reason='hip: one transport path (no split launches)')

    def constants_device(self, s, nq):
        """Device constants of s for a node count (uploaded once, kept in device_constants)."""
        key = (id(s), nq)
        dc = self.device_constants.get(key)
        if dc is None:
            dc = self.device_constants[key] = self.core.constants(self._context(s), nq)
        return dc

    def _context(self, s):
        hit = self._contexts.get(id(s))
        if hit is None or hit[0] is not s:
            hit = self._contexts[id(s)] = (s, _Context(s, self))
        return hit[1]

    # --------------------------------------------------------------------------------------------------- setup
    def setup(self, s, nq):
        """The constants dict `c` of the reference's setup (`source_solver.py:1243-1339`): Gauss-Legendre nodes on
        [0, 1], grid, profiles at the grid heights, extinction, sun table, ground pools (padded per cell as gd, gv),
        height, g, mix and albedo."""
        key = (id(s), nq)
        c = self.constants.get(key)
        if c is None:
            ctx = self._context(s)
            u, wu = np.polynomial.legendre.leggauss(nq)
            goff, gdir, gval, gm = ctx.given_pools
            ng = int(max(np.diff(goff), default=0))
            gd = np.zeros((ctx.geo.ns, ng, 3), 'f4')
            gv = np.zeros((ctx.geo.ns, ng, ctx.W), 'f4')
            for cell in range(ctx.geo.ns):
                a, b = goff[cell : cell + 2]
                gd[cell, : b - a] = gdir[a:b]
                gv[cell, : b - a] = gval[a:b]
            c = dict(domain=self.domain, L=self.L, W=ctx.W, height=np.float32(self.height), g=np.float32(self.g),
                     mix=np.float32(ctx.optics.mix[0]), albedo=float(ctx.optics.albedo[0]), u=(u + 1) / 2, wu=wu / 2,
                     h=ctx.geo.h.astype('f4'), chi=ctx.geo.chi.astype('f4'), hr=ctx.hr.astype('f4'),
                     ha=ctx.ha.astype('f4'), ext=ctx.optics.ext.astype('f4'), asca=ctx.optics.asca.astype('f4'),
                     sun_h=ctx.sun_h.astype('f4'), sun_col=np.asarray(s.geo.sun.col, 'f4'), gm=gm.astype('f4'),
                     gd=gd, gv=gv, goff=goff, gdir=gdir, gval=gval)
            if getattr(s, 'density_optics', None) is not None:
                c['density_optics'] = s.density_optics
            self.constants[key] = c
        return c

    # ---------------------------------------------------------------------------------------------------- launch
    def _device_bundle(self, bundle):
        db = getattr(bundle, '_hip_bundle', None)
        if db is None:
            db = DeviceBundle(_HostBundle(bundle), self.device)
            bundle._hip_bundle = db
        return db

    def _lanes(self, a, G, rows=None):
        """A field or boundary on the device with its last axis padded to G lanes (the last lane repeated)."""
        if torch.is_tensor(a):
            a = a.to(self.device, torch.float32)
            if a.shape[-1] < G:
                a = torch.cat([a, a[..., -1:].expand(*a.shape[:-1], G - a.shape[-1])], -1)
            return a.contiguous()
        return torch.from_numpy(np.ascontiguousarray(_pad_lanes(np.asarray(a, 'f4'), G))).to(self.device)

    def launch(self, bundle, s, field, first=False, phase=0, boundary=None, timed=False):
        """(moments, flux, query) of one bundle (`source_solver.py:2428-2456`): phase 0 the moments f32[cells, 4+K, W]
        and ground flux f32[cells, W]; phase 1 the flux only; phase 2 the per-ray query f32[rays, W, 4]. first: single
        scattering (the field is not read); boundary f[nc, W]: the ground term of the warm pass."""
        ctx = self._context(s)
        W = ctx.W
        G = lane_count(W)
        db = self._device_bundle(bundle)
        rows = ctx.rows
        fd = None
        if not first:
            fd = self._lanes(field, G)
            if tuple(fd.shape) != (ctx.geo.ns, rows, G):
                raise ValueError(f'E_SOURCE_SHAPE: {tuple(fd.shape[:2])} for [{ctx.geo.ns}, {rows}, W]')
        bd = None if boundary is None else self._lanes(boundary, G)
        if timed:
            start = torch.cuda.Event(enable_timing=True)
            start.record()
        give = (lambda t: t.cpu().numpy()) if self.xp is np else (lambda t: t.contiguous())
        if phase == 2:
            rad = self.core.radiance(db, ctx, fd, first, bd, 4)
            out = (None, None, give(rad[:, :W]))
            self.metrics['query_launches'] += 1
        else:
            moments, flux, _ = self.core.launch(db, ctx, fd, first, phase, bd)
            out = (None if moments is None else give(moments[..., :W]), give(flux[..., :W]), None)
        if timed:
            end = torch.cuda.Event(enable_timing=True)
            end.record()
            self.cuda_events.append(('first' if first else f'phase{phase}', start, end))
        self.metrics['launches'] += 1
        self.metrics['samples'] += db.samples
        self.memory()
        return out

    def transport(self, geometry_, s, field, first=False):
        """Phase-2 adapter (`source_solver.py:2458-2484`): one raw pack of query rays -> f8[n, W, 2, 4] with the four
        tensor components in lane 0. field: [cells, rows, 2 W] with the source in the even lanes, or empty."""
        if 'edges' not in geometry_:
            raise ValueError('E_SAMPLE_CACHE: fused path requires starts/directions/segments, not legacy point arrays')
        ctx = self._context(s)
        self.setup(s, geometry_['nq'])
        if field is None or np.size(self.host(field)) == 0:
            packed = np.zeros((ctx.geo.ns, ctx.rows, ctx.W), 'f4')
        else:
            packed = np.ascontiguousarray(self.host(field)[..., ::2], 'f4')
        _, _, value = self.launch(_RayPack(geometry_), s, packed, first, 2)
        value = self.host(value)
        out = np.zeros((len(value), ctx.W, 2, 4), 'f8')
        out[:, :, 0] = value
        return out
