"""The backend seam: `Kernel(backend='hip')` owns the device copies of bundles and constants and launches the HIP
kernels (the role of `src/sosfuse2/solve/source_solver.py:1109-2484` Kernel: setup, launch, transport).

Device memory is PyTorch's (one process, one GPU); every launch runs on torch's current stream. Lanes: a chunk of W
wavelengths runs on G = 1, 16 or 32 lanes per ray (W <= G), padded by repeating the last lane.
"""

import ctypes
import math

import numpy as np
import torch

from . import geometry, hip, layout

LANE_ROWS = hip.LANE_CONSTANTS


def lane_count(W):
    if W == 1:
        return 1
    if W <= 16:
        return 16
    if W <= 32:
        return 32
    raise ValueError(f'E_WAVELENGTHS: {W} lanes exceed 32; split the chunk')


def _pad_lanes(a, G):
    """Pad the last axis to G by repeating its last entry."""
    a = np.asarray(a)
    if a.shape[-1] == G:
        return a
    idx = np.r_[np.arange(a.shape[-1]), np.full(G - a.shape[-1], a.shape[-1] - 1)]
    return a[..., idx]


class DeviceBundle:
    """Device copy of a `geometry.Bundle` (the reference's upload_geometry, `solve/ray_geometry.py:256-287`)."""

    def __init__(self, b, device):
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(device)
        self.host = b
        self.groups, self.rays, self.nq = b.groups, b.rays, b.nq
        self.group_off = t(b.group_off, 'i8')
        self.starts = t(b.starts, 'f8')
        self.targets = t(b.targets.reshape(-1, 9), 'f4')
        self.ray_group = t(b.ray_group, 'i4')
        self.dirs = t(b.dirs, 'f8')
        self.dloc = t(b.dloc, 'f4')
        self.weights = t(b.weights, 'f4')
        self.edge_off = t(b.edge_off, 'i8')
        self.edges = t(b.edges, 'f8')
        self.lower = t(b.lower.view('i2'), 'i2')
        self.end_j = t(b.end_j, 'i4')
        self.end_f = t(b.end_f, 'f4')
        self.end_col = t(b.end_col, 'f4')
        self.hits = t(b.hits, 'u1')
        self.samples = b.samples
        self.plan = None
        self.struct = hip.Bundle(
            self.groups, self.rays, *(hip.ptr(x) for x in (
                self.group_off, self.starts, self.targets, self.ray_group, self.dirs, self.dloc, self.weights,
                self.edge_off, self.edges, self.lower, self.end_j, self.end_f, self.end_col, self.hits)), None, None,
            None)

    @classmethod
    def from_tensors(cls, groups, rays, nq, samples, t):
        """A device bundle from arrays already in HBM (device_geometry builds them there): t holds group_off i8,
        starts f8 [groups, 3], targets f4 [groups, 9], ray_group i4, dirs f8 [rays, 3], dloc f4 [rays, 3], weights
        f4, edge_off i8 [rays + 1], edges f8, lower i2, end_j i4, end_f f4, end_col f4 [rays, 6], hits u1."""
        self = cls.__new__(cls)
        self.host = None
        self.groups, self.rays, self.nq, self.samples = int(groups), int(rays), int(nq), int(samples)
        for k in ('group_off', 'starts', 'targets', 'ray_group', 'dirs', 'dloc', 'weights', 'edge_off', 'edges', 'lower',
                  'end_j', 'end_f', 'end_col', 'hits'):
            setattr(self, k, t[k].contiguous())
        self.plan = None
        self.struct = hip.Bundle(
            self.groups, self.rays, *(hip.ptr(x) for x in (
                self.group_off, self.starts, self.targets, self.ray_group, self.dirs, self.dloc, self.weights,
                self.edge_off, self.edges, self.lower, self.end_j, self.end_f, self.end_col, self.hits)), None, None,
            None)
        return self

    def nbytes(self):
        return sum(x.numel() * x.element_size() for x in (
            self.group_off, self.starts, self.targets, self.ray_group, self.dirs, self.dloc, self.weights,
            self.edge_off, self.edges, self.lower, self.end_j, self.end_f, self.end_col, self.hits))

    def part(self, g0, g1):
        """Sub-bundle of the groups [g0, g1): views of this bundle's per-group, per-ray and per-edge arrays, with
        the group, edge and ray-group offsets rebased to the part (new small tensors). Every ray of a group stays in
        its group's part, so transport and projection give the same values per ray and group as the whole bundle."""
        if not 0 <= g0 < g1 <= self.groups:
            raise ValueError(f'E_BUNDLE_PART: groups [{g0}, {g1}) outside [0, {self.groups})')
        go = self.group_off
        r0, r1 = int(go[g0].item()), int(go[g1].item())
        e0, e1 = int(self.edge_off[r0].item()), int(self.edge_off[r1].item())
        sub = DeviceBundle.__new__(DeviceBundle)
        sub.host = None
        sub.groups, sub.rays, sub.nq = g1 - g0, r1 - r0, self.nq
        sub.group_off = (go[g0 : g1 + 1] - r0).contiguous()
        sub.starts, sub.targets = self.starts[g0:g1], self.targets[g0:g1]
        sub.ray_group = (self.ray_group[r0:r1] - g0).contiguous()
        for k in ('dirs', 'dloc', 'weights', 'lower', 'end_j', 'end_f', 'end_col', 'hits'):  # per ray
            setattr(sub, k, getattr(self, k)[r0:r1])
        sub.edge_off = (self.edge_off[r0 : r1 + 1] - e0).contiguous()
        sub.edges = self.edges[e0:e1]
        # sampled intervals (longer than 1e-8 km) of the part's rays, times the node count
        sampled = (sub.edges[1:] - sub.edges[:-1]) > 1e-8
        if sub.rays > 1 and len(sampled):
            cross = sub.edge_off[1:-1] - 1  # pairs that join the last edge of a ray to the next ray's first
            sampled[cross[(cross >= 0) & (cross < len(sampled))]] = False
        sub.samples = int(sampled.sum().item()) * self.nq
        sub.plan = None
        sub.struct = hip.Bundle(
            sub.groups, sub.rays, *(hip.ptr(x) for x in (
                sub.group_off, sub.starts, sub.targets, sub.ray_group, sub.dirs, sub.dloc, sub.weights,
                sub.edge_off, sub.edges, sub.lower, sub.end_j, sub.end_f, sub.end_col, sub.hits)), None, None, None)
        sub.span = (g0, g1, r0, r1)
        return sub


INDEX_LIMIT = 2**31 - 1  # node and (node, corner) visit ids are int32 on the device
PART_NODES = 2**27  # quadrature nodes per plan part (4 x 2^27 visit ids stay far inside int32)


def split_groups(per_group, limit):
    """Split consecutive groups of per_group[i] nodes into ranges [g0, g1) of at most `limit` nodes each (a group
    larger than the limit alone in its range, within the int32 visit-id bound)."""
    per_group = np.asarray(per_group, np.int64)
    if limit < 1:
        raise ValueError('E_PART_NODES: the part size must be positive')
    out, g0, acc = [], 0, 0
    for g, n in enumerate(per_group):
        n = int(n)
        check_plan_size(n)
        if acc and n and acc + n > limit:
            out.append((g0, g))
            g0, acc = g, 0
        acc += n
    if len(per_group):
        out.append((g0, len(per_group)))
    return out


def check_plan_size(nodes):
    """Raise E_PLAN_SIZE when the 4 x nodes (node, corner) visit ids of a plan would not fit the device's int32 ids
    (they would wrap to negative values and index the records out of bounds)."""
    if 4 * int(nodes) > INDEX_LIMIT:
        raise OverflowError(f'E_PLAN_SIZE: {nodes} quadrature nodes give {4 * int(nodes)} (node, corner) visits, '
                            f'more than the {INDEX_LIMIT} int32 visit ids allow; use a coarser quadrature')
    return int(nodes)


class Plan:
    """Visit plan of a bundle (fixed by its geometry): every quadrature node's cell, mu, cos(phi) and direction; the
    nodes sorted by their base cell with work items of at most `node_chunk` nodes of one base cell (SH records,
    four corners per node); and the (node, corner) visits sorted by visited cell with items of at most `chunk`
    visits (ground records of first light)."""

    def __init__(self, kernel, dc, db, chunk=2048, node_chunk=1024):
        self.node_chunk = node_chunk
        self.arena = getattr(db, 'arena', None)
        dev = kernel.device
        st = kernel.stream()
        self._nodes_args = (dc.struct, db.struct, st)
        nc = dc.struct.nc
        cells = dc.struct.nh * nc
        self._cells = cells
        self._nc = nc
        self._chunk = chunk
        kept = getattr(db, 'kept', None) if self.arena is not None else None
        if kept is not None and kept['node_chunk'] == node_chunk:
            # a bundle part's node offsets, sorted nodes and work items, kept from its first plan: the plans of the
            # later launches skip the node data, the sort by base cell and the work items
            self.node_off = kept['node_off']
            db.struct.node_off_dev = hip.ptr(self.node_off)
            self.nodes = N = kept['nodes']
            self.node, self.sorted_ray, self.sorted_cp = kept['node'], kept['sorted_ray'], kept['sorted_cp']
            self.sh_cell, self.sh_v0, self.sh_v1 = kept['items']
            self.sh_items = len(self.sh_cell)
            self._node_cell = self._node_mu = self._node_cp = self._node_dir = None  # on demand (first light)
        else:
            counts = torch.empty(max(db.rays, 1), dtype=torch.int64, device=dev)
            hip.call('sos_plan_count', dc.struct, db.struct, hip.ptr(counts), st)
            self.node_off = torch.zeros(db.rays + 1, dtype=torch.int64, device=dev)
            self.node_off[1:] = torch.cumsum(counts[: db.rays], 0)
            db.struct.node_off_dev = hip.ptr(self.node_off)
            self.nodes = N = check_plan_size(self.node_off[-1].item())
            self._node_cell = None
            self._fill_nodes()
            # nodes by base cell (SH records), with each node's ray and cos(phi) in the same order
            self.node = torch.argsort(self.node_cell[:N], stable=True).to(torch.int32)
            ray_of = torch.repeat_interleave(torch.arange(db.rays, dtype=torch.int32, device=dev), counts[: db.rays])
            self.sorted_ray = ray_of[self.node.long()] if N else torch.zeros(1, dtype=torch.int32, device=dev)
            del ray_of
            self.sorted_cp = self.node_cp[self.node.long()]
            self.set_node_chunk(node_chunk)
            if self.arena is not None:
                db.kept = dict(node_chunk=node_chunk, node_off=self.node_off, nodes=N, node=self.node,
                               sorted_ray=self.sorted_ray, sorted_cp=self.sorted_cp,
                               items=(self.sh_cell, self.sh_v0, self.sh_v1))
        self.basis = None  # basis table [rays][KZ] (sos_plan_basis), built on first use (basis_table())
        self.vis = None  # (node, corner) visits by visited cell, built on first use (visit_items())
        self.visits = 4 * N
        self.records = self._buf('records', 4 * N, 'visits', dc.G)
        self.geom = None  # static node geometry of the warm consumer, built on first use (geometry())
        self.sun = None  # sun geometry of first light, built on first use (sun_geometry())

    def _fill_nodes(self):
        """Every node's cell (ir nc + j), mu, cos(phi) and direction (sos_plan_nodes)."""
        N, dev = self.nodes, self.node_off.device
        grid, bundle, st = self._nodes_args
        self._node_cell = torch.empty(max(N, 1), dtype=torch.int32, device=dev)
        self._node_mu = torch.empty(max(N, 1), dtype=torch.float32, device=dev)
        self._node_cp = torch.empty(max(N, 1), dtype=torch.float32, device=dev)
        self._node_dir = torch.empty((max(N, 1), 4), dtype=torch.float32, device=dev)
        hip.call('sos_plan_nodes', grid, bundle, hip.ptr(self.node_off), hip.ptr(self._node_cell),
                 hip.ptr(self._node_mu), hip.ptr(self._node_cp), hip.ptr(self._node_dir), st)

    @property
    def node_cell(self):
        if self._node_cell is None:
            self._fill_nodes()
        return self._node_cell

    @property
    def node_mu(self):
        if self._node_cell is None:
            self._fill_nodes()
        return self._node_mu

    @property
    def node_cp(self):
        if self._node_cell is None:
            self._fill_nodes()
        return self._node_cp

    @property
    def node_dir(self):
        if self._node_cell is None:
            self._fill_nodes()
        return self._node_dir

    def visit_items(self):
        """The (node, corner) visits sorted by visited cell and their work items of at most `chunk` visits (the
        ground records of first light read them), built once."""
        if self.vis is None:
            N, nc, dev = self.nodes, self._nc, self.node_off.device
            offs = torch.tensor([0, 1, nc, nc + 1], dtype=torch.int32, device=dev)
            vcell = (self.node_cell[:N, None] + offs[None, :]).reshape(-1)
            self.vis = torch.argsort(vcell, stable=True).to(torch.int32)
            self.item_cell, self.item_v0, self.item_v1 = _items(vcell, self._cells, self._chunk, dev)
            self.items = len(self.item_cell)
        return self

    def basis_table(self, kernel, dc, db):
        """[rays, KZ] Legendre part of every ray's SH basis row (sos_plan_basis), built once."""
        if self.basis is None:
            kz = (kernel.K + 15) // 16 * 16
            self.basis = self._buf('basis', db.rays, 'rays', kz)
            hip.call('sos_plan_basis', dc.struct, db.struct, hip.ptr(self.basis), kernel.stream())
        return self.basis

    def geometry(self, kernel, dc, db):
        """[nodes, 32] static node geometry (sos_plan_geometry), built once."""
        if self.geom is None:
            self.geom = self._buf('geom', self.nodes, 'nodes', 32)
            hip.call('sos_plan_geometry', dc.struct, db.struct, hip.ptr(self.node_off), hip.ptr(self.geom),
                     kernel.stream())
            db.struct.node_geom_dev = hip.ptr(self.geom)
        return self.geom

    def sun_geometry(self, kernel, dc, db):
        """[nodes, 12] sun geometry of first light (sos_plan_sun), built once."""
        if self.sun is None:
            self.sun = self._buf('sun', self.nodes, 'nodes', 12)
            hip.call('sos_plan_sun', dc.struct, db.struct, hip.ptr(self.node_off), hip.ptr(self.sun), kernel.stream())
            db.struct.node_sun_dev = hip.ptr(self.sun)
        return self.sun

    def set_node_chunk(self, node_chunk):
        """Work items of the SH records: runs of at most node_chunk nodes of one base cell (a tuning knob; the
        records do not depend on it)."""
        self.node_chunk = node_chunk
        self.sh_cell, self.sh_v0, self.sh_v1 = _items(self.node_cell[: self.nodes], self._cells, node_chunk,
                                                      self.node_off.device)
        self.sh_items = len(self.sh_cell)

    def records_for(self, G):
        """The records buffer [4 nodes, G] (reallocated when a chunk of another lane count uses the plan)."""
        if self.records.shape[1] != G:
            self.records = None
            self.records = self._buf('records', 4 * self.nodes, 'visits', G)
        return self.records

    def _buf(self, key, rows, kind, width):
        """An FP32 [rows, width] buffer. The plan of a bundle part takes it from the arena its parts share: one
        buffer per key, sized for the largest part (its nodes, visits or rays) and reused by every part, so the parts'
        plans, built one after another, do not allocate and free tens of GB each."""
        dev = self.node_off.device
        a = self.arena
        if a is None:
            return torch.empty((max(rows, 1), width), dtype=torch.float32, device=dev)
        need = max(rows, 1) * width
        buf = a['bufs'].get(key)
        if buf is None or buf.numel() < need:
            a['bufs'][key] = buf = None
            cap = {'nodes': a['nodes'], 'visits': 4 * a['nodes'], 'rays': a['rays']}[kind]
            buf = a['bufs'][key] = torch.empty(max(need, max(cap, 1) * width), dtype=torch.float32, device=dev)
        return buf[:need].view(max(rows, 1), width)


def _items(key, cells, chunk, dev):
    """Work items (cell, start, end) of at most `chunk` entries over the entries sorted by `key` (a cell index)."""
    per_cell = torch.bincount(key, minlength=cells)
    cell_off = torch.zeros(cells + 1, dtype=torch.int64, device=dev)
    cell_off[1:] = torch.cumsum(per_cell, 0)
    nchunk = (per_cell + chunk - 1) // chunk
    item_cell = torch.repeat_interleave(torch.arange(cells, device=dev, dtype=torch.int32), nchunk)
    first_item = torch.cumsum(nchunk, 0) - nchunk
    idx = torch.arange(len(item_cell), device=dev) - torch.repeat_interleave(first_item, nchunk)
    v0 = cell_off[:-1][item_cell.long()] + idx * chunk
    v1 = torch.minimum(v0 + chunk, cell_off[1:][item_cell.long()])
    return item_cell, v0, v1


class DeviceConstants:
    """Device copy of one context's constants for one node count (the reference's Kernel.setup + constants_device)."""

    def __init__(self, ctx, nq, L, G, device):
        c = ctx.constants(nq, padded=False)
        o = ctx.optics
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(device)
        self.G, self.L, self.nq = G, L, nq
        self.W = o.W
        lanes = np.zeros((LANE_ROWS, G), 'f4')
        lanes[0:5] = _pad_lanes(o.ext, G)
        lanes[5] = _pad_lanes(o.asca, G)
        lanes[6] = _pad_lanes(o.g, G)
        lanes[7] = _pad_lanes(o.mix, G)
        lanes[8] = _pad_lanes(o.albedo, G)
        lanes[9] = _pad_lanes(o.strato_ext, G)
        lanes[10] = _pad_lanes(o.strato_sca, G)
        lanes[11] = _pad_lanes(o.strato_g, G)
        self.lane = t(lanes, 'f4')
        self.gl_u, self.gl_w = t(c['u'], 'f8'), t(c['wu'], 'f8')
        self.h, self.chi, self.hr = t(c['h'], 'f4'), t(c['chi'], 'f4'), t(c['hr'], 'f4')
        self.ha, self.sd = t(c['ha'], 'f4'), t(c['strato_density'], 'f4')
        self.sun_h, self.sun_col = t(c['sun_h'], 'f4'), t(c['sun_col'], 'f4')
        given = getattr(ctx, 'given_pools', None)
        if given is not None:
            # pool offsets, directions, values [P, W] and per-cell tensor moments [cells, 4, W] supplied by the caller
            # (the reference's s.goff, s.gdir, s.gval, s.gm)
            goff, gdir, gval, gm = given
            P = len(gdir)
            self.goff, self.gdir = t(goff, 'i8'), t(gdir if P else np.zeros((1, 3)), 'f4')
            self.gval = t(_pad_lanes(gval, G) if P else np.zeros((1, G)), 'f4')
            self.gm = t(_pad_lanes(gm, G), 'f4')
            pool_max = int(np.diff(goff).max(initial=0))
        else:
            pools = ctx.pools
            pd = pool_device(pools, device)
            self.goff, self.gdir = pd['off'], pd['loc32']
            P = len(pools.loc)
            pool_max = int(np.diff(pools.off).max(initial=0))
            # the state-dependent pool values and tensor moments are computed on the device (sos_ground_pools, below)
            self.gval = torch.zeros((max(P, 1), G), dtype=torch.float32, device=device)
            self.gm = torch.zeros((ctx.geo.ns, 4, G), dtype=torch.float32, device=device)
        self.factor = t(_pad_lanes(ctx.eigen(), G), 'f4')
        # one asymmetry per launch (a single-state chunk): first light may share phase evaluations across lanes
        self.uniform_phase = bool(np.all(lanes[6] == lanes[6][0]) and np.all(lanes[11] == lanes[11][0]))
        # phase groups of the ground records: distinct (aerosol g, sulfate g) in lane order, None beyond four
        keys = []
        for w in range(G):
            key = (lanes[6][w], lanes[11][w] if ctx.strato else np.float32(0))
            if key not in keys:
                keys.append(key)
        self.phase_groups = None
        if len(keys) <= 4:
            self.phase_groups = (len(keys), (ctypes.c_float * 4)(*[k[0] for k in keys]),
                                 (ctypes.c_float * 4)(*[k[1] for k in keys]))
        p = c['profile']
        g = ctx.geo
        self.struct = hip.Grid(
            g.nh, g.nc, L, G, nq, int(c['strato']), float(c['top']), p['air_H'], p['aer_H'], p['o3_c'], p['o3_s'],
            p['no2_c'], p['no2_s'], p['st_c'], p['st_s'],
            *(hip.ptr(x) for x in (self.gl_u, self.gl_w, self.h, self.chi, self.hr, self.lane)),
            len(c['sun_h']), c['sun_col'].shape[1], hip.ptr(self.sun_h), hip.ptr(self.sun_col), hip.ptr(self.goff),
            hip.ptr(self.gdir), hip.ptr(self.gval), hip.ptr(self.gm), 0, 0, pool_max)
        set_profiles(self.struct, p['spec'], self, device)
        if P and given is None:
            hip.call('sos_ground_pools', self.struct, hip.ptr(pd['off']), hip.ptr(pd['column']), hip.ptr(pd['factor']),
                     hip.ptr(pd["loc"]), hip.ptr(self.gval), hip.ptr(self.gm),
                     torch.cuda.current_stream(device).cuda_stream)


def set_profiles(struct, spec, owner, device):
    """Fill a grid struct's profile kinds from a ProfileSpec (atmosphere.ProfileSpec); a table goes to the device as
    [6][n] FP32 rows, kept on `owner`."""
    for k in range(6):
        struct.profile_kind[k] = int(spec.kinds[k])
        struct.profile_a[k] = float(spec.a[k])
        struct.profile_b[k] = float(spec.b[k])
    owner.profile_table = None
    if spec.table is not None:
        rows = spec.table_array()
        owner.profile_table = torch.from_numpy(np.ascontiguousarray(rows, dtype='f4')).to(device)
        struct.profile_table_dev = hip.ptr(owner.profile_table)
        struct.profile_table_n = rows.shape[1]


_POOLS = {}


def pool_device(pools, device):
    """Device copy of a grid's ground pools (state-independent: CSR offsets, FP64 column vectors, geometric factors
    and directions, FP32 directions for the transport), made once per pools object."""
    key = (id(pools), str(device))
    hit = _POOLS.get(key)
    if hit is not None and hit[0] is pools:
        return hit[1]
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(device)
    P = len(pools.loc)
    d = dict(off=t(pools.off, 'i8'),
             loc32=t(pools.loc if P else np.zeros((1, 3)), 'f4'),
             loc=t(pools.loc if P else np.zeros((1, 3)), 'f8'),
             column=t(pools.column if P else np.zeros((1, 6)), 'f8'),
             factor=t(pools.factor if P else np.zeros(1), 'f8'))
    _POOLS[key] = (pools, d)
    return d


class Kernel:
    """HIP transport kernel of one solve chunk (backend 'hip' only: `source_solver.py:1113-1114` accepts 'numpy' and
    'cupy'; this build adds 'hip' and runs nothing else)."""

    backend = 'hip'
    CONTEXTS = 4  # contexts whose device constants stay cached

    def __init__(self, backend='hip', L=16, block_size=256, device='cuda', mixed=False, records=True):
        if backend != 'hip':
            raise ValueError(f"E_BACKEND: backend {backend!r}; this engine runs 'hip' only")
        layout.count(L)
        hip.mark_started()
        if not torch.cuda.is_available():
            raise RuntimeError('E_HIP_DEVICE: no GPU visible to torch')
        if mixed and not records:
            raise ValueError('E_MIXED: mixed precision runs the SH contraction as records (records=True)')
        hip.lib()
        self.L = L
        self.K = layout.count(L)
        self.block_size = block_size
        self.device = torch.device(device)
        self.mixed = mixed
        self.records = records  # SH by per-cell MFMA records (True) or per-ray register caches (False)
        self.uniform_phase = True  # use the shared-phase first light when a chunk's lanes share the asymmetry
        self.node_geometry = True  # warm records consumer on the plan's static node geometry
        self.xp = torch
        self._constants = {}
        self.metrics = dict(samples=0, launches=0)
        self.timed = None  # list of (name, start event, end event) while timing is on
        # launch shapes chosen by the autotuner (sosfuse2_amd.tune); results do not depend on them
        self.tuning = dict(blocks_per_cu=0, node_chunk=1024, sh_tile=3)
        self.counters = None  # device uint64 [2] while counting is on
        # bundles of more quadrature nodes than this run in parts of whole groups, one visit plan at a time (each
        # part's plan is built for its launch and freed after it): about 2^27 nodes x (232 + 16 G) bytes of plan,
        # node geometry and records per part, and int32 visit ids within a part
        self.part_nodes = PART_NODES

    def sync(self):
        torch.cuda.synchronize(self.device)

    def stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def host(self, x):
        return x.detach().cpu().numpy() if torch.is_tensor(x) else np.asarray(x)

    def constants(self, ctx, nq):
        """Device constants of a context for a node count, cached for the most recent CONTEXTS contexts (the cache
        holds the context, so an id is not reused while its entry lives)."""
        key = (id(ctx), nq)
        dc = self._constants.get(key)
        if dc is None:
            live = list(dict.fromkeys(k[0] for k in self._constants))
            if id(ctx) not in live and len(live) >= self.CONTEXTS:
                old = live[0]
                for k in [k for k in self._constants if k[0] == old]:
                    del self._constants[k]
            dc = DeviceConstants(ctx, nq, self.L, lane_count(ctx.W), self.device)
            self._constants[key] = (ctx, dc)
            return dc
        return dc[1]

    def bundle(self, b):
        return DeviceBundle(b, self.device)

    # ------------------------------------------------------------------------------------------------ launches
    def plane(self, ctx, field, dc):
        """Combined SH plane of a source field (m-major per cell): (FP32 plane [cells, K, G], None), or in mixed
        precision the split FP16 plane [2 (hi, lo), cells, KZ/4, G, 4] (K padded to KZ, quads of four K positions per
        lane) and its per-(cell, lane) power-of-two scale."""
        K = self.K
        if self.mixed:
            KZ = (K + 15) // 16 * 16
            out = torch.empty((2, ctx.geo.ns, KZ // 4, dc.G, 4), dtype=torch.float16, device=self.device)
        else:
            out = torch.empty((ctx.geo.ns, K, dc.G), dtype=torch.float32, device=self.device)
        scale = torch.empty((ctx.geo.ns, dc.G), dtype=torch.float32, device=self.device) if self.mixed else None
        hip.call('sos_make_plane', dc.struct, ctx.rows, hip.ptr(field), hip.ptr(dc.ha), hip.ptr(dc.sd),
                 int(self.mixed), hip.ptr(out), hip.ptr(scale), self.stream())
        return out, scale

    def plan(self, db, dc):
        if db.plan is None:
            db.plan = Plan(self, dc, db, node_chunk=self.tuning['node_chunk'])
        elif db.plan.node_chunk != self.tuning['node_chunk']:
            db.plan.set_node_chunk(self.tuning['node_chunk'])
        return db.plan

    def parts(self, db, dc):
        """[db] when its plan fits part_nodes, else sub-bundles of consecutive whole groups of at most part_nodes
        quadrature nodes each (cached on the bundle per part size)."""
        key = ('parts', self.part_nodes)
        cached = getattr(db, '_parts', None)
        if cached is not None and cached[0] == key:
            return cached[1]
        if db.plan is not None or db.rays == 0:
            db.nodes_total = db.plan.nodes if db.plan is not None else 0
            out = [db]
        else:
            counts = torch.empty(max(db.rays, 1), dtype=torch.int64, device=self.device)
            hip.call('sos_plan_count', dc.struct, db.struct, hip.ptr(counts), self.stream())
            cum = torch.zeros(db.rays + 1, dtype=torch.int64, device=self.device)
            cum[1:] = torch.cumsum(counts[: db.rays], 0)
            per_group = cum[db.group_off[1:]] - cum[db.group_off[:-1]]
            db.nodes_total = int(cum[-1].item())
            per_group = per_group.cpu().numpy()
            spans = split_groups(per_group, self.part_nodes)
            out = [db.part(g0, g1) for g0, g1 in spans]
            if len(out) == 1:
                out = [db]
            else:
                # the parts' plans share their large buffers, sized for the largest part
                arena = dict(nodes=int(max(per_group[g0:g1].sum() for g0, g1 in spans)),
                             rays=max(p.rays for p in out), bufs={})
                for p in out:
                    p.arena = arena
        db._parts = (key, out)
        return out

    def _event(self):
        if self.timed is None:
            return None
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def radiance(self, db, ctx, field, first=False, boundary=None, pairs=4, plane=None):
        """Per-ray radiance tensor [rays, G, pairs] (part by part for a bundle above part_nodes)."""
        dc = self.constants(ctx, db.nq)
        parts = self.parts(db, dc)
        if len(parts) == 1:
            return self._radiance(db, ctx, field, first, boundary, pairs, plane)
        if not first and plane is None:
            plane = self.plane(ctx, field, dc)
        out = torch.empty((db.rays, dc.G, pairs), dtype=torch.float32, device=self.device)
        for part in parts:
            _, _, r0, r1 = part.span
            self._radiance(part, ctx, field, first, boundary, pairs, plane, out[r0:r1])
            part.plan = None
        return out

    def _radiance(self, db, ctx, field, first=False, boundary=None, pairs=4, plane=None, out=None):
        dc = self.constants(ctx, db.nq)
        dc.struct.blocks_per_cu = self.tuning['blocks_per_cu']
        dc.struct.sh_tile = self.tuning['sh_tile']
        if out is None:
            out = torch.empty((db.rays, dc.G, pairs), dtype=torch.float32, device=self.device)
        flags = (hip.FIRST if first else 0) | (hip.BOUNDARY if boundary is not None else 0)
        records = None
        if first and dc.uniform_phase and self.uniform_phase:
            flags |= hip.UNIFORM_PHASE
        if first and self.records and dc.G > 1 and dc.phase_groups is not None:
            pl = self.plan(db, dc)
            ng, gg, sgg = dc.phase_groups
            e0 = self._event()
            records = pl.records_for(dc.G)
            pl.visit_items()
            hip.call('sos_ground_records', dc.struct, ng, gg, sgg, pl.items, hip.ptr(pl.item_cell),
                     hip.ptr(pl.item_v0), hip.ptr(pl.item_v1), hip.ptr(pl.vis), hip.ptr(pl.node_dir),
                     hip.ptr(records), self.stream())
            if e0 is not None:
                self.timed.append(('ground_records', e0, self._event()))
            flags |= hip.GROUND_RECORDS
            if pairs == 4 and self.node_geometry:
                pl.geometry(self, dc, db)
                pl.sun_geometry(self, dc, db)
                flags |= hip.NODE_GEOMETRY
        name = 'first' if first else ('warm' if boundary is not None else
                                      ('query' if pairs == 6 or db.groups != ctx.geo.nc else 'ground'))
        if not first:
            if plane is None:
                plane = self.plane(ctx, field, dc)
            if self.records:
                pl = self.plan(db, dc)
                e0 = self._event()
                records = pl.records_for(dc.G)
                basis = pl.basis_table(self, dc, db)
                hip.call('sos_sh_records', dc.struct, hip.ptr(plane[0]), hip.ptr(plane[1]), int(self.mixed),
                         pl.sh_items, hip.ptr(pl.sh_cell), hip.ptr(pl.sh_v0), hip.ptr(pl.sh_v1), hip.ptr(pl.node),
                         hip.ptr(pl.sorted_ray), hip.ptr(pl.sorted_cp), hip.ptr(basis), hip.ptr(records),
                         self.stream())
                if e0 is not None:
                    e1 = self._event()
                    self.timed.append(('sh_' + name, e0, e1))
                records = pl.records
                flags |= hip.RECORDS
                if pairs == 4 and self.node_geometry:
                    pl.geometry(self, dc, db)
                    flags |= hip.NODE_GEOMETRY
        rows = ctx.rows
        if not first and (flags & hip.RECORDS) and (flags & hip.NODE_GEOMETRY) and self.tuning.get('tensor_rows', True):
            # the node-geometry consumer reads only the four Rayleigh tensor rows of the field: a lane-major copy
            # [cells][lanes][4] gives each corner one 16-byte read with a 32-bit offset
            field = field[:, :4].transpose(1, 2).contiguous()
            flags |= hip.TENSOR4
        elif not first and (flags & hip.RECORDS) and self.tuning.get('tensor_rows', True) and field.shape[1] > 10:
            # with records the warm pass reads only the tensor rows of the field: a compact copy [cells][10][lanes]
            # keeps every corner's rows within 1.3 KB of each other instead of a full field row (4.5 KB at L = 32)
            # apart, so the transport's corner reads stay in few pages and cache lines
            field, rows = field[:, :10].contiguous(), 10
        e0 = self._event()
        hip.call('sos_transport', dc.struct, db.struct, flags, pairs, rows,
                 None if first else hip.ptr(field), None if first else hip.ptr(plane[0]), hip.ptr(boundary),
                 hip.ptr(records), hip.ptr(out), hip.ptr(self.counters), self.stream())
        if e0 is not None:
            self.timed.append((name, e0, self._event()))
        self.metrics['launches'] += 1
        self.metrics['samples'] += db.samples
        return out

    def launch(self, db, ctx, field, first=False, phase=0, boundary=None, plane=None):
        """(moments [groups, 4+K, G], flux [groups, G], query) of one bundle (`source_solver.py:2428-2456`).
        phase 0: moments and flux; 1: flux only; 2: the per-ray query [rays, G, 4]."""
        if phase == 2:
            return None, None, self.radiance(db, ctx, field, first, boundary, 4, plane)
        dc = self.constants(ctx, db.nq)
        flux = torch.empty((db.groups, dc.G), dtype=torch.float32, device=self.device)
        moments = None
        if phase == 0:
            moments = torch.empty((db.groups, 4 + self.K, dc.G), dtype=torch.float32, device=self.device)
        parts = self.parts(db, dc)
        if len(parts) > 1 and not first and plane is None:
            plane = self.plane(ctx, field, dc)
        for part in parts:
            rad = self._radiance(part, ctx, field, first, boundary, 4, plane)
            g0, g1 = part.span[:2] if part is not db else (0, db.groups)
            basis = self.plan(part, dc).basis_table(self, dc, part) if phase == 0 else None
            hip.call('sos_project', dc.struct, part.struct, hip.ptr(rad), int(phase == 1), hip.ptr(basis),
                     hip.ptr(None if moments is None else moments[g0:g1]), hip.ptr(flux[g0:g1]), self.stream())
            if part is not db:
                del rad, basis
                part.plan = None
        return moments, flux, None
